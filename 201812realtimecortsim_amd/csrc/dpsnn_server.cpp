// dpsnn_server.cpp -- request loop for simulation jobs on one GPU
// (continuous batching over the engine's C ABI).
//
// Builder threads turn submitted jobs into host-side networks
// (dpsnn_prepare: parameters, generator, delivery layout) while the host
// lists of built jobs not yet uploaded stay under a bound.  The scheduler
// thread admits built jobs first-fit in arrival order while their exact
// device footprint (dpsnn_device_bytes_needed) fits the HBM budget, and an
// uploader thread copies each admitted job to the device (dpsnn_upload, on
// the job's own stream) while the batch keeps stepping.  The scheduler adds
// uploaded jobs to its dpsnn_batch, steps the batch a few ms per round (all
// running jobs in the same launches, each at its own ms), and retires the
// jobs that reached their env_totalSimTime_ms: they leave the batch, their
// files are drained, their memory returns to the budget.  Jobs arriving
// while others run join the batch at the next round.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "dpsnn_hip.h"
#include "dpsnn_host.h"

namespace {

using Clock = std::chrono::steady_clock;

// Device memory of the batch itself (key scratch of the longest in-list at
// the largest chunk, items, graphs), kept out of the jobs' budget.
constexpr uint64_t kBatchReserve = uint64_t(11) << 29;  // 5.5 GiB

struct Job {
  std::string inidir, outdir, env;
  dpsnn_options opt{};
  dpsnn_job_info info{};
  dpsnn_net *net = nullptr;
  uint64_t hostBytes = 0;  // counted in the server's backlog until uploaded
  std::string message;
};

}  // namespace

struct dpsnn_server {
  dpsnn_server_options o{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cvBuild, cvSched, cvDone, cvUpload;
  std::deque<Job *> buildQ, waiting, uploadQ, ready;  // ready: uploaded, not yet in the batch
  std::vector<Job *> running;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<std::thread> builders;
  std::thread sched, uploader;
  bool stop = false;
  uint64_t nextId = 1, used = 0, budget = 0;
  // host bytes of built jobs not yet uploaded, and their bound: a fast
  // generator would otherwise build every queued job while the device is
  // busy and hold all their host lists at once
  uint64_t backlog = 0, backlogLimit = 0;
  uint64_t eventsDone = 0, eventsRunning = 0;
  uint32_t uploading = 0;                 // admitted, in uploadQ or being uploaded
  // bumped whenever admission could change (a job waits, memory or a
  // running slot frees up); the scheduler sleeps on waiting jobs until the
  // generation moves past the one it last tried
  uint64_t admitGen = 0, triedGen = ~uint64_t(0);
  double stepSeconds = 0, uploadSeconds = 0, admitSeconds = 0, accountSeconds = 0, retireSeconds = 0;
  uint32_t hostThreads = 0;
  dpsnn_batch *batch = nullptr;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void finish_locked(dpsnn_server *s, Job *j, int state, int status, const std::string &msg) {
  j->info.state = state;
  j->info.status = status;
  j->info.finished_s = s->now();
  if (!msg.empty()) {
    j->message = msg;
    std::snprintf(j->info.message, sizeof j->info.message, "%s", msg.c_str());
  }
  s->cvDone.notify_all();
}

void builder_loop(dpsnn_server *s) {
  for (;;) {
    Job *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cvBuild.wait(lk, [&] { return s->stop || (!s->buildQ.empty() && s->backlog < s->backlogLimit); });
      if (s->stop) return;
      j = s->buildQ.front();
      s->buildQ.pop_front();
      j->info.state = DPSNN_JOB_BUILDING;
    }
    dpsnn_options opt = j->opt;
    opt.env = j->env.empty() ? nullptr : j->env.c_str();
    opt.output_dir = j->outdir.empty() ? nullptr : j->outdir.c_str();
    if (!opt.host_threads) opt.host_threads = s->hostThreads;
    dpsnn_net *n = nullptr;
    const int rc = dpsnn_prepare(j->inidir.c_str(), &opt, &n);
    const std::string err = rc ? dpsnn_last_error() : "";
    const uint64_t need = rc ? 0 : dpsnn_device_bytes_needed(n);
    const uint64_t host = rc ? 0 : dpsnn_host_bytes(n);
    std::unique_lock<std::mutex> lk(s->mu);
    if (rc) {
      finish_locked(s, j, DPSNN_JOB_FAILED, rc, err);
    } else if (need > s->budget) {
      lk.unlock();
      dpsnn_destroy(n);
      lk.lock();
      finish_locked(s, j, DPSNN_JOB_FAILED, DPSNN_ERR_MEMORY, "job needs more device memory than the budget");
    } else {
      j->net = n;
      j->hostBytes = host;
      s->backlog += host;
      j->info.device_bytes = need;
      j->info.built_s = s->now();
      j->info.state = DPSNN_JOB_WAITING;
      s->waiting.push_back(j);
      s->admitGen++;
      s->cvSched.notify_one();
    }
  }
}

// Uploads admitted jobs one at a time, beside the scheduler's steps.
void uploader_loop(dpsnn_server *s) {
  for (;;) {
    Job *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cvUpload.wait(lk, [&] { return s->stop || !s->uploadQ.empty(); });
      if (s->stop) return;
      j = s->uploadQ.front();
      s->uploadQ.pop_front();
    }
    const Clock::time_point t0 = Clock::now();
    const int rc = dpsnn_upload(j->net);
    const std::string err = rc ? dpsnn_last_error() : "";
    if (rc) dpsnn_destroy(j->net);
    std::lock_guard<std::mutex> lk(s->mu);
    s->uploadSeconds += std::chrono::duration<double>(Clock::now() - t0).count();
    s->uploading--;
    s->backlog -= j->hostBytes;
    j->hostBytes = 0;
    s->cvBuild.notify_all();
    s->admitGen++;
    if (rc) {
      j->net = nullptr;
      s->used -= j->info.device_bytes;
      finish_locked(s, j, DPSNN_JOB_FAILED, rc, err);
    } else {
      s->ready.push_back(j);
    }
    s->cvSched.notify_one();
  }
}

void scheduler_loop(dpsnn_server *s) {
  const uint32_t round = s->o.steps_per_round ? s->o.steps_per_round : 10;
  for (;;) {
    std::vector<Job *> admit;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // with nothing running, wait for uploads, or for a change that could
      // admit a waiting job, rather than spin
      s->cvSched.wait(lk, [&] {
        return s->stop || !s->ready.empty() || !s->running.empty() ||
               (!s->waiting.empty() && !s->uploading && s->admitGen != s->triedGen);
      });
      if (s->stop) return;
      // first fit in arrival order; admitted jobs go to the uploader
      for (auto it = s->waiting.begin(); it != s->waiting.end();) {
        Job *j = *it;
        const bool room =
            !s->o.max_running || s->running.size() + s->uploading + s->ready.size() < s->o.max_running;
        if (room && s->used + j->info.device_bytes <= s->budget) {
          s->used += j->info.device_bytes;
          s->uploadQ.push_back(j);
          s->uploading++;
          it = s->waiting.erase(it);
          s->cvUpload.notify_one();
        } else {
          ++it;
        }
      }
      s->triedGen = s->admitGen;
      admit.assign(s->ready.begin(), s->ready.end());
      s->ready.clear();
    }
    const Clock::time_point ta = Clock::now();
    for (Job *j : admit) {
      const int rc = dpsnn_batch_add(s->batch, j->net);
      if (rc) {
        const std::string err = dpsnn_last_error();
        dpsnn_destroy(j->net);
        std::lock_guard<std::mutex> lk(s->mu);
        j->net = nullptr;
        s->used -= j->info.device_bytes;
        s->admitGen++;
        finish_locked(s, j, DPSNN_JOB_FAILED, rc, err);
        continue;
      }
      std::lock_guard<std::mutex> lk(s->mu);
      j->info.admitted_s = s->now();
      j->info.state = DPSNN_JOB_RUNNING;
      s->running.push_back(j);
    }
    const double admitDt = std::chrono::duration<double>(Clock::now() - ta).count();
    if (s->running.empty()) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->admitSeconds += admitDt;
      continue;
    }
    const Clock::time_point t0 = Clock::now();
    const int rc = dpsnn_batch_step(s->batch, round);
    const double dt = std::chrono::duration<double>(Clock::now() - t0).count();
    const std::string err = rc < 0 ? dpsnn_last_error() : "";
    uint64_t events = 0;
    double accountDt = 0, retireDt = 0;
    for (Job *j : std::vector<Job *>(s->running)) {
      const Clock::time_point tj = Clock::now();
      if (rc >= 0 && dpsnn_now_ms(j->net) < dpsnn_total_ms(j->net)) {
        dpsnn_stats st{};
        if (dpsnn_get_stats(j->net, &st) == 0) events += st.synaptic_events;
        accountDt += std::chrono::duration<double>(Clock::now() - tj).count();
        continue;
      }
      dpsnn_batch_remove(s->batch, j->net);
      int frc = rc < 0 ? rc : dpsnn_sync(j->net);
      dpsnn_stats fst{};
      if (!frc) frc = dpsnn_get_stats(j->net, &fst);
      const std::string ferr = frc ? (rc < 0 ? err : std::string(dpsnn_last_error())) : "";
      const bool drop = frc || !s->o.keep_finished;
      if (drop) dpsnn_destroy(j->net);
      retireDt += std::chrono::duration<double>(Clock::now() - tj).count();
      std::lock_guard<std::mutex> lk(s->mu);
      j->info.stats = fst;
      if (drop) {
        j->net = nullptr;
        s->used -= j->info.device_bytes;
      }
      s->running.erase(std::find(s->running.begin(), s->running.end(), j));
      s->admitGen++;
      if (frc) {
        finish_locked(s, j, DPSNN_JOB_FAILED, frc, ferr);
      } else {
        s->eventsDone += j->info.stats.synaptic_events;
        finish_locked(s, j, DPSNN_JOB_DONE, 0, "");
      }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->eventsRunning = events;
    s->stepSeconds += dt;
    s->admitSeconds += admitDt;
    s->accountSeconds += accountDt;
    s->retireSeconds += retireDt;
  }
}

bool job_over(const Job *j) { return j->info.state == DPSNN_JOB_DONE || j->info.state == DPSNN_JOB_FAILED; }

}  // namespace

extern "C" {

int dpsnn_server_create(const dpsnn_server_options *opt, dpsnn_server **out) {
  if (!out) return DPSNN_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<dpsnn_server> s(new dpsnn_server());
  if (opt) s->o = *opt;
  s->t0 = Clock::now();
  uint64_t freeB = 0, totalB = 0;
  int rc = dpsnn_device_memory(s->o.device, &freeB, &totalB);
  if (rc) return rc;
  uint64_t budget = s->o.hbm_budget_bytes ? s->o.hbm_budget_bytes : freeB / 10 * 9;
  budget = std::min(budget, freeB);
  if (budget <= kBatchReserve) return DPSNN_ERR_MEMORY;
  s->budget = budget - kBatchReserve;
  if ((rc = dpsnn_batch_create(s->o.device, &s->batch))) return rc;
  const uint64_t phys = uint64_t(sysconf(_SC_PHYS_PAGES)) * uint64_t(sysconf(_SC_PAGE_SIZE));
  s->backlogLimit = s->o.host_backlog_bytes ? s->o.host_backlog_bytes : std::min(uint64_t(32) << 30, phys / 4);
  const uint32_t nb = s->o.build_threads ? s->o.build_threads : 2;
  s->hostThreads = std::max(1u, dpsnn::default_threads() / nb);
  dpsnn_server *p = s.get();
  for (uint32_t i = 0; i < nb; ++i) s->builders.emplace_back(builder_loop, p);
  s->sched = std::thread(scheduler_loop, p);
  s->uploader = std::thread(uploader_loop, p);
  *out = s.release();
  return 0;
}

int dpsnn_server_submit(dpsnn_server *s, const char *inidir, const dpsnn_options *opt, uint64_t *id) {
  if (!s || !inidir) return DPSNN_ERR_ARG;
  std::unique_ptr<Job> j(new Job());
  j->inidir = inidir;
  if (opt) {
    j->opt = *opt;
    if (opt->env) j->env = opt->env;
    if (opt->output_dir) j->outdir = opt->output_dir;
  }
  j->opt.env = nullptr;
  j->opt.output_dir = nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  j->info.id = s->nextId++;
  j->info.state = DPSNN_JOB_QUEUED;
  j->info.submitted_s = s->now();
  if (id) *id = j->info.id;
  s->buildQ.push_back(j.get());
  s->jobs[j->info.id] = std::move(j);
  s->cvBuild.notify_one();
  return 0;
}

int dpsnn_server_job(dpsnn_server *s, uint64_t id, dpsnn_job_info *info) {
  if (!s || !info) return DPSNN_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DPSNN_ERR_ARG;
  *info = it->second->info;
  return 0;
}

int dpsnn_server_wait(dpsnn_server *s, uint64_t id, double timeout_s, dpsnn_job_info *info) {
  if (!s) return DPSNN_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DPSNN_ERR_ARG;
  Job *j = it->second.get();
  auto pred = [&] { return job_over(j); };
  if (timeout_s < 0)
    s->cvDone.wait(lk, pred);
  else
    s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (info) *info = j->info;
  return job_over(j) ? DPSNN_DONE : 0;
}

int dpsnn_server_wait_all(dpsnn_server *s, double timeout_s) {
  if (!s) return DPSNN_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto pred = [&] {
    for (auto &kv : s->jobs)
      if (!job_over(kv.second.get())) return false;
    return true;
  };
  if (timeout_s < 0)
    s->cvDone.wait(lk, pred);
  else
    s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  return pred() ? DPSNN_DONE : 0;
}

int dpsnn_server_take(dpsnn_server *s, uint64_t id, dpsnn_net **net) {
  if (!s || !net) return DPSNN_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return DPSNN_ERR_ARG;
  Job *j = it->second.get();
  if (j->info.state != DPSNN_JOB_DONE || !j->net) return DPSNN_ERR_ARG;
  *net = j->net;
  j->net = nullptr;
  s->used -= j->info.device_bytes;
  s->admitGen++;
  s->cvSched.notify_one();
  return 0;
}

int dpsnn_server_get_stats(dpsnn_server *s, dpsnn_server_stats *st) {
  if (!s || !st) return DPSNN_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = dpsnn_server_stats{};
  for (auto &kv : s->jobs) {
    const int state = kv.second->info.state;
    st->submitted++;
    if (state == DPSNN_JOB_DONE) st->finished++;
    if (state == DPSNN_JOB_FAILED) st->failed++;
    if (state == DPSNN_JOB_RUNNING) st->running++;
    if (state == DPSNN_JOB_WAITING) st->waiting++;
    if (state == DPSNN_JOB_BUILDING) st->building++;
    if (state == DPSNN_JOB_QUEUED) st->queued++;
  }
  st->synaptic_events = s->eventsDone + s->eventsRunning;
  st->device_bytes = s->used + kBatchReserve;
  st->elapsed_s = s->now();
  st->step_s = s->stepSeconds;
  st->upload_s = s->uploadSeconds;
  st->admit_s = s->admitSeconds;
  st->account_s = s->accountSeconds;
  st->retire_s = s->retireSeconds;
  return 0;
}

void dpsnn_server_destroy(dpsnn_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cvBuild.notify_all();
  s->cvSched.notify_all();
  s->cvUpload.notify_all();
  for (auto &t : s->builders) t.join();
  if (s->sched.joinable()) s->sched.join();
  if (s->uploader.joinable()) s->uploader.join();
  dpsnn_batch_destroy(s->batch);
  for (auto &kv : s->jobs)
    if (kv.second->net) dpsnn_destroy(kv.second->net);
  delete s;
}

}  // extern "C"

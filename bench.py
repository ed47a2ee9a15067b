"""Benchmark of the MI355X engine on DPSNN's per-millisecond LIFCA step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1]
    python bench.py --mode batch [--steps K] [--warmup W]   (configs[3])
    python bench.py --mode serve [--serve-window S]          (configs[4])
    python bench.py --mode copy                              (counter calibration)

One "step" = one simulated millisecond of the whole local network on one GPU
(the reference's completeTimeStep_ms).  Default workload: configs[1], a
24x24 grid of 1250-neuron columns (720,000 neurons, ~810 M synapses),
on-device Philox external drive, no raster files; synthetic network from
the reference generator (no external data).  Prints one JSON line:
  value         synaptic events delivered per wall second over the K timed
                steps (inputs resident in HBM; host I/O not in the region)
  ms_per_step   wall ms per simulated ms = real-time factor x 1000
  roofline      one step's device work (phase-1 and phase-2 kernels) against
                HBM bandwidth: algorithmic bytes = 25 B per synaptic event +
                72 B per neuron (+16 B per replayed external event), divided
                by its average time from HIP events on the network's stream;
                traffic = HBM bytes per step from the committed counter
                summary (profiles/step_<config>_traffic.json)
  cpu_baseline  the oracle (scalar C restatement of DPSNN) on the host, on a
                bounded sample (configs[0]'s network, a few tens of ms)
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_baseline(pkg, sample_ms):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    from helpers import OracleNet
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg)
    ora = OracleNet(Oracle(), cfg)
    # synaptic events = sum over delivered runs; count them from the spikes
    # and the out-degree per delay, exactly as delivered
    import numpy as np
    b = ora.backward()
    D = 20
    deg = np.zeros((ora.n, D), np.int64)
    np.add.at(deg, (b["pre"], b["delay"]), 1)
    hist = []
    events = 0
    t0 = time.perf_counter()
    for t in range(sample_ms):
        ora.fire()
        ids, _ = ora.last_spikes()
        hist.append(ids)
        for d in range(min(D, t + 1)):
            events += int(deg[hist[t - d], d].sum())
    el = time.perf_counter() - t0
    ora.close()
    return {"value": events / el, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"oracle (scalar C port of DPSNN) on configs[0]'s 5000-neuron network, "
                      f"{sample_ms} ms simulated after init, {el:.1f} s wall"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batch(pkg, args):
    """64 independent networks (grid side 2..12, own seeds, bath
    2-6 Hz per input, STDP on half) advancing in one batched step; then the
    same networks stepped one at a time for the same number of steps."""
    import tempfile
    W = importlib.import_module("201812realtimecortsim_amd.workloads")
    mix = W.batch_mix(count=args.batch_count)
    tmp = tempfile.mkdtemp(prefix="dpsnn_batch_")
    nets = []
    t0 = time.perf_counter()
    for i, spec in enumerate(mix):
        d = os.path.join(tmp, f"n{i:02d}")
        env = W.write_instance(d, spec)
        nets.append(pkg.Network(d, drive=pkg.DRIVE_PHILOX, philox_seed=7000 + i, env=env, drop_host_synapses=True))
        if i % 8 == 7 or i == len(mix) - 1:
            log(f"built {i + 1}/{len(mix)} networks, {time.perf_counter() - t0:.0f} s")
    init_s = time.perf_counter() - t0
    neurons = sum(n.n for n in nets)
    synapses = sum(n.stats()["synapses"] for n in nets)
    dev_bytes = sum(n.stats()["device_bytes"] for n in nets)
    b = pkg.Batch()
    for n in nets:
        b.add(n)
    b.step(args.warmup)
    b.sync()
    shape = b.tune(1) if not args.no_tune else b.shape()
    log(f"tuned shape {shape}")
    b.step(5)
    b.sync()

    def totals():
        st = [n.stats() for n in nets]
        return sum(x["synaptic_events"] for x in st), sum(x["spikes"] for x in st)

    e0, s0 = totals()
    b.profile(True)
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    kms, launches = b.kernel_time()
    e1, s1 = totals()
    batched = (e1 - e0) / wall
    log(f"batched: {1000 * wall / args.steps:.2f} ms per step, {batched:.3e} events/s")
    for n in list(nets):
        b.remove(n)
    b.close()
    pe, pw = 0, 0.0
    for n in nets:
        a = n.stats()["synaptic_events"]
        n.step(1)  # stand-alone scratch
        n.sync()
        a = n.stats()["synaptic_events"]
        t0 = time.perf_counter()
        n.step(args.steps)
        n.sync()
        pw += time.perf_counter() - t0
        pe += n.stats()["synaptic_events"] - a
    plain = pe / pw
    log(f"plain (one network at a time): {plain:.3e} events/s")
    for n in nets:
        n.close()
    per_launch = (25.0 * (e1 - e0) + 72.0 * neurons * args.steps) / args.steps
    achieved = per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": batched,
        "unit": "synaptic events/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"configs[3] batching: {len(mix)} independent LIFCA networks, grid side 2-12 "
                               f"columns x 1250 neurons, own seeds, bath 2-6 Hz/input, STDP on "
                               f"{sum(x['stdp'] for x in mix)}; one batched step = 1 ms of every network",
                   "networks": len(mix), "neurons": neurons, "synapses": synapses, "device_bytes": dev_bytes,
                   "drive": "philox", "init_seconds": init_s, "tuned_shape": shape,
                   "mean_rate_hz": (s1 - s0) / args.steps * 1000.0 / neurons},
        "batched_vs_plain": batched / plain,
        "plain_events_per_s": plain,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "batched step (graph of k_batch_gather + k_batch_integrate per chunk)",
                     "kernel_ms": kms, "launches": launches, "algorithmic_bytes_per_launch": per_launch},
        "cpu_baseline": None,
    }


def run_server(pkg, jobs, dirs, envs, args, max_running):
    """Replays the job arrivals against a server for the window; returns
    what it finished and delivered by the window's end."""
    srv = pkg.Server(hbm_budget_bytes=int(args.serve_budget_gb * 1e9), build_threads=args.build_threads,
                     steps_per_round=10, max_running=max_running)
    t0 = time.perf_counter()
    ids, k, last = [], 0, t0
    while True:
        now = time.perf_counter() - t0
        while k < len(jobs) and jobs[k]["arrival_s"] <= now:
            ids.append(srv.submit(dirs[k], drive=pkg.DRIVE_PHILOX, philox_seed=9000 + k, env=envs[k],
                                  drop_host_synapses=True))
            k += 1
        if now >= args.serve_window:
            break
        if time.perf_counter() - last >= 10.0:
            st = srv.stats()
            log(f"{'plain' if max_running == 1 else 'served'} t={now:5.0f} s: submitted {st['submitted']} "
                f"building {st['building']} waiting {st['waiting']} running {st['running']} "
                f"finished {st['finished']} events {st['synaptic_events']:.3e}")
            last = time.perf_counter()
        time.sleep(0.005)
    st = srv.stats()
    infos = [srv.job(j) for j in ids]
    srv.close()
    done = [i for i in infos if i["state"] == "done"]
    lat = np.array([i["finished_s"] - i["submitted_s"] for i in done]) if done else np.zeros(0)
    pct = (lambda q: float(np.percentile(lat, q)) if lat.size else None)
    return {"submitted": len(ids), "finished": len(done), "failed": sum(i["state"] == "failed" for i in infos),
            "running_at_end": st["running"], "waiting_at_end": st["waiting"], "building_at_end": st["building"],
            "events": st["synaptic_events"], "elapsed_s": st["elapsed_s"],
            "events_per_s": st["synaptic_events"] / st["elapsed_s"],
            "latency_p50_s": pct(50), "latency_p99_s": pct(99),
            "finished_sim_ms": sum(i["stats"]["steps"] for i in done),
            "step_s": st["step_s"], "upload_s": st["upload_s"]}


def bench_serve(pkg, args):
    """Poisson job arrivals (20 jobs/s for 120 s), each job a
    LIFCA network of 1-16 x 1-16 columns, 0.5-5 s simulated, STDP with
    probability 0.3, admitted under a 200 GB HBM budget with continuous
    batching; the same arrival stream served one job at a time is the plain
    comparison."""
    import tempfile
    W = importlib.import_module("201812realtimecortsim_amd.workloads")
    jobs = W.serve_jobs(rate_per_s=args.serve_rate, window_s=args.serve_window)
    tmp = tempfile.mkdtemp(prefix="dpsnn_serve_")
    dirs, envs = [], []
    for i, spec in enumerate(jobs):
        d = os.path.join(tmp, f"j{i:05d}")
        envs.append(W.write_instance(d, spec))
        dirs.append(d)
    # offered work at the saturated rate of these networks (~400 Hz, 1125
    # synapses per neuron)
    offered = sum(j["side"] ** 2 * 1250 * 0.40 * 1125 * j["duration_ms"] for j in jobs) / args.serve_window
    log(f"{len(jobs)} jobs over {args.serve_window:.0f} s; offered ~{offered:.3e} synaptic events/s")
    served = run_server(pkg, jobs, dirs, envs, args, 0)
    log(f"served: {served}")
    plain = run_server(pkg, jobs, dirs, envs, args, 1)
    log(f"plain: {plain}")
    return {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": served["events_per_s"],
        "unit": "synaptic events/s",
        "n_gpus": 1, "steps": None, "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"configs[4] serving: Poisson arrivals {args.serve_rate:g} jobs/s for "
                               f"{args.serve_window:g} s, grid side 1-16 x 1250 neurons, 0.5-5 s simulated, "
                               f"STDP p=0.3, {args.serve_budget_gb:g} GB HBM admission budget, continuous batching",
                   "jobs_arrived": len(jobs), "offered_events_per_s": offered, "drive": "philox"},
        "served": served,
        "plain": plain,
        "served_vs_plain": served["events_per_s"] / plain["events_per_s"] if plain["events_per_s"] else None,
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_copy(args):
    """A device-to-device copy of known size (256 MiB, five times): the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 256 * 1024 * 1024
    x = torch.empty(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes": n, "copies": 5}


def traffic_from_profile(config):
    """HBM bytes per step of this workload from the committed counter summary
    (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, scaled by the
    same counters read for a plain copy) -- measured by a profiled run of
    this build, not by this run -- with where and over which steps it was
    measured; or None."""
    path = os.path.join(ROOT, "profiles", f"step_{config}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        j = json.load(f)
    return {"bytes": j.get("hbm_bytes_per_step"),
            "source": f"profiles/step_{config}_traffic.json (counter passes over {j.get('workload', '?')})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--settle", type=int, default=100,
                    help="untimed ms run before the warm-up steps: the start-up transient of the "
                         "network (rates still rising) is not part of the timed regime")
    ap.add_argument("--config", default="c1")
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve", "copy"])
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-window", type=float, default=120.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--build-threads", type=int, default=4)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--drive", default="philox", choices=["philox", "replay"])
    ap.add_argument("--cpu-sample-ms", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs one network per GPU and this build measures one GPU; "
              "--gpus must be 1 (multi-GPU sharding is a design note, not a runnable mode)",
              file=sys.stderr)
        return 2
    pkg = importlib.import_module("201812realtimecortsim_amd")
    if args.mode == "batch":
        print(json.dumps(bench_batch(pkg, args)), flush=True)
        return 0
    if args.mode == "serve":
        print(json.dumps(bench_serve(pkg, args)), flush=True)
        return 0
    if args.mode == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return 0
    cfg = pkg.config_dir(args.config)
    total = args.settle + args.warmup + args.steps
    pkg.apply_env(cfg, env_totalSimTime_ms=max(total, 1), env_spikes_writeDisable=1)
    drive = pkg.DRIVE_PHILOX if args.drive == "philox" else pkg.DRIVE_REPLAY
    print(f"[bench] building {args.config} ...", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    net = pkg.Network(cfg, drive=drive, philox_seed=20181201)
    init_s = time.perf_counter() - t0
    st0 = net.stats()
    print(f"[bench] init {init_s:.1f} s: {st0['neurons']} neurons, {st0['synapses']} synapses, "
          f"{st0['device_bytes'] / 1e9:.1f} GB on device", file=sys.stderr, flush=True)
    net.step(args.settle)
    net.step(args.warmup)
    net.sync()
    a = net.stats()
    t_first = args.settle + args.warmup
    net.profile(True)
    t0 = time.perf_counter()
    net.step(args.steps)
    net.sync()
    wall = time.perf_counter() - t0
    b = net.stats()
    kms, launches = net.kernel_time(0)
    net.close()
    traffic = traffic_from_profile(args.config)
    events = b["synaptic_events"] - a["synaptic_events"]
    ext = b["external_events"] - a["external_events"]
    spikes = b["spikes"] - a["spikes"]
    neurons = st0["neurons"]
    per_launch = (25.0 * events + 72.0 * neurons * args.steps +
                  (16.0 * ext if drive == pkg.DRIVE_REPLAY else 0.0)) / max(args.steps, 1)
    achieved = per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    line = {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": events / wall,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "real_time_factor": (wall / (args.steps / 1000.0)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"configs[{args.config[1:]}] plain LIFCA step, "
                               f"{st0['neurons']} neurons, {st0['synapses']} synapses, "
                               f"{args.drive} external drive",
                   "neurons": neurons, "synapses": st0["synapses"], "drive": args.drive,
                   "spikes_per_step": spikes / args.steps,
                   "mean_rate_hz": spikes / args.steps * 1000.0 / neurons,
                   "init_seconds": init_s,
                   "timed_ms": [t_first, t_first + args.steps],
                   "phase1": ("stencil window (k_collect + k_sort)" if b["window_form"] else
                              "in-list gather (k_gather)") if t_first >= 32 else "mixed (steps before 32 gather)",
                   "graph_replayed_steps": b["graph_steps"] - a["graph_steps"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "one step's device work (phase 1 + k_integrate), HIP events on the "
                               "network's stream",
                     "kernel_ms": kms, "launches": launches,
                     "algorithmic_bytes_per_launch": per_launch},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(pkg, args.cpu_sample_ms)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

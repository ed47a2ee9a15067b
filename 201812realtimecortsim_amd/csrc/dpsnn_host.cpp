// dpsnn_host.cpp -- host side of the engine (see dpsnn_host.h).
//
// Numeric types follow the reference declaration by declaration, because
// the connectome and the replayed drive must come out bit-identical:
//   * ini values are parsed into float and widened into double fields;
//   * weights are float until quantized to int16 (weightType);
//   * delays are computed in float and rounded with roundr2i.
// This file is compiled without FP contraction.
#include "dpsnn_host.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <sys/stat.h>

namespace dpsnn {

// --------------------------------------------------------------------------
// ran3 / NormDev.  randdev.c:116-189 (rand3 with the author's |mj| fix),
// :477-502 (SetRandomSeed: idum = -seed then one draw, so seed 0 after the
// first initialisation only advances the stream), :584-606 (polar
// Box-Muller whose second value is cached across reseeds).
// --------------------------------------------------------------------------
int32_t Ran3::next_int() {
  constexpr int32_t kSeed = 161803398;
  if (idum_ < 0 || iff_ == 0) {
    iff_ = 1;
    int32_t mj = kSeed - (idum_ < 0 ? -idum_ : idum_);
    mj = mj < 0 ? -mj : mj;
    mj %= kBig;
    ma_[55] = mj;
    int32_t mk = 1;
    for (int32_t i = 1; i <= 54; ++i) {
      const int32_t ii = (21 * i) % 55;
      ma_[ii] = mk;
      mk = mj - mk;
      if (mk < 0) mk += kBig;
      mj = ma_[ii];
    }
    for (int round = 0; round < 4; ++round)
      for (int32_t i = 1; i <= 55; ++i) {
        ma_[i] -= ma_[1 + (i + 30) % 55];
        if (ma_[i] < 0) ma_[i] += kBig;
      }
    inext_ = 0;
    inextp_ = 31;
    idum_ = 1;
  }
  if (++inext_ == 56) inext_ = 1;
  if (++inextp_ == 56) inextp_ = 1;
  int32_t mj = ma_[inext_] - ma_[inextp_];
  if (mj < 0) mj += kBig;
  ma_[inext_] = mj;
  ++count_;
  return mj;
}

void Ran3::seed(int32_t s) {
  idum_ = -s;
  (void)next();
}

double Ran3::norm() {
  if (set_) {
    set_ = 0;
    return cached_;
  }
  double v1, v2, r;
  do {
    v1 = (double)(2.0 * next() - 1.0);
    v2 = (double)(2.0 * next() - 1.0);
    r = v1 * v1 + v2 * v2;
  } while (r >= 1.0 || r == 0.0);
  const double fac = (double)std::sqrt(-2.0 * std::log(r) / r);
  cached_ = v1 * fac;
  set_ = 1;
  return v2 * fac;
}

Ran3::State Ran3::state() const {
  State st{};
  st.idum = idum_; st.inext = inext_; st.inextp = inextp_; st.iff = iff_; st.set = set_;
  std::memcpy(st.ma, ma_, sizeof ma_);
  st.cached = cached_;
  st.count = count_;
  return st;
}

void Ran3::set_state(const State &st) {
  idum_ = st.idum; inext_ = st.inext; inextp_ = st.inextp; iff_ = st.iff; set_ = st.set;
  std::memcpy(ma_, st.ma, sizeof ma_);
  cached_ = st.cached;
  count_ = st.count;
}

// --------------------------------------------------------------------------
// erflib pieces: roundr2i (erflib.c:172-181) and the truncated Gaussian
// LUT of the bath (erflib.c:41-169).
// --------------------------------------------------------------------------
static uint32_t round_half_up(double r) {
  const double f = std::floor(r);
  return (f + 0.5 > r) ? (uint32_t)f : (uint32_t)(f + 1.0);
}

static double erfc_nr(double x) {  // erfcc, Numerical Recipes
  const double z = std::fabs(x), t = 1.0 / (1.0 + 0.5 * z);
  const double ans = t * std::exp(-z * z - 1.26551223 + t * (1.00002368 + t * (0.37409196 +
                     t * (0.09678418 + t * (-0.18628806 + t * (0.27886807 + t * (-1.13520398 +
                     t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}
static double cdf(double x, double mu, double sigma) {
  return 1 - 0.5 * erfc_nr((x - mu) / (std::sqrt(2.0) * sigma));
}
static double bisect_quantile(double p, double mu, double sigma, double a, double b) {
  if (p <= 0.0) return a;
  if (p >= 1.0) return b;
  double lo = a, hi = b;
  const double c0 = cdf(a, mu, sigma);
  for (int it = 0; (hi - lo) > 1e-6 && it < 256; ++it) {
    const double x = (hi + lo) * 0.5;
    if ((cdf(x, mu, sigma) - c0) / (cdf(b, mu, sigma) - c0) < p) lo = x; else hi = x;
  }
  return (hi + lo) * 0.5;
}
static void gaussian_lut(double *tab, int n, double mu, double sigma, double xmin, double xmax) {
  if (!(sigma > 0.0)) {
    for (int i = 0; i < n; ++i) tab[i] = mu;
    return;
  }
  const double dp = 1.0 / (double)n;
  double x1 = xmin;
  int i = 0;
  for (double p = dp; p <= 1.0; p += dp, ++i) {
    const double x0 = x1;
    x1 = bisect_quantile(p, mu, sigma, xmin, xmax);
    tab[i] = (x1 + x0) * 0.5;
  }
}

// --------------------------------------------------------------------------
// Parameters: DPSNN_getParameters.c:23-831 (env), :834-898 (column.txt
// sizes), :900-1053 (plasticity.txt) and DPSNN_LIFCAconnectome.cpp:16-131
// (connectivity.txt), :240-300 (bathEfficacy.txt), :302-398 (column.txt
// values), :400-487 (delay range).
// --------------------------------------------------------------------------
namespace {

struct EnvReader {
  std::string &err;
  const EnvMap *over;
  bool ok = true;
  const char *get(const char *name) const {
    if (over) {
      auto it = over->find(name);
      if (it != over->end()) return it->second.c_str();
    }
    return std::getenv(name);
  }
  uint32_t u32(const char *name) {
    const char *s = get(name);
    if (!s) {
      if (ok) err = std::string("missing env var ") + name;
      ok = false;
      return 0;
    }
    return (uint32_t)std::atoi(s);
  }
  int32_t flag(const char *name) {
    const uint32_t v = u32(name);
    if (ok && v > 1) {
      err = std::string("invalid value of ") + name;
      ok = false;
    }
    return (int32_t)v;
  }
};

// One ini line: whitespace tokens until a '#'-token, each through atof into
// float (the readers' float initValue[]).  Returns the token count.
int ini_tokens(char *line, std::vector<float> &vals) {
  int n = 0;
  char *save = nullptr;  // strtok_r: networks may be read on several threads at once
  for (char *tok = strtok_r(line, " \t\n", &save); tok; tok = strtok_r(nullptr, " \t\n", &save)) {
    if (tok[0] == '#') break;
    if ((size_t)n < vals.size()) vals[n] = (float)std::atof(tok);
    ++n;
  }
  return n;
}

FILE *open_ini(const std::string &dir, const char *name) {
  return std::fopen((dir + "/" + name).c_str(), "r");
}

bool read_columns(const std::string &dir, Params &p, std::string &err) {
  FILE *f = open_ini(dir, "column.txt");
  if (!f) { err = "cannot open column.txt"; return false; }
  std::vector<char> line(1 << 12);
  std::vector<float> v(12);
  uint32_t npc = 0, nsp = 0;
  while (std::fgets(line.data(), (int)line.size(), f)) {
    std::fill(v.begin(), v.end(), 0.0f);
    const int n = ini_tokens(line.data(), v);
    if (n <= 0) continue;
    if (n < 3) { err = "column.txt: a line needs at least 3 values"; std::fclose(f); return false; }
    const uint32_t s = (uint32_t)v[0];
    if (s >= kMaxSubPop) { err = "column.txt: sub-population id out of range"; std::fclose(f); return false; }
    SubPop &q = p.sp[s];
    q.count = (uint32_t)v[1];
    q.offset = s == 0 ? 0 : p.sp[s - 1].offset + p.sp[s - 1].count;
    q.JExt = v[2]; q.DJExt = v[3]; q.NuExt = v[4]; q.Tau = v[5]; q.Theta = v[6];
    q.H = v[7]; q.Tarp = v[8]; q.AlphaC = v[9]; q.TauC = v[10]; q.gC = v[11];
    npc += (uint32_t)v[1];
    ++nsp;
  }
  std::fclose(f);
  if (npc == 0) { err = "column.txt: neuronsPerCM <= 0"; return false; }
  p.pub.subPopNumber = nsp;
  p.pub.neuronsPerCM = npc;
  return true;
}

bool read_connectivity(const std::string &dir, Params &p, std::string &err) {
  FILE *f = open_ini(dir, "connectivity.txt");
  if (!f) { err = "cannot open connectivity.txt"; return false; }
  const uint32_t mx = p.maxModDeltaX, my = p.maxModDeltaY;
  const uint32_t cells = (mx + 1) * (my + 1);
  std::vector<char> line(10 * kMaxStencil * kMaxStencil + 64);
  std::vector<float> v(6 + kMaxStencil * kMaxStencil);
  std::vector<float> spec(size_t(kMaxSubPop) * kMaxSubPop * kMaxStencil * kMaxStencil, 0.0f);
  auto at = [&](uint32_t s, uint32_t t, uint32_t x, uint32_t y) -> float & {
    return spec[((size_t(s) * kMaxSubPop + t) * kMaxStencil + x) * kMaxStencil + y];
  };
  while (std::fgets(line.data(), (int)line.size(), f)) {
    std::fill(v.begin(), v.end(), 0.0f);
    const int n = ini_tokens(line.data(), v);
    if (n <= 0) continue;
    if ((uint32_t)(n - 6) != cells) {
      err = "connectivity.txt: number of stencil values does not match env_stencilDim";
      std::fclose(f);
      return false;
    }
    const uint32_t s = (uint32_t)v[0], t = (uint32_t)v[1];
    if (s >= p.pub.subPopNumber || t >= p.pub.subPopNumber) {
      err = "connectivity.txt: sub-population id out of range";
      std::fclose(f);
      return false;
    }
    p.sp[s].J[t] = v[2]; p.sp[s].DJ[t] = v[3]; p.sp[s].DMin[t] = v[4]; p.sp[s].DMax[t] = v[5];
    uint32_t k = 6;
    for (uint32_t y = 0; y <= my; ++y)
      for (uint32_t x = 0; x <= mx; ++x) at(s, t, x, y) = v[k++];
  }
  std::fclose(f);
  p.prob.assign(size_t(kMaxSubPop) * kMaxSubPop * kMaxStencil * kMaxStencil, 0.0f);
  auto pr = [&](uint32_t s, uint32_t t, uint32_t x, uint32_t y) -> float & {
    return p.prob[((size_t(s) * kMaxSubPop + t) * kMaxStencil + x) * kMaxStencil + y];
  };
  for (uint32_t s = 0; s < p.pub.subPopNumber; ++s)
    for (uint32_t t = 0; t < p.pub.subPopNumber; ++t) {
      if (p.pub.overallConnectivity == 0) {  // explicit stencil: mirror quadrant
        for (uint32_t x = 0; x <= mx; ++x)
          for (uint32_t y = 0; y <= my; ++y) {
            const float q = at(s, t, x, y);
            pr(s, t, mx + x, my + y) = q;
            pr(s, t, mx - x, my + y) = q;
            pr(s, t, mx - x, my - y) = q;
            pr(s, t, mx + x, my - y) = q;
          }
      } else {  // homogeneous
        for (uint32_t x = 0; x < p.stencilX; ++x)
          for (uint32_t y = 0; y < p.stencilY; ++y) pr(s, t, x, y) = at(s, t, 0, 0);
      }
    }
  return true;
}

bool read_bath(const std::string &dir, Params &p, std::string &err) {
  FILE *f = open_ini(dir, "bathEfficacy.txt");
  if (!f) { err = "cannot open bathEfficacy.txt"; return false; }
  p.bath.assign(size_t(kMaxStencil) * kMaxStencil, 0);
  std::vector<char> line(100000);
  uint32_t y = 0;
  while (std::fgets(line.data(), (int)line.size(), f)) {
    uint32_t x = 0, n = 0;
    char *save = nullptr;
    for (char *tok = strtok_r(line.data(), " \t\n", &save); tok; tok = strtok_r(nullptr, " \t\n", &save)) {
      if (tok[0] == '#') break;
      if (x < kMaxStencil && y < kMaxStencil) p.bath[size_t(x) * kMaxStencil + y] = (uint32_t)std::atoi(tok);
      ++x;
      ++n;
    }
    if (n == 0) continue;
    if (n != p.pub.stencilDim) {
      err = "bathEfficacy.txt: row length does not match env_stencilDim";
      std::fclose(f);
      return false;
    }
    ++y;
  }
  std::fclose(f);
  return true;
}

bool read_plasticity(const std::string &dir, Params &p, std::string &err) {
  dpsnn_parameters &q = p.pub;
  FILE *f = open_ini(dir, "plasticity.txt");
  if (!f) {  // defaults, DPSNN_getParameters.c:916-928
    q.antiCausalCoeff = -1.0f; q.STDPmultiplier = 0.001f; q.tauSTDP_ms = 20.0f;
    q.causalSTDPmaxAbsValue = 0.1f; q.antiCausalSTDPmaxAbsValue = 0.12f;
    q.tauDerivativeDecay_s = 1.0f; q.derivativeDrift = 0.0f;
    return true;
  }
  std::vector<char> line(200);
  std::vector<float> v(20, 0.0f);
  int n = 0;
  while (std::fgets(line.data(), (int)line.size(), f)) n = ini_tokens(line.data(), v);  // last line wins
  std::fclose(f);
  if (n == 0 || n >= 20) { err = "plasticity.txt: empty or too many values"; return false; }
  if (v[0] != 1.0f) { err = "plasticity.txt: unknown plasticity algorithm"; return false; }
  if (n != 8) { err = "plasticity.txt: expected 8 values"; return false; }
  q.antiCausalCoeff = v[1]; q.STDPmultiplier = v[2]; q.tauSTDP_ms = v[3];
  q.causalSTDPmaxAbsValue = v[4]; q.antiCausalSTDPmaxAbsValue = v[5];
  q.tauDerivativeDecay_s = v[6]; q.derivativeDrift = v[7];
  return true;
}

}  // namespace

bool parse_env_overrides(const char *text, EnvMap &out, std::string &err) {
  out.clear();
  if (!text) return true;
  std::string s(text);
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t end = s.find_first_of(";\n", pos);
    if (end == std::string::npos) end = s.size();
    std::string item = s.substr(pos, end - pos);
    const size_t a = item.find_first_not_of(" \t\r");
    if (a != std::string::npos) {
      item = item.substr(a, item.find_last_not_of(" \t\r") - a + 1);
      const size_t eq = item.find('=');
      if (eq == std::string::npos || eq == 0) {
        err = "environment override without name=value: " + item;
        return false;
      }
      out[item.substr(0, eq)] = item.substr(eq + 1);
    }
    pos = end + 1;
  }
  return true;
}

unsigned default_threads() {
  unsigned n = std::max(1u, std::thread::hardware_concurrency());
  if (const char *e = std::getenv("OMP_NUM_THREADS")) {  // the process's CPU share, where set
    const int k = std::atoi(e);
    if (k > 0) n = std::min(n, (unsigned)k);
  }
  return n;
}

bool read_params(const std::string &dir, Params &p, std::string &err, const EnvMap *over) {
  dpsnn_parameters &q = p.pub;
  q = dpsnn_parameters{};
  EnvReader env{err, over};
  q.overallConnectivity = env.u32("env_overallConnectivity");
  if (env.ok && q.overallConnectivity > 1) { err = "invalid env_overallConnectivity"; return false; }
  q.stencilDim = env.u32("env_stencilDim");
  if (!env.ok) return false;
  if (q.stencilDim == 0 || q.stencilDim > kMaxStencil) { err = "env_stencilDim out of range"; return false; }
  if (q.overallConnectivity == 0) {
    p.maxModDeltaX = p.maxModDeltaY = q.stencilDim - 1;
    p.stencilX = p.stencilY = 2 * p.maxModDeltaX + 1;
  } else {
    p.maxModDeltaX = p.maxModDeltaY = 0;
    p.stencilX = p.stencilY = q.stencilDim;
  }
  if (p.stencilX > kMaxStencil) { err = "stencil larger than 96 columns"; return false; }
  q.globH = env.u32("env_globH");
  if (env.ok && q.globH != 1) { err = "env_globH must be 1: one process's whole local network per device"; return false; }
  const uint32_t synGen = env.u32("env_synGen");
  if (env.ok && synGen != 2) { err = "unsupported env_synGen (only 2, simpleCorticalModule)"; return false; }
  if (!env.ok) return false;
  if (!read_columns(dir, p, err)) return false;
  q.globCFX = env.u32("env_CFX");
  q.globCFY = env.u32("env_CFY");
  if (!env.ok) return false;
  q.globCFT = q.globCFX * q.globCFY;
  if (q.globCFT == 0 || q.globCFT > 96 * 96 + 1) { err = "globCFT out of range"; return false; }
  q.globN = q.locN = q.neuronsPerCM * q.globCFT;
  if (q.globN >= (1u << 24)) { err = "more than 2^24 neurons per device network"; return false; }
  const char *ws = env.get("env_minSynWeight_f");
  if (!ws) { err = "missing env var env_minSynWeight_f"; return false; }
  q.minSynWeight_f = (float)std::atof(ws);
  if (q.minSynWeight_f >= 0 || q.minSynWeight_f < -6.0) { err = "env_minSynWeight_f out of range"; return false; }
  q.factorFloat_2_weightType = -32767 / q.minSynWeight_f;
  q.factorWeightType_2_Float = q.minSynWeight_f / -32767;
  q.tabNeuron = q.globN / q.globCFT;  // makeMaxCMFract = 1
  q.D = env.u32("env_D");
  if (env.ok && (q.D == 0 || q.D > kDelaySlots)) { err = "env_D out of range (1..80)"; return false; }
  q.totalSimTime_ms = env.u32("env_totalSimTime_ms");
  q.startPlasticity_ms = env.u32("env_startPlasticity_ms");
  q.stopPlasticity_ms = env.u32("env_stopPlasticity_ms");
  q.moduloSec = env.u32("env_moduloSec");
  q.startStatFiles_s = env.u32("env_startStatFiles_s");
  q.stopStatFiles_s = env.u32("env_stopStatFiles_s");
  q.spikes_writeDisable = env.flag("env_spikes_writeDisable");
  p.stdpEventDisable = env.flag("env_statSTDPevent_writeDisable");
  q.statSynPeriodicProbe_writeDisable = env.flag("env_statSynPeriodicProbe_writeDisable");
  q.statRatesPerPop_writeDisable = env.flag("env_statRatesPerPop_writeDisable");
  q.startRatesPerPop_s = env.u32("env_startRatesPerPop_s");
  q.stopRatesPerPop_s = env.u32("env_stopRatesPerPop_s");
  q.startSynPeriodicProbe_s = env.u32("env_startSynPeriodicProbe_s");
  q.stopSynPeriodicProbe_s = env.u32("env_stopSynPeriodicProbe_s");
  p.startSTDPevent_s = env.u32("env_startSTDPevent_s");
  p.stopSTDPevent_s = env.u32("env_stopSTDPevent_s");
  // optional (DPSNN_getParameters.c:754-760): 0 = no debug window
  const char *dbg = env.get("env_debugPrintEnable_ms");
  p.debugPrintEnable_ms = dbg ? (uint32_t)std::atoi(dbg) : 0;
  q.ratesSampling = env.u32("env_ratesSampling");
  if (!env.ok) return false;
  if (q.moduloSec == 0) { err = "env_moduloSec must be > 0"; return false; }
  if (q.ratesSampling == 0) { err = "env_ratesSampling must be > 0"; return false; }
  const char *gs = env.get("env_globalSeed");
  if (!gs) { err = "missing env var env_globalSeed"; return false; }
  if (std::atoi(gs) < 0) { err = "env_globalSeed < 0 asks for a time-based seed; give a fixed seed"; return false; }
  q.globalSeed = (uint32_t)std::atoi(gs);
  const char *hm = env.get("env_howManyOutputFiles");
  q.howManyOutputFiles = hm ? (uint32_t)std::atoi(hm) : 0;
  if (q.howManyOutputFiles > 1) q.howManyOutputFiles = 0;
  if (!read_plasticity(dir, p, err)) return false;
  if (q.tauSTDP_ms * 3 > kStdpSpan || q.tauSTDP_ms <= 0 || q.STDPmultiplier < 0) {
    err = "plasticity.txt: tauSTDP_ms must be in (0, 128/3] and STDPmultiplier >= 0";
    return false;
  }
  (void)env.u32("env_collectivesAlgo");
  (void)env.u32("env_pktLength");
  if (!env.ok) return false;
  if (!read_bath(dir, p, err) || !read_connectivity(dir, p, err)) return false;
  double dmin = 1e36, dmax = 0;
  for (uint32_t s = 0; s < q.subPopNumber; ++s)
    for (uint32_t t = 0; t < q.subPopNumber; ++t) {
      dmax = std::max(dmax, p.sp[s].DMax[t]);
      dmin = std::min(dmin, p.sp[s].DMin[t]);
    }
  q.delayMin = (uint32_t)dmin;
  q.delayMax = (uint32_t)dmax;
  return true;
}

// --------------------------------------------------------------------------
// Connectome generator: simpleCM_invokeIndivSynGenerators
// (DPSNN_LIFCAcorticalModuleGen.cpp:62-220) with generateTargetNeuList's
// Bernoulli branch (DPSNN_LIFCAconnectome.cpp:513-823, non-periodic edges),
// getEmptyRandomSynapses (:1079-1133), getRandomDelay_UNI (:299-338),
// the truncated-normal weight draw (:158-183) and generateSourceNeuCExt
// (:871-887).  The random stream is reseeded with the column index at each
// column's first neuron; the NormDev cache and the seed-0 quirk carry over.
// --------------------------------------------------------------------------
// External bath of neuron i of sub-population src, in column i / npc
// (generateSourceNeuCExt's mirrored index, DPSNN_LIFCAconnectome.cpp:871-887):
// its mean inter-event time and per-ms event cap.
void neuron_bath(const Params &p, uint32_t i, uint32_t src, Connectome &c) {
  const dpsnn_parameters &q = p.pub;
  const uint32_t col = i / q.neuronsPerCM;
  const float nu = (float)p.sp[src].NuExt;
  uint32_t bx = col % q.globCFX, by = col / q.globCFX;
  if (bx >= q.globCFX / 2) bx = -1 * (bx - q.globCFX + 1);
  if (by >= q.globCFY / 2) by = -1 * (by - q.globCFY + 1);
  if (bx >= q.stencilDim) bx = q.stencilDim - 1;
  if (by >= q.stencilDim) by = q.stencilDim - 1;
  const float cext = (float)p.bath[size_t(bx) * kMaxStencil + by];
  if (cext != 0) {
    c.invNuExt[i] = 1000.0 / (nu * cext);
    c.maxExt[i] = (uint32_t)std::ceil(20 / c.invNuExt[i]);
  } else {
    c.invNuExt[i] = 0;
    c.maxExt[i] = 0;
  }
}

namespace {

struct Target {
  uint32_t id;
  uint8_t sub;
};

uint32_t sub_pop_in_column(const Params &p, uint32_t inCM) {
  for (uint32_t s = 0; s < p.pub.subPopNumber; ++s)
    if (inCM >= p.sp[s].offset && inCM < p.sp[s].count + p.sp[s].offset) return s;
  return 0;
}

// getEmptyRandomSynapses (DPSNN_LIFCAconnectome.cpp:1079-1133) draws r and
// returns the first n with C_n >= r, or maxN + 1 if there is none, where
// P_1 = C_1 = q, P_{n+1} = P_n * (1 - q) and C_{n+1} = C_n + P_{n+1}, each
// rounded in double.  C_n depends on (q, n) only, and rounding keeps it
// non-decreasing, so that scan is a lower_bound over C_1..C_maxN accumulated
// once per (q, maxN) with the same operations in the same order: the same n
// for every r, in log2(maxN) compares instead of one dependent multiply-add
// per candidate neuron skipped.  Tables are built until they hold kMaxEntries
// values (large stencils with many distinct probabilities); the cells left
// over accumulate C_n as they go, as the reference does.
class GapTables {
 public:
  explicit GapTables(const Params &p)
      : ns_(p.pub.subPopNumber), sx_(p.stencilX), sy_(p.stencilY),
        slot_(size_t(ns_) * ns_ * sx_ * sy_, kNone) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> made;  // (prob bits, maxN) -> table
    for (uint32_t s = 0; s < ns_; ++s)
      for (uint32_t t = 0; t < ns_; ++t)
        for (uint32_t x = 0; x < sx_; ++x)
          for (uint32_t y = 0; y < sy_; ++y) {
            const float prob = p.prob_at(s, t, x, y);
            if (!(prob > 0.00001)) continue;  // the scan never runs for these
            const uint32_t maxN = p.sp[t].count + 1;
            uint32_t bits;
            std::memcpy(&bits, &prob, sizeof bits);
            const auto key = std::make_pair(bits, maxN);
            if (!made.count(key) && cum_.size() + maxN > kMaxEntries) continue;
            const auto ins = made.emplace(key, (uint32_t)start_.size());
            if (ins.second) {
              start_.push_back(cum_.size());
              const double q = (double)prob;
              double P = q, C = P;
              for (uint32_t n = 1; n <= maxN; ++n) {
                cum_.push_back(C);
                P *= 1 - q;
                C += P;
              }
            }
            slot_[index(s, t, x, y)] = ins.first->second;
          }
  }
  // C_1.. of source sub-population s, target t, stencil cell (x, y), or
  // nullptr for a cell without a table.
  const double *at(uint32_t s, uint32_t t, uint32_t x, uint32_t y) const {
    const uint32_t k = slot_[index(s, t, x, y)];
    return k == kNone ? nullptr : cum_.data() + start_[k];
  }

 private:
  static constexpr uint32_t kNone = ~0u;
  static constexpr size_t kMaxEntries = size_t(1) << 25;  // 256 MiB of doubles
  size_t index(uint32_t s, uint32_t t, uint32_t x, uint32_t y) const {
    return ((size_t(s) * ns_ + t) * sx_ + x) * sy_ + y;
  }
  uint32_t ns_, sx_, sy_;
  std::vector<uint32_t> slot_;
  std::vector<size_t> start_;
  std::vector<double> cum_;
};

// The gap to the next target of a scan whose remaining candidates number
// `rem` (the distance to the end of the target sub-population).  A gap of rem
// or more ends the scan wherever it lands, so it is returned as rem: only
// C_1..C_{rem-1} are searched.  The first few are checked in order (a dense
// stencil cell's gaps are short), the rest by binary search.
uint32_t geometric_gap(Ran3 &rng, const double *cum, float prob, uint32_t rem) {
  const double r = rng.next();
  const uint32_t lim = rem - 1;  // the largest gap that stays inside
  if (!cum) {  // no table: C_n accumulated as the reference does
    const double q = (double)prob;
    double P = q, C = P;
    for (uint32_t n = 1; n <= lim; ++n) {
      if (!(C < r)) return n;
      P *= 1 - q;
      C += P;
    }
    return rem;
  }
  const uint32_t head = std::min<uint32_t>(lim, 16);
  for (uint32_t i = 0; i < head; ++i)
    if (!(cum[i] < r)) return i + 1;
  return (uint32_t)(std::lower_bound(cum + head, cum + lim, r) - cum) + 1;
}

// getRandomDelay_UNI (DPSNN_LIFCAconnectome.cpp:299-338) for one (source,
// target) sub-population pair.  The float bounds of the delay levels depend
// on the pair only, so they are rounded once.  A draw is then the reference's
// product, sum and division in double, narrowed to float, scaled and rounded:
// with step > 0 and range >= 0 every one of these operations is monotone
// (IEEE rounding is), so the delay is a non-decreasing step function of
// ran3's integer draw mj.  Its rise points are found once by bisection on
// that same formula, and a draw looks its delay up among them.
struct DelayRule {
  float DMin = 0, range = 0, lo = 0, step = 0;
  uint32_t base = 0;          // the delay of mj = 0
  std::vector<int32_t> rise;  // rise[k]: the smallest mj whose delay is base + k + 1
  bool stepped = false;       // false: every draw evaluates the formula
  DelayRule() = default;
  DelayRule(const Params &p, uint32_t s, uint32_t t) {
    float dmin = (float)p.sp[s].DMin[t], dmax = (float)p.sp[s].DMax[t];
    lo = (float)p.pub.delayMin;
    const float hi = (float)p.pub.delayMax, levels = (float)kDelaySlots;
    step = (hi - lo) / levels;
    lo += step / 2;
    dmax = round_half_up((dmax - lo) / step) * step + lo;
    dmin = round_half_up((dmin - lo) / step) * step + lo;
    DMin = dmin;
    range = dmax - dmin;
    const int32_t last = Ran3::kBig - 1;
    if (!(step > 0) || !(range >= 0) || !(scaled(0) >= 0) || !(scaled(last) < 256)) return;
    base = round_half_up(scaled(0));
    for (uint32_t v = base + 1; v <= round_half_up(scaled(last)); ++v) {
      int32_t a = 0, b = last;  // value(b) >= v > value(a)
      while (b - a > 1) {
        const int32_t m = a + (b - a) / 2;
        (value(m) >= v ? b : a) = m;
      }
      rise.push_back(b);
    }
    stepped = true;
  }
  // the level before rounding, for the draw mj
  double scaled(int32_t mj) const {
    const double r = mj * (1.0 / (double)Ran3::kBig);
    const float level = ((r * range + DMin - lo) / step);
    return level * step;
  }
  uint32_t value(int32_t mj) const { return round_half_up(scaled(mj)); }
  uint32_t draw(Ran3 &rng) const { return at(rng.next_int()); }
  // value(mj), from the rise points where there are some
  uint32_t at(int32_t mj) const {
    if (!stepped) return value(mj);
    const uint32_t n = (uint32_t)rise.size();
    uint32_t k = (uint32_t)((uint64_t)mj * n / Ran3::kBig);  // near the answer
    while (k > 0 && mj < rise[k - 1]) --k;
    while (k < n && mj >= rise[k]) ++k;
    return base + k;
  }
};

// What every column's generation reads and no draw changes: the gap tables,
// the delay rule of each sub-population pair and each in-column index's
// sub-population.
struct ColumnRules {
  GapTables gaps;
  std::vector<DelayRule> delay;  // [source][target]
  std::vector<uint8_t> subOf;    // [index in column]
  explicit ColumnRules(const Params &p) : gaps(p) {
    const uint32_t ns = p.pub.subPopNumber;
    for (uint32_t s = 0; s < ns; ++s)
      for (uint32_t t = 0; t < ns; ++t) delay.emplace_back(p, s, t);
    for (uint32_t k = 0; k < p.pub.neuronsPerCM; ++k) subOf.push_back((uint8_t)sub_pop_in_column(p, k));
  }
};

}  // namespace

// Synapses of one column in generation order, with what the column-parallel
// resolution needs: the NormDev state at exit and, when the column was
// generated assuming a cached Gaussian on entry, which weight consumed it.
struct ColumnOut {
  std::vector<uint32_t> count;  // synapses per neuron of the column
  std::vector<uint32_t> post;
  std::vector<uint8_t> delay;
  std::vector<int16_t> w;
  int32_t exitSet = 0;
  double exitCached = 0;
  bool usedNorm = false;        // any NormDev call in the column
  int64_t patchAt = -1;         // synapse whose weight used the entry cache
  float patchMean = 0, patchSigma = 0;
  bool ok = true;
  std::string err;
  Ran3 exitRng;                 // the stream where the column ended
};

namespace {

// An empty column that keeps `old`'s array memory, for the next column.
ColumnOut recycled(ColumnOut &&old) {
  ColumnOut o;
  o.count = std::move(old.count);
  o.post = std::move(old.post);
  o.delay = std::move(old.delay);
  o.w = std::move(old.w);
  o.post.clear();
  o.delay.clear();
  o.w.clear();
  return o;
}

// Generates the neurons of column `col` from `rng` (already positioned at the
// column's first draw).  If `placeholderEntry`, the first NormDev value is
// the cached entry value, unknown yet: it is assumed to pass the weight
// bounds (it does unless |x| >= 1/DJ) and the weight is patched later.
void generate_column(const Params &p, const ColumnRules &rules, uint32_t col, Ran3 &rng, bool placeholderEntry,
                     Connectome &c, ColumnOut &out, bool writeNeuronInfo) {
  const dpsnn_parameters &q = p.pub;
  const uint32_t npc = q.neuronsPerCM, cfx = col % q.globCFX, cfy = col / q.globCFX;
  // the column of stencil cell (x, y), false if it lies off the grid
  auto cell_column = [&](uint32_t x, uint32_t y, uint32_t &colBase) {
    int32_t tx, ty;
    if (q.overallConnectivity == 0) {
      tx = (int32_t)cfx + (int32_t)x - (int32_t)p.maxModDeltaX;
      ty = (int32_t)cfy + (int32_t)y - (int32_t)p.maxModDeltaY;
    } else {
      tx = (int32_t)x;
      ty = (int32_t)y;
    }
    if (tx < 0 || tx >= (int32_t)q.globCFX || ty < 0 || ty >= (int32_t)q.globCFY) return false;
    colBase = (uint32_t)tx * npc + (uint32_t)ty * npc * q.globCFX;
    return true;
  };
  {
    // room for the expected synapse count plus a margin far above its
    // binomial spread: the arrays are neither regrown nor oversized
    double expect = 0;
    uint32_t colBase;
    for (uint32_t s = 0; s < q.subPopNumber; ++s)
      for (uint32_t ts = 0; ts < q.subPopNumber; ++ts)
        for (uint32_t x = 0; x < p.stencilX; ++x)
          for (uint32_t y = 0; y < p.stencilY; ++y) {
            const float prob = p.prob_at(s, ts, x, y);
            if (prob > 0 && cell_column(x, y, colBase))
              expect += (double)p.sp[s].count * prob * p.sp[ts].count;
          }
    const size_t cap = (size_t)(expect * 1.02) + 4096;
    out.post.reserve(cap);
    out.delay.reserve(cap);
    out.w.reserve(cap);
  }
  std::vector<Target> targets;
  targets.reserve(kMaxOutDegree);
  out.count.assign(npc, 0);
  bool pendingPlaceholder = placeholderEntry;
  for (uint32_t inCM = 0; inCM < npc; ++inCM) {
    const uint32_t i = col * npc + inCM;
    const uint32_t src = rules.subOf[inCM];
    if (writeNeuronInfo) c.subPop[i] = (uint8_t)src;
    targets.clear();
    for (uint32_t ts = 0; ts < q.subPopNumber; ++ts)
      for (uint32_t x = 0; x < p.stencilX; ++x)
        for (uint32_t y = 0; y < p.stencilY; ++y) {
          uint32_t colBase;
          if (!cell_column(x, y, colBase)) continue;
          const float prob = p.prob_at(src, ts, x, y);
          if (!(prob > 0)) continue;
          if (!(prob > 0.00001)) {
            out.ok = false;
            out.err = "connection probabilities in (0, 1e-5] take a branch the reference marks as incorrect; not supported";
            return;
          }
          const uint32_t first = p.sp[ts].offset, end = first + p.sp[ts].count;
          uint32_t cur = first - 1;  // wraps like the reference's uint32 arithmetic
          const double *cum = rules.gaps.at(src, ts, x, y);
          while ((cur += geometric_gap(rng, cum, prob, end - cur)) < end) {
            if (targets.size() + 1 >= kMaxOutDegree) {
              out.ok = false;
              out.err = "a neuron would get 7500 or more synapses (DSD__maxM)";
              return;
            }
            targets.push_back({colBase + cur, rules.subOf[cur]});
          }
        }
    for (const Target &tg : targets) {
      const uint32_t d = rules.delay[src * q.subPopNumber + tg.sub].draw(rng);
      const float mean = (float)p.sp[src].J[tg.sub];
      const float sigma = (float)(p.sp[src].DJ[tg.sub] * mean);
      float w;
      if (pendingPlaceholder) {
        // consumes the cached entry value: no draw; weight patched later
        pendingPlaceholder = false;
        rng.set_norm_state(0, 0.0);
        out.patchAt = (int64_t)out.post.size();
        out.patchMean = mean;
        out.patchSigma = sigma;
        w = mean;
      } else if (mean > 0.0) {
        do { w = (float)(mean + rng.norm() * sigma); } while (w < 0.0 || w >= 2.0 * mean);
      } else {
        do { w = (float)(mean + rng.norm() * sigma); } while (w <= 2.0 * mean || w > 0);
      }
      out.usedNorm = true;
      if (w < q.minSynWeight_f || w > -q.minSynWeight_f) {
        out.ok = false;
        out.err = "synaptic weight outside [env_minSynWeight_f, -env_minSynWeight_f]";
        return;
      }
      if (d >= kDelaySlots || d >= q.D) {
        out.ok = false;
        out.err = "a stored delay is not below env_D";
        return;
      }
      out.post.push_back(tg.id);
      out.delay.push_back((uint8_t)d);
      out.w.push_back((int16_t)(w * q.factorFloat_2_weightType));
      ++out.count[inCM];
    }
    if (!writeNeuronInfo) continue;
    neuron_bath(p, i, src, c);
  }
  out.exitSet = rng.norm_set();
  out.exitCached = rng.norm_cached();
  out.exitRng = rng;
}

}  // namespace

void connectome_neurons(const Params &p, Connectome &c) {
  const dpsnn_parameters &q = p.pub;
  const uint32_t N = q.globN;
  c.N = N;
  c.subPop.assign(N, 0);
  c.invNuExt.assign(N, 0.0);
  c.maxExt.assign(N, 0);
  c.bathLut.assign(size_t(q.subPopNumber) * kLutDepth, 0.0);
  for (uint32_t s = 0; s < q.subPopNumber; ++s) {  // initLUT, bath part
    const float J = (float)p.sp[s].JExt, DJ = (float)p.sp[s].DJExt;
    gaussian_lut(&c.bathLut[size_t(s) * kLutDepth], kLutDepth, J, J * DJ, 0.0, 2.0 * J);
  }
}

namespace {

// Sizes the backward list's five arrays to S entries.  Growing zero-fills
// (and first touches) them, so each array gets a thread of its own.
void grow_backward(Connectome &c, uint64_t S, unsigned threads) {
  std::function<void()> grow[] = {[&] { c.pre.resize(S); }, [&] { c.post.resize(S); },
                                  [&] { c.weight.resize(S); }, [&] { c.delay.resize(S); },
                                  [&] { c.kind.resize(S); }};
  const unsigned n = std::max(1u, std::min<unsigned>(threads, 5));
  std::vector<std::thread> pool;
  for (unsigned k = 0; k < n; ++k)
    pool.emplace_back([&, k]() {
      for (unsigned j = k; j < 5; j += n) grow[j]();
    });
  for (auto &t : pool) t.join();
}

// Runs fn(k) for k in [0, n) on up to `threads` threads, k handed out in order.
template <class Fn>
void for_each_index(uint32_t n, unsigned threads, Fn fn) {
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < std::max(1u, std::min<unsigned>(threads, n)); ++t)
    pool.emplace_back([&]() {
      for (uint32_t k; (k = next.fetch_add(1)) < n;) fn(k);
    });
  for (auto &t : pool) t.join();
}

// The backward list straight from the columns' generation-order synapses:
// inflate_backward's stable counting sort by delay with the columns as the
// chunks, which are contiguous neuron ranges in order.  Each column's arrays
// are released once scattered.
void inflate_columns(Connectome &c, std::vector<ColumnOut> &col, uint32_t npc, unsigned threads) {
  if (threads == 0) threads = default_threads();
  const uint32_t ncol = (uint32_t)col.size();
  uint64_t S = 0;
  for (const ColumnOut &o : col) S += o.post.size();
  c.S = S;
  grow_backward(c, S, threads);
  std::vector<uint64_t> cnt(size_t(ncol) * kDelaySlots, 0);
  for_each_index(ncol, threads, [&](uint32_t k) {
    uint64_t *cc = &cnt[size_t(k) * kDelaySlots];
    for (const uint8_t d : col[k].delay) ++cc[d];
  });
  // bucket-major, column-minor offsets
  uint64_t run = 0;
  uint32_t maxd = 0;
  for (uint32_t d = 0; d < kDelaySlots; ++d)
    for (uint32_t k = 0; k < ncol; ++k) {
      const uint64_t v = cnt[size_t(k) * kDelaySlots + d];
      if (v) maxd = d;
      cnt[size_t(k) * kDelaySlots + d] = run;
      run += v;
    }
  for_each_index(ncol, threads, [&](uint32_t k) {
    uint64_t *pos = &cnt[size_t(k) * kDelaySlots];
    const ColumnOut &o = col[k];
    uint64_t s = 0;
    for (uint32_t j = 0; j < npc; ++j) {
      const uint32_t i = k * npc + j;
      const uint8_t kind = c.subPop[i];
      for (const uint64_t e = s + o.count[j]; s < e; ++s) {
        const uint64_t b = pos[o.delay[s]]++;
        c.pre[b] = i;
        c.post[b] = o.post[s];
        c.delay[b] = o.delay[s];
        c.kind[b] = kind;
        c.weight[b] = o.w[s];
      }
    }
    col[k] = ColumnOut();
  });
  c.max_delay = maxd;
}

}  // namespace

bool generate_connectome(const Params &p, Ran3 &rng, Connectome &c, std::string &err, unsigned threads) {
  const dpsnn_parameters &q = p.pub;
  const uint32_t npc = q.neuronsPerCM, cols = q.globCFT;
  connectome_neurons(p, c);
  if (q.tabNeuron != npc) {
    err = "reseeding period (tabNeuron) differs from the column size";
    return false;
  }
  if (threads == 0) threads = default_threads();
  const ColumnRules rules(p);
  // column 0 continues the stream seeded with env_globalSeed: SetRandomSeed(0)
  // only advances it (DPSNN_LIFCAcorticalModuleGen.cpp:96-107, randdev.c:477-489)
  std::vector<ColumnOut> kept(cols);  // each column's synapses, in generation order
  {
    ColumnOut o;
    rng.seed(0);
    generate_column(p, rules, 0, rng, false, c, o, true);
    if (!o.ok) { err = o.err; return false; }
    kept[0] = std::move(o);
  }
  int32_t set = rng.norm_set();
  double cached = rng.norm_cached();
  // columns 1..: each restarts ran3 from its own index, while NormDev's
  // cached value carries over from the column before
  auto finish = [&](uint32_t col, ColumnOut &o) {
    set = o.exitSet;
    cached = o.exitCached;
    if (col == cols - 1) {
      // the simulation stream continues from where the last column ended
      rng = o.exitRng;
      rng.set_norm_state(set, cached);
    }
    kept[col] = std::move(o);
  };
  if (threads <= 2) {
    // one thread: each column from its true entry state, in order
    for (uint32_t col = 1; col < cols; ++col) {
      ColumnOut o;
      Ran3 r;
      r.seed((int32_t)col);
      r.set_norm_state(set, cached);
      generate_column(p, rules, col, r, false, c, o, true);
      if (!o.ok) { err = o.err; return false; }
      finish(col, o);
    }
    inflate_columns(c, kept, npc, threads);
    return true;
  }
  // several threads: both variants of a column (entry cache empty / full)
  // are generated ahead by a pool, at most `window` columns past the last
  // one resolved; this thread picks and patches them in column order
  struct Variants {
    ColumnOut a, b;
    int done = 0;
  };
  std::vector<Variants> var(cols);
  std::vector<ColumnOut> spare;  // variants not taken, their memory reused
  const uint32_t jobs = 2 * (cols - 1), window = threads + 2;
  uint32_t nextJob = 0, resolved = 0;
  bool stop = false;
  std::mutex mu;
  std::condition_variable cv;
  auto worker = [&]() {
    for (;;) {
      uint32_t job;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || nextJob >= jobs || 1 + nextJob / 2 <= resolved + window; });
        if (stop || nextJob >= jobs) return;
        job = nextJob++;
        if (!spare.empty()) {
          (job & 1 ? var[1 + job / 2].b : var[1 + job / 2].a) = std::move(spare.back());
          spare.pop_back();
        }
      }
      const uint32_t col = 1 + job / 2;
      const bool variantB = job & 1;
      Ran3 r;
      r.seed((int32_t)col);
      r.set_norm_state(0, 0.0);
      generate_column(p, rules, col, r, variantB, c, variantB ? var[col].b : var[col].a, !variantB);
      {
        std::lock_guard<std::mutex> lk(mu);
        ++var[col].done;
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (unsigned th = 0; th < std::min<unsigned>(threads, jobs); ++th) pool.emplace_back(worker);
  auto drain = [&]() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    for (auto &th : pool) th.join();
  };
  for (uint32_t col = 1; col < cols; ++col) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return var[col].done == 2; });
    }
    ColumnOut *o = set ? &var[col].b : &var[col].a;
    if (set && o->ok) {
      if (!o->usedNorm) {
        // nothing drew a Gaussian: the cache passes through unchanged
        o->exitSet = 1;
        o->exitCached = cached;
      } else {
        const float w = (float)(o->patchMean + cached * o->patchSigma);
        const bool inRange = o->patchMean > 0.0 ? !(w < 0.0 || w >= 2.0 * o->patchMean)
                                                : !(w <= 2.0 * o->patchMean || w > 0);
        if (inRange && !(w < q.minSynWeight_f || w > -q.minSynWeight_f)) {
          o->w[o->patchAt] = (int16_t)(w * q.factorFloat_2_weightType);
        } else {
          // the cached value is rejected by the weight bounds: regenerate
          // the column sequentially from the true entry state
          ColumnOut redo;
          Ran3 r;
          r.seed((int32_t)col);
          r.set_norm_state(1, cached);
          generate_column(p, rules, col, r, false, c, redo, false);
          var[col].b = std::move(redo);
        }
      }
    }
    if (!o->ok) {
      err = o->err;
      drain();
      return false;
    }
    finish(col, *o);
    ColumnOut unused = recycled(std::move(set ? var[col].a : var[col].b));
    var[col].a = ColumnOut();
    var[col].b = ColumnOut();
    {
      std::lock_guard<std::mutex> lk(mu);
      spare.push_back(std::move(unused));
      resolved = col;
    }
    cv.notify_all();
  }
  drain();
  inflate_columns(c, kept, npc, threads);
  return true;
}

void inflate_backward(Connectome &c, const std::vector<uint64_t> &gen_off, const std::vector<uint32_t> &g_post,
                      const std::vector<uint8_t> &g_delay, const std::vector<int16_t> &g_w, unsigned threads) {
  // inflate: stable counting sort by delay into the backward list,
  // DPSNN_localNet_init.c:651-680 (neuron chunks in parallel, each writing
  // its own sub-range of every delay bucket, so the order stays stable)
  const uint32_t N = c.N;
  if (threads == 0) threads = default_threads();
  const uint64_t S = g_post.size();
  c.S = S;
  grow_backward(c, S, threads);
  const unsigned chunks = std::max(1u, std::min<unsigned>(threads, (N + 4095) / 4096));
  std::vector<uint64_t> cnt(size_t(chunks) * kDelaySlots, 0);
  auto chunk_lo = [&](unsigned k) { return (uint32_t)((uint64_t)N * k / chunks); };
  {
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < chunks; ++k)
      pool.emplace_back([&, k]() {
        uint64_t *cc = &cnt[size_t(k) * kDelaySlots];
        for (uint64_t s = gen_off[chunk_lo(k)]; s < gen_off[chunk_lo(k + 1)]; ++s) ++cc[g_delay[s]];
      });
    for (auto &t : pool) t.join();
  }
  // bucket-major, chunk-minor offsets
  uint64_t run = 0;
  uint32_t maxd = 0;
  for (uint32_t d = 0; d < kDelaySlots; ++d)
    for (unsigned k = 0; k < chunks; ++k) {
      const uint64_t v = cnt[size_t(k) * kDelaySlots + d];
      if (v) maxd = d;
      cnt[size_t(k) * kDelaySlots + d] = run;
      run += v;
    }
  {
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < chunks; ++k)
      pool.emplace_back([&, k]() {
        uint64_t *pos = &cnt[size_t(k) * kDelaySlots];
        for (uint32_t i = chunk_lo(k); i < chunk_lo(k + 1); ++i)
          for (uint64_t s = gen_off[i]; s < gen_off[i + 1]; ++s) {
            const uint64_t b = pos[g_delay[s]]++;
            c.pre[b] = i;
            c.post[b] = g_post[s];
            c.delay[b] = g_delay[s];
            c.kind[b] = c.subPop[i];
            c.weight[b] = g_w[s];
          }
      });
    for (auto &t : pool) t.join();
  }
  c.max_delay = maxd;
}

// --------------------------------------------------------------------------
// Delivery layout.  The reference delivers, at step t, for d = D-1 down to
// 0, the spikes of the step t-d in ascending id, each through the run of its
// synapses that starts at offsetByDelay[pre][d] (DPSNN_localNet_init.c:
// 702-723) and continues while the source id matches -- findAndGetNextSyn
// (DPSNN_synapse.c:48-136) reads one entry past its bucket, so a run that
// ends a bucket continues into the next bucket when that bucket starts with
// the same source.  Events reach each target's input list in that order,
// which is also the tie order of its stable time sort.  Here each target
// gets the list of (delay of the run, source, synapse) in exactly that
// order; at step t an item is live when its source fired in step t-delay.
// --------------------------------------------------------------------------
void build_delivery(const Params &p, Connectome &c, bool with_incoming, unsigned threads) {
  const uint32_t N = c.N, D = std::min<uint32_t>(p.pub.D, kDelaySlots);
  const uint64_t S = c.S;
  if (threads == 0) threads = default_threads();
  // bucket boundaries of the backward list, which is sorted by delay
  std::vector<uint64_t> bstart(kDelaySlots + 1, 0);
  for (uint32_t d = 0; d <= kDelaySlots; ++d)
    bstart[d] = (uint64_t)(std::lower_bound(c.delay.begin(), c.delay.end(), d,
                                            [](uint8_t a, uint32_t b) { return a < b; }) - c.delay.begin());
  // The reference's delivery order: buckets by descending delay, in each the
  // runs of one source, a run possibly spilling over into the next bucket.
  // It is cut into `parts` contiguous pieces at run starts; piece k's items
  // for a target then follow those of the pieces before it.  A cut is
  // (di, s): the run starting at s in bucket D - 1 - di; (D, 0) is the end.
  struct Cut {
    uint32_t di;
    uint64_t s;
  };
  // one piece per thread, as long as the pieces' per-target positions
  // (8 bytes per piece and target) stay within 1 GiB
  const unsigned parts = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>(threads, (uint64_t(1) << 30) / (8 * uint64_t(std::max(N, 1u)))));
  std::vector<Cut> cut(parts + 1);
  auto bucket_start = [&](uint32_t di) { return di < D ? Cut{di, bstart[D - 1 - di]} : Cut{D, 0}; };
  cut[0] = bucket_start(0);
  cut[parts] = Cut{D, 0};
  const uint64_t total = bstart[D];
  for (unsigned k = 1; k < parts; ++k) {
    uint64_t v = total * k / parts;
    uint32_t di = 0;
    for (; di < D; ++di) {
      const uint32_t d = D - 1 - di;
      if (v < bstart[d + 1] - bstart[d]) break;
      v -= bstart[d + 1] - bstart[d];
    }
    Cut at = bucket_start(di);
    if (di < D) {
      const uint32_t d = D - 1 - di;
      uint64_t s = bstart[d] + v;
      while (s > bstart[d] && s < bstart[d + 1] && c.pre[s] == c.pre[s - 1]) ++s;  // to a run start
      at = s < bstart[d + 1] ? Cut{di, s} : bucket_start(di + 1);
    }
    const Cut &prev = cut[k - 1];
    if (at.di < prev.di || (at.di == prev.di && at.s < prev.s)) at = prev;
    cut[k] = at;
  }
  // Visits piece k's delivered (run delay, source, synapse, target) in order.
  auto visit = [&](unsigned k, auto &&f) {
    const Cut a = cut[k], b = cut[k + 1];
    for (uint32_t di = a.di; di < D && di <= b.di; ++di) {
      const uint32_t d = D - 1 - di;
      uint64_t s = di == a.di ? a.s : bstart[d];
      const uint64_t se = di == b.di ? b.s : bstart[d + 1];
      while (s < se) {
        const uint32_t src = c.pre[s];
        uint64_t e = s;
        while (e < S && c.pre[e] == src) {  // the run, possibly spilling over
          f(d, src, e, c.post[e]);
          ++e;
        }
        s = e < bstart[d + 1] ? e : bstart[d + 1];
      }
    }
  };
  // per piece and target: the count, then where the piece's next item goes
  std::vector<uint64_t> pos(size_t(parts) * N, 0);
  // turns the counts into positions in a list laid out by offsets `off`
  auto place = [&](std::vector<uint64_t> &off) {
    off.assign(N + 1, 0);
    for (uint32_t j = 0; j < N; ++j) {
      uint64_t at = off[j];
      for (unsigned k = 0; k < parts; ++k) {
        const uint64_t n = pos[size_t(k) * N + j];
        pos[size_t(k) * N + j] = at;
        at += n;
      }
      off[j + 1] = at;
    }
  };
  for_each_index(parts, threads, [&](uint32_t k) {
    uint64_t *cnt = &pos[size_t(k) * N];
    visit(k, [&](uint32_t, uint32_t, uint64_t, uint32_t tgt) { ++cnt[tgt]; });
  });
  place(c.in_off);
  c.max_in = 0;
  for (uint32_t j = 0; j < N; ++j) c.max_in = std::max<uint32_t>(c.max_in, (uint32_t)(c.in_off[j + 1] - c.in_off[j]));
  const uint64_t items = c.in_off[N];
  {
    std::thread a([&] { c.in_pre_d.resize(items); }), b([&] { c.in_w.resize(items); });
    c.in_syn.resize(items);
    a.join();
    b.join();
  }
  for_each_index(parts, threads, [&](uint32_t k) {
    uint64_t *at = &pos[size_t(k) * N];
    visit(k, [&](uint32_t d, uint32_t src, uint64_t e, uint32_t tgt) {
      const uint64_t i = at[tgt]++;
      c.in_pre_d[i] = src | (d << 24);
      c.in_w[i] = c.weight[e];
      c.in_syn[i] = (uint32_t)e;
    });
  });
  if (!with_incoming) return;
  // incoming synapses per target in backward-list order (LTP list,
  // DPSNN_localNet_init.c:225-237): the list cut into `parts` equal pieces
  std::fill(pos.begin(), pos.end(), 0);
  auto piece = [&](uint32_t k) { return S * k / parts; };
  for_each_index(parts, threads, [&](uint32_t k) {
    uint64_t *cnt = &pos[size_t(k) * N];
    for (uint64_t s = piece(k); s < piece(k + 1); ++s) ++cnt[c.post[s]];
  });
  place(c.inc_off);
  c.inc_syn.resize(S);
  for_each_index(parts, threads, [&](uint32_t k) {
    uint64_t *at = &pos[size_t(k) * N];
    for (uint64_t s = piece(k); s < piece(k + 1); ++s) c.inc_syn[at[c.post[s]]++] = (uint32_t)s;
  });
}

// --------------------------------------------------------------------------
// Stencil-window layout of the in-lists (see dpsnn_host.h).  The in-lists
// are those of build_delivery: per target, items in delivery insertion
// order, i.e. run delay descending, then source ascending.
// --------------------------------------------------------------------------
void build_window(const Params &p, const Connectome &c, uint32_t unitsPerCol, uint32_t maxEntries,
                  WindowLayout &L, unsigned threads) {
  L = WindowLayout{};
  const uint32_t N = c.N, npc = p.pub.neuronsPerCM, ncols = p.pub.globCFT;
  if (N == 0 || c.in_off.size() != size_t(N) + 1 || c.in_pre_d.size() != c.in_off[N]) {
    L.why = "the in-lists are not on the host";
    return;
  }
  if (npc == 0 || uint64_t(npc) * ncols != N) {
    L.why = "neurons are not whole columns";
    return;
  }
  if (threads == 0) threads = default_threads();
  threads = std::max(1u, std::min<unsigned>(threads, ncols));
  auto parallel = [&](uint32_t count, auto &&body) {
    std::vector<std::thread> pool;
    std::atomic<uint32_t> next{0};
    for (unsigned th = 0; th < threads; ++th)
      pool.emplace_back([&]() {
        for (uint32_t i; (i = next.fetch_add(1)) < count;) body(i);
      });
    for (auto &t : pool) t.join();
  };
  // stencil of each column: the distinct columns its targets' sources lie in
  std::vector<std::vector<uint32_t>> stencil(ncols);
  std::vector<uint32_t> maxD(ncols, 0);
  parallel(ncols, [&](uint32_t col) {
    std::vector<uint8_t> seen(ncols, 0);
    uint32_t md = 0;
    for (uint64_t k = c.in_off[uint64_t(col) * npc]; k < c.in_off[uint64_t(col + 1) * npc]; ++k) {
      const uint32_t pd = c.in_pre_d[k];
      seen[(pd & 0xFFFFFFu) / npc] = 1;
      md = std::max(md, pd >> 24);
    }
    for (uint32_t s = 0; s < ncols; ++s)
      if (seen[s]) stencil[col].push_back(s);
    maxD[col] = md;
  });
  uint32_t slots = 1, passes = 1;
  for (uint32_t col = 0; col < ncols; ++col) {
    slots = std::max<uint32_t>(slots, (uint32_t)stencil[col].size());
    passes = std::max(passes, maxD[col] + 1);
  }
  if (uint64_t(slots) * npc + 1 > maxEntries || uint64_t(slots) * npc >= 0xFFFFu) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "a stencil window of %u columns x %u neurons does not fit", slots, npc);
    L.why = buf;
    return;
  }
  L.slots = slots;
  L.npc = npc;
  L.passes = passes;
  L.unitsPerCol = std::max<uint32_t>(std::max<uint32_t>(unitsPerCol, 1), (npc + kWindowMaxUnit - 1) / kWindowMaxUnit);
  L.unitsPerCol = std::min(L.unitsPerCol, npc);
  L.unit = (npc + L.unitsPerCol - 1) / L.unitsPerCol;
  L.unitsPerCol = (npc + L.unit - 1) / L.unit;
  L.units = ncols * L.unitsPerCol;
  L.cols.assign(size_t(ncols) * slots, -1);
  for (uint32_t col = 0; col < ncols; ++col)
    for (uint32_t k = 0; k < stencil[col].size(); ++k) L.cols[size_t(col) * slots + k] = (int32_t)stencil[col][k];
  const uint32_t W = kWindowWaves, C = kWindowChunk;
  // segment bounds of each target of a unit: cum[q][d] = items of run
  // delay >= d (the in-list is run delay descending)
  auto unit_bounds = [&](uint32_t u, uint32_t &j0, uint32_t &m, std::vector<uint32_t> &cum) {
    const uint32_t col = u / L.unitsPerCol, part = u % L.unitsPerCol;
    j0 = col * npc + part * L.unit;
    m = std::min(L.unit, (col + 1) * npc - j0);
    cum.assign(size_t(m) * (passes + 1), 0);
    for (uint32_t q = 0; q < m; ++q) {
      const uint64_t b = c.in_off[j0 + q], e = c.in_off[j0 + q + 1];
      uint32_t *cq = &cum[size_t(q) * (passes + 1)];
      for (uint64_t k = b; k < e; ++k) ++cq[c.in_pre_d[k] >> 24];
      for (uint32_t d = passes; d-- > 0;) cq[d] += cq[d + 1];
    }
  };
  const size_t streams = size_t(L.units) * passes * W;
  L.chunk.assign(streams + 1, 0);
  parallel(L.units, [&](uint32_t u) {
    uint32_t j0, m;
    std::vector<uint32_t> cum;
    unit_bounds(u, j0, m, cum);
    for (uint32_t pass = 0; pass < passes; ++pass) {
      const uint32_t d = passes - 1 - pass;
      for (uint32_t q = 0; q < m; ++q) {
        const uint32_t len = cum[size_t(q) * (passes + 1) + d] - cum[size_t(q) * (passes + 1) + d + 1];
        L.chunk[(size_t(u) * passes + pass) * W + q % W + 1] += (len + C - 1) / C;
      }
    }
  });
  for (size_t s = 0; s < streams; ++s) L.chunk[s + 1] += L.chunk[s];
  if (L.chunk[streams] * C >= (uint64_t(1) << 32)) {  // the device indexes items with 32 bits
    L.why = "more than 2^32 window items";
    L.chunk.clear();
    return;
  }
  L.item.assign(L.chunk[streams] * C, (uint32_t)(slots * npc) << 16);
  L.desc.assign(L.chunk[streams], 0);
  parallel(L.units, [&](uint32_t u) {
    uint32_t j0, m;
    std::vector<uint32_t> cum;
    unit_bounds(u, j0, m, cum);
    const uint32_t col = j0 / npc;
    std::vector<int32_t> slotOf(ncols, -1);
    for (uint32_t k = 0; k < stencil[col].size(); ++k) slotOf[stencil[col][k]] = (int32_t)k;
    for (uint32_t pass = 0; pass < passes; ++pass) {
      const uint32_t d = passes - 1 - pass;
      for (uint32_t w = 0; w < W; ++w) {
        uint64_t ch = L.chunk[(size_t(u) * passes + pass) * W + w];
        for (uint32_t q = w; q < m; q += W) {
          const uint32_t *cq = &cum[size_t(q) * (passes + 1)];
          // run delays above d come first in the in-list
          const uint64_t b = c.in_off[j0 + q] + cq[d + 1], e = c.in_off[j0 + q] + cq[d];
          uint64_t at = ch * C;
          for (uint64_t k = b; k < e; ++k, ++at) {
            const uint32_t src = c.in_pre_d[k] & 0xFFFFFFu, scol = src / npc;
            const uint32_t widx = (uint32_t)slotOf[scol] * npc + (src - scol * npc);
            L.item[at] = widx << 16 | (uint16_t)c.in_w[k];
          }
          const uint64_t nch = (e - b + C - 1) / C;
          for (uint64_t x = 0; x < nch; ++x) L.desc[ch + x] = (uint16_t)(q | (x == 0 ? kWindowFirst : 0u));
          ch += nch;
        }
      }
    }
  });
  L.ok = true;
}

// --------------------------------------------------------------------------
// External drive replay: regularDynamicTimeStep's bath loop
// (DPSNN_neuron_sim.c:604-640) with getPoissonLapse (:844-851) -- a LUT
// index draw then a lapse draw per event, lapses-only past the per-ms cap --
// in neuron order, every step, from the network's single ran3 stream.
// --------------------------------------------------------------------------
void ExternalReplay::init(const Connectome &c, Ran3 *rng) {
  c_ = &c;
  rng_ = rng;
  assigned_.assign(c.N, 0.0);
}

void ExternalReplay::generate(uint32_t t, ExternalStep &out) {
  const Connectome &c = *c_;
  const double next = (double)((double)t + 1.0);
  out.off.resize(c.N + 1);
  out.time.clear();
  out.value.clear();
  out.off[0] = 0;
  for (uint32_t i = 0; i < c.N; ++i) {
    const double inv = c.invNuExt[i];
    if (inv > 0 && assigned_[i] < next) {
      const double *lut = &c.bathLut[size_t(c.subPop[i]) * kLutDepth];
      const uint32_t cap = c.maxExt[i];
      uint32_t n = 0;
      double a = assigned_[i];
      auto lapse = [&]() {
        double u = (double)1.0 - rng_->next();
        if (u < 0) u = 0;
        return (-logf((float)u) * inv);
      };
      do {
        out.time.push_back(a);
        out.value.push_back(lut[(uint32_t)(rng_->next() * kLutDepth)]);
        ++n;
        a += lapse();
      } while (a < next && n < cap);
      if (n >= cap && a < next) {
        do { a += lapse(); } while (a < next);
      }
      assigned_[i] = a;
    }
    out.off[i + 1] = (uint32_t)out.time.size();
  }
}

// --------------------------------------------------------------------------
// Output files.  Names as statBasicClass::openFile (DPSNN_stat.c:55-107);
// raster lines "%d %.7g" (DPSNN_stat.c:275-296); ratesPerPop lines as
// statSpikingRatesPerPopClass::write (DPSNN_stat.c:225-273), with the
// per-population counting of DPSNN_localNet_sim.c:227-253.
// --------------------------------------------------------------------------
bool StatWriter::init(const Params &p, const std::string &dir) {
  p_ = &p;
  dir_ = dir;
  perPop_.assign(size_t(p.pub.globCFT) * p.pub.subPopNumber, 0);
  buf_.resize(1 << 16);
  return true;
}

bool StatWriter::window(uint32_t sec) const {
  const dpsnn_parameters &q = p_->pub;
  return sec >= q.startStatFiles_s && sec <= q.stopStatFiles_s && sec % q.moduloSec == 0;
}

FILE *StatWriter::open_file(const char *prefix, uint32_t sec, const char *mode) {
  const dpsnn_parameters &q = p_->pub;
  char name[512];
  if (q.howManyOutputFiles == 1)
    std::snprintf(name, sizeof name, "%s-sec%u-C%u-H%u-h%03u.dat", prefix, sec, q.globCFT, q.globH, 0u);
  else if (sec == 0)
    std::snprintf(name, sizeof name, "%s-start-C%u-H%u-h%u.dat", prefix, q.globCFT, q.globH, 0u);
  else if (sec % 2 == 0)
    std::snprintf(name, sizeof name, "%s-lasteven-C%u-H%u-h%u.dat", prefix, q.globCFT, q.globH, 0u);
  else
    std::snprintf(name, sizeof name, "%s-lastodd-C%u-H%u-h%u.dat", prefix, q.globCFT, q.globH, 0u);
  return std::fopen((dir_ + "/" + name).c_str(), mode);
}

bool StatWriter::resume(const State &s, uint32_t sec) {
  close();
  if (s.perPop.size() != perPop_.size()) return false;
  perPop_ = s.perPop;
  if (s.spikesOpen && !(spikes_ = open_file("spikes", sec, "a"))) return false;
  if (s.ratesOpen && !(rates_ = open_file("ratesPerPop", sec, "a"))) return false;
  return true;
}

void StatWriter::flush() {
  if (spikes_) std::fflush(spikes_);
  if (rates_) std::fflush(rates_);
}

bool StatWriter::step(uint32_t t, const uint32_t *ids, const double *te, uint32_t n,
                      const uint32_t *popCounts) {
  const dpsnn_parameters &q = p_->pub;
  const uint32_t sec = t / 1000;
  const bool inWindow = window(sec);
  const bool ratesSec = sec >= q.startRatesPerPop_s && sec <= q.stopRatesPerPop_s;
  if (t % 1000 == 0 && inWindow && !dir_.empty()) {
    if (!q.spikes_writeDisable) {
      if (spikes_) std::fclose(spikes_);
      if (!(spikes_ = open_file("spikes", sec))) return false;
    }
    if (!q.statRatesPerPop_writeDisable && ratesSec) {
      if (rates_) std::fclose(rates_);
      if (!(rates_ = open_file("ratesPerPop", sec))) return false;
    }
  }
  if (spikes_ && inWindow) {
    for (uint32_t k = 0; k < n; ++k) {
      const int len = std::snprintf(buf_.data(), buf_.size(), "%d %.7g\n", (int)ids[k], te[k]);
      std::fwrite(buf_.data(), 1, (size_t)len, spikes_);
    }
  }
  for (size_t k = 0; k < perPop_.size(); ++k) perPop_[k] += popCounts[k];
  if (rates_ && inWindow && ratesSec && t % q.ratesSampling == 0) {
    std::fprintf(rates_, "%d ", (int)(t - q.ratesSampling));
    for (uint32_t col = 0; col < q.globCFT; ++col)
      for (uint32_t s = 0; s < q.subPopNumber; ++s)
        std::fprintf(rates_, "%.7g ",
                     ((float)(perPop_[size_t(col) * q.subPopNumber + s]) * 1000.0) /
                         ((float)(p_->sp[s].count) * (float)(q.ratesSampling)));
    std::fprintf(rates_, " \n");
    std::fill(perPop_.begin(), perPop_.end(), 0);
  }
  if (t % 1000 == 999 && inWindow) {
    if (spikes_) { std::fclose(spikes_); spikes_ = nullptr; }
    if (rates_ && ratesSec) { std::fclose(rates_); rates_ = nullptr; }
  }
  return true;
}

// --------------------------------------------------------------------------
// STDP tables, initLongTermPlasticity DPSNN_neuron_sim.c:45-92.
// --------------------------------------------------------------------------
void stdp_tables(const dpsnn_parameters &q, std::vector<float> &causal, std::vector<float> &anti) {
  causal.assign(kStdpSpan, 0.0f);
  anti.assign(kStdpSpan, 0.0f);
  causal[0] = q.causalSTDPmaxAbsValue * q.STDPmultiplier;
  causal[1] = causal[0];
  anti[0] = q.antiCausalCoeff * q.antiCausalSTDPmaxAbsValue * q.STDPmultiplier;
  for (uint32_t i = 1; i < kStdpSpan; ++i) causal[i] = (1 - 1 / q.tauSTDP_ms) * causal[i - 1];
  for (uint32_t i = 1; i < kStdpSpan; ++i) anti[i] = (1 - 1 / q.tauSTDP_ms) * anti[i - 1];
}

// --------------------------------------------------------------------------
// STDPevent file (statSTDPeventClass::write, DPSNN_stat.c:451-470): one line
// "%8u %7d %7d %2.7f" (ms, pre, post, contribution) per non-zero LTP
// contribution at the fired scan (causalSTDP_ms_ofBackwardSynapses,
// DPSNN_neuron_sim.c:139-177: neurons ascending, each one's incoming
// synapses in backward-list order) and per non-zero LTD contribution at
// delivery (synapseSetLastActiveTime_antiCausalSTDP, :207-277, in the
// delivery loop's order DPSNN_localNet_sim.c:837-999: run delays
// descending, sources ascending, each source's run continued while the
// source matches).  The contributions depend on spike times only, so they
// are restated here from the spikes of each step.
// --------------------------------------------------------------------------
bool StdpEventWriter::init(const Params &p, const Connectome &c, const std::string &dir) {
  p_ = &p;
  c_ = &c;
  dir_ = dir;
  stdp_tables(p.pub, causal_, anti_);
  const uint32_t N = c.N, D = std::min<uint32_t>(p.pub.D, kDelaySlots);
  D_ = D;
  if (c.pre.size() != c.S || c.inc_off.size() != size_t(N) + 1) return false;
  // first synapse of each (source, delay) run: the backward list is sorted
  // by (delay, source)
  first_.assign(size_t(N) * D, UINT32_MAX);
  for (uint64_t s = c.S; s-- > 0;)
    if (c.delay[s] < D) first_[size_t(c.pre[s]) * D + c.delay[s]] = (uint32_t)s;
  lastSpike_.assign(N, -100000);  // DPSNN_localNet_init.c:217
  lastAct_.assign(c.S, 0);
  ring_.assign(D, {});
  buf_.resize(256);
  return true;
}

bool StdpEventWriter::writes(uint32_t t) const {
  const dpsnn_parameters &q = p_->pub;
  const uint32_t sec = t / 1000, dbg = p_->debugPrintEnable_ms;
  return fp_ && t >= dbg && (dbg == 0 || t <= dbg + 20) && sec >= q.startStatFiles_s && sec <= q.stopStatFiles_s &&
         sec % q.moduloSec == 0 && sec >= p_->startSTDPevent_s && sec <= p_->stopSTDPevent_s;
}

void StdpEventWriter::line(uint32_t t, uint32_t pre, uint32_t post, float v) {
  const int len = std::snprintf(buf_.data(), buf_.size(), "%8u %7d %7d %2.7f \n", t, (int)pre, (int)post, v);
  std::fwrite(buf_.data(), 1, (size_t)len, fp_);
}

bool StdpEventWriter::step(uint32_t t, const uint32_t *ids, uint32_t n) {
  const dpsnn_parameters &q = p_->pub;
  const Connectome &c = *c_;
  const uint32_t sec = t / 1000;
  if (t % 1000 == 0 && sec >= q.startStatFiles_s && sec <= q.stopStatFiles_s && sec % q.moduloSec == 0 &&
      sec >= p_->startSTDPevent_s && sec <= p_->stopSTDPevent_s && !dir_.empty()) {
    if (fp_) std::fclose(fp_);
    char name[512];
    if (q.howManyOutputFiles == 1)
      std::snprintf(name, sizeof name, "STDPevent-sec%u-C%u-H%u-h%03u.dat", sec, q.globCFT, q.globH, 0u);
    else if (sec == 0)
      std::snprintf(name, sizeof name, "STDPevent-start-C%u-H%u-h%u.dat", q.globCFT, q.globH, 0u);
    else
      std::snprintf(name, sizeof name, "STDPevent-%s-C%u-H%u-h%u.dat", sec % 2 == 0 ? "lasteven" : "lastodd",
                    q.globCFT, q.globH, 0u);
    if (!(fp_ = std::fopen((dir_ + "/" + name).c_str(), "w"))) return false;
  }
  const bool plastic = t >= q.startPlasticity_ms && t <= q.stopPlasticity_ms;
  const bool out = writes(t);
  std::vector<uint32_t> &cur = ring_[t % D_];
  cur.assign(ids, ids + n);
  // the fired scan: last emitted spike, LTP
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t i = ids[k];
    lastSpike_[i] = (int32_t)t;
    if (!plastic) continue;
    for (uint64_t x = c.inc_off[i]; x < c.inc_off[i + 1]; ++x) {
      const uint32_t s = c.inc_syn[x];
      const int32_t last = lastAct_[s];
      if (last == 0) continue;
      const int dt = (int)t - 1 - last;
      const float f = dt < (int)kStdpSpan ? causal_[dt] : 0.0f;
      if (out && f != 0.0f && t != 0) line(t, c.pre[s], c.post[s], f);
    }
  }
  // delivery: LTD
  if (plastic) {
    for (uint32_t d = D_; d-- > 0;) {
      if (d > t) continue;
      const std::vector<uint32_t> &src = ring_[(t - d) % D_];
      for (uint32_t pre : src) {
        uint32_t s = first_[size_t(pre) * D_ + d];
        if (s == UINT32_MAX) continue;
        for (; s < c.S && c.pre[s] == pre; ++s) {
          const uint32_t post = c.post[s];
          lastAct_[s] = (int32_t)t;
          const int dt = (int)t - lastSpike_[post];
          const float f = dt < (int)kStdpSpan ? anti_[dt] : 0.0f;
          if (out && f != 0.0f) line(t, pre, post, f);
        }
      }
    }
  }
  if (fp_) std::fflush(fp_);
  if (t % 1000 == 999 && fp_ && sec >= q.startStatFiles_s && sec <= q.stopStatFiles_s && sec % q.moduloSec == 0) {
    std::fclose(fp_);
    fp_ = nullptr;
  }
  return true;
}

void StdpEventWriter::close() {
  if (fp_) std::fclose(fp_);
  fp_ = nullptr;
}

void StatWriter::close() {
  if (spikes_) std::fclose(spikes_);
  if (rates_) std::fclose(rates_);
  spikes_ = rates_ = nullptr;
}

}  // namespace dpsnn

namespace dpsnn {

// --------------------------------------------------------------------------
// Perseo .ini files (localNetClass::writeIniFiles, DPSNN_writeIniFiles.c:
// 31-158) at world size 1 (loc_h 0, every column local).  Per column (x
// outer, y inner): PerseoConFile gets one line per (stencil cell, target
// sub-population, source sub-population) whose connection probability is
// positive -- over the reference's whole compile-time stencil range, the
// target column being the cell's offset from the source (explicit stencil)
// or the cell itself (homogeneous) when it lies on the grid; PerseoModFile
// gets one line per sub-population with the column's external bath
// (generateSourceNeuCExt, DPSNN_LIFCAconnectome.cpp:871-887).
// --------------------------------------------------------------------------
bool write_ini_files(const Params &p, const std::string &outdir, std::string &err) {
  const dpsnn_parameters &q = p.pub;
  const std::string con = outdir + "/PerseoConFile_h0.txt", mod = outdir + "/PerseoModFile_h0.txt";
  FILE *fc = std::fopen(con.c_str(), "w");
  FILE *fm = fc ? std::fopen(mod.c_str(), "w") : nullptr;
  if (!fc || !fm) {
    if (fc) std::fclose(fc);
    err = "cannot create the Perseo files in " + outdir;
    return false;
  }
  const uint32_t nsub = q.subPopNumber;
  for (uint32_t cx = 0; cx < q.globCFX; ++cx)
    for (uint32_t cy = 0; cy < q.globCFY; ++cy) {
      const uint32_t src = cx + cy * q.globCFX;
      for (uint32_t sx = 0; sx < kMaxStencil; ++sx)
        for (uint32_t sy = 0; sy < kMaxStencil; ++sy) {
          const int32_t tx = q.overallConnectivity == 0 ? (int32_t)(cx + sx) - (int32_t)p.maxModDeltaX : (int32_t)sx;
          const int32_t ty = q.overallConnectivity == 0 ? (int32_t)(cy + sy) - (int32_t)p.maxModDeltaY : (int32_t)sy;
          if (tx < 0 || tx >= (int32_t)q.globCFX || ty < 0 || ty >= (int32_t)q.globCFY) continue;
          const uint32_t tgt = (uint32_t)tx + q.globCFX * (uint32_t)ty;
          for (uint32_t ts = 0; ts < nsub; ++ts)
            for (uint32_t ss = 0; ss < nsub; ++ss) {
              const float prob = p.prob_at(ss, ts, sx, sy);
              if (!(prob > 0.0f)) continue;
              const SubPop &s = p.sp[ss];
              std::fprintf(fc, " %5d %5d     %g   %g   %g   '%s'   %8g   %g\n", (int)(nsub * tgt + ts),
                           (int)(nsub * src + ss), (double)prob, s.DMin[ts] + 0.5, s.DMax[ts] + 0.5, "Fixed",
                           s.J[ts], s.DJ[ts]);
            }
        }
      uint32_t bx = cx, by = cy;
      if (bx >= q.globCFX / 2) bx = -1 * (bx - q.globCFX + 1);
      if (by >= q.globCFY / 2) by = -1 * (by - q.globCFY + 1);
      if (bx >= q.stencilDim) bx = q.stencilDim - 1;
      if (by >= q.stencilDim) by = q.stencilDim - 1;
      const float cext = (float)p.bath[size_t(bx) * kMaxStencil + by];
      for (uint32_t ss = 0; ss < nsub; ++ss) {
        const SubPop &s = p.sp[ss];
        std::fprintf(fm, " %6d   %g   %g   %g   %7g   %g   %g   %g   %g   %g   %4g   %4g   %d \n", (int)s.count, s.JExt,
                     s.DJExt, (double)cext, s.NuExt, s.Tau, s.Theta, s.H, s.Tarp, s.AlphaC, s.TauC, s.gC, 0);
      }
    }
  const bool ok = std::fclose(fc) == 0 && std::fclose(fm) == 0;
  if (!ok) err = "writing the Perseo files in " + outdir + " failed";
  return ok;
}

}  // namespace dpsnn

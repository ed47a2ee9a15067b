// dpsnn_engine.hip -- MI355X engine for DPSNN's per-millisecond LIFCA step.
//
// One network = one process's whole local network on one GPU.  Step t (one
// reference completeTimeStep_ms, DPSNN_localNet_sim.c:22-307) is ONE kernel:
//
//   k_step         per target neuron: streams the neuron's in-list (delivery
//                  insertion order), keeps the items whose source fired in
//                  step t-delay, orders them by arrival time (wave-wide
//                  bucket sort in LDS on exact 64-bit keys), merges the
//                  external bath (internal first on ties) and runs
//                  LIFCADynamic per event in FP64, one lane per neuron.
//                  Replaces delivery + addInputSpike +
//                  regularDynamicTimeStep + LIFCADynamic (:837-999,
//                  DPSNN_neuron_sim.c:289-300, :522-824).
//                  Its tail is the fired scan of step t+1: each wave ballots
//                  its 64 neurons' spikes into one word of the spike bitmap
//                  of slot t+1, stores their emission times in the slot's Te
//                  history, the last-spike ms and the per-(column, subPop)
//                  counts, and applies STDP's LTP to the incoming synapses of
//                  the neurons that fired.  That replaces the fired scan +
//                  compaction + world-size-1 axonal exchange + ring insert
//                  (:103-148, :620-806, DPSNN_spike.c:209-239) and
//                  causalSTDP_ms (DPSNN_neuron_sim.c:94-177).
//   k_compact      (only when a raster is wanted) ascending spike list of
//                  slot t+1 appended to the device raster buffer.
//   k_commit       (STDP, ms % 1000 == 999) the per-second weight commit.
//
// The device holds no per-step event queue: a target pulls its live events
// from its static in-list and the spike history (bitmap + Te per slot), so
// the only per-step traffic is the in-list stream, the history lookups and
// the neuron state.  Built with -ffp-contract=off: the reference's FP64
// update is second order and unfused, and must round where it rounds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dpsnn_exp.h"
#include "dpsnn_hip.h"
#include "dpsnn_host.h"

namespace dpsnn {

constexpr int kBlock = 256;      // integrate / scan block (4 waves)
constexpr uint32_t kPad = 0xFFFFFFFFu;
constexpr int kExtRing = 8;      // replayed external steps in flight
constexpr uint32_t kMaxInList = 3900;  // per-neuron in-list (12-bit key field, LDS)

static thread_local std::string g_err;
static int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(DPSNN_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kWave = 64;
constexpr double kNotFired = -1.0;  // Te history entry of a neuron that did not fire (spike times are >= 0)

// XCD-aware persistent walk over work items [0, total).  The dispatcher deals
// workgroups to the 8 XCDs round-robin (workgroup b on XCD b % 8), and each
// XCD has its own L2: the blocks of one XCD walk one contiguous eighth of
// the items, so neighbouring neurons, which share most of their sources,
// read the spike history through the same L2.  gridDim.x is a multiple of 8.
constexpr uint32_t kXcds = 8;
struct XcdWalk {
  uint32_t first, end, stride;
};
__device__ inline XcdWalk xcd_walk(uint32_t total) {
  const uint32_t x = blockIdx.x % kXcds, lb = blockIdx.x / kXcds, per = gridDim.x / kXcds;
  const uint32_t begin = (uint32_t)((uint64_t)total * x / kXcds);
  const uint32_t end = (uint32_t)((uint64_t)total * (x + 1) / kXcds);
  return XcdWalk{begin + lb, end, per};
}

// XCD-ordered work queue over items [0, total): XCD x (= blockIdx.x % 8, as
// the dispatcher deals workgroups) owns the x-th contiguous eighth, and its
// waves take the items in order from the XCD's counter, so the items in
// flight on one XCD stay a narrow window (neighbouring neurons, shared
// sources, one L2).  An XCD whose eighth is exhausted helps the next ones.
// A wave takes a run of consecutive items per counter update (guided: a
// share of what is left of the range for each of the XCD's waves, between
// kGrabMin and kGrabMax, so that counter updates are few while much is left
// and the tail stays balanced); next_item returns the first item and the
// run's end (*last), or `total` when none is left.  Whole wave; ctr holds
// kXcds counters that are zero at the launch.  Each counter sits on its own
// 128-byte line (ctr[x * kCtrStride]); an exhausted counter is recognised by
// a plain read before any atomic.  The counters live in device-coherent
// memory, where each update costs a round trip: with runs of 4, the 24x24-column network's
// phase 1 was 9% slower than with the guided runs, with runs of 1 twice as
// slow.
constexpr uint32_t kGrabMin = 4, kGrabMax = 16;
constexpr uint32_t kCtrStride = 32;
__device__ inline uint32_t next_item(uint32_t *ctr, uint32_t total, uint32_t &x, uint32_t home, uint32_t *last) {
  const uint32_t wavesPerXcd = max(1u, gridDim.x * (blockDim.x / 64u) / kXcds);
  for (;;) {
    const uint32_t begin = (uint32_t)((uint64_t)total * x / kXcds);
    const uint32_t end = (uint32_t)((uint64_t)total * (x + 1) / kXcds);
    uint32_t *c = ctr + x * kCtrStride;
    uint32_t k = end - begin, g = kGrabMin;
    if ((threadIdx.x & 63) == 0) {
      k = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (begin + k < end) {
        g = min(kGrabMax, max(kGrabMin, (end - begin - k) / (2u * wavesPerXcd)));
        k = atomicAdd(c, g);
      }
    }
    k = __builtin_amdgcn_readfirstlane(k);  // lane 0's values, known uniform (scalar registers)
    g = __builtin_amdgcn_readfirstlane(g);
    if (begin + k < end) {
      *last = min(end, begin + k + g);
      return begin + k;
    }
    x = (x + 1) % kXcds;
    if (x == home) return total;
  }
}

// LDS written by some lanes of a wave and read by others: order the wave's
// accesses around this point.  The one-wave workgroups of phase 1 use it in
// place of __syncthreads(), whose workgroup-scope fence would also wait for
// every global load in flight (the records loaded ahead for the next target).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int kWaveLTP = 256;    // STDP kernels: 4 waves per block

// Per-sub-population LIFCA parameters, in this order.
enum { kTau, kTheta, kH, kTarp, kAlphaC, kTauC, kGC, kInvTau, kInvTauC, kParams };

struct DevNet {
  uint32_t N, NW, RS, nsub, npc, ncols, cap, drive;
  uint32_t denseMax;       // phase 1 takes its fast path for up to this many live items (<= kDense * 64)
  float w2f;
  uint64_t philoxSeed;
  const uint8_t *subPop;
  const double *sp;        // [nsub][kParams]
  const double *lut;       // [nsub][256]
  const double *invNuExt;
  const uint32_t *maxExt;
  double *v, *c, *Tr, *Te, *assigned;
  uint64_t *philoxCtr;
  uint64_t *bits;          // [RS][NW], RS a power of two
  double *teHist;          // [RS][N] emission time, or kNotFired
  const uint64_t *inOff;   // [N+1] start of each neuron's in-list
  const uint32_t *inPD;    // pre | delay << 24, in delivery insertion order (padded by one wave)
  const int16_t *inW;      // (padded by one wave)
  uint64_t *keys;          // sorted event ticks, per tile of 64 neurons [cap/kRun][64][kRun];
                           // from step kPackT on tick << 16 | weight
  int16_t *keyW;           // their weights before step kPackT, same layout
  uint16_t *keyPos;        // STDP, general path: their in-list positions, same layout
  uint32_t *keyCount;      // [grid][64] or [N/64][64]
  uint32_t *workCtr;       // [kXcds][kCtrStride] phase-1 work counters (zero between steps)
  unsigned long long *counters;  // [kCounterSlots][kCounterStride]: [0] events, [1] external, [2] spikes
  // STDP (null when plasticity never runs).  Synapse state is kept in
  // target-major order: synapse k of neuron j's incoming list lives at
  // incOff[j] + k; in-list items name it by k (inSyn).
  int32_t *lastSpike;      // [N] last emission step (ms) of each neuron
  const uint64_t *incOff;  // [N+1]
  int16_t *wT;
  float *tdT, *tdPre;      // timeDerivative; tdPre: before the decay (probes)
  int32_t *lastT;          // lastActivationTime
  const uint8_t *kindT;    // presynaptic sub-population
  const uint16_t *inSyn;   // per in-list item
  const float *causal, *anti;        // [128] each
  const float *maxSyn, *revMax;      // [nsub][nsub]
  float f2w, drift, decayMul;
  int decayReset;
  // Stencil-window form of phase 1 (k_collect + k_sort; null when the
  // network has no window layout): the in-lists regrouped per unit of
  // targets of one column, per pass (run delay, descending) and per wave,
  // in padded chunks of kWindowChunk items (window index << 16 | weight).
  const uint32_t *winItem;
  const uint16_t *winDesc;   // per chunk: its target's index in the unit
  const uint64_t *winChunk;  // [units][passes][kWindowWaves] + 1: first chunk
  const int32_t *winCols;    // [columns][slots]: source column of each stencil slot, or -1
  uint32_t winSlots, winPasses, winUnit, winUnitsPerCol, winUnits;
  uint32_t *winCtr;          // [kXcds][kCtrStride] unit counters (zero between steps)
  uint32_t *bigList;         // [N] targets k_sort leaves to k_sort_big
  uint32_t *bigCount;        // their count (zero between steps)
  uint32_t sortCap;          // k_sort keeps targets of up to this many records (<= kSortCap)
  uint32_t fuseSort;         // 1: k_collect sorts each unit's targets itself (no k_sort)
  // Steps replayed from a captured graph read their t here (the kernel
  // argument is then ignored); null: t is the argument.
  const uint32_t *tDev;
};

// The step's counters are spread over kCounterSlots cache lines (a block
// adds to the line of blockIdx % kCounterSlots; the host sums them), so
// that the per-wave updates of one step do not all queue on one line.
constexpr uint32_t kCounterSlots = 16, kCounterStride = 16;
__device__ inline unsigned long long *counter_slot(unsigned long long *counters) {
  return counters + (blockIdx.x % kCounterSlots) * kCounterStride;
}

// --------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter-based stream for the
// on-device external drive.  Key = (seed, neuron), counter = call index.
// --------------------------------------------------------------------------
__device__ inline void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
    const uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    c[0] = h1 ^ c[1] ^ k0;
    c[1] = l1;
    c[2] = h0 ^ c[3] ^ k1;
    c[3] = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct PhiloxStream {
  uint32_t k0, k1;
  uint64_t ctr;
  double spare;  // second uniform of the last block
  int have;
  __device__ double next() {
    if (have) {
      have = 0;
      return spare;
    }
    uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0x44505300u};
    philox4x32(c, k0, k1);
    ++ctr;
    const uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)c[1];
    const uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)c[3];
    spare = (double)(b & ((1ull << 53) - 1)) * 0x1p-53;
    have = 1;
    return (double)(a & ((1ull << 53) - 1)) * 0x1p-53;
  }
};

// a / b, correctly rounded (as IEEE division), for a normal b with
// y = RN(1/b) and a / b within the normal range: q0 = RN(a y) is within two
// ulps of a / b; one correction q1 = RN(q0 + (a - b q0) y) makes it faithful;
// the second, q2 = RN(q1 + r1 y) with the exact residual r1 = a - b q1, is
// RN(a / b) (Markstein's theorem: q faithful, y = RN(1/b)).  Five FMA-unit
// operations instead of the general division's scaled Newton sequence.
__device__ inline double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = fma(fma(-q0, b, a), y, q0);
  return fma(fma(-q1, b, a), y, q1);
}

// --------------------------------------------------------------------------
// LIFCADynamic, DPSNN_neuron_sim.c:664-824 (second-order, non-Perseo path).
// The divisions by Tau and TauC use their reciprocals (div_rn); the results
// are the IEEE quotients.
// --------------------------------------------------------------------------
struct Lifca {
  double v, c, Tr, Te;
  double Tau, Theta, H, Tarp, AlphaC, TauC, gC, invTau, invTauC;
  bool fired;
  // Both of the reference's branches (out of / in the refractory period)
  // decay c by the same expression; the out-of-refractory one also moves v.
  // The lanes of a wave take different branches, so v's update is computed
  // for every lane and kept where the neuron is out of its refractory
  // period: the same operations in the same order, without divergent
  // control flow.  Only the refractory exit (exp) stays a branch; it is rare.
  __device__ inline void event(double t, double I) {
    double deltaT = t - Tr;
    const double TFLES = t - Te;
    const bool active = TFLES > Tarp;
    if (active && Tr - Te < Tarp) {
      c *= dpsnn_exp(div_rn(-(deltaT - TFLES + Tarp), TauC, invTauC));
      deltaT = TFLES - Tarp;
    }
    const double dtHalf = deltaT / 2;
    const double rmHalf = div_rn(dtHalf, Tau, invTau);
    const double rc = div_rn(deltaT, TauC, invTauC);
    double w = v - v * rmHalf - gC * c * dtHalf;
    w = w - w * rmHalf - gC * c * dtHalf;
    w += I;
    c = c * (1 - rc + rc * rc / 2);
    const bool fire = active && w >= Theta;
    v = fire ? H : (active ? w : v);
    c = fire ? c + AlphaC : c;
    Te = fire ? t : Te;
    fired = fired || fire;
    Tr = t;
  }
};

// --------------------------------------------------------------------------
// Fired scan of step t1 = t + 1, run by the wave that integrated 64
// consecutive neurons j0..j0+63 (j0 % 64 == 0) in step t: one bitmap word,
// the emission times of the neurons that fired, their last-spike ms
// (setLastEmittedSpikeTime_ms, DPSNN_neuron_init.c:200-202) and the
// per-(column, subPop) counts.  Every word of the slot and every neuron's
// entry in the slot's Te history (its emission time, or kNotFired) is
// written by exactly one wave, so the slot needs no clearing.  Whole wave,
// uniform.
// --------------------------------------------------------------------------
__device__ __attribute__((always_inline)) inline uint64_t publish_spikes(const DevNet &P, uint32_t t1, uint32_t j, bool f, double te,
                                          uint32_t *popCnt, int count) {
  const uint64_t mask = __ballot(f);
  const uint32_t slot = t1 & (P.RS - 1);
  if ((threadIdx.x & 63) == 0 && (j >> 6) < P.NW) {
    P.bits[(size_t)slot * P.NW + (j >> 6)] = mask;
    if (mask && count) atomicAdd(counter_slot(P.counters) + 2, (unsigned long long)__popcll(mask));
  }
  if (j < P.N) P.teHist[(size_t)slot * P.N + j] = f ? te : kNotFired;
  if (f && P.lastSpike) P.lastSpike[j] = (int32_t)t1;
  if (popCnt && mask) {
    // one counter update per distinct (column, sub-population) among the
    // wave's fired lanes (64 consecutive neurons span at most a few)
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t key = f ? (j / P.npc) * P.nsub + P.subPop[j] : 0u;
    uint64_t pending = mask;
    while (pending) {
      const uint32_t lead = (uint32_t)__ffsll((long long)pending) - 1;
      const uint32_t k = (uint32_t)__shfl((int)key, (int)lead);
      const uint64_t same = __ballot(f && key == k) & pending;
      if (lane == lead) atomicAdd(&popCnt[k], (uint32_t)__popcll(same));
      pending &= ~same;
    }
  }
  return mask;
}

// --------------------------------------------------------------------------
// STDP.  LTP (causalSTDP_ms_ofBackwardSynapses, DPSNN_neuron_sim.c:94-177):
// when the scan of step t finds that neuron j fired (in step t-1's
// dynamics), each incoming synapse with lastActivationTime != 0 gets
// timeDerivative += causal[t - 1 - last] (0 past the 128 ms table).  One
// wave per firing neuron, its incoming synapses contiguous.
// --------------------------------------------------------------------------
__device__ __attribute__((always_inline)) inline void ltp_neuron(const DevNet &P, uint32_t t, uint32_t j) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t end = P.incOff[j + 1];
  constexpr int U = 4;  // independent synapses per lane per round: loads first, then stores
  for (uint64_t s0 = P.incOff[j] + lane; s0 < end; s0 += U * kWave) {
    int32_t last[U];
    float td[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t s = s0 + (uint64_t)u * kWave;
      last[u] = s < end ? P.lastT[s] : 0;
      td[u] = s < end ? P.tdT[s] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t s = s0 + (uint64_t)u * kWave;
      if (s < end) {
        float f = 0.0f;
        if (last[u] != 0) {
          const int dt = (int)t - 1 - last[u];
          if (dt < kStdpSpan) f = P.causal[dt];
        }
        P.tdT[s] = td[u] + f;
      }
    }
  }
}

// LTP of the scan at step t for the neurons of one bitmap word (whole wave).
__device__ inline void ltp_word(const DevNet &P, uint32_t t, uint64_t mask, uint32_t j0) {
  while (mask) {
    const uint32_t b = (uint32_t)__ffsll((long long)mask) - 1;
    mask &= mask - 1;
    ltp_neuron(P, t, j0 + b);
  }
}

// Stand-alone LTP of the scan at step t: one wave per neuron that fired
// (so the sweeps of different neurons overlap), the bitmap of slot t telling
// which.  Runs after the step's phase 2 (and after a commit at ms % 1000 ==
// 999, which must come first).
__global__ __launch_bounds__(kWaveLTP) void k_ltp(DevNet P, uint32_t t) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  const uint64_t *bits = P.bits + (size_t)(t & (P.RS - 1)) * P.NW;
  for (uint32_t j = wave; j < P.N; j += waves)
    if ((bits[j >> 6] >> (j & 63)) & 1ull) ltp_neuron(P, t, j);
}

// longTermPlasticity (DPSNN_localNet_sim.c:312-414), once per simulated
// second at ms % 1000 == 999: exc->exc synapses take w + td (+drift),
// clipped to [0, 2J] in float and re-quantized to int16; td then decays
// (or resets when tauDerivativeDecay_s == 1).
__global__ __launch_bounds__(kWaveLTP) void k_commit(DevNet P) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t j = wave; j < P.N; j += waves) {
    const uint32_t tgt = P.subPop[j];
    for (uint64_t s = P.incOff[j] + lane; s < P.incOff[j + 1]; s += 64) {
      const uint32_t src = P.kindT[s];
      const float maxSyn = P.maxSyn[src * P.nsub + tgt];
      const float rev = P.revMax[src * P.nsub + tgt];
      float clamp = 0.0f;
      if (maxSyn > 0.0 && rev > 0.0) {
        float w = (float)(P.wT[s] * P.w2f);
        float td = P.tdT[s];
        td += P.drift;
        w = w + td;
        if (w > maxSyn) {
          clamp = (w - maxSyn);
          td = td - clamp;
          w = maxSyn;
        }
        if (w < 0.0) {
          clamp = w;
          td = td - clamp;
          w = 0.0;
        }
        P.wT[s] = (int16_t)(w * P.f2w);
        P.tdT[s] = td;
      }
      if (P.tdPre) P.tdPre[s] = P.tdT[s];
      if (P.decayReset)
        P.tdT[s] = 0.0;
      else
        P.tdT[s] = (P.tdT[s] - P.drift + clamp) * P.decayMul;
    }
  }
}

// After a commit the in-list items pick up their synapse's new weight.
__global__ __launch_bounds__(kWaveLTP) void k_refresh_weights(DevNet P) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t j = wave; j < P.N; j += waves) {
    const uint64_t base = P.incOff[j];
    for (uint64_t it = P.inOff[j] + lane; it < P.inOff[j + 1]; it += 64)
      const_cast<int16_t *>(P.inW)[it] = P.wT[base + P.inSyn[it]];
  }
}

// --------------------------------------------------------------------------
// k_compact: single workgroup; ascending ids of slot t's bitmap with their
// emission times, appended at *rasterCount.
// --------------------------------------------------------------------------
constexpr int kCompactThreads = 1024;
__global__ __launch_bounds__(kCompactThreads) void k_compact(DevNet P, uint32_t t, uint32_t *ids,
                                                             double *te, uint32_t *rasterCount,
                                                             uint32_t *stepCount) {
  __shared__ uint32_t part[kCompactThreads];
  const uint32_t slot = t & (P.RS - 1);
  const uint64_t *w = P.bits + (size_t)slot * P.NW;
  const uint32_t per = (P.NW + kCompactThreads - 1) / kCompactThreads;
  const uint32_t w0 = threadIdx.x * per, w1 = min(P.NW, w0 + per);
  uint32_t n = 0;
  for (uint32_t k = w0; k < w1; ++k) n += __popcll(w[k]);
  part[threadIdx.x] = n;
  __syncthreads();
  for (int off = 1; off < kCompactThreads; off <<= 1) {  // inclusive scan
    const uint32_t add = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  const uint32_t base = *rasterCount;
  uint32_t pos = base + part[threadIdx.x] - n;
  for (uint32_t k = w0; k < w1; ++k) {
    uint64_t m = w[k];
    while (m) {
      const int b = __ffsll((long long)m) - 1;
      m &= m - 1;
      const uint32_t id = k * 64 + (uint32_t)b;
      ids[pos] = id;
      te[pos] = P.teHist[(size_t)slot * P.N + id];
      ++pos;
    }
  }
  __syncthreads();
  if (threadIdx.x == kCompactThreads - 1) {
    *rasterCount = base + part[kCompactThreads - 1];
    *stepCount = part[kCompactThreads - 1];
  }
}

// --------------------------------------------------------------------------
// k_step: the delivery + ordering + dynamics of one step.
//
// Event key.  An event of step t arrives at k = t + frac(Te_pre) (the
// reference's double, DPSNN_neuron_sim.c:297, :579) and the reference's
// stable merge sort orders equal k by insertion order, which is the order
// of the neuron's in-list.  For t >= 1 every k lies in [t, 2t], so
// tick = bits(k) - bits(t) is exact and order-preserving (< 2^52), and
//     key = tick << 12 | position-in-in-list
// is unique and sorts exactly as (k, insertion order).  The key also
// recovers k (bits(t) + tick) and the synapse (position).
//
// Phase 1, one neuron per wave, all 64 lanes: stream the in-list
// (coalesced), keep the live items (source fired in step t - delay: one bit
// of the spike history), compact them into LDS in in-list order, load their
// emission times and form the keys, bucket the keys by the top bits of the
// tick (LDS counters), rank each key within its bucket, and store the
// sorted ticks with their weights to the scratch.
// Phase 2, one lane per neuron of the tile: merge the sorted internal
// events with the external bath (internal first on equal time) and run
// LIFCADynamic for each, in FP64.
// --------------------------------------------------------------------------
struct ExtView {  // replayed external events of this step
  const uint32_t *off;
  const double *time, *value;
};


__device__ inline uint64_t lane_mask_lt() {
  const uint32_t lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// External bath of one neuron in step t, one event at a time in time order:
// head is the next event's time (+inf when none is pending); take() returns
// its value and moves on.
constexpr double kNoEvent = __builtin_huge_val();
struct ExtReplayCursor {
  const double *time, *value;
  uint32_t x, xe;
  double head;
  __device__ inline void start() { head = x < xe ? time[x] : kNoEvent; }
  __device__ inline double take() {
    const double v = value[x++];
    head = x < xe ? time[x] : kNoEvent;
    return v;
  }
  __device__ inline void finish() {}
};

// The reference's Poisson bath (DPSNN_neuron_sim.c:604-640, :844-851) from
// the Philox stream: the first due event is always drawn (do-while), the
// following ones while fewer than the per-ms cap; past the cap only the
// lapses are drawn.
struct ExtPhiloxCursor {
  double a, inv, nextF;
  const double *lut;
  uint32_t cap, nExt;
  bool onE;
  PhiloxStream R;
  __device__ inline bool due() const { return onE && a < nextF && (nExt == 0 || nExt < cap); }
  __device__ inline double at() const { return a; }
  __device__ inline double take() {
    const double val = lut[(uint32_t)(R.next() * kLutDepth)];
    ++nExt;
    const double u = 1.0 - R.next();
    a += -logf((float)u) * inv;
    return val;
  }
  __device__ inline void finish() {
    if (onE && nExt >= cap && a < nextF) {
      do {
        const double u = 1.0 - R.next();
        a += -logf((float)u) * inv;
      } while (a < nextF);
    }
  }
};

constexpr int kQueue = 4;  // Philox bath events per lane drawn ahead of the merge
typedef __attribute__((address_space(3))) double lds_f64;

// The Philox bath with the step's first events drawn ahead (fill, all lanes
// of the wave together, before the merge): up to kQueue (time, value) pairs
// per lane in LDS (qt/qv: the lane's column of [kQueue][64]); events past
// the queue are drawn from the cursor as they come.  The Philox stream is
// consumed in the same order either way, so the events are the same.
struct ExtQueuedCursor {
  ExtPhiloxCursor xp;
  lds_f64 *qt, *qv;  // LDS-typed: a generic pointer here compiled to flat loads, which wait on global loads
  uint32_t qi, qn;
  double head;  // time of the next due event, or kNoEvent
  __device__ inline void fill(bool on) {
    qn = 0;
    for (;;) {
      const bool d = on && qn < (uint32_t)kQueue && xp.due();
      if (!__ballot(d)) break;
      if (d) {
        qt[qn * kWave] = xp.at();
        qv[qn * kWave] = xp.take();
        ++qn;
      }
    }
    qi = 0;
    head = qn ? qt[0] : (on && xp.due() ? xp.at() : kNoEvent);
  }
  __device__ inline double take() {
    double v;
    if (qi < qn) {
      v = qv[qi * kWave];
      ++qi;
    } else {
      v = xp.take();
    }
    head = qi < qn ? qt[qi * kWave] : (xp.due() ? xp.at() : kNoEvent);
    return v;
  }
  __device__ inline void finish() { xp.finish(); }
};

constexpr int kChunk = 8;  // internal events per lane per LDS-staged chunk
// From step kPackT on, a scratch entry carries its weight in the low 16 bits
// (tick << 16 | weight): for t >= 32 every tick is at most 2^47 (t + frac
// stays within [t, t + 1] and 2^(52 - floor(log2 t)) <= 2^47).  Earlier
// steps keep the weight in its own array.
constexpr uint32_t kPackT = 32;
constexpr int kRun = 16;   // events per neuron run in the blocked scratch
// LDS of one wave in phase 2: one chunk of 64 lanes' ticks, a row of
// kStageRow per event (the wave has read all of a chunk before it stores the
// next one over it), then the bath queue (times, values: [kQueue][64]
// doubles each).  A row is 64 ticks plus 8 of padding: the transposing
// stores write 8 rows x 8 neurons per instruction, which a 64-wide row would
// put on one bank per neuron (8-way conflicts); 72 spreads them over all 64
// banks.
constexpr uint32_t kStageRow = kWave + 8;
constexpr size_t kStageKBytes = kChunk * kStageRow * sizeof(uint64_t);
constexpr size_t kStageBytes = kStageKBytes + 2 * kQueue * kWave * sizeof(double);

// A buffer view of `bytes` (< 2^31) bytes at base (raw: stride 0, byte
// offsets; the gfx9 descriptor word).  An access at kDropOffset is out of
// range: a load returns zero and a store is dropped, without touching
// memory and without a branch.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kDropOffset = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t buffer_view(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
// A pointer that is the same in every lane of the wave, as a scalar value
// (a buffer view needs its base in scalar registers).
template <class T>
__device__ inline T *wave_uniform(T *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return (T *)(((uint64_t)hi << 32) | lo);
}

// Chunk c (events 8c..8c+7 of every lane) from the tile's blocked scratch
// [i / kRun][q][i % kRun]: each neuron's 8 ticks are one 64-byte piece of
// its run, loaded by the wave in 8 coalesced instructions (instruction r:
// neurons 8r..8r+7) and transposed through LDS; each lane's 8 weights are
// one 16-byte load.  Only the pieces of neurons with events in the chunk
// (bit q of act) are fetched; the others read as zero from an
// out-of-range offset.  Entries past a neuron's count are never consumed.
static_assert(kChunk == 8 && kRun % kChunk == 0, "chunks are whole pieces of a run");
__device__ inline void load_chunk(const __amdgpu_buffer_rsrc_t &tileK, const int16_t *tileW, uint32_t c,
                                  uint64_t act, uint64_t (&rk)[kChunk], uint4 &rw, bool packed) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t blk = (c * kChunk / kRun) * (kWave * kRun) + (c * kChunk) % kRun;
  const uint64_t mine = act >> (lane >> 3);  // bit 8r: neuron 8r + lane / 8 has events in chunk c
#pragma unroll
  for (int r = 0; r < kChunk; ++r) {
    const uint32_t off = (blk + (uint32_t)(r * 8 + (lane >> 3)) * kRun + (lane & 7)) * 8u;
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(tileK, ((mine >> (8 * r)) & 1u) ? off : kDropOffset, 0, 0);
    rk[r] = (uint64_t)v.x | ((uint64_t)v.y << 32);
  }
  if (!packed) rw = *(const uint4 *)(tileW + blk + (size_t)lane * kRun);
}
__device__ inline void store_chunk(uint64_t *stageK, const uint64_t (&rk)[kChunk]) {
  const uint32_t lane = threadIdx.x & 63, e = lane & 7;
#pragma unroll
  for (int r = 0; r < kChunk; ++r) stageK[e * kStageRow + (uint32_t)r * 8 + (lane >> 3)] = rk[r];
}

// Phase 2, one lane per neuron: lane q integrates neuron j0 + q (q < m) over
// its n sorted internal events (the tile's blocked scratch tileK/tileW:
// ticks relative to bits(t) and int16 weights) merged with its external
// bath, LIFCADynamic per event in FP64.  The events are staged through LDS
// (stage, kStageBytes) a chunk ahead.  Returns the lane's external event
// count and whether (and when) its neuron fired.  Whole wave, uniform.
// PK 1: every entry carries its weight (t >= kPackT), PK 2: t decides at
// run time.  The packed-only form needs 16 fewer VGPRs (128: 4 waves/SIMD
// instead of 3).
template <bool REPLAY, int PK = 2>
__device__ __attribute__((always_inline)) inline uint32_t integrate_wave(const DevNet &P, uint32_t t, const ExtView &X, uint32_t j0, uint32_t m,
                                          const uint64_t *tileK, const int16_t *tileW, uint32_t n, uint8_t *stage,
                                          bool &fired, double &te) {
  const uint32_t lane = threadIdx.x & 63;
  const bool on = lane < m;
  const uint32_t j = j0 + lane;
  if (!on) n = 0;
  uint32_t nmax = n;
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, (uint32_t)__shfl_xor((int)nmax, o));
  const double tF = (double)t, nextF = (double)(tF + 1.0);
  const uint64_t tBits = dpsnn_asu64(tF);
  Lifca L;
  uint32_t sp = 0;
  if (on) {
    sp = P.subPop[j];
    const double *prm = P.sp + (size_t)sp * kParams;
    L.v = P.v[j]; L.c = P.c[j]; L.Tr = P.Tr[j]; L.Te = P.Te[j];
    L.Tau = prm[kTau]; L.Theta = prm[kTheta]; L.H = prm[kH]; L.Tarp = prm[kTarp];
    L.AlphaC = prm[kAlphaC]; L.TauC = prm[kTauC]; L.gC = prm[kGC];
    L.invTau = prm[kInvTau]; L.invTauC = prm[kInvTauC];
  }
  L.fired = false;
  ExtReplayCursor xr{};
  xr.head = kNoEvent;
  ExtQueuedCursor xq{};
  if (on) {
    if (REPLAY) {
      xr = ExtReplayCursor{X.time, X.value, X.off[j], X.off[j + 1], kNoEvent};
      xr.start();
    } else {
      ExtPhiloxCursor &xp = xq.xp;
      xp.a = P.assigned[j];
      xp.inv = P.invNuExt[j];
      xp.nextF = nextF;
      xp.lut = P.lut + (size_t)sp * kLutDepth;
      xp.cap = P.maxExt[j];
      xp.nExt = 0;
      xp.onE = xp.inv > 0;
      xp.R = PhiloxStream{(uint32_t)P.philoxSeed, (uint32_t)(P.philoxSeed >> 32) ^ j, P.philoxCtr[j], 0.0, 0};
    }
  }
  if (!REPLAY) {
    lds_f64 *q = (lds_f64 *)(stage + kStageKBytes);
    xq.qt = q + lane;
    xq.qv = q + kQueue * kWave + lane;
    xq.fill(on);
  }
  // one merge loop with a single LIFCADynamic site: internal events from
  // the staged chunk, the bath event instead when it is strictly earlier
  // (internal first on equal times), and the rest of the bath once the
  // lane's internal events are used up
  const uint32_t chunks = (nmax + kChunk - 1) / kChunk;
  const uint32_t iters = chunks ? chunks : 1u;
  uint64_t *stageK = (uint64_t *)stage;  // [kChunk][kStageRow]
  uint64_t rk[kChunk];
  uint4 wCur = make_uint4(0, 0, 0, 0), wNext = wCur;
  const bool packed = PK == 2 ? t >= kPackT : PK == 1;
  const __amdgpu_buffer_rsrc_t tileR = buffer_view(wave_uniform(tileK), P.cap * kWave * (uint32_t)sizeof(uint64_t));
  if (chunks) {
    load_chunk(tileR, tileW, 0, __ballot(n > 0), rk, wCur, packed);
    store_chunk(stageK, rk);
    wave_sync();
  }
  for (uint32_t c = 0; c < iters; ++c) {
    const uint64_t *curK = stageK;
    const bool more = c + 1 < chunks;
    if (more) load_chunk(tileR, tileW, c + 1, __ballot(n > (c + 1) * kChunk), rk, wNext, packed);
    uint64_t wlo = (uint64_t)wCur.x | ((uint64_t)wCur.y << 32), whi = (uint64_t)wCur.z | ((uint64_t)wCur.w << 32);
    const uint32_t left = n > c * kChunk ? min((uint32_t)kChunk, n - c * kChunk) : 0u;
    const bool last = c * kChunk + left >= n;
    uint32_t e = 0;
    // the next internal event's entry is read as soon as the previous one is
    // taken, so that the LDS read overlaps the event's dynamics
    uint64_t kv = left ? curK[lane] : 0ull;
    for (;;) {
      const bool hasI = e < left;
      const double kI = dpsnn_asf64(tBits + (packed ? kv >> 16 : kv));
      // the bath event goes first when strictly earlier than the internal
      // one; with no internal event left in this chunk, only on the last
      // chunk (otherwise the next chunk is staged first)
      const double lim = hasI ? kI : (last ? kNoEvent : -kNoEvent);
      const double hx = REPLAY ? xr.head : xq.head;
      double at, val;
      if (hx < lim) {
        at = hx;
        val = REPLAY ? xr.take() : xq.take();
      } else if (hasI) {
        const int16_t w16 = (int16_t)(uint16_t)(packed ? kv : wlo);
        wlo = (wlo >> 16) | (whi << 48);
        whi >>= 16;
        at = kI;
        val = (double)(((float)w16) * P.w2f);
        ++e;
        if (e < left) kv = curK[e * kStageRow + lane];
      } else {
        break;
      }
      L.event(at, val);
    }
    if (more) {
      store_chunk(stageK, rk);
      wCur = wNext;
    }
    wave_sync();
  }
  uint32_t nExt = 0;
  if (on) {
    if (REPLAY) {
      nExt = xr.xe - X.off[j];
    } else {
      xq.finish();
      nExt = xq.xp.nExt;
      P.assigned[j] = xq.xp.a;
      P.philoxCtr[j] = xq.xp.R.ctr;
    }
    P.v[j] = L.v; P.c[j] = L.c; P.Tr[j] = L.Tr; P.Te[j] = L.Te;
  }
  fired = on && L.fired;
  te = L.Te;
  return nExt;
}

constexpr uint64_t kNoKey = ~0ull;

// Bucket of a key: the top log2(B) bits of its tick (tick < 2^(52 - e),
// e = floor(log2 t)), clamped.
template <int B>
__device__ inline uint32_t key_bucket(uint64_t tick, int shift) {
  const uint64_t b = tick >> shift;
  return b < (uint64_t)B ? (uint32_t)b : (uint32_t)(B - 1);
}

// *(base + byteOff): a 32-bit byte offset from a uniform base, so that a
// load needs one address register (the spike history is under 4 GiB:
// checked at build).
template <class T>
__device__ inline T load_at(const T *base, uint32_t byteOff) {
  return *(const T *)((const char *)base + byteOff);
}

// Byte offsets of source `pre`'s entries of slot s: its emission time in the
// Te history ([RS][N] doubles) and the 32-bit word of the fired bitmap
// ([RS][NW] u64, read as [RS][2NW] u32) that holds its bit (bit pre & 31).
// N < 2^24 (checked at build), so the products are 24-bit multiplies.
__device__ inline uint32_t te_offset(const DevNet &P, uint32_t s, uint32_t pre) {
  return (__umul24(s, P.N) + pre) * 8u;
}
__device__ inline uint32_t bit_offset(const DevNet &P, uint32_t s, uint32_t pre) {
  return (__umul24(s, 2u * P.NW) + (pre >> 5)) * 4u;
}

// LTD, synapseSetLastActiveTime_antiCausalSTDP (DPSNN_neuron_sim.c:
// 207-277): every synapse delivered to neuron j in step t gets last = t and
// td += anti[t - lastSpike_j], one increment per target, taken before step
// t's dynamics move lastSpike.
__device__ inline float ltd_increment(const DevNet &P, uint32_t t, uint32_t j) {
  const int dt = (int)t - P.lastSpike[j];
  return dt < kStdpSpan ? P.anti[dt] : 0.0f;
}

// Applies it to the n synapses whose in-list positions phase 1 left in the
// blocked scratch (pos = neuron j's [0][q][0]).  A synapse is delivered at
// most once per step (refractory >= 1 ms), so the updates are independent:
// rounds of 4 per lane issue the position loads, the synapse-index loads,
// the value loads, then the stores.  Whole wave.
__device__ __attribute__((always_inline)) inline void ltd_neuron(const DevNet &P, uint32_t t, uint32_t j, float a, uint32_t n, const uint16_t *pos) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t beg = P.inOff[j], sbase = P.incOff[j];
  constexpr int U = 4;
  for (uint32_t i0 = lane; i0 < n; i0 += U * kWave) {
    uint32_t p[U], sy[U];
    float td[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + (uint32_t)u * kWave;
      p[u] = i < n ? pos[(size_t)(i / kRun) * (kWave * kRun) + (i % kRun)] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) sy[u] = i0 + (uint32_t)u * kWave < n ? P.inSyn[beg + p[u]] : 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) td[u] = i0 + (uint32_t)u * kWave < n ? P.tdT[sbase + sy[u]] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + (uint32_t)u * kWave < n) {
        P.lastT[sbase + sy[u]] = (int32_t)t;
        P.tdT[sbase + sy[u]] = td[u] + a;
      }
  }
}

constexpr int kDense = 16;  // fast path: up to kDense * 64 live keys per neuron, kDense per lane in registers
constexpr int kSlowK = 8;   // register batches (of 64 items) of the general path

// Fast path, rounds one and two: the live items of neuron j's in-list
// (beg, len), compacted into rec[0, n) in in-list order as
// item | position << 32 | synapse << 44 (with plastic, the item's synapse
// in j's incoming list, for LTD); returns n (uniform).  K batches of 64
// in-list words are loaded at once, then the bitmap words of their sources
// (a target's sources are a few neighbouring columns, so these mostly hit
// the CU's L1).  The in-list arrays are padded by one wave, so the words
// past len that the last batch loads exist; they are never live.
template <int K>
__device__ __attribute__((always_inline)) inline uint32_t compact_live(const DevNet &P, uint32_t t, uint64_t beg, uint32_t len, uint64_t *rec,
                                        int plastic) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t *bw = (const uint32_t *)P.bits;
  uint32_t n = 0;
  for (uint32_t base = 0; base < len; base += (uint32_t)K * kWave) {
    uint32_t it[K], wb[K], sy[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (base + k * kWave < len) {
        it[k] = P.inPD[beg + base + k * kWave + lane];
        sy[k] = plastic ? P.inSyn[beg + base + k * kWave + lane] : 0u;
      }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (base + k * kWave < len)
        wb[k] = load_at(bw, bit_offset(P, (t - (it[k] >> 24)) & (P.RS - 1), it[k] & 0xFFFFFFu));
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (base + k * kWave < len) {
        const uint32_t idx = base + k * kWave + lane;
        const bool live = idx < len && (it[k] >> 24) <= t && ((wb[k] >> (it[k] & 31u)) & 1u);
        const uint64_t m = __ballot(live);
        if (live) {
          const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
          rec[n + below] = (uint64_t)it[k] | ((uint64_t)idx << 32) | ((uint64_t)sy[k] << 44);
        }
        n += (uint32_t)__popcll(m);
      }
  }
  return n;
}

// Keys of in-list items base + k*64 + lane (k < K) of neuron j; kNoKey for
// items that are absent or not live (the source did not fire in step
// t - delay: its bit in that slot's fired bitmap is clear).  With `first`,
// also counts the live keys per bucket (LDS) and keeps their weights in LDS
// by position (ldsW).  Three rounds of K independent loads: the in-list
// words; the bitmap words of their sources (a target's sources are a few
// neighbouring columns, so these mostly hit the CU's L1); then, for live
// items only, the emission times and weights.  A source's bitmap bit is set
// exactly when its Te history entry is an emission time (publish_spikes
// writes both), so the emission-time loads go to L2 only for delivered
// events.
template <int K, int B>
__device__ __attribute__((always_inline)) inline uint32_t gather_keys(const DevNet &P, uint32_t t, uint32_t j, uint64_t beg, uint32_t len,
                                       uint32_t base, uint64_t (&key)[K], uint32_t *bcnt, int16_t *ldsW,
                                       int shift, uint64_t tBits, bool first, int plastic) {
  static_assert(K <= 32, "live items are kept in one 32-bit mask");
  const uint32_t lane = threadIdx.x & 63;
  const double tF = (double)t;
  const uint32_t *bw = (const uint32_t *)P.bits;
  uint32_t it[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t idx = base + k * kWave + lane;
    it[k] = idx < len ? P.inPD[beg + idx] : kPad;
  }
  uint32_t wb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t d = it[k] >> 24, pre = it[k] & 0xFFFFFFu;
    wb[k] = 0u;
    if (it[k] != kPad && d <= t) wb[k] = load_at(bw, bit_offset(P, (t - d) & (P.RS - 1), pre));
  }
  uint32_t live = 0u;
#pragma unroll
  for (int k = 0; k < K; ++k) live |= ((wb[k] >> (it[k] & 31u)) & 1u) << k;
  double te[K];
  int16_t w[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t d = it[k] >> 24, pre = it[k] & 0xFFFFFFu;
    te[k] = kNotFired;
    if ((live >> k) & 1u) {
      te[k] = load_at(P.teHist, te_offset(P, (t - d) & (P.RS - 1), pre));
      if (first) w[k] = P.inW[beg + base + k * kWave + lane];
    }
  }
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool on = te[k] >= 0.0;
    key[k] = kNoKey;
    if (on) {
      const uint32_t idx = base + k * kWave + lane;
      const double kk = tF + (te[k] - trunc(te[k]));  // modf, exact
      const uint64_t tick = dpsnn_asu64(kk) - tBits;
      key[k] = (tick << 12) | idx;
      if (first) {
        atomicAdd(&bcnt[key_bucket<B>(tick, shift)], 1u);
        ldsW[idx] = w[k];
      }
    }
    n += (uint32_t)__popcll(__ballot(on));
  }
  return n;
}

template <int K, int B>
__device__ __attribute__((always_inline)) inline void scatter_keys(const uint64_t (&key)[K], uint32_t *bcnt, uint64_t *sorted, int shift) {
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (key[k] != kNoKey) sorted[atomicAdd(&bcnt[key_bucket<B>(key[k] >> 12, shift)], 1u)] = key[k];
}

// LDS of one phase-1 wave: sort buffer (u64) and weights by position (i16)
// for the longest in-list, then the bucket counters.
__host__ __device__ inline size_t phase1_lds(uint32_t cap, int B) {
  return (size_t)cap * sizeof(uint64_t) + (size_t)cap * sizeof(int16_t) + (size_t)(B + 1) * sizeof(uint32_t);
}

// Offset of event p of a neuron in the tile's blocked scratch, from the
// neuron's [0][q][0].
__device__ inline size_t run_offset(uint32_t p) { return (size_t)(p / kRun) * (kWave * kRun) + (p % kRun); }

// Exclusive scan of the B bucket counters in place (R = B / 64 per lane);
// bcnt[B] receives the total.  Whole wave.
template <int B>
__device__ __attribute__((always_inline)) inline void bucket_scan(uint32_t *bcnt) {
  constexpr int R = B / kWave;
  const uint32_t lane = threadIdx.x & 63;
  uint32_t c[R], sum = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) { c[r] = bcnt[lane * R + r]; sum += c[r]; }
  uint32_t incl = sum;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += v;
  }
  uint32_t run = incl - sum;
  wave_sync();
#pragma unroll
  for (int r = 0; r < R; ++r) { bcnt[lane * R + r] = run; run += c[r]; }
  if (lane == kWave - 1) bcnt[B] = run;
}

// The n keys in sorted[] are scattered into their buckets (bcnt[b] = end of
// bucket b).  Within a bucket each key's place is its rank (keys are
// unique): the count of the bucket's smaller keys.  Calls put(key, place)
// once per key.  Lanes take the keys in bucket order, so
// lanes of one bucket read the same words together; the reads are
// independent of each other.  Whole wave.
template <int B, class Put>
__device__ __attribute__((always_inline)) inline void rank_in_buckets(const uint64_t *sorted, const uint32_t *bcnt,
                                                                      uint32_t n, int shift, Put &&put) {
  for (uint32_t i = threadIdx.x & 63; i < n; i += kWave) {
    const uint64_t x = sorted[i];
    const uint32_t b = key_bucket<B>(x >> 12, shift);
    const uint32_t lo = b ? bcnt[b - 1] : 0u, hi = bcnt[b];
    uint32_t r = 0, y = lo;
    for (; y + 4 <= hi; y += 4) {  // four independent reads per round
      const uint64_t a0 = sorted[y], a1 = sorted[y + 1], a2 = sorted[y + 2], a3 = sorted[y + 3];
      r += (a0 < x ? 1u : 0u) + (a1 < x ? 1u : 0u) + (a2 < x ? 1u : 0u) + (a3 < x ? 1u : 0u);
    }
    for (; y < hi; ++y) r += sorted[y] < x ? 1u : 0u;
    put(x, lo + r);
  }
}

// Phase 1 for neuron j, whole wave: live keys of its in-list, bucket-sorted
// in LDS (lds: sorted[cap], ldsW[cap], bcnt[B + 1]), stored in arrival order
// as ticks relative to bits(t) (outK) with their weights (outW) and, with
// plastic, their in-list positions (outP, for LTD); returns n.
// The scratch of a tile of 64 neurons is blocked by runs of kRun events:
// event i of neuron q at [i / kRun][q][i % kRun] (a neuron's keys fill whole
// 128-byte lines), and outK/outW point at neuron q's [0][q][0].
// Fast path (n <= denseMax): the live items are compacted into sorted[]
// (compact_live), then each lane takes live items p = lane, lane + 64, ...:
// emission time and weight loads for delivered events only, keys in
// registers, bucket counts, scatter.  Otherwise the general path gathers
// the in-list in register batches of kSlowK * 64 items, twice.
// With plastic, the step's LTD of neuron j's delivered synapses is applied
// here too: inline on the fast path, from the stored positions (ltd_neuron)
// on the general path.
template <int K, int B>
__device__ __attribute__((always_inline)) inline uint32_t gather_sort_neuron(const DevNet &P, uint32_t t, uint32_t j, int plastic, uint8_t *lds,
                                              uint32_t cap, uint64_t *outK, int16_t *outW, uint16_t *outP) {
  static_assert(B % kWave == 0, "bucket count must be a multiple of 64");
  constexpr int LOG2B = B == 64 ? 6 : B == 128 ? 7 : B == 256 ? 8 : B == 512 ? 9 : 10;
  uint64_t *sorted = (uint64_t *)lds;
  int16_t *ldsW = (int16_t *)(lds + (size_t)cap * sizeof(uint64_t));
  uint32_t *bcnt = (uint32_t *)(lds + (size_t)cap * (sizeof(uint64_t) + sizeof(int16_t)));
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t tBits = dpsnn_asu64((double)t);
  const double tF = (double)t;
  const int e = t ? (31 - __clz((int)t)) : 0;
  const int shift = 52 - LOG2B - e;
  constexpr uint32_t kSuper = (uint32_t)kSlowK * kWave;
  const uint64_t beg = P.inOff[j];
  const uint32_t len = (uint32_t)(P.inOff[j + 1] - beg);
  for (uint32_t b = lane; b <= (uint32_t)B; b += kWave) bcnt[b] = 0;
  uint32_t n = 0;
  if (t > 0) n = compact_live<K>(P, t, beg, len, sorted, plastic);
  wave_sync();
  const bool fast = n <= P.denseMax;
  uint64_t key[kDense];
  if (fast) {
    uint64_t rc[kDense];
#pragma unroll
    for (int r = 0; r < kDense; ++r)
      if ((uint32_t)r * kWave < n) {
        const uint32_t p = r * kWave + lane;
        rc[r] = sorted[p < n ? p : 0u];
      }
    double te[kDense];
    int16_t w[kDense];
    float td[kDense];
#pragma unroll
    for (int r = 0; r < kDense; ++r)
      if ((uint32_t)r * kWave < n) {
        const uint32_t it = (uint32_t)rc[r], idx = (uint32_t)(rc[r] >> 32) & 0xFFFu;
        te[r] = load_at(P.teHist, te_offset(P, (t - (it >> 24)) & (P.RS - 1), it & 0xFFFFFFu));
        w[r] = P.inW[beg + idx];
      }
    if (plastic) {
      // LTD (synapseSetLastActiveTime_antiCausalSTDP): each delivered
      // synapse gets last = t and td += a; a synapse is delivered at most
      // once per step (refractory >= 1 ms), so the updates are independent
      const uint64_t sbase = P.incOff[j];
      const float a = ltd_increment(P, t, j);
#pragma unroll
      for (int r = 0; r < kDense; ++r)
        if ((uint32_t)r * kWave < n) td[r] = P.tdT[sbase + (uint32_t)(rc[r] >> 44)];
#pragma unroll
      for (int r = 0; r < kDense; ++r)
        if ((uint32_t)r * kWave < n && r * kWave + lane < n) {
          const uint64_t s = sbase + (uint32_t)(rc[r] >> 44);
          P.lastT[s] = (int32_t)t;
          P.tdT[s] = td[r] + a;
        }
    }
#pragma unroll
    for (int r = 0; r < kDense; ++r)
      if ((uint32_t)r * kWave < n) {
        const uint32_t idx = (uint32_t)(rc[r] >> 32) & 0xFFFu;
        const double kk = tF + (te[r] - trunc(te[r]));  // modf, exact
        const uint64_t tick = dpsnn_asu64(kk) - tBits;
        key[r] = (tick << 12) | idx;
        if (r * kWave + lane < n) {
          atomicAdd(&bcnt[key_bucket<B>(tick, shift)], 1u);
          ldsW[idx] = w[r];
        }
      }
  } else {
    n = 0;
    uint64_t k8[kSlowK];
    for (uint32_t base = 0; base < len; base += kSuper)
      n += gather_keys<kSlowK, B>(P, t, j, beg, len, base, k8, bcnt, ldsW, shift, tBits, true, plastic);
  }
  wave_sync();
  bucket_scan<B>(bcnt);
  wave_sync();
  // scatter into buckets (order inside a bucket is fixed below); the general
  // path gathers its batches again (L2 hits)
  if (fast) {
#pragma unroll
    for (int r = 0; r < kDense; ++r)
      if ((uint32_t)r * kWave < n && r * kWave + lane < n)
        sorted[atomicAdd(&bcnt[key_bucket<B>(key[r] >> 12, shift)], 1u)] = key[r];
  } else {
    uint64_t k8[kSlowK];
    for (uint32_t base = 0; base < len; base += kSuper) {
      gather_keys<kSlowK, B>(P, t, j, beg, len, base, k8, bcnt, ldsW, shift, tBits, false, 0);
      scatter_keys<kSlowK, B>(k8, bcnt, sorted, shift);
    }
  }
  wave_sync();
  rank_in_buckets<B>(sorted, bcnt, n, shift, [&](uint64_t x, uint32_t p) {
    const size_t at = run_offset(p);
    const int16_t wv = ldsW[x & 0xFFFu];
    if (t >= kPackT) {
      outK[at] = ((x >> 12) << 16) | (uint16_t)wv;
    } else {
      outK[at] = x >> 12;
      outW[at] = wv;
    }
    if (plastic && !fast) outP[at] = (uint16_t)(x & 0xFFFu);
  });
  wave_sync();
  if (plastic && !fast) ltd_neuron(P, t, j, ltd_increment(P, t, j), n, outP);
  return n;
}

// What the tail of step t does for the scan of step t+1: where its counts
// go (null: not collected) and whether LTP runs there.
struct NextScan {
  uint32_t *popCnt;
  int ltp;
  int count;  // step t+1 is simulated: its scan counts the spikes
};

// LDS of one wave of the fused step: phase 1's sort, then phase 2's stage.
__host__ __device__ inline size_t step_lds(uint32_t cap, int B) {
  const size_t p1 = phase1_lds(cap, B);
  return p1 > kStageBytes ? p1 : kStageBytes;
}

// One tile (neurons 64*tile .. 64*tile+63) of step t, one wave: phase 1 for
// each neuron into the wave's blocked scratch wkeys/wW[cap/8][64][8] and
// wcnt[64],
// phase 2 one lane per neuron, then the fired scan of step t+1.  Returns the
// tile's delivered events (uniform) and adds external events to *exSum (per
// lane).
template <int K, int B>
__device__ __attribute__((always_inline)) inline unsigned long long step_tile(const DevNet &P, uint32_t t, const ExtView &X, int ltd,
                                               const NextScan &S, uint32_t tile, uint32_t cap, uint8_t *lds,
                                               uint64_t *wkeys, int16_t *wW, uint16_t *wP, uint32_t *wcnt,
                                               unsigned long long &exSum) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t j0 = tile * kWave;
  const uint32_t m = min((uint32_t)kWave, P.N - j0);
  unsigned long long evSum = 0;
  for (uint32_t q = 0; q < m; ++q) {
    const uint32_t n =
        gather_sort_neuron<K, B>(P, t, j0 + q, ltd, lds, cap, wkeys + q * kRun, wW + q * kRun, wP + q * kRun);
    if (lane == 0) wcnt[q] = n;
    evSum += n;
  }
  __threadfence_block();
  __syncthreads();
  // ---------------- phase 2 ----------------
  const uint32_t n = lane < m ? wcnt[lane] : 0u;
  bool f = false;
  double te = 0.0;
  if (P.drive == DPSNN_DRIVE_REPLAY)
    exSum += integrate_wave<true>(P, t, X, j0, m, wkeys, wW, n, lds, f, te);
  else
    exSum += integrate_wave<false>(P, t, X, j0, m, wkeys, wW, n, lds, f, te);
  // ---------------- fired scan of step t+1 ----------------
  const uint64_t mask = publish_spikes(P, t + 1, j0 + lane, f, te, S.popCnt, S.count);
  if (S.ltp) ltp_word(P, t + 1, mask, j0);
  __syncthreads();
  return evSum;
}

// Fused step of one network: each wave owns tiles blockIdx.x, +gridDim.x, ...
template <int K, int B>
__global__ __launch_bounds__(kWave) void k_step(DevNet P, uint32_t t, ExtView X, int ltd, NextScan S) {
  extern __shared__ uint64_t lds[];
  uint64_t *wkeys = P.keys + (size_t)blockIdx.x * kWave * P.cap;
  int16_t *wW = P.keyW + (size_t)blockIdx.x * kWave * P.cap;
  uint16_t *wP = P.keyPos ? P.keyPos + (size_t)blockIdx.x * kWave * P.cap : nullptr;
  uint32_t *wcnt = P.keyCount + (size_t)blockIdx.x * kWave;
  unsigned long long evSum = 0, exSum = 0;
  const XcdWalk w = xcd_walk(P.NW);
  for (uint32_t tile = w.first; tile < w.end; tile += w.stride)
    evSum += step_tile<K, B>(P, t, X, ltd, S, tile, P.cap, (uint8_t *)lds, wkeys, wW, wP, wcnt, exSum);
  for (int o = 32; o > 0; o >>= 1) exSum += __shfl_down(exSum, o);
  if (threadIdx.x == 0) {
    atomicAdd(counter_slot(P.counters) + 0, evSum);
    atomicAdd(counter_slot(P.counters) + 1, exSum);
  }
}

// Per-network, per-step arguments of a batched launch.
struct BatchItem {
  uint32_t t;
  int ltd;
  NextScan S;
  ExtView X;
  uint32_t tileBegin;  // first global tile of this network
  uint32_t net;        // index into the batch's DevNet array
};

// Batched step: many independent networks (each at its own ms) in one
// persistent launch.  Global tile g belongs to the network whose
// [tileBegin, tileBegin + NW) holds it; each wave walks g = blockIdx.x,
// +gridDim.x, ... with scratch sized for the batch's longest in-list.
template <int K, int B>
__global__ __launch_bounds__(kWave) void k_batch_step(const DevNet *nets, const BatchItem *items, uint32_t count,
                                                      uint32_t totalTiles, uint32_t cap, uint64_t *keys,
                                                      int16_t *keyW, uint16_t *keyPos, uint32_t *keyCount) {
  extern __shared__ uint64_t lds[];
  uint64_t *wkeys = keys + (size_t)blockIdx.x * kWave * cap;
  int16_t *wW = keyW + (size_t)blockIdx.x * kWave * cap;
  uint16_t *wP = keyPos + (size_t)blockIdx.x * kWave * cap;
  uint32_t *wcnt = keyCount + (size_t)blockIdx.x * kWave;
  const XcdWalk w = xcd_walk(totalTiles);
  for (uint32_t g = w.first; g < w.end; g += w.stride) {
    uint32_t lo = 0, hi = count;  // last item with tileBegin <= g
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (items[mid].tileBegin <= g) lo = mid; else hi = mid;
    }
    const BatchItem &it = items[lo];
    const DevNet &P = nets[it.net];
    unsigned long long exSum = 0;
    const unsigned long long ev = step_tile<K, B>(P, it.t, it.X, it.ltd, it.S, g - it.tileBegin, cap,
                                                  (uint8_t *)lds, wkeys, wW, wP, wcnt, exSum);
    for (int o = 32; o > 0; o >>= 1) exSum += __shfl_down(exSum, o);
    if (threadIdx.x == 0) {
      atomicAdd(counter_slot(P.counters) + 0, ev);
      if (exSum) atomicAdd(counter_slot(P.counters) + 1, exSum);
    }
  }
}

// Split form: k_gather sorts every neuron's live events into its own rows
// of network-wide buffers (one wave per neuron), then k_integrate runs
// phase 2 one lane per neuron and the fired scan of step t+1.  Each kernel
// gets its own register budget.
template <int K, int B>
__global__ __launch_bounds__(kWave) void k_gather(DevNet P, uint32_t t, int ltd) {
  extern __shared__ uint64_t lds[];
  if (P.tDev) t = *P.tDev;
  unsigned long long evSum = 0;
  const uint32_t home = blockIdx.x % kXcds;
  uint32_t x = home;
  for (uint32_t j0, j1; (j0 = next_item(P.workCtr, P.N, x, home, &j1)) < P.N;)
    for (uint32_t j = j0; j < j1; ++j) {
      const size_t at = (size_t)(j >> 6) * P.cap * kWave + (j & 63) * kRun;  // the tile's blocked scratch
      const uint32_t n = gather_sort_neuron<K, B>(P, t, j, ltd, (uint8_t *)lds, P.cap, P.keys + at, P.keyW + at,
                                                  ltd ? P.keyPos + at : nullptr);
      if (threadIdx.x == 0) P.keyCount[j] = n;
      evSum += n;
    }
  if (threadIdx.x == 0) atomicAdd(counter_slot(P.counters) + 0, evSum);
}

// ---------------------------------------------------------------------------
// Stencil-window form of phase 1 (steps t >= kPackT without LTD).
//
// k_collect: one workgroup of kWindowWaves waves per unit (a range of the
// targets of one column), handed out by an XCD-ordered queue.  For each
// run delay d, in the in-lists' order (descending), the workgroup stages
// the emission times of slot t - d of the column's stencil columns into
// LDS, already turned into packed ticks (bits(t + frac(Te)) - bits(t)) << 16,
// or kNoKey where the source did not fire.  Its waves then stream the
// unit's items of delay d (window index << 16 | weight, coalesced chunks):
// an item is live when its window entry holds a tick, and its record
// (tick << 16 | weight) is appended to its target's run in the blocked
// scratch.  Each target's records come out in the order of its live
// in-list items, so a record's place in the run breaks equal ticks exactly
// as the in-list position does.  No per-event global load: the only
// per-event memory access besides the item stream is the LDS lookup.
//
// k_sort: one wave per target orders its records by (tick, place) in LDS
// (bucket counts, scatter, rank within bucket) and writes them back in
// place, as phase 2 reads them.
// ---------------------------------------------------------------------------
constexpr int kCollectThreads = kWindowWaves * kWave;
constexpr uint32_t kWindowMinNeurons = 1u << 16;  // the window form by default from this size
// Targets per unit: enough items per pass to amortise the pass's window
// load, units small enough for a short queue tail (1250-neuron columns: 4
// units of 313; measured best on grids of 8x8, 12x12 and 24x24 columns)
constexpr uint32_t kWindowUnitTargets = 320;
constexpr uint32_t kWindowMaxEntries = 12 * 1024;  // window entries: kWinPerThread per thread of k_collect

// Block-wide work queue: the next unit for the whole workgroup (one unit
// per counter update), XCD-ordered as next_item.  sh: two LDS words.
__device__ inline uint32_t next_unit(uint32_t *ctr, uint32_t total, uint32_t &x, uint32_t home, uint32_t *sh) {
  __syncthreads();  // every wave has read the previous hand-out
  if (threadIdx.x == 0) {
    uint32_t got = total;
    for (;;) {
      const uint32_t begin = (uint32_t)((uint64_t)total * x / kXcds);
      const uint32_t end = (uint32_t)((uint64_t)total * (x + 1) / kXcds);
      uint32_t *c = ctr + x * kCtrStride;
      if (begin + __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < end) {
        const uint32_t k = atomicAdd(c, 1u);
        if (begin + k < end) {
          got = begin + k;
          break;
        }
      }
      x = (x + 1) % kXcds;
      if (x == home) break;
    }
    sh[0] = got;
    sh[1] = x;
  }
  __syncthreads();
  x = sh[1];
  return sh[0];
}

// LDS of k_collect: the window (entries + 1 u64), the per-target record
// counts and run offsets of a unit plus one slot for padding chunks, two
// queue words.
__host__ __device__ inline size_t collect_lds(uint32_t entries, uint32_t unit) {
  return ((size_t)entries + 1) * sizeof(uint64_t) + ((size_t)unit + 1) * sizeof(uint2) + 2 * sizeof(uint32_t);
}

// Window entries per thread of k_collect (W <= kWinPerThread * 1024): each
// thread keeps its entries' history offsets in registers and loads all of
// a pass's emission times at once.
constexpr int kWinPerThread = 12;
constexpr int kCollectBatches = 4;  // 64-item batches per round of loads (the next round is in flight too)

template <int B, int CH>
__device__ __attribute__((always_inline)) inline void sort_records(uint32_t t, uint32_t n, uint8_t *lds, uint32_t cap,
                                                                   uint64_t *out);
__device__ inline uint64_t *record_runs(const DevNet &P, uint32_t j);
constexpr uint32_t kSortCap = 768;
constexpr int kSortChunks = (int)(kSortCap / kWave);
constexpr int kFusedSortB = 256;  // buckets of the sort inside k_collect
constexpr int kFusedSortChunks = 8;  // records per lane it holds in registers (more: read twice)
constexpr bool kFusedSortDefault = false;
// one wave's sort area inside k_collect, 16-byte aligned (it holds u64s)
__host__ __device__ inline size_t fused_sort_stride() { return (phase1_lds(kSortCap, kFusedSortB) + 15) / 16 * 16; }

__global__ __launch_bounds__(kCollectThreads) void k_collect(DevNet P, uint32_t t) {
  extern __shared__ uint64_t lds[];
  if (P.tDev) t = *P.tDev;
  const uint32_t W = P.winSlots * P.npc;  // entry W: the padding items' (never live)
  uint64_t *win = lds;
  // per target of the unit (+ padding's slot): records so far (.x) and the
  // element offset of its run in the unit's scratch (.y)
  uint2 *ct = (uint2 *)(win + W + 1);
  uint32_t *sh = (uint32_t *)(ct + P.winUnit + 1);
  const uint32_t pad = P.winUnit;  // the target index of padding chunks past a stream's end
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const uint32_t field = (lane >> 3) * 4;  // this lane's chunk's field in a batch's run-start table
  const bool chunkEnd = (lane & 7) == 7;
  const double tF = (double)t;
  const uint64_t tBits = dpsnn_asu64(tF);
  unsigned long long evSum = 0;
  const uint32_t home = blockIdx.x % kXcds;
  uint32_t x = home;
  for (;;) {
    const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)next_unit(P.winCtr, P.winUnits, x, home, sh));
    if (u >= P.winUnits) break;
    const uint32_t col = u / P.winUnitsPerCol, part = u - col * P.winUnitsPerCol;
    const uint32_t j0 = col * P.npc + part * P.winUnit;
    const uint32_t m = min(P.winUnit, (col + 1) * P.npc - j0);
    // the unit's records: tile (j >> 6) of the blocked scratch relative to
    // the unit's first tile, so that a record's offset fits 32 bits
    uint64_t *unitKeys = P.keys + (size_t)(j0 >> 6) * P.cap * kWave;
    const uint32_t j0lo = j0 & 63, tileElems = P.cap * kWave;
    const __amdgpu_buffer_rsrc_t keyRsrc = buffer_view(unitKeys, ((j0lo + m + 63) >> 6) * tileElems * 8u);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      const uint32_t jj = j0lo + i;
      ct[i] = make_uint2(0u, (jj >> 6) * tileElems + (jj & 63) * kRun);
    }
    if (threadIdx.x == 0) {
      win[W] = kNoKey;  // the fused sort's LDS may have overlaid it
      ct[pad] = make_uint2(0u, 0u);
    }
    // this thread's window entries e = tid + r * blockDim: the history
    // offset of their source (or ~0u: no source column in that slot)
    uint32_t off[kWinPerThread];
#pragma unroll
    for (int r = 0; r < kWinPerThread; ++r) {
      const uint32_t e = threadIdx.x + (uint32_t)r * blockDim.x;
      off[r] = ~0u;
      if (e < W) {
        const uint32_t k = e / P.npc;
        const int32_t sc = P.winCols[(size_t)col * P.winSlots + k];
        if (sc >= 0) off[r] = (uint32_t)sc * P.npc + (e - k * P.npc);
      }
    }
    for (uint32_t pass = 0; pass < P.winPasses; ++pass) {
      // emission times of the pass's slot (run delay d: slot t - d)
      const double *slot = P.teHist + (size_t)((t - (P.winPasses - 1 - pass)) & (P.RS - 1)) * P.N;
      double wv[kWinPerThread];
#pragma unroll
      for (int r = 0; r < kWinPerThread; ++r) wv[r] = off[r] != ~0u ? slot[off[r]] : kNotFired;
      __syncthreads();  // the previous pass's lookups are done
#pragma unroll
      for (int r = 0; r < kWinPerThread; ++r) {
        const uint32_t e = threadIdx.x + (uint32_t)r * blockDim.x;
        const double v = wv[r];
        if (e < W) win[e] = v >= 0.0 ? (dpsnn_asu64(tF + (v - trunc(v))) - tBits) << 16 : kNoKey;  // modf, exact
      }
      __syncthreads();
      const uint64_t *co = P.winChunk + ((size_t)u * P.winPasses + pass) * kWindowWaves + wave;
      const uint32_t cb = (uint32_t)co[0], ce = (uint32_t)co[1];
      // The wave's targets come in order, each one's chunks contiguous, the
      // first flagged.  A live item's place is the count of its target at
      // the start of the pass (read where the target's first chunk is) or,
      // for the target carried over from the previous batch, the carried
      // count, plus the live items before it in its run of chunks: the
      // live lanes below it (mbcnt) less those below its run's first lane,
      // taken from that lane in one permute.
      uint32_t carry = 0;
      uint32_t it[kCollectBatches], q[kCollectBatches], itN[kCollectBatches], qN[kCollectBatches];
      // the stream's items and descriptors, addressed by 32-bit offsets.
      // Every load is issued (lanes past the stream's end read its last
      // chunk and take padding instead), and records are stored through a
      // buffer view of the unit's scratch with an out-of-range offset where
      // an item is not live (the store is dropped): a round issues a fixed
      // number of memory operations, so the next round's loads stay in
      // flight while this one is processed.
      const uint32_t *sItem = P.winItem + (size_t)cb * kWindowChunk;
      const uint16_t *sDesc = P.winDesc + cb;
      const uint32_t last = ce - cb - 1;  // used only when the stream is not empty
      auto load = [&](uint32_t c0, uint32_t (&I)[kCollectBatches], uint32_t (&Q)[kCollectBatches]) {
#pragma unroll
        for (int v = 0; v < kCollectBatches; ++v) {
          const uint32_t rel = c0 - cb + (uint32_t)v * kWindowChunk + (lane >> 3);
          const uint32_t at = min(rel, last);
          const uint32_t i = __builtin_nontemporal_load(
              (const uint32_t *)((const char *)sItem + (at * kWindowChunk + (lane & 7)) * 4u));
          const uint32_t d = (uint32_t)load_at(sDesc, at * 2u);
          I[v] = rel <= last ? i : W << 16;
          Q[v] = rel <= last ? d : (pad | kWindowFirst);
        }
      };
      if (cb < ce) load(cb, it, q);
      for (uint32_t c0 = cb; c0 < ce; c0 += kWindowChunk * kCollectBatches) {
        const uint32_t cn = c0 + kWindowChunk * kCollectBatches;
        load(cn, itN, qN);  // past the end on the last round: the last chunk again, unused
        uint64_t r[kCollectBatches];
        uint2 base[kCollectBatches];
#pragma unroll
        for (int v = 0; v < kCollectBatches; ++v) {
          r[v] = win[it[v] >> 16];
          base[v] = ct[q[v] & kWindowTarget];
        }
#pragma unroll
        for (int v = 0; v < kCollectBatches; ++v) {
          if (c0 + (uint32_t)v * kWindowChunk >= ce) break;  // uniform
          const uint32_t qd = q[v];
          const bool live = r[v] != kNoKey;  // never on padding
          const bool first = qd >= kWindowFirst;
          const uint64_t liveM = __ballot(live);
          const uint64_t startM = __ballot(first) & 0x0101010101010101ull;  // chunks' first lanes
          // the run-start chunk of every chunk of the batch (4-bit fields):
          // the last chunk at or before it that is a target's first; 0 before
          // the first such chunk (the carried target's run)
          uint32_t T = 0;
          for (uint64_t sm = startM; sm; sm &= sm - 1) {
            const uint32_t k = (uint32_t)__builtin_ctzll(sm) >> 3;
            const uint32_t fill = ~0u << (4 * k);
            T = (T & ~fill) | ((k * 0x11111111u) & fill);
          }
          const uint32_t pre =
              __builtin_amdgcn_mbcnt_hi((uint32_t)(liveM >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)liveM, 0u));
          const uint32_t val = (first ? base[v].x : carry) - pre;
          const uint32_t p = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(((T >> field) & 15u) << 5), (int)val) + pre;
          const uint32_t at = base[v].y + (p / kRun) * (kWave * kRun) + p % kRun;
          const uint64_t rec = r[v] | (it[v] & 0xFFFFu);
          __builtin_amdgcn_raw_buffer_store_b64(u32x2{(uint32_t)rec, (uint32_t)(rec >> 32)}, keyRsrc,
                                                live ? at * 8u : kDropOffset, 0, 0);
          const uint32_t after = p + (live ? 1u : 0u);
          // a run's count grows along it: the largest of its chunks' last
          // lanes is the count after this batch
          if (chunkEnd) atomicMax(&ct[qd & kWindowTarget].x, after);
          carry = (uint32_t)__builtin_amdgcn_readlane((int)after, 63);
        }
#pragma unroll
        for (int v = 0; v < kCollectBatches; ++v) {
          it[v] = itN[v];
          q[v] = qN[v];
        }
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      P.keyCount[j0 + i] = ct[i].x;
      evSum += ct[i].x;
    }
    if (P.fuseSort) {
      // Sort the unit's targets while their records are recent: wave w
      // takes targets w + 16 k (k < 64, as m <= 1024), its count held in
      // lane k, since the waves' sort areas overlay the window and counts.
      const uint32_t mi = wave + kWindowWaves * lane;
      const uint32_t held = mi < m ? ct[mi].x : 0u;
      __syncthreads();
      uint8_t *area = (uint8_t *)lds + (size_t)wave * fused_sort_stride();
      for (uint32_t k = 0; wave + k * kWindowWaves < m; ++k) {
        const uint32_t j = j0 + wave + k * kWindowWaves;
        const uint32_t nrec = (uint32_t)__builtin_amdgcn_readlane((int)held, (int)k);
        if (nrec > P.sortCap) {
          if (lane == 0) P.bigList[atomicAdd(P.bigCount, 1u)] = j;
          continue;
        }
        sort_records<kFusedSortB, kFusedSortChunks>(t, nrec, area, kSortCap, record_runs(P, j));
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) evSum += __shfl_down(evSum, o);
  if (lane == 0 && evSum) atomicAdd(counter_slot(P.counters) + 0, evSum);
}

// Orders neuron j's n records (tick << 16 | weight, in live in-list order)
// in place at out (neuron j's [0][q][0] of the blocked scratch).  Sort key
// tick << 12 | place: unique, ordered as (arrival time, in-list position).
// LDS: phase1_lds(cap, B), cap >= n.  Up to CH * 64 records are held in
// registers; more are read again from out.  Whole wave.
template <int B, int CH>
__device__ __attribute__((always_inline)) inline void sort_records(uint32_t t, uint32_t n, uint8_t *lds, uint32_t cap,
                                                                   uint64_t *out) {
  constexpr int LOG2B = B == 64 ? 6 : B == 128 ? 7 : B == 256 ? 8 : B == 512 ? 9 : 10;
  uint64_t *sorted = (uint64_t *)lds;
  int16_t *ldsW = (int16_t *)(lds + (size_t)cap * sizeof(uint64_t));
  uint32_t *bcnt = (uint32_t *)(lds + (size_t)cap * (sizeof(uint64_t) + sizeof(int16_t)));
  const uint32_t lane = threadIdx.x & 63;
  const int shift = 52 - LOG2B - (31 - __clz((int)t));
  for (uint32_t b = lane; b <= (uint32_t)B; b += kWave) bcnt[b] = 0;
  wave_sync();
  const bool fast = n <= (uint32_t)CH * kWave;
  uint64_t key[CH];
  if (fast) {
    uint64_t rec[CH];
#pragma unroll
    for (int r = 0; r < CH; ++r)
      if ((uint32_t)r * kWave < n) {
        const uint32_t p = r * kWave + lane;
        rec[r] = out[run_offset(p < n ? p : 0u)];
      }
#pragma unroll
    for (int r = 0; r < CH; ++r)
      if ((uint32_t)r * kWave < n) {
        const uint32_t p = r * kWave + lane;
        key[r] = ((rec[r] >> 16) << 12) | p;
        if (p < n) {
          ldsW[p] = (int16_t)(uint16_t)rec[r];
          atomicAdd(&bcnt[key_bucket<B>(rec[r] >> 16, shift)], 1u);
        }
      }
  } else {
    for (uint32_t p = lane; p < n; p += kWave) {
      const uint64_t rc = out[run_offset(p)];
      ldsW[p] = (int16_t)(uint16_t)rc;
      atomicAdd(&bcnt[key_bucket<B>(rc >> 16, shift)], 1u);
    }
  }
  wave_sync();
  bucket_scan<B>(bcnt);
  wave_sync();
  if (fast) {
#pragma unroll
    for (int r = 0; r < CH; ++r)
      if ((uint32_t)r * kWave < n && r * kWave + lane < n)
        sorted[atomicAdd(&bcnt[key_bucket<B>(key[r] >> 12, shift)], 1u)] = key[r];
  } else {
    for (uint32_t p = lane; p < n; p += kWave) {
      const uint64_t k = ((out[run_offset(p)] >> 16) << 12) | p;
      sorted[atomicAdd(&bcnt[key_bucket<B>(k >> 12, shift)], 1u)] = k;
    }
  }
  wave_sync();  // every record has been read: the runs are rewritten in place
  rank_in_buckets<B>(sorted, bcnt, n, shift, [&](uint64_t x, uint32_t p) {
    out[run_offset(p)] = ((x >> 12) << 16) | (uint16_t)ldsW[x & 0xFFFu];
  });
  wave_sync();
}

// k_sort sizes each wave's LDS for kSortCap records, far above a step's
// live events per target (a few hundred at the saturated rate) but below
// the longest in-list, so that more waves fit a CU; a target with more
// records is listed for k_sort_big, whose LDS holds the longest in-list.

__device__ inline uint64_t *record_runs(const DevNet &P, uint32_t j) {
  return P.keys + (size_t)(j >> 6) * P.cap * kWave + (j & 63) * kRun;
}

// One wave per block, targets j = blockIdx.x + k * gridDim.x: each target's
// records are its own lines, so there is no locality for a work queue to
// keep (its counter round trips measured slower than this static order).
template <int B>
__global__ __launch_bounds__(kWave) void k_sort(DevNet P, uint32_t t) {
  extern __shared__ uint64_t lds[];
  if (P.tDev) t = *P.tDev;
  for (uint32_t j = blockIdx.x; j < P.N; j += gridDim.x) {
    const uint32_t n = P.keyCount[j];
    if (n > P.sortCap) {
      if ((threadIdx.x & 63) == 0) P.bigList[atomicAdd(P.bigCount, 1u)] = j;
      continue;
    }
    sort_records<B, kSortChunks>(t, n, (uint8_t *)lds, kSortCap, record_runs(P, j));
  }
}

// The targets k_sort left (more than kSortCap records), one wave each;
// bigCount is cleared by the next k_integrate.  Runs after every k_collect.
template <int B>
__global__ __launch_bounds__(kWave) void k_sort_big(DevNet P, uint32_t t) {
  extern __shared__ uint64_t lds[];
  if (P.tDev) t = *P.tDev;
  if (blockIdx.x == 0 && threadIdx.x < kXcds) P.winCtr[threadIdx.x * kCtrStride] = 0;  // for the next k_collect
  const uint32_t count = *P.bigCount;
  for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t j = P.bigList[i];
    sort_records<B, kDense>(t, P.keyCount[j], (uint8_t *)lds, P.cap, record_runs(P, j));
  }
}

// Blocks of kBlock threads: each wave holds 64 consecutive neurons starting
// at a multiple of 64, as publish_spikes requires.
template <int PK>
__global__ __launch_bounds__(kBlock) void k_integrate(DevNet P, uint32_t t, ExtView X, NextScan S) {
  __shared__ uint8_t stage[kBlock / kWave][kStageBytes];
  if (P.tDev) t = *P.tDev;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (blockIdx.x == 0 && threadIdx.x < kXcds) P.workCtr[threadIdx.x * kCtrStride] = 0;  // for the next k_gather
  if (blockIdx.x == 0 && threadIdx.x == 0 && P.bigCount) *P.bigCount = 0;                 // for the next k_sort
  const uint32_t j0 = blockIdx.x * kBlock + wave * kWave;
  const uint32_t m = j0 < P.N ? min((uint32_t)kWave, P.N - j0) : 0u;
  const uint32_t n = lane < m ? P.keyCount[j0 + lane] : 0u;
  const size_t at = (size_t)(min(j0, P.N) >> 6) * P.cap * kWave;  // the wave's tile
  bool f = false;
  double te = 0.0;
  unsigned long long ex;
  if (P.drive == DPSNN_DRIVE_REPLAY)
    ex = integrate_wave<true, PK>(P, t, X, j0, m, P.keys + at, P.keyW + at, n, stage[wave], f, te);
  else
    ex = integrate_wave<false, PK>(P, t, X, j0, m, P.keys + at, P.keyW + at, n, stage[wave], f, te);
  publish_spikes(P, t + 1, j0 + lane, f, te, S.popCnt, S.count);  // LTP follows in k_ltp
  for (int o = 32; o > 0; o >>= 1) ex += __shfl_down(ex, o);
  if (lane == 0 && ex) atomicAdd(counter_slot(P.counters) + 1, ex);
}

// The item of global tile g: the last one with tileBegin <= g.
__device__ inline uint32_t batch_item_of(const BatchItem *items, uint32_t count, uint32_t g) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (items[mid].tileBegin <= g) lo = mid; else hi = mid;
  }
  return lo;
}

// Split form of a batched step over the global tiles [g0, g1) (one chunk of
// the members' concatenated tiles): k_batch_gather runs phase 1 one wave per
// neuron into the chunk's blocked scratch (tile g at (g - g0)), then
// k_batch_integrate runs phase 2, one wave per tile, and the fired scan of
// each member's next step.
template <int K, int B>
__global__ __launch_bounds__(kWave) void k_batch_gather(const DevNet *nets, const BatchItem *items, uint32_t count,
                                                        uint32_t g0, uint32_t g1, uint32_t cap, uint64_t *keys,
                                                        int16_t *keyW, uint16_t *keyPos, uint32_t *keyCount,
                                                        uint32_t *workCtr) {
  extern __shared__ uint64_t lds[];
  const uint32_t total = (g1 - g0) * kWave, home = blockIdx.x % kXcds;
  uint32_t x = home;
  // delivered events are summed per member and added to its counter when
  // the wave moves to another member (one update per run, not per neuron)
  unsigned long long *evCtr = nullptr;
  unsigned long long evSum = 0;
  for (uint32_t u0, u1; (u0 = next_item(workCtr, total, x, home, &u1)) < total;)
  for (uint32_t u = u0; u < u1; ++u) {
    const uint32_t g = g0 + u / kWave, q = u % kWave;
    const BatchItem &it = items[batch_item_of(items, count, g)];
    const DevNet &P = nets[it.net];
    const uint32_t j = (g - it.tileBegin) * kWave + q;
    if (j >= P.N) continue;
    const size_t at = (size_t)(u / kWave) * cap * kWave + q * kRun;
    const uint32_t n = gather_sort_neuron<K, B>(P, it.t, j, it.ltd, (uint8_t *)lds, cap, keys + at, keyW + at,
                                                it.ltd ? keyPos + at : nullptr);
    if (P.counters != evCtr) {
      if (threadIdx.x == 0 && evSum) atomicAdd(counter_slot(evCtr), evSum);
      evCtr = P.counters;
      evSum = 0;
    }
    evSum += n;
    if (threadIdx.x == 0) keyCount[u] = n;
  }
  if (threadIdx.x == 0 && evSum) atomicAdd(counter_slot(evCtr), evSum);
}

template <int PK>
__global__ __launch_bounds__(kBlock) void k_batch_integrate(const DevNet *nets, const BatchItem *items,
                                                            uint32_t count, uint32_t g0, uint32_t g1, uint32_t cap,
                                                            const uint64_t *keys, const int16_t *keyW,
                                                            const uint32_t *keyCount, uint32_t *workCtr) {
  __shared__ uint8_t stage[kBlock / kWave][kStageBytes];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (blockIdx.x == 0 && threadIdx.x < kXcds) workCtr[threadIdx.x * kCtrStride] = 0;  // for the next chunk's gather
  const uint32_t g = g0 + blockIdx.x * (kBlock / kWave) + wave;
  if (g >= g1) return;  // whole wave
  const BatchItem &it = items[batch_item_of(items, count, g)];
  const DevNet &P = nets[it.net];
  const uint32_t j0 = (g - it.tileBegin) * kWave;
  const uint32_t m = min((uint32_t)kWave, P.N - j0);
  const uint32_t u = (g - g0) * kWave;
  const uint32_t n = lane < m ? keyCount[u + lane] : 0u;
  const size_t at = (size_t)(g - g0) * cap * kWave;
  bool f = false;
  double te = 0.0;
  unsigned long long ex;
  if (P.drive == DPSNN_DRIVE_REPLAY)
    ex = integrate_wave<true, PK>(P, it.t, it.X, j0, m, keys + at, keyW + at, n, stage[wave], f, te);
  else
    ex = integrate_wave<false, PK>(P, it.t, it.X, j0, m, keys + at, keyW + at, n, stage[wave], f, te);
  publish_spikes(P, it.t + 1, j0 + lane, f, te, it.S.popCnt, it.S.count);  // LTP follows in k_batch_ltp
  for (int o = 32; o > 0; o >>= 1) ex += __shfl_down(ex, o);
  if (lane == 0 && ex) atomicAdd(counter_slot(P.counters) + 1, ex);
}

// LTP of the members' next scans over global tiles [g0, g1), one wave per
// neuron that fired (item S.ltp: the member's next scan runs LTP).
__global__ __launch_bounds__(kWaveLTP) void k_batch_ltp(const DevNet *nets, const BatchItem *items, uint32_t count,
                                                        uint32_t g0, uint32_t g1) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t u = wave; u < (g1 - g0) * kWave; u += waves) {
    const uint32_t g = g0 + u / kWave;
    const BatchItem &it = items[batch_item_of(items, count, g)];
    if (!it.S.ltp) continue;
    const DevNet &P = nets[it.net];
    const uint32_t j = (g - it.tileBegin) * kWave + u % kWave;
    if (j >= P.N) continue;
    const uint64_t *bits = P.bits + (size_t)((it.t + 1) & (P.RS - 1)) * P.NW;
    if ((bits[j >> 6] >> (j & 63)) & 1ull) ltp_neuron(P, it.t + 1, j);
  }
}

// The step counter of graph-replayed steps: set before a replay, advanced
// by the last node of each step.
__global__ void k_set_t(uint32_t *t, uint32_t v) { *t = v; }
__global__ void k_tick(uint32_t *t) { *t += 1u; }

// Fills n doubles with v.
__global__ void k_fill_f64(double *p, double v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Launch shapes: K = in-list chunks of 64 items held in registers per
// pass, B = arrival-time buckets of the LDS sort.  Every shape computes the
// same result bit for bit; they differ only in speed.
using StepFn = void (*)(DevNet, uint32_t, ExtView, int, NextScan);
using GatherFn = void (*)(DevNet, uint32_t, int);
using BatchFn = void (*)(const DevNet *, const BatchItem *, uint32_t, uint32_t, uint32_t, uint64_t *, int16_t *,
                         uint16_t *, uint32_t *);
using BatchGatherFn = void (*)(const DevNet *, const BatchItem *, uint32_t, uint32_t, uint32_t, uint32_t, uint64_t *,
                               int16_t *, uint16_t *, uint32_t *, uint32_t *);
using SortFn = void (*)(DevNet, uint32_t);
struct KernelShape {
  int K, B;
  StepFn step;
  GatherFn gather;
  BatchFn batch;
  BatchGatherFn batchGather;
  SortFn sort, sortBig;
};
#define DPSNN_SHAPE(k, b) \
  {k, b, k_step<k, b>, k_gather<k, b>, k_batch_step<k, b>, k_batch_gather<k, b>, k_sort<b>, k_sort_big<b>}
static const KernelShape kShapes[] = {
    DPSNN_SHAPE(8, 64),  DPSNN_SHAPE(8, 256),  DPSNN_SHAPE(8, 1024),
    DPSNN_SHAPE(16, 64), DPSNN_SHAPE(16, 256), DPSNN_SHAPE(16, 1024),
    DPSNN_SHAPE(24, 64), DPSNN_SHAPE(24, 256), DPSNN_SHAPE(24, 1024),
};
#undef DPSNN_SHAPE
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

static uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

// Phase-1 waves (one-wave workgroups) resident per CU for a kernel and its
// LDS, from the runtime's occupancy calculator.
static uint32_t resident_waves(const void *kernel, size_t lds) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kWave, lds) != hipSuccess || blocks < 1)
    blocks = 1;
  return (uint32_t)blocks;
}

static int find_shape(int K, int B) {
  for (int i = 0; i < kNumShapes; ++i)
    if (kShapes[i].K == K && kShapes[i].B == B) return i;
  return -1;
}

static size_t lds_bytes(uint32_t cap, int B) { return step_lds(cap, B); }

}  // namespace dpsnn

// ============================================================================
// Host engine
// ============================================================================
using namespace dpsnn;

struct ExtSlot {
  uint32_t *h_off = nullptr;
  double *h_time = nullptr, *h_value = nullptr;
  size_t capEv = 0;
  uint32_t *d_off = nullptr;
  double *d_time = nullptr, *d_value = nullptr;
  hipEvent_t done = nullptr;
  bool pending = false;
};

struct dpsnn_net {
  Params par;
  Connectome con;
  Ran3 rng;
  ExternalReplay replay;
  ExternalStep extStep;
  StatWriter writer;
  bool writing = false;
  StdpEventWriter stdpWriter;  // the STDPevent file (host restatement from the spikes)
  bool stdpEvents = false;
  dpsnn_options opt{};
  hipStream_t stream = nullptr;
  DevNet dev{};
  std::vector<void *> allocs;
  uint64_t devBytes = 0;
  // history
  uint32_t t = 0;
  // external ring
  ExtSlot ext[kExtRing];
  uint32_t extNext = 0;
  // raster drain
  uint32_t *d_rasterIds = nullptr;
  double *d_rasterTe = nullptr;
  uint32_t *d_rasterCount = nullptr;  // [1]
  uint32_t *d_stepCount = nullptr;    // [drain]
  uint32_t *d_popCnt = nullptr;       // [drain][ncols*nsub]
  uint32_t rasterCap = 0, drainEvery = 100, pendingSteps = 0, pendingFrom = 0;
  bool wantRaster = false;
  std::vector<uint32_t> h_ids, h_steps, h_pop;
  std::vector<double> h_te;
  // last spikes
  std::vector<uint32_t> lastIds, aheadIds;
  std::vector<double> lastTe, aheadTe;
  uint32_t aheadMs = UINT32_MAX;
  double initSeconds = 0;
  uint64_t items = 0;
  uint32_t grid = 0;
  size_t ldsBytes = 0;
  bool stdp = false;         // plasticity runs at some step of [0, total)
  bool uploaded = false;     // build_device done (dpsnn_upload)
  std::string outDir;        // owned copy of opt.output_dir
  uint32_t stdpBlocks = 0;
  int shape = 0;             // kShapes index
  struct dpsnn_batch *batch = nullptr;  // stepped by this batch (its stream)
  hipStream_t ownStream = nullptr;
  bool split = false;        // k_gather + k_integrate instead of the fused k_step
  uint32_t gatherGrid = 0;
  std::vector<uint16_t> inSynLocal;
  // stencil-window phase 1 (k_collect + k_sort)
  WindowLayout window;       // host staging until upload
  bool winOn = false;        // the device holds the layout
  uint32_t winGrid = 0, sortGrid = 0, sortBigGrid = 0;
  size_t winLds = 0;
  // profiling
  bool profile = false;
  struct ProfEv {
    hipEvent_t a, b;
    uint32_t steps;  // steps between a and b
  };
  std::vector<ProfEv> profEv;
  size_t profUsed = 0;
  // graph-replayed steps (kGraphSteps consecutive plain steps per replay)
  hipGraphExec_t stepGraph = nullptr;
  const uint64_t *graphKeys = nullptr;  // the key scratch the graph was captured with
  uint32_t *d_t = nullptr;
  bool graphs = true;        // DPSNN_GRAPH=0: plain launches only
  uint64_t graphSteps = 0;   // steps replayed from the graph so far
  bool collect = false;  // raster / rates are being produced
};

template <class T>
static int dev_alloc(dpsnn_net *n, T **p, size_t count) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess)
    return fail(DPSNN_ERR_MEMORY, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  n->allocs.push_back(*p);
  n->devBytes += bytes;
  return 0;
}
template <class T>
static int upload(dpsnn_net *n, T *d, const T *h, size_t count) {
  HIP_TRY(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, n->stream));
  return 0;
}

static void free_all(dpsnn_net *n) {
  if (n->stream) hipStreamSynchronize(n->stream);
  for (void *p : n->allocs) hipFree(p);
  n->allocs.clear();
  for (auto &s : n->ext) {
    if (s.h_off) hipHostFree(s.h_off);
    if (s.h_time) hipHostFree(s.h_time);
    if (s.h_value) hipHostFree(s.h_value);
    if (s.d_off) hipFree(s.d_off);
    if (s.d_time) hipFree(s.d_time);
    if (s.d_value) hipFree(s.d_value);
    if (s.done) hipEventDestroy(s.done);
    s = ExtSlot{};
  }
  for (auto &e : n->profEv) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  if (n->stepGraph) hipGraphExecDestroy(n->stepGraph);
  n->stepGraph = nullptr;
  n->profEv.clear();
  if (n->stream) hipStreamDestroy(n->stream);
  n->stream = nullptr;
}

// Uploads the stencil-window layout and sizes k_collect (one workgroup of
// kCollectThreads per CU: its LDS holds a whole window) and k_sort.
static int upload_window(dpsnn_net *n, int cus) {
  const WindowLayout &L = n->window;
  DevNet &P = n->dev;
  size_t lds = collect_lds(L.slots * L.npc, L.unit);
  int maxLds = 0;
  HIP_TRY(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, n->opt.device));
  if (lds > (size_t)maxLds) return 0;  // the gather form stays
  // the sort fused into k_collect: its waves' sort areas overlay the window
  const KernelShape &K = kShapes[n->shape];
  const size_t fusedLds = std::max(lds, (size_t)kWindowWaves * fused_sort_stride());
  // (DPSNN_FUSED_SORT=0/1 overrides the default, for tests and timing)
  const char *fe = getenv("DPSNN_FUSED_SORT");
  const bool fuse = fe ? atoi(fe) != 0 : kFusedSortDefault;
  P.fuseSort = fuse && K.B == kFusedSortB && L.unit <= (uint32_t)kCollectThreads && fusedLds <= (size_t)maxLds;
  if (P.fuseSort) lds = fusedLds;
  // the device's limit, not this network's size: networks of other sizes
  // launch the same kernel
  if (lds > (size_t)64 * 1024)
    HIP_TRY(hipFuncSetAttribute((const void *)k_collect, hipFuncAttributeMaxDynamicSharedMemorySize, maxLds));
  uint32_t *d_item;
  uint16_t *d_desc;
  uint64_t *d_chunk;
  int32_t *d_cols;
  int rc;
  if ((rc = dev_alloc(n, &d_item, L.item.size())) || (rc = upload(n, d_item, L.item.data(), L.item.size())) ||
      (rc = dev_alloc(n, &d_desc, L.desc.size())) || (rc = upload(n, d_desc, L.desc.data(), L.desc.size())) ||
      (rc = dev_alloc(n, &d_chunk, L.chunk.size())) || (rc = upload(n, d_chunk, L.chunk.data(), L.chunk.size())) ||
      (rc = dev_alloc(n, &d_cols, L.cols.size())) || (rc = upload(n, d_cols, L.cols.data(), L.cols.size())) ||
      (rc = dev_alloc(n, &P.winCtr, kXcds * kCtrStride)) || (rc = dev_alloc(n, &P.bigList, P.N)) ||
      (rc = dev_alloc(n, &P.bigCount, 1)))
    return rc;
  HIP_TRY(hipMemsetAsync(P.winCtr, 0, sizeof(uint32_t) * kXcds * kCtrStride, n->stream));
  HIP_TRY(hipMemsetAsync(P.bigCount, 0, sizeof(uint32_t), n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));  // the host staging is dropped after upload
  P.winItem = d_item; P.winDesc = d_desc; P.winChunk = d_chunk; P.winCols = d_cols;
  P.winSlots = L.slots; P.winPasses = L.passes; P.winUnit = L.unit; P.winUnitsPerCol = L.unitsPerCol;
  P.winUnits = L.units;
  // DPSNN_SORT_CAP lowers k_sort's limit, so that tests send targets through
  // k_sort_big
  P.sortCap = kSortCap;
  if (const char *e = getenv("DPSNN_SORT_CAP")) P.sortCap = std::min<uint32_t>(kSortCap, (uint32_t)atoi(e));
  n->winLds = lds;
  n->winGrid = round_up(std::min<uint32_t>(L.units, (uint32_t)cus), kXcds);
  n->sortGrid = round_up(
      std::min<uint32_t>(P.N, (uint32_t)cus * resident_waves((const void *)K.sort, phase1_lds(kSortCap, K.B))), kXcds);
  n->sortBigGrid = std::min<uint32_t>(P.N, (uint32_t)cus * 4);
  n->winOn = true;
  return 0;
}

static int build_device(dpsnn_net *n) {
  Connectome &c = n->con;
  const dpsnn_parameters &q = n->par.pub;
  DevNet &P = n->dev;
  P.N = c.N;
  P.NW = (c.N + 63) / 64;
  // live slots t-maxDelay..t plus the one step t+1 writes, rounded to a
  // power of two (slot = ms & (RS - 1))
  P.RS = 1;
  while (P.RS < c.max_delay + 2) P.RS <<= 1;
  // phase 1 addresses the spike history with 32-bit byte offsets and 24-bit
  // neuron indices
  if (c.N >= (1u << 24)) return fail(DPSNN_ERR_CONFIG, "%u neurons; the engine handles fewer than 2^24", c.N);
  if (uint64_t(P.RS) * c.N * sizeof(double) > (1ull << 32))
    return fail(DPSNN_ERR_CONFIG, "spike history (%u slots x %u neurons) larger than 4 GiB", P.RS, c.N);
  P.nsub = q.subPopNumber;
  P.npc = q.neuronsPerCM;
  P.ncols = q.globCFT;
  // longest in-list, rounded to a whole number of scratch runs
  P.cap = (std::max<uint32_t>(c.max_in, 1) + (kRun - 1)) / kRun * kRun;
  // DPSNN_PHASE1_DENSE_MAX lowers the fast path's limit (0: every neuron on
  // the general path), so that tests can compare the two paths
  P.denseMax = kDense * kWave;
  if (const char *e = getenv("DPSNN_PHASE1_DENSE_MAX")) P.denseMax = std::min<uint32_t>(P.denseMax, (uint32_t)atoi(e));
  P.drive = (uint32_t)n->opt.drive;
  P.w2f = q.factorWeightType_2_Float;
  P.philoxSeed = n->opt.philox_seed;
  int rc;
  const uint32_t N = c.N;
  // per-subPop parameters and LUT
  std::vector<double> sp(size_t(P.nsub) * kParams, 0.0);
  for (uint32_t s = 0; s < P.nsub; ++s) {
    const SubPop &x = n->par.sp[s];
    double *o = &sp[size_t(s) * kParams];
    o[kTau] = x.Tau; o[kTheta] = x.Theta; o[kH] = x.H; o[kTarp] = x.Tarp;
    o[kAlphaC] = x.AlphaC; o[kTauC] = x.TauC; o[kGC] = x.gC;
    // div_rn needs normal divisors with normal reciprocals
    for (double d : {x.Tau, x.TauC})
      if (!(std::fabs(d) >= 0x1p-1021 && std::fabs(d) <= 0x1p1021))
        return fail(DPSNN_ERR_CONFIG, "sub-population %u: time constant %g outside the supported range", s, d);
    o[kInvTau] = 1.0 / x.Tau;
    o[kInvTauC] = 1.0 / x.TauC;
  }
  double *d_sp, *d_lut, *d_inv;
  uint8_t *d_sub;
  uint32_t *d_maxExt;
  if ((rc = dev_alloc(n, &d_sp, sp.size())) || (rc = upload(n, d_sp, sp.data(), sp.size()))) return rc;
  if ((rc = dev_alloc(n, &d_lut, c.bathLut.size())) || (rc = upload(n, d_lut, c.bathLut.data(), c.bathLut.size()))) return rc;
  if ((rc = dev_alloc(n, &d_inv, N)) || (rc = upload(n, d_inv, c.invNuExt.data(), N))) return rc;
  if ((rc = dev_alloc(n, &d_sub, N)) || (rc = upload(n, d_sub, c.subPop.data(), N))) return rc;
  if ((rc = dev_alloc(n, &d_maxExt, N)) || (rc = upload(n, d_maxExt, c.maxExt.data(), N))) return rc;
  P.sp = d_sp; P.lut = d_lut; P.invNuExt = d_inv; P.subPop = d_sub; P.maxExt = d_maxExt;
  // neuron state, initNeuralKind DPSNN_neuron_init.c:99-115
  std::vector<double> v(N), zeros(N, 0.0), te(N, -1000.0);
  for (uint32_t i = 0; i < N; ++i) v[i] = n->par.sp[c.subPop[i]].H;
  if ((rc = dev_alloc(n, &P.v, N)) || (rc = upload(n, P.v, v.data(), N))) return rc;
  if ((rc = dev_alloc(n, &P.c, N)) || (rc = upload(n, P.c, zeros.data(), N))) return rc;
  if ((rc = dev_alloc(n, &P.Tr, N)) || (rc = upload(n, P.Tr, zeros.data(), N))) return rc;
  if ((rc = dev_alloc(n, &P.Te, N)) || (rc = upload(n, P.Te, te.data(), N))) return rc;
  if ((rc = dev_alloc(n, &P.assigned, N)) || (rc = upload(n, P.assigned, zeros.data(), N))) return rc;
  if ((rc = dev_alloc(n, &P.philoxCtr, N))) return rc;
  HIP_TRY(hipMemsetAsync(P.philoxCtr, 0, sizeof(uint64_t) * N, n->stream));
  // history
  if ((rc = dev_alloc(n, &P.bits, size_t(P.RS) * P.NW))) return rc;
  HIP_TRY(hipMemsetAsync(P.bits, 0, sizeof(uint64_t) * size_t(P.RS) * P.NW, n->stream));
  if ((rc = dev_alloc(n, &P.teHist, size_t(P.RS) * N))) return rc;
  hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, n->stream, P.teHist, kNotFired, uint64_t(P.RS) * N);
  HIP_TRY(hipGetLastError());
  // in-lists (CSR in delivery insertion order)
  if (c.max_in >= kMaxInList)
    return fail(DPSNN_ERR_CONFIG, "a neuron has %u in-list items; the engine handles up to %u", c.max_in,
                kMaxInList - 1);
  n->items = c.in_off[N];
  {
    uint64_t *d_off;
    uint32_t *d_pd;
    int16_t *d_w;
    const uint64_t items = c.in_off[N];
    if ((rc = dev_alloc(n, &d_off, N + 1)) || (rc = upload(n, d_off, c.in_off.data(), N + 1))) return rc;
    // one wave of padding: phase 1 loads whole waves of in-list words
    if ((rc = dev_alloc(n, &d_pd, items + kWave)) || (rc = upload(n, d_pd, c.in_pre_d.data(), items))) return rc;
    HIP_TRY(hipMemsetAsync(d_pd + items, 0, sizeof(uint32_t) * kWave, n->stream));
    if ((rc = dev_alloc(n, &d_w, items)) || (rc = upload(n, d_w, c.in_w.data(), items))) return rc;
    P.inOff = d_off; P.inPD = d_pd; P.inW = d_w;
    HIP_TRY(hipStreamSynchronize(n->stream));  // host staging is dropped after create
  }
  // STDP state, target-major (see DevNet)
  if (n->stdp) {
    const uint64_t S = c.S;
    std::vector<int16_t> wT(S);
    std::vector<uint8_t> kindT(S);
    for (uint64_t k = 0; k < S; ++k) {
      wT[k] = c.weight[c.inc_syn[k]];
      kindT[k] = c.kind[c.inc_syn[k]];
    }
    uint64_t *d_incOff;
    uint8_t *d_kind;
    uint16_t *d_inSyn;
    float *d_causal, *d_anti, *d_max, *d_rev;
    if ((rc = dev_alloc(n, &d_incOff, N + 1)) || (rc = upload(n, d_incOff, c.inc_off.data(), N + 1))) return rc;
    if ((rc = dev_alloc(n, &P.wT, S)) || (rc = upload(n, P.wT, wT.data(), S))) return rc;
    if ((rc = dev_alloc(n, &d_kind, S)) || (rc = upload(n, d_kind, kindT.data(), S))) return rc;
    if ((rc = dev_alloc(n, &d_inSyn, c.in_off[N] + kWave)) ||  // padded like the in-list words
        (rc = upload(n, d_inSyn, n->inSynLocal.data(), c.in_off[N])))
      return rc;
    HIP_TRY(hipMemsetAsync(d_inSyn + c.in_off[N], 0, sizeof(uint16_t) * kWave, n->stream));
    if ((rc = dev_alloc(n, &P.tdT, S)) || (rc = dev_alloc(n, &P.lastT, S))) return rc;
    HIP_TRY(hipMemsetAsync(P.tdT, 0, 4 * S, n->stream));
    HIP_TRY(hipMemsetAsync(P.lastT, 0, 4 * S, n->stream));
    if (!q.statSynPeriodicProbe_writeDisable) {
      if ((rc = dev_alloc(n, &P.tdPre, S))) return rc;
    }
    std::vector<int32_t> lastSpike(N, -100000);  // DPSNN_localNet_init.c:217
    if ((rc = dev_alloc(n, &P.lastSpike, N)) || (rc = upload(n, P.lastSpike, lastSpike.data(), N))) return rc;
    // tables, initLongTermPlasticity DPSNN_neuron_sim.c:45-92
    std::vector<float> causal, anti;
    stdp_tables(q, causal, anti);
    std::vector<float> maxSyn(size_t(P.nsub) * P.nsub), rev(size_t(P.nsub) * P.nsub);
    for (uint32_t a = 0; a < P.nsub; ++a)
      for (uint32_t b = 0; b < P.nsub; ++b) {
        maxSyn[a * P.nsub + b] = 2.0 * n->par.sp[a].J[b];  // source a -> target b
        rev[a * P.nsub + b] = 2.0 * n->par.sp[b].J[a];
      }
    if ((rc = dev_alloc(n, &d_causal, kStdpSpan)) || (rc = upload(n, d_causal, causal.data(), kStdpSpan))) return rc;
    if ((rc = dev_alloc(n, &d_anti, kStdpSpan)) || (rc = upload(n, d_anti, anti.data(), kStdpSpan))) return rc;
    if ((rc = dev_alloc(n, &d_max, maxSyn.size())) || (rc = upload(n, d_max, maxSyn.data(), maxSyn.size()))) return rc;
    if ((rc = dev_alloc(n, &d_rev, rev.size())) || (rc = upload(n, d_rev, rev.data(), rev.size()))) return rc;
    P.incOff = d_incOff; P.kindT = d_kind; P.inSyn = d_inSyn;
    P.causal = d_causal; P.anti = d_anti; P.maxSyn = d_max; P.revMax = d_rev;
    P.f2w = q.factorFloat_2_weightType;
    P.drift = q.derivativeDrift;
    P.decayReset = q.tauDerivativeDecay_s == 1.0 ? 1 : 0;
    P.decayMul = (1 - 1 / q.tauDerivativeDecay_s);
    HIP_TRY(hipStreamSynchronize(n->stream));
  }
  // launch shapes.  The split form (k_gather, one wave per neuron, then
  // k_integrate) is the default; the fused k_step (each wave whole tiles of
  // 64 neurons) is kept as an option (DPSNN_SPLIT=0).
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, n->opt.device));
  cus = std::max(cus, 1);
  const uint32_t tiles = P.NW;
  int K = c.max_in <= 8 * 64 ? 8 : (c.max_in <= 16 * 64 ? 16 : 24), B = 256;
  if (const char *e = std::getenv("DPSNN_CHUNKS")) K = std::atoi(e);
  if (const char *e = std::getenv("DPSNN_BUCKETS")) B = std::atoi(e);
  n->shape = find_shape(K, B);
  if (n->shape < 0) return fail(DPSNN_ERR_ARG, "no kernel shape with %d register chunks and %d buckets", K, B);
  n->ldsBytes = lds_bytes(P.cap, B);
  const uint32_t wavesPerCu = std::max<uint32_t>(1, std::min<uint32_t>(16, (uint32_t)((160u << 10) / n->ldsBytes)));
  n->grid = round_up(std::min<uint32_t>(tiles, (uint32_t)cus * wavesPerCu), kXcds);
  n->stdpBlocks = std::min<uint32_t>((N + 3) / 4, (uint32_t)cus * 16);
  n->split = true;
  if (const char *e = std::getenv("DPSNN_SPLIT")) n->split = std::atoi(e) != 0;
  if (const char *e = std::getenv("DPSNN_GRAPH")) n->graphs = std::atoi(e) != 0;
  n->gatherGrid = round_up(
      std::min<uint32_t>(N, (uint32_t)cus * resident_waves((const void *)kShapes[n->shape].gather,
                                                               phase1_lds(P.cap, kShapes[n->shape].B))),
      kXcds);
  if (n->window.ok && (rc = upload_window(n, cus))) return rc;
  // the key scratch of stand-alone steps is allocated at the first
  // dpsnn_step (a network stepped only by a batch uses the batch's)
  if ((rc = dev_alloc(n, &P.workCtr, kXcds * kCtrStride))) return rc;
  HIP_TRY(hipMemsetAsync(P.workCtr, 0, sizeof(uint32_t) * kXcds * kCtrStride, n->stream));
  if ((rc = dev_alloc(n, &P.counters, kCounterSlots * kCounterStride))) return rc;
  HIP_TRY(hipMemsetAsync(P.counters, 0, sizeof(unsigned long long) * kCounterSlots * kCounterStride, n->stream));
  // raster / rates drain buffers
  n->drainEvery = n->opt.drain_every_ms ? n->opt.drain_every_ms : 100;
  const size_t popPer = size_t(P.ncols) * P.nsub;
  if ((rc = dev_alloc(n, &n->d_popCnt, popPer * n->drainEvery))) return rc;
  HIP_TRY(hipMemsetAsync(n->d_popCnt, 0, sizeof(uint32_t) * popPer * n->drainEvery, n->stream));
  n->wantRaster = (n->writing && !q.spikes_writeDisable) || n->opt.keep_spikes || n->stdpEvents;
  n->collect = n->writing || n->opt.keep_spikes;
  if (n->wantRaster) {
    n->rasterCap = N * std::min<uint32_t>(n->drainEvery, 1000) + 1;
    if ((rc = dev_alloc(n, &n->d_rasterIds, n->rasterCap)) || (rc = dev_alloc(n, &n->d_rasterTe, n->rasterCap)) ||
        (rc = dev_alloc(n, &n->d_rasterCount, 1)) || (rc = dev_alloc(n, &n->d_stepCount, n->drainEvery)))
      return rc;
    HIP_TRY(hipMemsetAsync(n->d_rasterCount, 0, sizeof(uint32_t), n->stream));
  }
  // external replay slots
  if (P.drive == DPSNN_DRIVE_REPLAY) {
    for (auto &s : n->ext) {
      HIP_TRY(hipHostMalloc((void **)&s.h_off, sizeof(uint32_t) * (N + 1), 0));
      HIP_TRY(hipMalloc((void **)&s.d_off, sizeof(uint32_t) * (N + 1)));
      HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    }
  }
  HIP_TRY(hipStreamSynchronize(n->stream));
  return 0;
}

static int ensure_ext_cap(ExtSlot &s, size_t need) {
  if (need <= s.capEv) return 0;
  const size_t cap = need + need / 2 + 1024;
  if (s.h_time) hipHostFree(s.h_time);
  if (s.h_value) hipHostFree(s.h_value);
  if (s.d_time) hipFree(s.d_time);
  if (s.d_value) hipFree(s.d_value);
  HIP_TRY(hipHostMalloc((void **)&s.h_time, sizeof(double) * cap, 0));
  HIP_TRY(hipHostMalloc((void **)&s.h_value, sizeof(double) * cap, 0));
  HIP_TRY(hipMalloc((void **)&s.d_time, sizeof(double) * cap));
  HIP_TRY(hipMalloc((void **)&s.d_value, sizeof(double) * cap));
  s.capEv = cap;
  return 0;
}

// Copies the pending raster / rates out of device memory and writes files.
static int drain(dpsnn_net *n) {
  if (n->opt.keep_spikes && n->aheadMs != UINT32_MAX && n->aheadMs + 1 == n->t) {
    n->lastIds.swap(n->aheadIds);
    n->lastTe.swap(n->aheadTe);
    n->aheadMs = UINT32_MAX;
  }
  if (n->pendingSteps == 0) return 0;
  const DevNet &P = n->dev;
  const size_t popPer = size_t(P.ncols) * P.nsub;
  const uint32_t steps = n->pendingSteps;
  n->h_pop.resize(popPer * steps);
  HIP_TRY(hipMemcpyAsync(n->h_pop.data(), n->d_popCnt, sizeof(uint32_t) * popPer * steps,
                         hipMemcpyDeviceToHost, n->stream));
  uint32_t total = 0;
  if (n->wantRaster) {
    n->h_steps.resize(steps);
    HIP_TRY(hipMemcpyAsync(n->h_steps.data(), n->d_stepCount, sizeof(uint32_t) * steps,
                           hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipMemcpyAsync(&total, n->d_rasterCount, sizeof(uint32_t), hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    n->h_ids.resize(total);
    n->h_te.resize(total);
    if (total) {
      HIP_TRY(hipMemcpyAsync(n->h_ids.data(), n->d_rasterIds, sizeof(uint32_t) * total,
                             hipMemcpyDeviceToHost, n->stream));
      HIP_TRY(hipMemcpyAsync(n->h_te.data(), n->d_rasterTe, sizeof(double) * total,
                             hipMemcpyDeviceToHost, n->stream));
    }
    HIP_TRY(hipMemsetAsync(n->d_rasterCount, 0, sizeof(uint32_t), n->stream));
  }
  HIP_TRY(hipMemsetAsync(n->d_popCnt, 0, sizeof(uint32_t) * popPer * steps, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  uint32_t pos = 0;
  for (uint32_t k = 0; k < steps; ++k) {
    const uint32_t cnt = n->wantRaster ? n->h_steps[k] : 0;
    if (n->writing &&
        !n->writer.step(n->pendingFrom + k, n->h_ids.data() + pos, n->h_te.data() + pos, cnt,
                        n->h_pop.data() + popPer * k))
      return fail(DPSNN_ERR_IO, "writing output files in %s failed", n->opt.output_dir);
    if (n->stdpEvents && !n->stdpWriter.step(n->pendingFrom + k, n->h_ids.data() + pos, cnt))
      return fail(DPSNN_ERR_IO, "writing the STDPevent file in %s failed", n->opt.output_dir);
    // last_spikes reports the scan of the last completed step (ms t - 1);
    // the record of ms t, already produced, waits for the next step
    const uint32_t m = n->pendingFrom + k;
    if (n->opt.keep_spikes && (m + 1 == n->t || m == n->t)) {
      std::vector<uint32_t> &ids = m + 1 == n->t ? n->lastIds : n->aheadIds;
      std::vector<double> &te = m + 1 == n->t ? n->lastTe : n->aheadTe;
      ids.assign(n->h_ids.begin() + pos, n->h_ids.begin() + pos + cnt);
      te.assign(n->h_te.begin() + pos, n->h_te.begin() + pos + cnt);
      if (m == n->t) n->aheadMs = m;
    }
    pos += cnt;
  }
  n->pendingSteps = 0;
  // the files hold every drained step once a drain returns (dpsnn_sync and
  // the end of the run read them complete)
  if (n->writing) n->writer.flush();
  return 0;
}

// Downloads the target-major STDP state into backward-list order.
static int synapse_state(dpsnn_net *n, int16_t *w, float *td, int32_t *last, const float *tdSrc = nullptr) {
  const Connectome &c = n->con;
  const uint64_t S = c.S;
  std::vector<int16_t> wT(S);
  std::vector<float> tdT(S);
  std::vector<int32_t> lastT(S);
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(wT.data(), n->dev.wT, 2 * S, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(tdT.data(), tdSrc ? tdSrc : n->dev.tdT, 4 * S, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(lastT.data(), n->dev.lastT, 4 * S, hipMemcpyDeviceToHost));
  for (uint64_t k = 0; k < S; ++k) {
    const uint32_t b = c.inc_syn[k];
    if (w) w[b] = wT[k];
    if (td) td[b] = tdT[k];
    if (last) last[b] = lastT[k];
  }
  return 0;
}

// synPeriodicProbe_A/B (DPSNN_localNet_sim.c:361-364, :394-397 with the
// record format of DPSNN_synapse.h:63-91): every synapse before and after
// the commit, backward-list order.
static int write_probe(dpsnn_net *n, const char *prefix, uint32_t sec, const float *tdSrc) {
  const Connectome &c = n->con;
  const dpsnn_parameters &q = n->par.pub;
  std::vector<int16_t> w(c.S);
  std::vector<float> td(c.S);
  int rc = synapse_state(n, w.data(), td.data(), nullptr, tdSrc);
  if (rc) return rc;
  char name[512];
  if (q.howManyOutputFiles == 1)
    std::snprintf(name, sizeof name, "%s/%s-sec%u-C%u-H%u-h%03u.dat", n->opt.output_dir, prefix, sec, q.globCFT,
                  q.globH, 0u);
  else if (sec == 0)
    std::snprintf(name, sizeof name, "%s/%s-start-C%u-H%u-h%u.dat", n->opt.output_dir, prefix, q.globCFT, q.globH, 0u);
  else
    std::snprintf(name, sizeof name, "%s/%s-%s-C%u-H%u-h%u.dat", n->opt.output_dir, prefix,
                  sec % 2 == 0 ? "lasteven" : "lastodd", q.globCFT, q.globH, 0u);
  FILE *f = std::fopen(name, "w");
  if (!f) return fail(DPSNN_ERR_IO, "cannot open %s", name);
  std::vector<char> line(256);
  for (uint64_t s = 0; s < c.S; ++s) {
    const float wf = (float)(w[s] * q.factorWeightType_2_Float);
    const int len = std::snprintf(line.data(), line.size(), "%d %d %d %d %d %f %f\n", (int)c.pre[s], (int)c.post[s],
                                  (int)(c.post[s] / c.N), (int)c.delay[s], (int)c.kind[s], wf, td[s]);
    std::fwrite(line.data(), 1, (size_t)len, f);
  }
  std::fclose(f);
  return 0;
}

static int commit(dpsnn_net *n, uint32_t t) {
  const dpsnn_parameters &q = n->par.pub;
  const uint32_t sec = t / 1000;
  const bool probes = n->writing && !q.statSynPeriodicProbe_writeDisable && sec >= q.startStatFiles_s &&
                      sec <= q.stopStatFiles_s && sec % q.moduloSec == 0 && sec >= q.startSynPeriodicProbe_s &&
                      sec <= q.stopSynPeriodicProbe_s;
  int rc;
  if (probes && (rc = write_probe(n, "synPeriodicProbe_A", sec, nullptr))) return rc;
  hipLaunchKernelGGL(k_commit, dim3(n->stdpBlocks), dim3(kWaveLTP), 0, n->stream, n->dev);
  hipLaunchKernelGGL(k_refresh_weights, dim3(n->stdpBlocks), dim3(kWaveLTP), 0, n->stream, n->dev);
  HIP_TRY(hipGetLastError());
  if (probes && (rc = write_probe(n, "synPeriodicProbe_B", sec, n->dev.tdPre))) return rc;
  return 0;
}

static int prof_begin(dpsnn_net *n) {
  if (!n->profile) return 0;
  if (n->profUsed == n->profEv.size()) {
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    n->profEv.push_back({a, b, 0});
  }
  HIP_TRY(hipEventRecord(n->profEv[n->profUsed].a, n->stream));
  return 0;
}
static int prof_end(dpsnn_net *n, uint32_t steps = 1) {
  if (!n->profile) return 0;
  HIP_TRY(hipEventRecord(n->profEv[n->profUsed].b, n->stream));
  n->profEv[n->profUsed].steps = steps;
  n->profUsed++;
  return 0;
}

// Opens the raster/rates record of ms m (records are contiguous) and returns
// its index in the drain buffers; drains first when they are full.
static int open_record(dpsnn_net *n, uint32_t m, uint32_t *idx) {
  int rc;
  if (n->pendingSteps == n->drainEvery && (rc = drain(n))) return rc;
  if (n->pendingSteps == 0) n->pendingFrom = m;
  *idx = n->pendingSteps++;
  return 0;
}

static int compact(dpsnn_net *n, uint32_t m, uint32_t rec) {
  if (n->wantRaster)
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(kCompactThreads), 0, n->stream, n->dev, m, n->d_rasterIds,
                       n->d_rasterTe, n->d_rasterCount, n->d_stepCount + rec);
  return 0;
}

static bool plastic_at(const dpsnn_net *n, uint32_t t) {
  const dpsnn_parameters &q = n->par.pub;
  return n->stdp && t >= q.startPlasticity_ms && t <= q.stopPlasticity_ms;
}

// Key scratch of stand-alone steps, blocked by 8 events per neuron:
// [grid][cap/8][64][8] (fused) or [N/64][cap/8][64][8] (split).
static int ensure_scratch(dpsnn_net *n) {
  DevNet &P = n->dev;
  if (P.keys) return 0;
  const size_t slots = n->split ? size_t(P.NW) * kWave : size_t(n->grid) * kWave;
  int rc;
  if ((rc = dev_alloc(n, &P.keys, slots * P.cap)) || (rc = dev_alloc(n, &P.keyW, slots * P.cap)) ||
      (rc = dev_alloc(n, &P.keyCount, slots)))
    return rc;
  if (n->stdp && (rc = dev_alloc(n, &P.keyPos, slots * P.cap))) return rc;
  return 0;
}

// Key scratch of the allocation-free step (dpsnn_step_noalloc): one static
// array per device, held from the library's load on that device until the
// process ends.  Every allocation-free step on a device uses the same
// array, so such steps of two networks that overlap in time on one device
// race on it.  Nothing in it needs to start at zero: each step's phase 1
// writes every count and key its phase 2 reads.
constexpr size_t kStaticScratchBytes = size_t(128) << 20;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static void *g_scratchPtr[kMaxDevices];

// The device address of `scratch` on the current device (set up once per
// device: the name is a host-side stand-in, not a device address).
static int static_scratch_ptr(int device, uint8_t **out) {
  if (device < 0 || device >= kMaxDevices) return fail(DPSNN_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchPtr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DPSNN_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratchPtr[device] = p;
  }
  *out = (uint8_t *)g_scratchPtr[device];
  return 0;
}

// n's device view with its key scratch carved from the static array, in
// the layout ensure_scratch allocates; refuses networks whose scratch is
// larger than the array.
static int static_scratch_view(dpsnn_net *n, DevNet &P) {
  P = n->dev;
  const size_t slots = n->split ? size_t(P.NW) * kWave : size_t(n->grid) * kWave;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t kb = al(slots * P.cap * sizeof(uint64_t)), wb = al(slots * P.cap * sizeof(int16_t)),
               cb = al(slots * sizeof(uint32_t)), pb = n->stdp ? al(slots * P.cap * sizeof(uint16_t)) : 0;
  const size_t need = kb + wb + cb + pb;
  if (need > kStaticScratchBytes)
    return fail(DPSNN_ERR_ARG, "the network's step needs %zu bytes of key scratch; the allocation-free step holds %zu",
                need, kStaticScratchBytes);
  uint8_t *base;
  int rc = static_scratch_ptr(n->opt.device, &base);
  if (rc) return rc;
  P.keys = (uint64_t *)base;
  P.keyW = (int16_t *)(base + kb);
  P.keyCount = (uint32_t *)(base + kb + wb);
  P.keyPos = n->stdp ? (uint16_t *)(base + kb + wb + cb) : nullptr;
  return 0;
}

// What one step of a network needs around its kernel launch.
struct StepPlan {
  uint32_t t = 0;
  int ltd = 0;
  bool commitNow = false, ltpNext = false, recordNext = false;
  uint32_t rec = 0;
  NextScan S{nullptr, 0, 0};
  ExtView X{nullptr, nullptr, nullptr};
  ExtSlot *slot = nullptr;
};

// Host work before step t's kernel: raster/rates records, the replayed
// bath of step t, and the flags of the step.
static int pre_step(dpsnn_net *n, StepPlan &sp) {
  DevNet &P = n->dev;
  const uint32_t t = n->t, next = t + 1;
  const uint32_t T = n->par.pub.totalSimTime_ms;
  const size_t popPer = size_t(P.ncols) * P.nsub;
  int rc;
  sp = StepPlan{};
  sp.t = t;
  if (n->collect && t == 0) {
    uint32_t rec0;
    if ((rc = open_record(n, 0, &rec0)) || (rc = compact(n, 0, rec0))) return rc;
  }
  sp.ltd = plastic_at(n, t);
  sp.commitNow = sp.ltd && t % 1000 == 999;
  sp.ltpNext = next < T && plastic_at(n, next);
  sp.recordNext = n->collect && next < T;
  sp.S = NextScan{nullptr, sp.ltpNext && !sp.commitNow, next < T};
  if (sp.recordNext) {
    if ((rc = open_record(n, next, &sp.rec))) return rc;
    sp.S.popCnt = n->d_popCnt + popPer * sp.rec;
  }
  if (P.drive == DPSNN_DRIVE_REPLAY) {
    ExtSlot &s = n->ext[n->extNext];
    sp.slot = &s;
    n->extNext = (n->extNext + 1) % kExtRing;
    if (s.pending) HIP_TRY(hipEventSynchronize(s.done));
    n->replay.generate(t, n->extStep);
    const size_t ne = n->extStep.time.size();
    if ((rc = ensure_ext_cap(s, ne))) return rc;
    std::memcpy(s.h_off, n->extStep.off.data(), sizeof(uint32_t) * (P.N + 1));
    if (ne) {
      std::memcpy(s.h_time, n->extStep.time.data(), sizeof(double) * ne);
      std::memcpy(s.h_value, n->extStep.value.data(), sizeof(double) * ne);
    }
    HIP_TRY(hipMemcpyAsync(s.d_off, s.h_off, sizeof(uint32_t) * (P.N + 1), hipMemcpyHostToDevice, n->stream));
    if (ne) {
      HIP_TRY(hipMemcpyAsync(s.d_time, s.h_time, sizeof(double) * ne, hipMemcpyHostToDevice, n->stream));
      HIP_TRY(hipMemcpyAsync(s.d_value, s.h_value, sizeof(double) * ne, hipMemcpyHostToDevice, n->stream));
    }
    sp.X = ExtView{s.d_off, s.d_time, s.d_value};
  }
  return 0;
}

// Host work after step t's kernel: the per-second STDP commit (and the LTP
// of the scan of t+1, which follows it), the raster of t+1.
static int post_step(dpsnn_net *n, StepPlan &sp) {
  int rc;
  const uint32_t next = sp.t + 1;
  if (sp.commitNow) {
    if ((rc = commit(n, sp.t))) return rc;
    if (sp.ltpNext) hipLaunchKernelGGL(k_ltp, dim3(n->stdpBlocks), dim3(kWaveLTP), 0, n->stream, n->dev, next);
  }
  if (sp.recordNext && (rc = compact(n, next, sp.rec))) return rc;
  if (sp.slot) {
    HIP_TRY(hipEventRecord(sp.slot->done, n->stream));
    sp.slot->pending = true;
  }
  HIP_TRY(hipGetLastError());
  n->t = next;
  if (n->opt.keep_spikes) return drain(n);
  return 0;
}

// Step t of one network: delivery + dynamics of t, and the fired scan of
// t+1 (which the reference runs at the head of step t+1; nothing reads the
// scan's outputs in between).  The scan of step 0 finds no spikes.  P is
// the network's device view with the key scratch the kernels use (the
// network's own, or the allocation-free step's static one).
// Whether step sp of n runs phase 1 in the window form: split form, a
// window layout on the device, packed records (t >= kPackT), no LTD.
static bool window_step(const dpsnn_net *n, const StepPlan &sp) {
  return n->split && n->winOn && sp.t >= kPackT && !sp.ltd;
}

// The launches of a window-form step of n on `stream`, with the device view
// P (the network's own, or one whose key scratch is a batch's).
static void launch_window_step(const dpsnn_net *n, const DevNet &P, const StepPlan &sp, hipStream_t stream) {
  const KernelShape &K = kShapes[n->shape];
  hipLaunchKernelGGL(k_collect, dim3(n->winGrid), dim3(kCollectThreads), n->winLds, stream, P, sp.t);
  if (!P.fuseSort)
    hipLaunchKernelGGL(K.sort, dim3(n->sortGrid), dim3(kWave), phase1_lds(kSortCap, K.B), stream, P, sp.t);
  hipLaunchKernelGGL(K.sortBig, dim3(n->sortBigGrid), dim3(kWave), phase1_lds(P.cap, K.B), stream, P, sp.t);
  const uint32_t blocks = (P.N + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_integrate<1>, dim3(blocks), dim3(kBlock), 0, stream, P, sp.t, sp.X, sp.S);
  if (sp.S.ltp) hipLaunchKernelGGL(k_ltp, dim3(n->stdpBlocks), dim3(kWaveLTP), 0, stream, P, sp.t + 1);
}

static int one_step(dpsnn_net *n, const DevNet &P) {
  StepPlan sp;
  int rc;
  if ((rc = pre_step(n, sp)) || (rc = prof_begin(n))) return rc;
  const KernelShape &K = kShapes[n->shape];
  if (window_step(n, sp)) {
    launch_window_step(n, P, sp, n->stream);
  } else if (n->split) {
    hipLaunchKernelGGL(K.gather, dim3(n->gatherGrid), dim3(kWave), n->ldsBytes, n->stream, P, sp.t, sp.ltd);
    const uint32_t blocks = (P.N + kBlock - 1) / kBlock;
    if (sp.t >= kPackT)
      hipLaunchKernelGGL(k_integrate<1>, dim3(blocks), dim3(kBlock), 0, n->stream, P, sp.t, sp.X, sp.S);
    else
      hipLaunchKernelGGL(k_integrate<2>, dim3(blocks), dim3(kBlock), 0, n->stream, P, sp.t, sp.X, sp.S);
    if (sp.S.ltp) hipLaunchKernelGGL(k_ltp, dim3(n->stdpBlocks), dim3(kWaveLTP), 0, n->stream, P, sp.t + 1);
  } else {
    hipLaunchKernelGGL(K.step, dim3(n->grid), dim3(kWave), n->ldsBytes, n->stream, P, sp.t, sp.X, sp.ltd, sp.S);
  }
  if ((rc = prof_end(n))) return rc;
  return post_step(n, sp);
}

// ----------------------------------------------------------------------------
// Graph-replayed steps.  kGraphSteps consecutive plain steps are captured
// once into a HIP graph whose kernels read the step's t from the device
// counter d_t (set before each replay, advanced by a k_tick node after each
// step), so that the same graph replays any run of such steps.  Steps
// qualify when nothing the host does around a step is needed: split form,
// Philox drive, no raster or rates collected, no plasticity, t >= kPackT
// (the packed scratch), and the next step still inside the run.
// ----------------------------------------------------------------------------
constexpr uint32_t kGraphSteps = 10;

static bool graph_run_ok(const dpsnn_net *n, uint32_t steps) {
  const uint32_t t = n->t, T = n->par.pub.totalSimTime_ms;
  if (!n->graphs || !n->split || n->dev.drive != DPSNN_DRIVE_PHILOX || n->collect) return false;
  if (t < kPackT || (uint64_t)t + steps >= T) return false;
  for (uint32_t k = 0; k <= steps; ++k)
    if (plastic_at(n, t + k)) return false;
  return true;
}

// Launches of one plain step in the graph form (t from P.tDev).
static void launch_graph_step(dpsnn_net *n, const DevNet &G) {
  const KernelShape &K = kShapes[n->shape];
  if (n->winOn) {
    hipLaunchKernelGGL(k_collect, dim3(n->winGrid), dim3(kCollectThreads), n->winLds, n->stream, G, 0u);
    if (!G.fuseSort)
      hipLaunchKernelGGL(K.sort, dim3(n->sortGrid), dim3(kWave), phase1_lds(kSortCap, K.B), n->stream, G, 0u);
    hipLaunchKernelGGL(K.sortBig, dim3(n->sortBigGrid), dim3(kWave), phase1_lds(G.cap, K.B), n->stream, G, 0u);
  } else {
    hipLaunchKernelGGL(K.gather, dim3(n->gatherGrid), dim3(kWave), n->ldsBytes, n->stream, G, 0u, 0);
  }
  const uint32_t blocks = (G.N + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_integrate<1>, dim3(blocks), dim3(kBlock), 0, n->stream, G, 0u, ExtView{nullptr, nullptr, nullptr},
                     NextScan{nullptr, 0, 1});
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, n->stream, n->d_t);
}

static int capture_step_graph(dpsnn_net *n, const DevNet &P) {
  if (n->stepGraph) {
    HIP_TRY(hipGraphExecDestroy(n->stepGraph));
    n->stepGraph = nullptr;
  }
  if (!n->d_t) {
    int rc = dev_alloc(n, &n->d_t, 1);
    if (rc) return rc;
  }
  DevNet G = P;
  G.tDev = n->d_t;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
  for (uint32_t k = 0; k < kGraphSteps; ++k) launch_graph_step(n, G);
  const hipError_t e = hipStreamEndCapture(n->stream, &g);
  if (e != hipSuccess) return fail(DPSNN_ERR_DEVICE, "graph capture of the step failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&n->stepGraph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) {
    n->stepGraph = nullptr;
    return fail(DPSNN_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
  }
  n->graphKeys = P.keys;
  return 0;
}

// kGraphSteps steps from n->t by one graph replay.
static int graph_steps(dpsnn_net *n, const DevNet &P) {
  int rc;
  if ((!n->stepGraph || n->graphKeys != P.keys) && (rc = capture_step_graph(n, P))) return rc;
  hipLaunchKernelGGL(k_set_t, dim3(1), dim3(1), 0, n->stream, n->d_t, n->t);
  if ((rc = prof_begin(n))) return rc;
  HIP_TRY(hipGraphLaunch(n->stepGraph, n->stream));
  if ((rc = prof_end(n, kGraphSteps))) return rc;
  HIP_TRY(hipGetLastError());
  n->t += kGraphSteps;
  n->graphSteps += kGraphSteps;
  return 0;
}

// Up to nsteps steps with the device view P (graph replays where the steps
// qualify and `graphs` allows them).
static int run_steps(dpsnn_net *n, const DevNet &P, uint32_t nsteps, bool graphs) {
  int rc;
  for (uint32_t k = 0; k < nsteps;) {
    if (n->t >= n->par.pub.totalSimTime_ms) {
      rc = drain(n);
      return rc ? rc : DPSNN_DONE;
    }
    if (graphs && nsteps - k >= kGraphSteps && graph_run_ok(n, kGraphSteps)) {
      if ((rc = graph_steps(n, P))) return rc;
      k += kGraphSteps;
      continue;
    }
    if ((rc = one_step(n, P))) return rc;
    ++k;
  }
  return 0;
}

// ============================================================================
// C ABI
// ============================================================================
// --------------------------------------------------------------------------
// Device connectome generator (dpsnn_options.generator = device): the
// distributions of the reference's generator (generate_column;
// DPSNN_LIFCAcorticalModuleGen.cpp:62-220, DPSNN_LIFCAconnectome.cpp:
// 1079-1133) drawn from Philox4x32-10, one thread per source neuron, instead
// of the sequential ran3 stream: for each target sub-population and stencil
// cell with connection probability q, targets at geometric gaps over the
// cell's column (smallest n >= 1 with 1 - (1 - q)^n >= r, at most span + 1,
// as the reference's linear search), then per target a uniform delay
// (uniform_delay's level arithmetic) and a truncated Gaussian weight (the
// reference's rejection bounds), int16-quantised.  Pass 0 counts each
// source's synapses, pass 1 writes them at their offsets in generation
// order; the host then sorts them into the backward list as for the
// reference's.  Statistically equivalent, not bit-exact.
// --------------------------------------------------------------------------
struct GenView {
  uint32_t N, npc, CFX, CFY, nsub, sx, sy, mdx, mdy, overall, D;
  float f2w, minW, delayLo, delayHi;
  uint64_t seed;
  const float *prob;                 // [nsub][nsub][sx][sy]
  const float *J, *DJ, *DMin, *DMax;  // [nsub][nsub], source-major
  const uint32_t *subOff, *subCnt;   // [nsub], within a column
  uint32_t *count;
  const uint64_t *off;
  uint32_t *post;
  uint8_t *delay;
  int16_t *w;
  uint32_t *error;                   // bit 0: q in (0, 1e-5]; 1: 7500 or more
                                     // synapses; 2: weight bound; 3: delay >= env_D;
                                     // 4: no weight within the draw bounds
};

__device__ inline uint32_t gen_round_half_up(double r) {
  const double f = floor(r);
  return (f + 0.5 > r) ? (uint32_t)f : (uint32_t)(f + 1.0);
}

template <int PASS>
__global__ __launch_bounds__(256) void k_generate(GenView G) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.N) return;
  const uint32_t col = i / G.npc, inCM = i - col * G.npc, cfx = col % G.CFX, cfy = col / G.CFX;
  uint32_t src = 0;
  for (uint32_t s = 0; s < G.nsub; ++s)
    if (inCM >= G.subOff[s] && inCM < G.subOff[s] + G.subCnt[s]) src = s;
  PhiloxStream A{(uint32_t)G.seed, (uint32_t)(G.seed >> 32) ^ i, 0ull, 0.0, 0};
  PhiloxStream Bw{(uint32_t)G.seed ^ 0x5bd1e995u, (uint32_t)(G.seed >> 32) ^ i, 0ull, 0.0, 0};
  const uint64_t at0 = PASS ? G.off[i] : 0ull;
  uint32_t n = 0, err = 0;
  for (uint32_t ts = 0; ts < G.nsub; ++ts)
    for (uint32_t x = 0; x < G.sx; ++x)
      for (uint32_t y = 0; y < G.sy; ++y) {
        int32_t tx, ty;
        if (G.overall == 0) {
          tx = (int32_t)cfx + (int32_t)x - (int32_t)G.mdx;
          ty = (int32_t)cfy + (int32_t)y - (int32_t)G.mdy;
        } else {
          tx = (int32_t)x;
          ty = (int32_t)y;
        }
        if (tx < 0 || tx >= (int32_t)G.CFX || ty < 0 || ty >= (int32_t)G.CFY) continue;
        const float q = G.prob[((src * G.nsub + ts) * G.sx + x) * G.sy + y];
        if (!(q > 0.0f)) continue;
        if (!(q > 0.00001f)) {
          err |= 1u;
          continue;
        }
        const uint32_t first = G.subOff[ts], end = first + G.subCnt[ts];
        const uint32_t colBase = (uint32_t)tx * G.npc + (uint32_t)ty * G.npc * G.CFX;
        uint32_t cur = first - 1;  // wraps as the reference's uint32 arithmetic
        const uint32_t span = end - cur;
        const double lq = log1p(-(double)q);
        const float mean = G.J[src * G.nsub + ts], sigma = G.DJ[src * G.nsub + ts] * mean;
        for (;;) {
          const double r = A.next();
          double g = q >= 1.0f ? 1.0 : ceil(log1p(-r) / lq);
          g = fmin(fmax(g, 1.0), (double)span + 1.0);
          cur += (uint32_t)g;
          if (cur >= end) break;
          if (n + 1 >= kMaxOutDegree) {
            err |= 2u;
            break;
          }
          if (PASS) {
            // delay: uniform_delay's levels
            float DMin = G.DMin[src * G.nsub + ts], DMax = G.DMax[src * G.nsub + ts];
            float lo = G.delayLo, hi = G.delayHi;
            const float step = (hi - lo) / (float)kDelaySlots;
            hi -= step / 2;
            lo += step / 2;
            DMax = gen_round_half_up((DMax - lo) / step) * step + lo;
            DMin = gen_round_half_up((DMin - lo) / step) * step + lo;
            const float level = (float)((Bw.next() * (DMax - DMin) + DMin - lo) / step);
            const uint32_t d = gen_round_half_up(level * step);
            // weight: Gaussian (Box-Muller) within the reference's bounds
            // (at most 64 draws: a bound the draws cannot meet, such as a
            // zero mean, is reported instead of looping)
            float wv = mean;
            bool inBounds = false;
            for (int tries = 0; tries < 64 && !inBounds; ++tries) {
              const double u1 = 1.0 - Bw.next(), u2 = Bw.next();
              const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
              wv = (float)(mean + z * sigma);
              inBounds = mean > 0.0f ? !(wv < 0.0f || wv >= 2.0f * mean) : !(wv <= 2.0f * mean || wv > 0.0f);
            }
            if (!inBounds) err |= 16u;
            if (wv < G.minW || wv > -G.minW) err |= 4u;
            if (d >= kDelaySlots || d >= G.D) err |= 8u;
            G.post[at0 + n] = colBase + cur;
            G.delay[at0 + n] = (uint8_t)min(d, kDelaySlots - 1);
            G.w[at0 + n] = (int16_t)(wv * G.f2w);
          }
          ++n;
        }
      }
  if (!PASS) G.count[i] = n;
  if (err) atomicOr(G.error, err);
}

// The connectome of n drawn on its device (k_generate), then sorted into the
// backward list on the host.
static int generate_on_device(dpsnn_net *n, std::string &err) {
  const Params &p = n->par;
  const dpsnn_parameters &q = p.pub;
  Connectome &c = n->con;
  connectome_neurons(p, c);
  const uint32_t N = c.N, npc = q.neuronsPerCM, nsub = q.subPopNumber;
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t inCM = i % npc;
    uint32_t src = 0;
    for (uint32_t s = 0; s < nsub; ++s)
      if (inCM >= p.sp[s].offset && inCM < p.sp[s].offset + p.sp[s].count) src = s;
    c.subPop[i] = (uint8_t)src;
    neuron_bath(p, i, src, c);
  }
  std::vector<float> prob(size_t(nsub) * nsub * p.stencilX * p.stencilY), J(nsub * nsub), DJ(nsub * nsub),
      DMin(nsub * nsub), DMax(nsub * nsub);
  std::vector<uint32_t> off(nsub), cnt(nsub);
  for (uint32_t s = 0; s < nsub; ++s) {
    off[s] = p.sp[s].offset;
    cnt[s] = p.sp[s].count;
    for (uint32_t t = 0; t < nsub; ++t) {
      J[s * nsub + t] = (float)p.sp[s].J[t];
      DJ[s * nsub + t] = (float)p.sp[s].DJ[t];
      DMin[s * nsub + t] = (float)p.sp[s].DMin[t];
      DMax[s * nsub + t] = (float)p.sp[s].DMax[t];
      for (uint32_t x = 0; x < p.stencilX; ++x)
        for (uint32_t y = 0; y < p.stencilY; ++y)
          prob[((s * nsub + t) * p.stencilX + x) * p.stencilY + y] = p.prob_at(s, t, x, y);
    }
  }
  HIP_TRY(hipSetDevice(n->opt.device));
  hipStream_t st;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::vector<void *> bufs;
  auto fin = [&](int rc) {
    (void)hipStreamSynchronize(st);
    for (void *b : bufs) (void)hipFree(b);
    (void)hipStreamDestroy(st);
    return rc;
  };
  auto alloc = [&](void **ptr, size_t bytes) -> bool {
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) return false;
    bufs.push_back(*ptr);
    return true;
  };
  auto put = [&](void **ptr, const void *src, size_t bytes) -> bool {
    return alloc(ptr, bytes) && hipMemcpyAsync(*ptr, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  };
  GenView G{};
  G.N = N; G.npc = npc; G.CFX = q.globCFX; G.CFY = q.globCFY; G.nsub = nsub;
  G.sx = p.stencilX; G.sy = p.stencilY; G.mdx = p.maxModDeltaX; G.mdy = p.maxModDeltaY;
  G.overall = q.overallConnectivity; G.D = q.D;
  G.f2w = q.factorFloat_2_weightType; G.minW = q.minSynWeight_f;
  G.delayLo = (float)q.delayMin; G.delayHi = (float)q.delayMax;
  G.seed = n->opt.generator_seed;
  void *dProb, *dJ, *dDJ, *dDMin, *dDMax, *dOff, *dCnt, *dCount, *dErr;
  if (!put(&dProb, prob.data(), 4 * prob.size()) || !put(&dJ, J.data(), 4 * J.size()) ||
      !put(&dDJ, DJ.data(), 4 * DJ.size()) || !put(&dDMin, DMin.data(), 4 * DMin.size()) ||
      !put(&dDMax, DMax.data(), 4 * DMax.size()) || !put(&dOff, off.data(), 4 * nsub) ||
      !put(&dCnt, cnt.data(), 4 * nsub) || !alloc(&dCount, 4 * (size_t)N) || !alloc(&dErr, 4))
    return fin(fail(DPSNN_ERR_MEMORY, "device generator: out of device memory"));
  G.prob = (const float *)dProb; G.J = (const float *)dJ; G.DJ = (const float *)dDJ;
  G.DMin = (const float *)dDMin; G.DMax = (const float *)dDMax;
  G.subOff = (const uint32_t *)dOff; G.subCnt = (const uint32_t *)dCnt;
  G.count = (uint32_t *)dCount; G.error = (uint32_t *)dErr;
  if (hipMemsetAsync(G.error, 0, 4, st) != hipSuccess) return fin(fail(DPSNN_ERR_DEVICE, "device generator: memset failed"));
  const dim3 grid((N + 255) / 256);
  hipLaunchKernelGGL(k_generate<0>, grid, dim3(256), 0, st, G);
  if (hipGetLastError() != hipSuccess) return fin(fail(DPSNN_ERR_DEVICE, "device generator: launch failed"));
  std::vector<uint32_t> counts(N);
  uint32_t errBits = 0;
  if (hipMemcpyAsync(counts.data(), dCount, 4 * (size_t)N, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&errBits, dErr, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fin(fail(DPSNN_ERR_DEVICE, "device generator: count pass failed"));
  std::vector<uint64_t> gen_off(N + 1, 0);
  for (uint32_t i = 0; i < N; ++i) gen_off[i + 1] = gen_off[i] + counts[i];
  const uint64_t S = gen_off[N];
  void *dGenOff, *dPost, *dDelay, *dW;
  if (!put(&dGenOff, gen_off.data(), 8 * (size_t)N) || !alloc(&dPost, 4 * S) || !alloc(&dDelay, S) ||
      !alloc(&dW, 2 * S))
    return fin(fail(DPSNN_ERR_MEMORY, "device generator: out of device memory for %llu synapses",
                    (unsigned long long)S));
  G.off = (const uint64_t *)dGenOff; G.post = (uint32_t *)dPost; G.delay = (uint8_t *)dDelay; G.w = (int16_t *)dW;
  hipLaunchKernelGGL(k_generate<1>, grid, dim3(256), 0, st, G);
  if (hipGetLastError() != hipSuccess) return fin(fail(DPSNN_ERR_DEVICE, "device generator: launch failed"));
  std::vector<uint32_t> g_post(S);
  std::vector<uint8_t> g_delay(S);
  std::vector<int16_t> g_w(S);
  if (hipMemcpyAsync(g_post.data(), dPost, 4 * S, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(g_delay.data(), dDelay, S, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(g_w.data(), dW, 2 * S, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&errBits, dErr, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fin(fail(DPSNN_ERR_DEVICE, "device generator: fill pass failed"));
  fin(0);
  if (errBits & 1u) { err = "connection probabilities in (0, 1e-5] are not supported"; return DPSNN_ERR_CONFIG; }
  if (errBits & 2u) { err = "a neuron would get 7500 or more synapses (DSD__maxM)"; return DPSNN_ERR_CONFIG; }
  if (errBits & 4u) { err = "synaptic weight outside [env_minSynWeight_f, -env_minSynWeight_f]"; return DPSNN_ERR_CONFIG; }
  if (errBits & 8u) { err = "a stored delay is not below env_D"; return DPSNN_ERR_CONFIG; }
  if (errBits & 16u) { err = "synaptic weights cannot be drawn within their bounds (zero mean?)"; return DPSNN_ERR_CONFIG; }
  inflate_backward(c, gen_off, g_post, g_delay, g_w, n->opt.host_threads);
  return 0;
}

extern "C" {

const char *dpsnn_last_error(void) { return g_err.c_str(); }

int dpsnn_get_parameters(const char *inidir, dpsnn_parameters *out) {
  if (!inidir || !out) return fail(DPSNN_ERR_ARG, "null argument");
  Params p;
  std::string err;
  if (!read_params(inidir, p, err)) return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  *out = p.pub;
  return 0;
}

int dpsnn_write_ini_files(const char *inidir, const char *outdir, const char *env) {
  if (!inidir || !outdir) return fail(DPSNN_ERR_ARG, "null argument");
  Params p;
  std::string err;
  EnvMap over;
  if (env && !parse_env_overrides(env, over, err)) return fail(DPSNN_ERR_ARG, "%s", err.c_str());
  if (!read_params(inidir, p, err, env ? &over : nullptr)) return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  if (!write_ini_files(p, outdir, err)) return fail(DPSNN_ERR_IO, "%s", err.c_str());
  return 0;
}

int dpsnn_prepare(const char *inidir, const dpsnn_options *opt, dpsnn_net **out) {
  if (!inidir || !out) return fail(DPSNN_ERR_ARG, "null argument");
  *out = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<dpsnn_net> n(new dpsnn_net());
  if (opt) n->opt = *opt;
  if (n->opt.drive != DPSNN_DRIVE_REPLAY && n->opt.drive != DPSNN_DRIVE_PHILOX)
    return fail(DPSNN_ERR_ARG, "unknown drive mode %d", n->opt.drive);
  std::string err;
  EnvMap over;
  if (!parse_env_overrides(n->opt.env, over, err) || !read_params(inidir, n->par, err, &over))
    return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  n->opt.env = nullptr;  // not kept past the call
  if (n->opt.output_dir) {
    n->outDir = n->opt.output_dir;
    n->opt.output_dir = n->outDir.c_str();
  }
  const dpsnn_parameters &q = n->par.pub;
  n->stdp = q.startPlasticity_ms <= q.stopPlasticity_ms && q.startPlasticity_ms < q.totalSimTime_ms;
  if (n->stdp)
    for (uint32_t s = 0; s < q.subPopNumber; ++s)
      if (!(n->par.sp[s].Tarp >= 1.0))
        // a source could then fire in consecutive steps and one synapse be
        // delivered twice in a step (own bucket + spill-over run)
        return fail(DPSNN_ERR_CONFIG, "STDP on device needs refractory periods >= 1 ms");
  // localNet init: seed with env_globalSeed (DPSNN_localNet_init.c:119),
  // generate, then continue the stream from SetRandomSeed(loc_h=0) (:283)
  n->rng.seed((int32_t)q.globalSeed);
  if (n->opt.generator == DPSNN_GENERATOR_DEVICE) {
    const int grc = generate_on_device(n.get(), err);
    if (grc) return grc == DPSNN_ERR_CONFIG ? fail(DPSNN_ERR_CONFIG, "%s", err.c_str()) : grc;
  } else if (n->opt.generator != DPSNN_GENERATOR_HOST) {
    return fail(DPSNN_ERR_ARG, "unknown generator %u", n->opt.generator);
  } else if (!generate_connectome(n->par, n->rng, n->con, err, n->opt.host_threads)) {
    return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  }
  build_delivery(n->par, n->con, /*with_incoming=*/n->stdp, n->opt.host_threads);
  n->rng.seed(0);
  // DPSNN_REPLAY_STREAM=k > 0: the replayed bath continues from
  // SetRandomSeed(k) instead, the stream process loc_h = k would draw --
  // the reference's own drive with another stream, to measure its
  // run-to-run spread; unset or 0 is the reference's stream
  if (const char *e = std::getenv("DPSNN_REPLAY_STREAM"))
    if (const int k = std::atoi(e)) n->rng.seed(k);
  if (n->stdp) {
    // each in-list item names its synapse by its rank in the target's
    // incoming list (backward-list order)
    const Connectome &c = n->con;
    std::vector<uint32_t> rank(c.S);
    for (uint32_t j = 0; j < c.N; ++j) {
      if (c.inc_off[j + 1] - c.inc_off[j] > 65535)
        return fail(DPSNN_ERR_CONFIG, "neuron %u has more than 65535 incoming synapses", j);
      for (uint64_t k = c.inc_off[j]; k < c.inc_off[j + 1]; ++k) rank[c.inc_syn[k]] = (uint32_t)(k - c.inc_off[j]);
    }
    n->inSynLocal.resize(c.in_off[c.N]);
    for (uint64_t it = 0; it < c.in_off[c.N]; ++it) n->inSynLocal[it] = (uint16_t)rank[c.in_syn[it]];
  }
  if (n->con.max_in >= kMaxInList)
    return fail(DPSNN_ERR_CONFIG, "a neuron has %u in-list items; the engine handles up to %u", n->con.max_in,
                kMaxInList - 1);
  // stencil-window phase 1: on for large networks; DPSNN_WINDOW=1 forces
  // it on (tests of small networks), 0 off.  A connectome whose stencil
  // window does not fit the LDS keeps the gather form.
  {
    bool want = n->con.N >= kWindowMinNeurons;
    if (const char *e = std::getenv("DPSNN_WINDOW")) want = std::atoi(e) != 0;
    if (want) {
      // DPSNN_WINDOW_UNITS: units wanted per step instead (tuning; results do
      // not change)
      const char *ue = getenv("DPSNN_WINDOW_UNITS");
      const uint32_t unitsPerCol =
          ue && atoi(ue) > 0 ? std::max<uint32_t>(1, ((uint32_t)atoi(ue) + q.globCFT - 1) / q.globCFT)
                             : std::max<uint32_t>(1, (q.neuronsPerCM + kWindowUnitTargets - 1) / kWindowUnitTargets);
      build_window(n->par, n->con, unitsPerCol, kWindowMaxEntries, n->window, n->opt.host_threads);
    }
  }
  if (n->opt.drive == DPSNN_DRIVE_REPLAY) n->replay.init(n->con, &n->rng);
  n->writing = n->opt.output_dir && n->opt.output_dir[0];
  if (n->opt.drop_host_synapses && n->stdp && n->writing && !q.statSynPeriodicProbe_writeDisable)
    return fail(DPSNN_ERR_CONFIG, "synPeriodicProbe files need the host synapse list (drop_host_synapses is set)");
  if (n->writing && !n->writer.init(n->par, n->opt.output_dir))
    return fail(DPSNN_ERR_IO, "cannot write output files in %s", n->opt.output_dir);
  n->stdpEvents = n->writing && n->stdp && !n->par.stdpEventDisable;
  if (n->stdpEvents) {
    if (n->opt.drop_host_synapses)
      return fail(DPSNN_ERR_CONFIG, "the STDPevent file needs the host synapse list (drop_host_synapses is set)");
    if (!n->stdpWriter.init(n->par, n->con, n->opt.output_dir))
      return fail(DPSNN_ERR_CONFIG, "cannot set up the STDPevent writer");
  }
  n->initSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = n.release();
  return 0;
}

uint64_t dpsnn_device_bytes_needed(const dpsnn_net *n) {
  if (!n) return 0;
  if (n->uploaded) return n->devBytes;
  const Connectome &c = n->con;
  const dpsnn_parameters &q = n->par.pub;
  const uint64_t N = c.N, NW = (N + 63) / 64, items = c.in_off.empty() ? 0 : c.in_off[N];
  uint64_t RS = 1;
  while (RS < c.max_delay + 2) RS <<= 1;
  auto r = [](uint64_t b) { return (std::max<uint64_t>(b, 1) + 255) / 256 * 256; };
  uint64_t b = r(8 * q.subPopNumber * kParams) + r(8 * c.bathLut.size()) + r(8 * N) + r(N) + r(4 * N);
  b += 5 * r(8 * N) + r(8 * N);                                // state, Philox counters
  b += r(8 * RS * NW) + r(8 * RS * N);                         // spike history
  b += r(8 * (N + 1)) + r(4 * (items + kWave)) + r(2 * items);  // in-lists
  if (n->stdp) {
    const uint64_t S = c.S;
    b += r(8 * (N + 1)) + r(2 * S) + r(S) + r(2 * items) + 2 * r(4 * S) + r(4 * N);
    if (!q.statSynPeriodicProbe_writeDisable) b += r(4 * S);
    b += 2 * r(4 * kStdpSpan) + 2 * r(4 * q.subPopNumber * q.subPopNumber);
  }
  const uint64_t drainEvery = n->opt.drain_every_ms ? n->opt.drain_every_ms : 100;
  b += r(8 * kCounterSlots * kCounterStride) + r(4 * kXcds * kCtrStride) + r(4 * (uint64_t)q.globCFT * q.subPopNumber * drainEvery);
  const bool raster = (n->writing && !q.spikes_writeDisable) || n->opt.keep_spikes;
  if (raster) {
    const uint64_t cap = N * std::min<uint64_t>(drainEvery, 1000) + 1;
    b += r(4 * cap) + r(8 * cap) + r(4) + r(4 * drainEvery);
  }
  if (n->opt.drive == DPSNN_DRIVE_REPLAY) b += kExtRing * r(4 * (N + 1));
  const WindowLayout &L = n->window;
  if (L.ok)
    b += r(4 * L.item.size()) + r(2 * L.desc.size()) + r(8 * L.chunk.size()) + r(4 * L.cols.size()) +
         r(4 * kXcds * kCtrStride) + r(4 * N) + r(4);
  return b;
}

uint64_t dpsnn_host_bytes(const dpsnn_net *n) {
  if (!n) return 0;
  auto v = [](const auto &x) { return uint64_t(x.capacity()) * sizeof(x[0]); };
  const Connectome &c = n->con;
  const WindowLayout &L = n->window;
  return v(c.pre) + v(c.post) + v(c.delay) + v(c.kind) + v(c.weight) + v(c.in_off) + v(c.in_pre_d) + v(c.in_w) +
         v(c.in_syn) + v(c.inc_off) + v(c.inc_syn) + v(n->inSynLocal) + v(L.cols) + v(L.item) + v(L.desc) +
         v(L.chunk);
}

int dpsnn_upload(dpsnn_net *n) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (n->uploaded) return 0;
  auto t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipSetDevice(n->opt.device));
  HIP_TRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  n->ownStream = n->stream;
  int rc = build_device(n);
  if (rc) {
    free_all(n);
    return rc;
  }
  n->uploaded = true;
  // the device holds the in-lists; the host keeps the backward list for
  // dumps (unless asked not to) and drops the delivery staging
  std::vector<uint32_t>().swap(n->con.in_pre_d);
  std::vector<int16_t>().swap(n->con.in_w);
  std::vector<uint32_t>().swap(n->con.in_syn);
  std::vector<uint16_t>().swap(n->inSynLocal);
  n->window = WindowLayout{};
  if (n->opt.drop_host_synapses) {
    std::vector<uint32_t>().swap(n->con.pre);
    std::vector<uint32_t>().swap(n->con.post);
    std::vector<uint8_t>().swap(n->con.delay);
    std::vector<uint8_t>().swap(n->con.kind);
    std::vector<int16_t>().swap(n->con.weight);
    std::vector<uint64_t>().swap(n->con.inc_off);
    std::vector<uint32_t>().swap(n->con.inc_syn);
  }
  n->initSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

int dpsnn_create(const char *inidir, const dpsnn_options *opt, dpsnn_net **out) {
  if (!out) return fail(DPSNN_ERR_ARG, "null argument");
  dpsnn_net *n = nullptr;
  int rc = dpsnn_prepare(inidir, opt, &n);
  if (rc) return rc;
  if ((rc = dpsnn_upload(n))) {
    std::string keep = g_err;
    dpsnn_destroy(n);
    g_err = keep;
    return rc;
  }
  *out = n;
  return 0;
}

int dpsnn_step(dpsnn_net *n, uint32_t nsteps) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (n->batch) return fail(DPSNN_ERR_ARG, "network is a member of a batch: step it through dpsnn_batch_step");
  if (!n->uploaded) return fail(DPSNN_ERR_ARG, "network not uploaded (dpsnn_upload)");
  HIP_TRY(hipSetDevice(n->opt.device));
  int rc0 = ensure_scratch(n);
  if (rc0) return rc0;
  return run_steps(n, n->dev, nsteps, true);
}

int dpsnn_step_noalloc(dpsnn_net *n, uint32_t nsteps) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (n->batch) return fail(DPSNN_ERR_ARG, "network is a member of a batch: step it through dpsnn_batch_step");
  if (!n->uploaded) return fail(DPSNN_ERR_ARG, "network not uploaded (dpsnn_upload)");
  if (n->dev.drive == DPSNN_DRIVE_REPLAY)
    // the replayed bath's staging buffers grow with the step's event count
    return fail(DPSNN_ERR_ARG, "the allocation-free step runs the Philox drive only");
  HIP_TRY(hipSetDevice(n->opt.device));
  DevNet P;
  int rc = static_scratch_view(n, P);
  if (rc) return rc;
  // graph instantiation allocates: plain launches only
  return run_steps(n, P, nsteps, false);
}

int dpsnn_sync(dpsnn_net *n) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  HIP_TRY(hipStreamSynchronize(n->stream));
  return drain(n);
}

uint32_t dpsnn_now_ms(const dpsnn_net *n) { return n ? n->t : 0; }
uint32_t dpsnn_total_ms(const dpsnn_net *n) { return n ? n->par.pub.totalSimTime_ms : 0; }

int dpsnn_get_state(dpsnn_net *n, double *v, double *c, double *Tr, double *Te, uint8_t *fired) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  const size_t N = n->dev.N;
  HIP_TRY(hipStreamSynchronize(n->stream));
  if (v) HIP_TRY(hipMemcpy(v, n->dev.v, N * 8, hipMemcpyDeviceToHost));
  if (c) HIP_TRY(hipMemcpy(c, n->dev.c, N * 8, hipMemcpyDeviceToHost));
  if (Tr) HIP_TRY(hipMemcpy(Tr, n->dev.Tr, N * 8, hipMemcpyDeviceToHost));
  if (Te) HIP_TRY(hipMemcpy(Te, n->dev.Te, N * 8, hipMemcpyDeviceToHost));
  if (fired) {
    // neurons that fired in the last completed step: bitmap of slot t
    std::vector<uint64_t> w(n->dev.NW);
    const size_t slot = n->t & (n->dev.RS - 1);
    HIP_TRY(hipMemcpy(w.data(), n->dev.bits + slot * n->dev.NW, 8 * w.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) fired[i] = (uint8_t)((w[i >> 6] >> (i & 63)) & 1u);
  }
  return 0;
}

int dpsnn_last_spikes(dpsnn_net *n, uint32_t *ids, double *te, uint32_t *count) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (!n->opt.keep_spikes) return fail(DPSNN_ERR_ARG, "network created without keep_spikes");
  int rc = dpsnn_sync(n);
  if (rc) return rc;
  const uint32_t cnt = (uint32_t)n->lastIds.size();
  if (count) *count = cnt;
  if (ids) std::memcpy(ids, n->lastIds.data(), sizeof(uint32_t) * cnt);
  if (te) std::memcpy(te, n->lastTe.data(), sizeof(double) * cnt);
  return 0;
}

uint64_t dpsnn_num_synapses(const dpsnn_net *n) { return n ? n->con.S : 0; }

int dpsnn_get_synapses(dpsnn_net *n, uint32_t *pre, uint32_t *post, uint8_t *delay, int16_t *weight,
                       float *tderiv, int32_t *last_act) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (n->opt.drop_host_synapses) return fail(DPSNN_ERR_ARG, "the host synapse list was dropped (drop_host_synapses)");
  const Connectome &c = n->con;
  if (pre) std::memcpy(pre, c.pre.data(), 4 * c.S);
  if (post) std::memcpy(post, c.post.data(), 4 * c.S);
  if (delay) std::memcpy(delay, c.delay.data(), c.S);
  if (n->stdp) return synapse_state(n, weight, tderiv, last_act);
  if (weight) std::memcpy(weight, c.weight.data(), 2 * c.S);
  if (tderiv) std::memset(tderiv, 0, 4 * c.S);
  if (last_act) std::memset(last_act, 0, 4 * c.S);
  return 0;
}

int dpsnn_write_synapse_dump(dpsnn_net *n, const char *path);
int dpsnn_get_stats(dpsnn_net *n, dpsnn_stats *out) {
  if (!n || !out) return fail(DPSNN_ERR_ARG, "null argument");
  unsigned long long slots[kCounterSlots * kCounterStride], cnt[4] = {0, 0, 0, 0};
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(slots, n->dev.counters, sizeof slots, hipMemcpyDeviceToHost));
  for (uint32_t k = 0; k < kCounterSlots; ++k)
    for (int i = 0; i < 4; ++i) cnt[i] += slots[k * kCounterStride + i];
  out->steps = n->t;
  out->synaptic_events = cnt[0];
  out->external_events = cnt[1];
  out->spikes = cnt[2];
  out->synapses = n->con.S;
  out->delivery_items = n->items;
  out->neurons = n->con.N;
  out->max_in_items = n->con.max_in;
  out->init_seconds = n->initSeconds;
  out->device_bytes = n->devBytes;
  out->window_form = n->winOn ? 1u : 0u;
  out->reserved = 0;
  out->graph_steps = n->graphSteps;
  return 0;
}

void *dpsnn_stream(dpsnn_net *n) { return n ? (void *)n->stream : nullptr; }

int dpsnn_profile(dpsnn_net *n, int enable) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  HIP_TRY(hipStreamSynchronize(n->stream));
  n->profile = enable != 0;
  n->profUsed = 0;
  return 0;
}

int dpsnn_kernel_time(dpsnn_net *n, int kernel, double *avg_ms, uint64_t *launches) {
  if (!n) return fail(DPSNN_ERR_ARG, "null network");
  if (kernel != 0) return fail(DPSNN_ERR_ARG, "kernel 0 (k_step) is the only timed kernel");
  HIP_TRY(hipStreamSynchronize(n->stream));
  double total = 0;
  uint64_t steps = 0;
  for (size_t k = 0; k < n->profUsed; ++k) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, n->profEv[k].a, n->profEv[k].b));
    total += ms;
    steps += n->profEv[k].steps;
  }
  if (avg_ms) *avg_ms = steps ? total / (double)steps : 0.0;
  if (launches) *launches = steps;
  n->profUsed = 0;
  return 0;
}

void dpsnn_destroy(dpsnn_net *n) {
  if (!n) return;
  if (n->batch) dpsnn_batch_remove(n->batch, n);
  drain(n);
  n->writer.close();
  n->stdpWriter.close();
  free_all(n);
  delete n;
}

}  // extern "C"

// ============================================================================
// Checkpoint / restore of a network between steps: the device-resident
// state (neurons, bath cursors, Philox counters, spike history, counters,
// STDP synapse state) and the host stream state (ran3 with its NormDev
// cache, the replayed bath's next times, the open output files and the
// counts of the open rates window).  The connectome is not stored: restore
// rebuilds it from the same ini files (the generator is deterministic) and
// checks the sizes recorded in the file.
// ============================================================================
namespace {
constexpr char kCkptMagic[8] = {'D', 'P', 'S', 'N', 'N', 'C', 'K', '1'};
struct CkptHeader {
  char magic[8];
  uint32_t N, RS, NW, drive, stdp, t, nsub, ncols, writer, spikesOpen, ratesOpen, pad;
  uint64_t S, items;
};

struct CkptFile {
  FILE *f = nullptr;
  ~CkptFile() {
    if (f) std::fclose(f);
  }
};

template <class T>
int ck_write(FILE *f, const T *h, size_t n) {
  if (n && std::fwrite(h, sizeof(T), n, f) != n) return fail(DPSNN_ERR_IO, "checkpoint write failed");
  return 0;
}
template <class T>
int ck_read(FILE *f, T *h, size_t n) {
  if (n && std::fread(h, sizeof(T), n, f) != n) return fail(DPSNN_ERR_IO, "checkpoint file truncated");
  return 0;
}
template <class T>
int ck_dev_write(FILE *f, const T *d, size_t n) {
  std::vector<T> h(n);
  HIP_TRY(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return ck_write(f, h.data(), n);
}
template <class T>
int ck_dev_read(FILE *f, T *d, size_t n) {
  std::vector<T> h(n);
  int rc = ck_read(f, h.data(), n);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(d, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// The device arrays of the state, in file order (the STDP ones last).
template <class Fn>
int ck_device_arrays(dpsnn_net *n, Fn &&fn) {
  DevNet &P = n->dev;
  const size_t N = P.N, S = n->con.S;
  int rc;
  if ((rc = fn(P.v, N)) || (rc = fn(P.c, N)) || (rc = fn(P.Tr, N)) || (rc = fn(P.Te, N)) ||
      (rc = fn(P.assigned, N)) || (rc = fn(P.philoxCtr, N)) || (rc = fn(P.bits, size_t(P.RS) * P.NW)) ||
      (rc = fn(P.teHist, size_t(P.RS) * N)) || (rc = fn(P.counters, size_t(kCounterSlots) * kCounterStride)))
    return rc;
  if (n->stdp && ((rc = fn(P.lastSpike, N)) || (rc = fn(P.wT, S)) || (rc = fn(P.tdT, S)) || (rc = fn(P.lastT, S))))
    return rc;
  return 0;
}
}  // namespace

extern "C" {

int dpsnn_checkpoint(dpsnn_net *n, const char *path) {
  if (!n || !path) return fail(DPSNN_ERR_ARG, "null argument");
  if (!n->uploaded) return fail(DPSNN_ERR_ARG, "network not uploaded (dpsnn_upload)");
  if (n->batch) return fail(DPSNN_ERR_ARG, "network is a member of a batch: remove it first");
  if (n->stdpEvents) return fail(DPSNN_ERR_ARG, "a network writing the STDPevent file cannot be checkpointed");
  HIP_TRY(hipSetDevice(n->opt.device));
  int rc = dpsnn_sync(n);  // every step's files are written up to now
  if (rc) return rc;
  n->writer.flush();
  CkptFile cf;
  if (!(cf.f = std::fopen(path, "wb"))) return fail(DPSNN_ERR_IO, "cannot open %s", path);
  const StatWriter::State ws = n->writer.state();
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, sizeof h.magic);
  h.N = n->dev.N; h.RS = n->dev.RS; h.NW = n->dev.NW; h.drive = n->dev.drive; h.stdp = n->stdp;
  h.t = n->t; h.nsub = n->dev.nsub; h.ncols = n->dev.ncols; h.writer = n->writing;
  h.spikesOpen = ws.spikesOpen; h.ratesOpen = ws.ratesOpen;
  h.S = n->con.S; h.items = n->items;
  if ((rc = ck_write(cf.f, &h, 1))) return rc;
  if ((rc = ck_device_arrays(n, [&](auto *d, size_t cnt) { return ck_dev_write(cf.f, d, cnt); }))) return rc;
  const Ran3::State rs = n->rng.state();
  if ((rc = ck_write(cf.f, &rs, 1))) return rc;
  if (n->dev.drive == DPSNN_DRIVE_REPLAY && (rc = ck_write(cf.f, n->replay.next_times().data(), h.N))) return rc;
  if (n->writing && (rc = ck_write(cf.f, ws.perPop.data(), ws.perPop.size()))) return rc;
  if (std::fflush(cf.f) != 0) return fail(DPSNN_ERR_IO, "checkpoint write failed");
  return 0;
}

int dpsnn_restore(const char *inidir, const dpsnn_options *opt, const char *path, dpsnn_net **out) {
  if (!inidir || !path || !out) return fail(DPSNN_ERR_ARG, "null argument");
  *out = nullptr;
  CkptFile cf;
  if (!(cf.f = std::fopen(path, "rb"))) return fail(DPSNN_ERR_IO, "cannot open %s", path);
  CkptHeader h{};
  int rc = ck_read(cf.f, &h, 1);
  if (rc) return rc;
  if (std::memcmp(h.magic, kCkptMagic, sizeof h.magic) != 0)
    return fail(DPSNN_ERR_IO, "%s is not a network checkpoint", path);
  dpsnn_net *n = nullptr;
  if ((rc = dpsnn_create(inidir, opt, &n))) return rc;
  std::unique_ptr<dpsnn_net, void (*)(dpsnn_net *)> guard(n, dpsnn_destroy);
  if (h.N != n->dev.N || h.S != n->con.S || h.items != n->items || h.RS != n->dev.RS || h.NW != n->dev.NW ||
      h.nsub != n->dev.nsub || h.ncols != n->dev.ncols || h.stdp != (uint32_t)n->stdp)
    return fail(DPSNN_ERR_CONFIG, "the checkpoint was taken on another network (sizes differ)");
  if (h.drive != n->dev.drive) return fail(DPSNN_ERR_CONFIG, "the checkpoint was taken with another drive");
  if (h.t > n->par.pub.totalSimTime_ms)
    return fail(DPSNN_ERR_CONFIG, "the checkpoint is at %u ms, past env_totalSimTime_ms", h.t);
  HIP_TRY(hipSetDevice(n->opt.device));
  HIP_TRY(hipStreamSynchronize(n->stream));
  if ((rc = ck_device_arrays(n, [&](auto *d, size_t cnt) { return ck_dev_read(cf.f, d, cnt); }))) return rc;
  Ran3::State rs{};
  if ((rc = ck_read(cf.f, &rs, 1))) return rc;
  n->rng.set_state(rs);
  if (h.drive == DPSNN_DRIVE_REPLAY && (rc = ck_read(cf.f, n->replay.next_times().data(), h.N))) return rc;
  if (h.writer) {
    StatWriter::State ws;
    ws.perPop.resize(size_t(h.ncols) * h.nsub);
    ws.spikesOpen = h.spikesOpen;
    ws.ratesOpen = h.ratesOpen;
    if ((rc = ck_read(cf.f, ws.perPop.data(), ws.perPop.size()))) return rc;
    if (n->writing && !n->writer.resume(ws, h.t / 1000))
      return fail(DPSNN_ERR_IO, "cannot reopen the output files in %s", n->opt.output_dir);
  }
  n->t = h.t;
  *out = guard.release();
  return 0;
}

}  // extern "C"

// ============================================================================
// Batching: many independent networks advance in one persistent launch per
// step (k_batch_step).  Members keep their own state, history, in-lists,
// counters and output files; the batch owns the stream, the per-step item
// ring and the key scratch.
// ============================================================================
struct dpsnn_batch {
  int device = 0;
  int cus = 1;
  hipStream_t stream = nullptr;
  std::vector<dpsnn_net *> nets;
  bool dirty = true;
  DevNet *d_nets = nullptr;
  size_t netsCap = 0;
  static constexpr int kRing = 4;
  BatchItem *h_items[kRing] = {};
  BatchItem *d_items[kRing] = {};
  hipEvent_t done[kRing] = {};
  bool pending[kRing] = {};
  size_t itemsCap = 0;
  int ringNext = 0;
  uint64_t *keys = nullptr;
  int16_t *keyW = nullptr;
  uint16_t *keyPos = nullptr;
  uint32_t *keyCount = nullptr;
  uint32_t *workCtr = nullptr;  // [kXcds], zero between launches
  size_t scratchSlots = 0;
  uint32_t scratchCap = 0;
  uint32_t totalTiles = 0;   // of all members
  bool anyPlastic = false;   // some member has STDP (the split form then runs k_batch_ltp)
  uint32_t chunkTiles = 0;   // split form: tiles per chunk (scratch bound)
  int shape = -1;
  uint32_t wavesPerCu = 0;
  bool split = true;         // k_batch_gather + k_batch_integrate per chunk; else k_batch_step
  bool useGraph = true;      // replay each step's launches from a captured HIP graph
  struct Graph {
    hipGraphExec_t exec = nullptr;
    uint32_t count = 0, tiles = 0, chunkTiles = 0, wpc = 0;
    int shape = -1;
    bool split = false, packed = false;
  } graphs[kRing];
  bool profile = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t evUsed = 0;
  std::vector<StepPlan> plans;
  std::vector<dpsnn_net *> active;
  std::vector<uint32_t> windowed;  // this step's members (indices into active) in the window form
};

// Scratch the split form may use for one chunk of tiles.
constexpr size_t kBatchScratchBytes = size_t(5) << 30;  // also holds a 16 x 16-column member in the window form

namespace {

uint32_t batch_cap(const dpsnn_batch *b) {
  uint32_t cap = 1;
  for (auto *n : b->nets) cap = std::max(cap, n->dev.cap);
  return cap;
}

uint32_t lds_waves(size_t lds) {
  return std::max<uint32_t>(1, std::min<uint32_t>(16, (uint32_t)((160u << 10) / lds)));
}

void batch_free_items(dpsnn_batch *b) {
  for (int r = 0; r < dpsnn_batch::kRing; ++r) {
    if (b->h_items[r]) hipHostFree(b->h_items[r]);
    if (b->d_items[r]) hipFree(b->d_items[r]);
    b->h_items[r] = nullptr;
    b->d_items[r] = nullptr;
  }
  b->itemsCap = 0;
}

// Device copy of the members' DevNet structs, scratch for the longest
// in-list at the largest grid any shape uses, item ring for every member.
int batch_prepare(dpsnn_batch *b) {
  if (!b->dirty) return 0;
  HIP_TRY(hipStreamSynchronize(b->stream));
  // captured graphs name the buffers below, which may move
  for (auto &g : b->graphs) {
    if (g.exec) HIP_TRY(hipGraphExecDestroy(g.exec));
    g = dpsnn_batch::Graph{};
  }
  const size_t count = b->nets.size();
  if (count > b->netsCap) {
    if (b->d_nets) HIP_TRY(hipFree(b->d_nets));
    b->d_nets = nullptr;
    HIP_TRY(hipMalloc((void **)&b->d_nets, sizeof(DevNet) * count));
    b->netsCap = count;
  }
  std::vector<DevNet> h(count);
  for (size_t i = 0; i < count; ++i) h[i] = b->nets[i]->dev;
  if (count) HIP_TRY(hipMemcpy(b->d_nets, h.data(), sizeof(DevNet) * count, hipMemcpyHostToDevice));
  if (count > b->itemsCap) {
    batch_free_items(b);
    const size_t cap = count + 8;
    for (int r = 0; r < dpsnn_batch::kRing; ++r) {
      HIP_TRY(hipHostMalloc((void **)&b->h_items[r], sizeof(BatchItem) * cap, 0));
      HIP_TRY(hipMalloc((void **)&b->d_items[r], sizeof(BatchItem) * cap));
      b->pending[r] = false;
    }
    b->itemsCap = cap;
  }
  const uint32_t cap = batch_cap(b);
  b->totalTiles = 0;
  b->anyPlastic = false;
  for (auto *n : b->nets) {
    b->totalTiles += n->dev.NW;
    b->anyPlastic = b->anyPlastic || n->stdp;
  }
  const size_t perTile = (size_t)cap * kWave * (sizeof(uint64_t) + 2 * sizeof(int16_t));
  b->chunkTiles = (uint32_t)std::max<size_t>(1, std::min<size_t>(b->totalTiles, kBatchScratchBytes / perTile));
  const size_t fusedSlots = (size_t)round_up((uint32_t)b->cus * lds_waves(lds_bytes(cap, 64)), kXcds) * kWave;
  const size_t slots = std::max(fusedSlots, (size_t)b->chunkTiles * kWave);
  if (cap > b->scratchCap || slots > b->scratchSlots) {
    if (b->keys) HIP_TRY(hipFree(b->keys));
    if (b->keyW) HIP_TRY(hipFree(b->keyW));
    if (b->keyPos) HIP_TRY(hipFree(b->keyPos));
    if (b->keyCount) HIP_TRY(hipFree(b->keyCount));
    b->keys = nullptr;
    b->keyW = nullptr;
    b->keyPos = nullptr;
    b->keyCount = nullptr;
    hipError_t e = hipMalloc((void **)&b->keys, sizeof(uint64_t) * slots * cap);
    if (e != hipSuccess) return fail(DPSNN_ERR_MEMORY, "batch key scratch: %s", hipGetErrorString(e));
    e = hipMalloc((void **)&b->keyW, sizeof(int16_t) * slots * cap);
    if (e != hipSuccess) return fail(DPSNN_ERR_MEMORY, "batch weight scratch: %s", hipGetErrorString(e));
    e = hipMalloc((void **)&b->keyPos, sizeof(uint16_t) * slots * cap);
    if (e != hipSuccess) return fail(DPSNN_ERR_MEMORY, "batch position scratch: %s", hipGetErrorString(e));
    HIP_TRY(hipMalloc((void **)&b->keyCount, sizeof(uint32_t) * slots));
    b->scratchCap = cap;
    b->scratchSlots = slots;
  }
  if (b->shape < 0) {
    const int K = cap <= 8 * 64 ? 8 : (cap <= 16 * 64 ? 16 : 24);
    b->shape = find_shape(K, 256);
  }
  b->dirty = false;
  return 0;
}

uint32_t batch_waves(const dpsnn_batch *b, size_t lds) {
  uint32_t wpc = lds_waves(lds);
  if (b->wavesPerCu) wpc = std::min(wpc, b->wavesPerCu);
  return wpc;
}

// The device work of one batched step: the items of ring slot r, then either
// the persistent fused kernel or, chunk by chunk, phase 1 and phase 2
// (packed: every member is at t >= kPackT, so phase 2 runs its packed-only
// instance).
int batch_launch(dpsnn_batch *b, int r, uint32_t count, uint32_t tiles, bool packed) {
  HIP_TRY(hipMemcpyAsync(b->d_items[r], b->h_items[r], sizeof(BatchItem) * count, hipMemcpyHostToDevice, b->stream));
  const KernelShape &K = kShapes[b->shape];
  const uint32_t cap = b->scratchCap;
  if (!b->split) {
    const size_t lds = lds_bytes(cap, K.B);
    const uint32_t grid = round_up(std::min<uint32_t>(tiles, (uint32_t)b->cus * batch_waves(b, lds)), kXcds);
    hipLaunchKernelGGL(K.batch, dim3(grid), dim3(kWave), lds, b->stream, b->d_nets, b->d_items[r], count, tiles, cap,
                       b->keys, b->keyW, b->keyPos, b->keyCount);
  } else {
    const size_t lds = phase1_lds(cap, K.B);
    for (uint32_t g0 = 0; g0 < tiles; g0 += b->chunkTiles) {
      const uint32_t g1 = std::min(tiles, g0 + b->chunkTiles);
      const uint32_t grid = round_up(
          std::min<uint32_t>((g1 - g0) * kWave,
                             (uint32_t)b->cus * std::min(batch_waves(b, lds), resident_waves((const void *)K.batchGather, lds))),
          kXcds);
      hipLaunchKernelGGL(K.batchGather, dim3(grid), dim3(kWave), lds, b->stream, b->d_nets, b->d_items[r], count, g0,
                         g1, cap, b->keys, b->keyW, b->keyPos, b->keyCount, b->workCtr);
      const uint32_t blocks = (g1 - g0 + kBlock / kWave - 1) / (kBlock / kWave);
      if (packed)
        hipLaunchKernelGGL(k_batch_integrate<1>, dim3(blocks), dim3(kBlock), 0, b->stream, b->d_nets, b->d_items[r],
                           count, g0, g1, cap, b->keys, b->keyW, b->keyCount, b->workCtr);
      else
        hipLaunchKernelGGL(k_batch_integrate<2>, dim3(blocks), dim3(kBlock), 0, b->stream, b->d_nets, b->d_items[r],
                           count, g0, g1, cap, b->keys, b->keyW, b->keyCount, b->workCtr);
      if (b->anyPlastic)
        hipLaunchKernelGGL(k_batch_ltp, dim3(std::min<uint32_t>((g1 - g0) * (kWave * kWave / kWaveLTP), b->cus * 16u)),
                           dim3(kWaveLTP), 0, b->stream, b->d_nets, b->d_items[r], count, g0, g1);
    }
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

void batch_drop_graphs(dpsnn_batch *b) {
  for (auto &g : b->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = dpsnn_batch::Graph{};
  }
}

// The same work replayed from a graph captured once per ring slot and launch
// configuration (the per-step arguments live in the slot's items).
int batch_graph_launch(dpsnn_batch *b, int r, uint32_t count, uint32_t tiles, bool packed) {
  dpsnn_batch::Graph &g = b->graphs[r];
  const uint32_t wpc = b->wavesPerCu;
  if (!g.exec || g.count != count || g.tiles != tiles || g.chunkTiles != b->chunkTiles || g.wpc != wpc ||
      g.shape != b->shape || g.split != b->split || g.packed != packed) {
    if (g.exec) HIP_TRY(hipGraphExecDestroy(g.exec));
    g = dpsnn_batch::Graph{};
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_launch(b, r, count, tiles, packed);
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(DPSNN_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      g.exec = nullptr;
      return fail(DPSNN_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
    g.count = count;
    g.tiles = tiles;
    g.chunkTiles = b->chunkTiles;
    g.wpc = wpc;
    g.shape = b->shape;
    g.split = b->split;
    g.packed = packed;
  }
  HIP_TRY(hipGraphLaunch(g.exec, b->stream));
  return 0;
}

// One batched step of every member that has steps left; *any = false when
// none has.
int batch_one_step(dpsnn_batch *b, bool *any) {
  b->active.clear();
  for (auto *n : b->nets)
    if (n->t < n->par.pub.totalSimTime_ms) b->active.push_back(n);
  *any = !b->active.empty();
  if (!*any) return 0;
  int rc;
  if ((rc = batch_prepare(b))) return rc;
  const int r = b->ringNext;
  b->ringNext = (r + 1) % dpsnn_batch::kRing;
  if (b->pending[r]) HIP_TRY(hipEventSynchronize(b->done[r]));
  b->plans.resize(b->active.size());
  b->windowed.clear();
  uint32_t tiles = 0, count = 0;
  bool packed = true;
  for (size_t i = 0; i < b->active.size(); ++i) {
    dpsnn_net *n = b->active[i];
    StepPlan &sp = b->plans[i];
    if ((rc = pre_step(n, sp))) return rc;
    // a member whose step runs in the window form, and whose tiles fit the
    // batch's key scratch, steps on its own launches (below)
    if (window_step(n, sp) && (size_t)n->dev.NW * kWave <= b->scratchSlots) {
      b->windowed.push_back((uint32_t)i);
      continue;
    }
    packed = packed && sp.t >= kPackT;
    const uint32_t idx = (uint32_t)(std::find(b->nets.begin(), b->nets.end(), n) - b->nets.begin());
    b->h_items[r][count++] = BatchItem{sp.t, sp.ltd, sp.S, sp.X, tiles, idx};
    tiles += n->dev.NW;
  }
  if (b->profile) {
    if (b->evUsed == b->ev.size()) {
      hipEvent_t x, y;
      HIP_TRY(hipEventCreate(&x));
      HIP_TRY(hipEventCreate(&y));
      b->ev.emplace_back(x, y);
    }
    HIP_TRY(hipEventRecord(b->ev[b->evUsed].first, b->stream));
  }
  if (count && b->useGraph) {
    if ((rc = batch_graph_launch(b, r, count, tiles, packed))) return rc;
  } else if (count) {
    if ((rc = batch_launch(b, r, count, tiles, packed))) return rc;
  }
  // window-form members, one after another on the batch's stream and key
  // scratch (its cap is the largest member's, so their tiles fit it)
  for (uint32_t i : b->windowed) {
    const dpsnn_net *n = b->active[i];
    DevNet P = n->dev;
    P.keys = b->keys;
    P.keyW = b->keyW;
    P.keyCount = b->keyCount;
    P.cap = b->scratchCap;
    P.tDev = nullptr;
    launch_window_step(n, P, b->plans[i], b->stream);
  }
  HIP_TRY(hipGetLastError());
  if (b->profile) HIP_TRY(hipEventRecord(b->ev[b->evUsed++].second, b->stream));
  HIP_TRY(hipEventRecord(b->done[r], b->stream));
  b->pending[r] = true;
  for (size_t i = 0; i < b->active.size(); ++i)
    if ((rc = post_step(b->active[i], b->plans[i]))) return rc;
  return 0;
}

int batch_kernel_ms(dpsnn_batch *b, double *avg, uint64_t *launches) {
  HIP_TRY(hipStreamSynchronize(b->stream));
  double total = 0;
  for (size_t k = 0; k < b->evUsed; ++k) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[k].first, b->ev[k].second));
    total += ms;
  }
  if (avg) *avg = b->evUsed ? total / (double)b->evUsed : 0.0;
  if (launches) *launches = b->evUsed;
  b->evUsed = 0;
  return 0;
}

}  // namespace

extern "C" {

int dpsnn_batch_create(int device, dpsnn_batch **out) {
  if (!out) return fail(DPSNN_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<dpsnn_batch> b(new dpsnn_batch());
  b->device = device;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
  b->cus = std::max(b->cus, 1);
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  for (int r = 0; r < dpsnn_batch::kRing; ++r) HIP_TRY(hipEventCreateWithFlags(&b->done[r], hipEventDisableTiming));
  HIP_TRY(hipMalloc((void **)&b->workCtr, sizeof(uint32_t) * kXcds * kCtrStride));
  HIP_TRY(hipMemset(b->workCtr, 0, sizeof(uint32_t) * kXcds * kCtrStride));
  if (const char *e = std::getenv("DPSNN_GRAPH")) b->useGraph = std::atoi(e) != 0;
  if (const char *e = std::getenv("DPSNN_SPLIT")) b->split = std::atoi(e) != 0;
  *out = b.release();
  return 0;
}

int dpsnn_batch_add(dpsnn_batch *b, dpsnn_net *n) {
  if (!b || !n) return fail(DPSNN_ERR_ARG, "null argument");
  if (n->batch) return fail(DPSNN_ERR_ARG, "network is already a member of a batch");
  if (n->opt.device != b->device) return fail(DPSNN_ERR_ARG, "network and batch are on different devices");
  if (!n->uploaded) return fail(DPSNN_ERR_ARG, "network not uploaded (dpsnn_upload)");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(n->stream));
  n->ownStream = n->stream;
  n->stream = b->stream;
  n->batch = b;
  b->nets.push_back(n);
  b->dirty = true;
  return 0;
}

int dpsnn_batch_remove(dpsnn_batch *b, dpsnn_net *n) {
  if (!b || !n) return fail(DPSNN_ERR_ARG, "null argument");
  auto it = std::find(b->nets.begin(), b->nets.end(), n);
  if (it == b->nets.end() || n->batch != b) return fail(DPSNN_ERR_ARG, "network is not a member of this batch");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->nets.erase(it);
  n->stream = n->ownStream;
  n->ownStream = nullptr;
  n->batch = nullptr;
  b->dirty = true;
  return 0;
}

uint32_t dpsnn_batch_size(const dpsnn_batch *b) { return b ? (uint32_t)b->nets.size() : 0; }

int dpsnn_batch_step(dpsnn_batch *b, uint32_t nsteps) {
  if (!b) return fail(DPSNN_ERR_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  for (uint32_t k = 0; k < nsteps; ++k) {
    bool any = false;
    int rc = batch_one_step(b, &any);
    if (rc) return rc;
    if (!any) {
      for (auto *n : b->nets)
        if ((rc = drain(n))) return rc;
      return DPSNN_DONE;
    }
  }
  return 0;
}

int dpsnn_batch_sync(dpsnn_batch *b) {
  if (!b) return fail(DPSNN_ERR_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->stream));
  int rc;
  for (auto *n : b->nets)
    if ((rc = drain(n))) return rc;
  return 0;
}

int dpsnn_batch_get_shape(const dpsnn_batch *b, dpsnn_batch_shape *s) {
  if (!b || !s) return fail(DPSNN_ERR_ARG, "null argument");
  const int i = b->shape < 0 ? find_shape(24, 256) : b->shape;
  s->chunks = kShapes[i].K;
  s->buckets = kShapes[i].B;
  s->waves_per_cu = (int)b->wavesPerCu;
  s->ms_per_step = 0;
  s->split = b->split ? 1 : 0;
  return 0;
}

int dpsnn_batch_set_shape(dpsnn_batch *b, const dpsnn_batch_shape *s) {
  if (!b || !s) return fail(DPSNN_ERR_ARG, "null argument");
  const int i = find_shape(s->chunks, s->buckets);
  if (i < 0) return fail(DPSNN_ERR_ARG, "no kernel shape with %d register chunks and %d buckets", s->chunks, s->buckets);
  if (s->waves_per_cu < 0 || s->waves_per_cu > 16) return fail(DPSNN_ERR_ARG, "waves_per_cu must be in 0..16");
  b->shape = i;
  b->wavesPerCu = (uint32_t)s->waves_per_cu;
  b->split = s->split != 0;
  return 0;
}

int dpsnn_batch_profile(dpsnn_batch *b, int enable) {
  if (!b) return fail(DPSNN_ERR_ARG, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->profile = enable != 0;
  b->evUsed = 0;
  return 0;
}

int dpsnn_batch_kernel_time(dpsnn_batch *b, double *avg_ms, uint64_t *launches) {
  if (!b) return fail(DPSNN_ERR_ARG, "null batch");
  return batch_kernel_ms(b, avg_ms, launches);
}

int dpsnn_batch_tune(dpsnn_batch *b, uint32_t trial_steps, dpsnn_batch_shape *chosen) {
  if (!b) return fail(DPSNN_ERR_ARG, "null batch");
  if (trial_steps == 0) trial_steps = 1;
  HIP_TRY(hipSetDevice(b->device));
  const bool wasProfiling = b->profile;
  int rc;
  if ((rc = batch_prepare(b))) return rc;
  const uint32_t cap = batch_cap(b);
  double best = 1e300;
  int bestShape = b->shape;
  uint32_t bestWaves = b->wavesPerCu;
  bool bestSplit = b->split, any = true;
  // every candidate computes the same result; the trial steps are real steps
  for (int split = 1; split >= 0 && any; --split)
    for (int i = 0; i < kNumShapes && any; ++i) {
      const uint32_t top = lds_waves(split ? phase1_lds(cap, kShapes[i].B) : lds_bytes(cap, kShapes[i].B));
      for (uint32_t w : {0u, std::max<uint32_t>(1, top / 2)}) {
        b->shape = i;
        b->wavesPerCu = w;
        b->split = split != 0;
        HIP_TRY(hipStreamSynchronize(b->stream));
        b->profile = true;
        b->evUsed = 0;
        for (uint32_t k = 0; k < trial_steps && any; ++k)
          if ((rc = batch_one_step(b, &any))) return rc;
        double ms = 0;
        uint64_t launches = 0;
        if ((rc = batch_kernel_ms(b, &ms, &launches))) return rc;
        if (launches && ms < best) {
          best = ms;
          bestShape = i;
          bestWaves = w;
          bestSplit = b->split;
        }
        if (!any) break;
      }
    }
  b->shape = bestShape;
  b->wavesPerCu = bestWaves;
  b->split = bestSplit;
  b->profile = wasProfiling;
  if (chosen) {
    chosen->chunks = kShapes[bestShape].K;
    chosen->buckets = kShapes[bestShape].B;
    chosen->waves_per_cu = (int)bestWaves;
    chosen->ms_per_step = best < 1e300 ? best : 0.0;
    chosen->split = bestSplit ? 1 : 0;
  }
  return 0;
}

void dpsnn_batch_destroy(dpsnn_batch *b) {
  if (!b) return;
  hipSetDevice(b->device);
  while (!b->nets.empty()) dpsnn_batch_remove(b, b->nets.back());
  (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  batch_free_items(b);
  if (b->d_nets) (void)hipFree(b->d_nets);
  if (b->keys) hipFree(b->keys);
  if (b->keyW) hipFree(b->keyW);
  if (b->keyCount) (void)hipFree(b->keyCount);
  if (b->keyPos) (void)hipFree(b->keyPos);
  if (b->workCtr) (void)hipFree(b->workCtr);
  for (int r = 0; r < dpsnn_batch::kRing; ++r)
    if (b->done[r]) hipEventDestroy(b->done[r]);
  for (auto &e : b->ev) {
    hipEventDestroy(e.first);
    hipEventDestroy(e.second);
  }
  hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"

// ============================================================================
// Host-only access and diagnostics
// ============================================================================
struct dpsnn_host_net {
  Params par;
  Connectome con;
  Ran3 rng;
  ExternalReplay replay;
  ExternalStep step;
  uint32_t t = 0;
  WindowLayout window;
  uint32_t windowUnitsPerCol = 0;
};

namespace {
int write_dump(const Connectome &c, float w2f, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(DPSNN_ERR_IO, "cannot open %s", path);
  std::vector<char> line(256);
  for (uint64_t s = 0; s < c.S; ++s) {
    const float w = (float)(c.weight[s] * w2f);
    const int len = std::snprintf(line.data(), line.size(), "%d %d %d %d %d %f %f\n", (int)c.pre[s],
                                  (int)c.post[s], (int)(c.post[s] / c.N), (int)c.delay[s], (int)c.kind[s],
                                  w, 0.0f);
    std::fwrite(line.data(), 1, (size_t)len, f);
  }
  std::fclose(f);
  return 0;
}
__global__ void k_exp(const double *x, double *y, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = dpsnn_exp(x[i]);
}
__global__ void k_div(const double *a, const double *b, const double *y, double *q, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = div_rn(a[i], b[i], y[i]);
}
}  // namespace

extern "C" {

int dpsnn_host_build(const char *inidir, dpsnn_host_net **out) {
  if (!inidir || !out) return fail(DPSNN_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<dpsnn_host_net> h(new dpsnn_host_net());
  std::string err;
  if (!read_params(inidir, h->par, err)) return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  h->rng.seed((int32_t)h->par.pub.globalSeed);
  const bool verbose = std::getenv("DPSNN_VERBOSE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  if (!generate_connectome(h->par, h->rng, h->con, err)) return fail(DPSNN_ERR_CONFIG, "%s", err.c_str());
  auto t1 = std::chrono::steady_clock::now();
  build_delivery(h->par, h->con);
  auto t2 = std::chrono::steady_clock::now();
  if (verbose)
    std::fprintf(stderr, "[dpsnn] generate %.2f s, delivery layout %.2f s\n",
                 std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  h->rng.seed(0);
  h->replay.init(h->con, &h->rng);
  *out = h.release();
  return 0;
}
uint32_t dpsnn_host_num_neurons(const dpsnn_host_net *h) { return h ? h->con.N : 0; }
uint64_t dpsnn_host_num_synapses(const dpsnn_host_net *h) { return h ? h->con.S : 0; }
uint64_t dpsnn_host_num_items(const dpsnn_host_net *h) { return h ? h->con.in_off[h->con.N] : 0; }
int dpsnn_host_get_synapses(const dpsnn_host_net *h, uint32_t *pre, uint32_t *post, uint8_t *delay,
                            int16_t *weight) {
  if (!h) return fail(DPSNN_ERR_ARG, "null argument");
  const Connectome &c = h->con;
  if (pre) std::memcpy(pre, c.pre.data(), 4 * c.S);
  if (post) std::memcpy(post, c.post.data(), 4 * c.S);
  if (delay) std::memcpy(delay, c.delay.data(), c.S);
  if (weight) std::memcpy(weight, c.weight.data(), 2 * c.S);
  return 0;
}
int dpsnn_host_get_inlists(const dpsnn_host_net *h, uint64_t *off, uint32_t *pre_delay, int16_t *weight,
                           uint32_t *syn) {
  if (!h) return fail(DPSNN_ERR_ARG, "null argument");
  const Connectome &c = h->con;
  const uint64_t items = c.in_off[c.N];
  if (off) std::memcpy(off, c.in_off.data(), 8 * (c.N + 1));
  if (pre_delay) std::memcpy(pre_delay, c.in_pre_d.data(), 4 * items);
  if (weight) std::memcpy(weight, c.in_w.data(), 2 * items);
  if (syn) std::memcpy(syn, c.in_syn.data(), 4 * items);
  return 0;
}
int dpsnn_host_window(dpsnn_host_net *h, uint32_t units_per_col, uint32_t dims[6], uint64_t sizes[4],
                      uint32_t *item, uint16_t *desc, uint64_t *chunk, int32_t *cols) {
  if (!h || !dims || !sizes) return fail(DPSNN_ERR_ARG, "null argument");
  WindowLayout &L = h->window;
  if (!L.ok || h->windowUnitsPerCol != units_per_col) {
    build_window(h->par, h->con, units_per_col, kWindowMaxEntries, L);
    h->windowUnitsPerCol = units_per_col;
    if (!L.ok) return fail(DPSNN_ERR_CONFIG, "no stencil-window layout: %s", L.why.c_str());
  }
  const uint32_t d[6] = {L.slots, L.npc, L.passes, L.unit, L.unitsPerCol, L.units};
  std::memcpy(dims, d, sizeof d);
  sizes[0] = L.item.size(); sizes[1] = L.desc.size(); sizes[2] = L.chunk.size(); sizes[3] = L.cols.size();
  if (item) std::memcpy(item, L.item.data(), 4 * L.item.size());
  if (desc) std::memcpy(desc, L.desc.data(), 2 * L.desc.size());
  if (chunk) std::memcpy(chunk, L.chunk.data(), 8 * L.chunk.size());
  if (cols) std::memcpy(cols, L.cols.data(), 4 * L.cols.size());
  return 0;
}
int dpsnn_host_get_external(const dpsnn_host_net *h, double *invNuExt, uint32_t *maxCount) {
  if (!h) return fail(DPSNN_ERR_ARG, "null argument");
  if (invNuExt) std::memcpy(invNuExt, h->con.invNuExt.data(), 8 * h->con.N);
  if (maxCount) std::memcpy(maxCount, h->con.maxExt.data(), 4 * h->con.N);
  return 0;
}
int dpsnn_host_write_synapse_dump(const dpsnn_host_net *h, const char *path) {
  if (!h || !path) return fail(DPSNN_ERR_ARG, "null argument");
  return write_dump(h->con, h->par.pub.factorWeightType_2_Float, path);
}
int dpsnn_host_external_step(dpsnn_host_net *h, uint32_t *off, double *time, double *value, uint64_t cap,
                             uint64_t *count) {
  if (!h) return fail(DPSNN_ERR_ARG, "null argument");
  h->replay.generate(h->t++, h->step);
  const uint64_t ne = h->step.time.size();
  if (count) *count = ne;
  if (off) std::memcpy(off, h->step.off.data(), 4 * (size_t(h->con.N) + 1));
  if (ne > cap) return fail(DPSNN_ERR_ARG, "external buffer holds %llu events, %llu needed",
                            (unsigned long long)cap, (unsigned long long)ne);
  if (time) std::memcpy(time, h->step.time.data(), 8 * ne);
  if (value) std::memcpy(value, h->step.value.data(), 8 * ne);
  return 0;
}
void dpsnn_host_destroy(dpsnn_host_net *h) { delete h; }

int dpsnn_write_synapse_dump(dpsnn_net *n, const char *path) {
  if (!n || !path) return fail(DPSNN_ERR_ARG, "null argument");
  if (n->opt.drop_host_synapses) return fail(DPSNN_ERR_ARG, "the host synapse list was dropped (drop_host_synapses)");
  if (!n->stdp) return write_dump(n->con, n->par.pub.factorWeightType_2_Float, path);
  // plastic network: the device's current weights and derivatives
  const Connectome &c = n->con;
  std::vector<int16_t> w(c.S);
  std::vector<float> td(c.S);
  int rc = synapse_state(n, w.data(), td.data(), nullptr);
  if (rc) return rc;
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(DPSNN_ERR_IO, "cannot open %s", path);
  std::vector<char> line(256);
  const float w2f = n->par.pub.factorWeightType_2_Float;
  for (uint64_t s = 0; s < c.S; ++s) {
    const int len = std::snprintf(line.data(), line.size(), "%d %d %d %d %d %f %f\n", (int)c.pre[s], (int)c.post[s],
                                  (int)(c.post[s] / c.N), (int)c.delay[s], (int)c.kind[s], (float)(w[s] * w2f), td[s]);
    std::fwrite(line.data(), 1, (size_t)len, f);
  }
  std::fclose(f);
  return 0;
}

int dpsnn_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
  HIP_TRY(hipSetDevice(device));
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 0;
}

int dpsnn_device_exp(int device, const double *x, double *y, uint64_t n) {
  if (!x || !y) return fail(DPSNN_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(device));
  double *dx = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc((void **)&dx, 8 * std::max<uint64_t>(n, 1)));
  HIP_TRY(hipMalloc((void **)&dy, 8 * std::max<uint64_t>(n, 1)));
  HIP_TRY(hipMemcpy(dx, x, 8 * n, hipMemcpyHostToDevice));
  const uint64_t blocks = (n + 255) / 256;
  if (n) hipLaunchKernelGGL(k_exp, dim3((uint32_t)blocks), dim3(256), 0, 0, dx, dy, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(y, dy, 8 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipFree(dx));
  HIP_TRY(hipFree(dy));
  return 0;
}

int dpsnn_device_div(int device, const double *a, const double *b, double *q, uint64_t n) {
  if (!a || !b || !q) return fail(DPSNN_ERR_ARG, "null argument");
  HIP_TRY(hipSetDevice(device));
  std::vector<double> y(n);
  for (uint64_t i = 0; i < n; ++i) y[i] = 1.0 / b[i];  // RN(1/b), as the engine's set-up computes it
  double *d = nullptr;
  const size_t bytes = 8 * std::max<uint64_t>(n, 1);
  HIP_TRY(hipMalloc((void **)&d, 4 * bytes));
  HIP_TRY(hipMemcpy(d, a, 8 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + n, b, 8 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + 2 * n, y.data(), 8 * n, hipMemcpyHostToDevice));
  const uint64_t blocks = (n + 255) / 256;
  if (n) hipLaunchKernelGGL(k_div, dim3((uint32_t)blocks), dim3(256), 0, 0, d, d + n, d + 2 * n, d + 3 * n, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(q, d + 3 * n, 8 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipFree(d));
  return 0;
}

}  // extern "C"

// dpsnn_host.h -- host side of the engine: DPSNN's parameter/ini interface,
// its ran3 random stream, the LIFCA connectome generator, the delivery
// layout the device consumes, the replayed external drive and the
// raster/rates writers.  Everything here runs once at network creation or
// streams alongside the device step; nothing here is on the device path.
#pragma once
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "dpsnn_hip.h"

namespace dpsnn {

constexpr uint32_t kMaxSubPop = 27;      // DSD__subPopTot
constexpr uint32_t kMaxStencil = 96;     // makeStencil_Max
constexpr uint32_t kDelaySlots = 80;     // DSD__maxD: stored delays < 80
constexpr uint32_t kLutDepth = 256;      // ANALOG_DEPTH
constexpr uint32_t kStdpSpan = 128;      // DSD__maxSTDPspan_ms
constexpr uint32_t kMaxOutDegree = 7500; // DSD__maxM (miniM)

// One sub-population's parameters (neuSubPopParamStruct,
// DPSNN_parameters.h:177-194).  Values pass through float on the way in.
struct SubPop {
  uint32_t count = 0, offset = 0;
  double DMin[kMaxSubPop] = {}, DMax[kMaxSubPop] = {}, J[kMaxSubPop] = {},
         DJ[kMaxSubPop] = {};
  double JExt = 0, DJExt = 0, NuExt = 0, Tau = 0, Theta = 0, H = 0, Tarp = 0,
         AlphaC = 0, TauC = 0, gC = 0;
};

struct Params {
  dpsnn_parameters pub{};
  uint32_t maxModDeltaX = 0, maxModDeltaY = 0, stencilX = 0, stencilY = 0;
  SubPop sp[kMaxSubPop];
  std::vector<float> prob;        // [src][tgt][x][y] over kMaxStencil^2
  std::vector<uint32_t> bath;     // [x][y]
  // STDPevent file settings (env_statSTDPevent_writeDisable,
  // env_start/stopSTDPevent_s, env_debugPrintEnable_ms)
  int32_t stdpEventDisable = 1;
  uint32_t startSTDPevent_s = 0, stopSTDPevent_s = 0, debugPrintEnable_ms = 0;
  float prob_at(uint32_t s, uint32_t t, uint32_t x, uint32_t y) const {
    return prob[((size_t(s) * kMaxSubPop + t) * kMaxStencil + x) * kMaxStencil + y];
  }
};

// name -> value pairs that take precedence over the process environment
// (several networks of one process, each with its own env_* settings).
using EnvMap = std::map<std::string, std::string>;
// "name=value" items separated by ';' or newlines.
bool parse_env_overrides(const char *text, EnvMap &out, std::string &err);

// Reads env vars + ini files exactly as getParameters and
// describeConnectome do.  Returns false and sets err on a config error.
bool read_params(const std::string &dir, Params &p, std::string &err, const EnvMap *over = nullptr);

// ran3 + NormDev with the reference's persistent-state semantics
// (randdev.c:116-189, :477-502, :584-606), one instance per network.
class Ran3 {
 public:
  static constexpr int32_t kBig = 1000000000;
  // one draw: the integer mj in [0, kBig), or next() = mj / kBig
  int32_t next_int();
  double next() { return next_int() * (1.0 / (double)kBig); }
  void seed(int32_t s);
  double norm();
  uint64_t draws() const { return count_; }
  // NormDev's cached second value (set, value)
  int32_t norm_set() const { return set_; }
  double norm_cached() const { return cached_; }
  void set_norm_state(int32_t set, double cached) { set_ = set; cached_ = cached; }
  // the whole stream position (checkpoints)
  struct State {
    int32_t idum, inext, inextp, ma[56], iff, set;
    double cached;
    uint64_t count;
  };
  State state() const;
  void set_state(const State &s);
 private:
  int32_t idum_ = -77531, inext_ = 0, inextp_ = 0, ma_[56] = {}, iff_ = 0;
  int32_t set_ = 0;
  double cached_ = 0;
  uint64_t count_ = 0;
};

// The connectome in the reference's two orders plus what the device needs.
struct Connectome {
  uint32_t N = 0;
  uint64_t S = 0;
  // backward (delivery-ordered) list: stable sort of generation order by
  // delay (inflateIntermToBackwardSynList, DPSNN_localNet_init.c:631-740)
  std::vector<uint32_t> pre, post;
  std::vector<uint8_t> delay, kind;
  std::vector<int16_t> weight;
  // per-neuron
  std::vector<uint8_t> subPop;
  std::vector<double> invNuExt;
  std::vector<uint32_t> maxExt;
  std::vector<double> bathLut;    // [subPop][256]
  // per-target in-lists in delivery insertion order:
  // (delay desc, pre asc, position in the pre's run asc)
  std::vector<uint64_t> in_off;   // [N+1]
  std::vector<uint32_t> in_pre_d; // pre | delay << 24
  std::vector<int16_t> in_w;
  std::vector<uint32_t> in_syn;   // backward-list index of the synapse
  uint32_t max_in = 0;
  uint32_t max_delay = 0;         // largest stored delay
  // per-target incoming synapses (backward-list order), for LTP
  std::vector<uint64_t> inc_off;
  std::vector<uint32_t> inc_syn;
};

// Host threads used when a caller asks for 0: the hardware threads, capped
// by OMP_NUM_THREADS where it is set.
unsigned default_threads();

// threads = 0: default_threads().
// The per-neuron part of a connectome (sub-population, bath rate and cap,
// the bath LUT) before any synapse; generate_connectome starts with it.
void connectome_neurons(const Params &p, Connectome &c);
// Bath rate and per-ms cap of neuron i (sub-population src), in c.
void neuron_bath(const Params &p, uint32_t i, uint32_t src, Connectome &c);
// The backward list from synapses in generation order (source-major; source
// i's are [gen_off[i], gen_off[i + 1])): a stable counting sort by delay.
void inflate_backward(Connectome &c, const std::vector<uint64_t> &gen_off, const std::vector<uint32_t> &g_post,
                      const std::vector<uint8_t> &g_delay, const std::vector<int16_t> &g_w, unsigned threads = 0);
bool generate_connectome(const Params &p, Ran3 &rng, Connectome &c, std::string &err,
                         unsigned threads = 0);
// with_incoming: also build the per-target incoming synapse lists (STDP).
void build_delivery(const Params &p, Connectome &c, bool with_incoming = true, unsigned threads = 0);

// The in-lists regrouped for the stencil-window form of phase 1.  Every
// source of a target lies in a few columns around the target's column (its
// stencil).  A unit is a range of at most kWindowMaxUnit targets of one
// column; for each run delay d (a pass, descending as in the in-lists),
// the emission times of the unit's stencil columns in slot t - d fit the
// device's LDS, and the unit's items of that delay are streamed against
// them.  Within a (unit, pass), wave w of the workgroup owns the unit's
// targets q = w, w + kWindowWaves, ... and streams their segments of delay
// d in target order, each padded to whole chunks of kWindowChunk items.
//   item:  window index << 16 | (uint16) weight, window index =
//          slot * npc + (source - first neuron of the slot's column);
//          padding items carry window index slots * npc (never live)
//   desc:  per chunk, the index of its target within the unit, plus
//          kWindowFirst on the target's first chunk of the stream
//   chunk: [units][passes][kWindowWaves] + 1, first chunk of each stream
//   cols:  [columns][slots] source column of each stencil slot, or -1
// A target's live items of all passes, taken in stream order, are its live
// in-list items in in-list order.
constexpr uint32_t kWindowWaves = 16;
constexpr uint32_t kWindowChunk = 8;
constexpr uint32_t kWindowMaxUnit = 1024;
constexpr uint32_t kWindowFirst = 0x8000;   // desc flag: the target's first chunk in the stream
constexpr uint32_t kWindowTarget = 0x7FFF;  // desc mask: the target's index within the unit
struct WindowLayout {
  bool ok = false;
  std::string why;               // why the layout was not built
  uint32_t slots = 0, npc = 0, passes = 0, unit = 0, unitsPerCol = 0, units = 0;
  std::vector<int32_t> cols;
  std::vector<uint32_t> item;
  std::vector<uint16_t> desc;
  std::vector<uint64_t> chunk;
};
// Builds the layout from c's in-lists with at most `maxSlots` stencil slots
// (window entries slots * npc + 1 <= maxEntries) and units of
// ceil(npc / unitsPerCol) targets; sets ok = false (and why) when the
// connectome does not fit.
void build_window(const Params &p, const Connectome &c, uint32_t unitsPerCol, uint32_t maxEntries,
                  WindowLayout &L, unsigned threads = 0);

// Host-side replay of the external Poisson drive (DPSNN_neuron_sim.c:
// 604-640, :844-865): the reference's ran3 draws in neuron order, step by
// step, recorded as per-neuron (time, value) lists.
struct ExternalStep {
  std::vector<uint32_t> off;   // [N+1]
  std::vector<double> time, value;
};
class ExternalReplay {
 public:
  void init(const Connectome &c, Ran3 *rng);
  void generate(uint32_t t, ExternalStep &out);
  // next bath event time of each neuron (checkpoints)
  std::vector<double> &next_times() { return assigned_; }
 private:
  const Connectome *c_ = nullptr;
  Ran3 *rng_ = nullptr;
  std::vector<double> assigned_;
};

// STDP contribution tables (initLongTermPlasticity, DPSNN_neuron_sim.c:
// 45-92): causal (LTP) and anti-causal (LTD), kStdpSpan entries each.
void stdp_tables(const dpsnn_parameters &q, std::vector<float> &causal, std::vector<float> &anti);

// The STDPevent file (DPSNN_stat.c:451-470): every non-zero LTP and LTD
// contribution, restated on the host from each step's spikes (the
// contributions depend on spike times only).  Needs the host backward list
// and incoming lists.  step() is called once per ms, in order, with the
// neurons whose spikes the fired scan of that ms found.
class StdpEventWriter {
 public:
  bool init(const Params &p, const Connectome &c, const std::string &dir);
  bool step(uint32_t t, const uint32_t *ids, uint32_t n);
  void close();
  ~StdpEventWriter() { close(); }
 private:
  bool writes(uint32_t t) const;
  void line(uint32_t t, uint32_t pre, uint32_t post, float v);
  const Params *p_ = nullptr;
  const Connectome *c_ = nullptr;
  std::string dir_;
  FILE *fp_ = nullptr;
  uint32_t D_ = 1;
  std::vector<float> causal_, anti_;
  std::vector<uint32_t> first_;              // [N][D] first synapse of each run
  std::vector<int32_t> lastSpike_, lastAct_;
  std::vector<std::vector<uint32_t>> ring_;  // spikes of the last D ms
  std::vector<char> buf_;
};

// Raster and per-population rate files (DPSNN_stat.c:55-107, :225-296).
class StatWriter {
 public:
  bool init(const Params &p, const std::string &dir);
  // Called once per step, in order, with that step's compacted spikes and
  // per-(column, subPop) counts.
  bool step(uint32_t t, const uint32_t *ids, const double *te, uint32_t n,
            const uint32_t *popCounts);
  void close();
  void flush();
  // Checkpoints: the counts of the open rates window and which of the
  // second's files are open; resume reopens those files for appending.
  struct State {
    std::vector<uint32_t> perPop;
    int spikesOpen = 0, ratesOpen = 0;
  };
  State state() const { return State{perPop_, spikes_ != nullptr, rates_ != nullptr}; }
  bool resume(const State &s, uint32_t sec);
  ~StatWriter() { close(); }
 private:
  FILE *open_file(const char *prefix, uint32_t sec, const char *mode = "w");
  bool window(uint32_t sec) const;
  const Params *p_ = nullptr;
  std::string dir_;
  FILE *spikes_ = nullptr, *rates_ = nullptr;
  std::vector<uint32_t> perPop_;
  std::vector<char> buf_;
};

// Perseo .ini files of the network (localNetClass::writeIniFiles,
// DPSNN_writeIniFiles.c:31-158): PerseoConFile_h0.txt, PerseoModFile_h0.txt.
bool write_ini_files(const Params &p, const std::string &outdir, std::string &err);

}  // namespace dpsnn

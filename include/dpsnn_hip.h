/* dpsnn_hip.h -- C ABI of the MI355X engine for DPSNN's per-millisecond
 * LIFCA simulation step (one process's whole local network on one GPU).
 *
 * The engine replaces, behind this boundary, the reference's step path:
 *   localNetClass::completeTimeStep_ms   DPSNN_localNet_sim.c:22-307
 *   (fired scan, compaction, axonal spike exchange at world size 1, delay
 *   ring, delivery, per-neuron event ordering, external Poisson drive,
 *   LIFCA dynamics, STDP, raster and per-population rates)
 * while keeping its parameter interface (env vars + column.txt,
 * connectivity.txt, bathEfficacy.txt, plasticity.txt read by
 *   getParameters                        DPSNN_getParameters.c:23-831)
 * and its output files (spikes-*.dat, ratesPerPop-*.dat written as
 *   statSpikesClass::write               DPSNN_stat.c:275-296
 *   statSpikingRatesPerPopClass::write   DPSNN_stat.c:225-273).
 *
 * All entry points return 0 on success and a negative dpsnn_status on
 * failure (the reference prints and calls exit(0) instead); the message of
 * the last failure on the calling thread is dpsnn_last_error().
 * Plain pointers and sizes only: host buffers are caller-owned.
 */
#ifndef DPSNN_HIP_H
#define DPSNN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DPSNN_OK = 0,
  DPSNN_DONE = 1,            /* dpsnn_step: programmed time reached */
  DPSNN_ERR_CONFIG = -1,     /* bad env var / ini file (reference: exit(0)) */
  DPSNN_ERR_DEVICE = -2,     /* HIP runtime error */
  DPSNN_ERR_MEMORY = -3,     /* host or device allocation failed */
  DPSNN_ERR_ARG = -4,        /* invalid argument */
  DPSNN_ERR_OVERFLOW = -5,   /* a per-step bound was exceeded on device */
  DPSNN_ERR_IO = -6          /* output file could not be written */
} dpsnn_status;

/* External Poisson drive of each neuron (DPSNN_neuron_sim.c:604-640). */
typedef enum {
  DPSNN_DRIVE_REPLAY = 0,    /* the reference's own ran3 stream, generated on
                                the host in neuron order and replayed on the
                                device: bit-exact with the reference */
  DPSNN_DRIVE_PHILOX = 1     /* same rates and amplitudes from an on-device
                                counter-based Philox4x32-10 stream:
                                statistically equivalent, not bit-exact */
} dpsnn_drive;

typedef struct {
  int device;                /* HIP device ordinal */
  int drive;                 /* dpsnn_drive */
  uint64_t philox_seed;      /* seed of the Philox drive */
  const char *output_dir;    /* where spikes-*.dat / ratesPerPop-*.dat go;
                                NULL or "" = no files */
  uint32_t drain_every_ms;   /* raster/rates device->host drain period
                                (0 = default 100 ms) */
  int keep_spikes;           /* 1 = keep each step's spike list readable
                                through dpsnn_last_spikes */
  const char *env;           /* "env_name=value" items separated by ';' or
                                newlines that take precedence over the
                                process environment (several networks with
                                their own settings in one process); NULL =
                                the environment only */
  int drop_host_synapses;    /* 1 = free the host copy of the synapse list
                                once the device holds it (large batches and
                                served jobs): dpsnn_get_synapses, synapse
                                dumps and synPeriodicProbe files are then
                                unavailable */
  uint32_t host_threads;     /* host threads of the connectome generator and
                                delivery layout (0 = the hardware threads,
                                capped by OMP_NUM_THREADS where set).  The
                                generator draws each column once on 1 or 2
                                threads and both of its entry variants on
                                3 or more; the output is the same */
  uint32_t generator;        /* dpsnn_generator: who draws the connectome */
  uint64_t generator_seed;   /* seed of the device generator */
} dpsnn_options;

/* Connectome generator (DPSNN_LIFCAcorticalModuleGen.cpp:62-220). */
typedef enum {
  DPSNN_GENERATOR_HOST = 0,    /* the reference's generator and ran3 stream on
                                  the host: bit-exact with the reference */
  DPSNN_GENERATOR_DEVICE = 1   /* the same distributions (connection
                                  probabilities per stencil cell, geometric
                                  target gaps, uniform delays, truncated
                                  Gaussian weights) drawn on the device from
                                  Philox4x32-10 keyed by generator_seed and
                                  the source neuron: statistically
                                  equivalent, not bit-exact.  dpsnn_prepare
                                  then uses the device named in the options */
} dpsnn_generator;

/* Parameters as getParameters returns them (the subset this path reads;
 * field names follow struct DPSNN_parameters, DPSNN_parameters.h:72-146). */
typedef struct {
  uint32_t globH, locN, globN, D, totalSimTime_ms, globCFX, globCFY, globCFT,
      neuronsPerCM, subPopNumber, tabNeuron, moduloSec, ratesSampling,
      globalSeed, howManyOutputFiles, startPlasticity_ms, stopPlasticity_ms,
      startStatFiles_s, stopStatFiles_s, startRatesPerPop_s, stopRatesPerPop_s,
      stencilDim, overallConnectivity, delayMin, delayMax;
  int32_t spikes_writeDisable, statRatesPerPop_writeDisable;
  float minSynWeight_f, factorFloat_2_weightType, factorWeightType_2_Float;
  float STDPmultiplier, antiCausalCoeff, tauSTDP_ms, causalSTDPmaxAbsValue,
      antiCausalSTDPmaxAbsValue, tauDerivativeDecay_s, derivativeDrift;
  int32_t statSynPeriodicProbe_writeDisable;
  uint32_t startSynPeriodicProbe_s, stopSynPeriodicProbe_s;
} dpsnn_parameters;

typedef struct dpsnn_net dpsnn_net;

/* Counters of what the device did (cumulative since creation). */
typedef struct {
  uint64_t steps;            /* simulated ms */
  uint64_t spikes;           /* spikes found by the fired scans so far
                                (the raster's spike count) */
  uint64_t synaptic_events;  /* delivered synaptic events (sum of runs) */
  uint64_t external_events;  /* external bath events integrated */
  uint64_t synapses;         /* synapses in the connectome */
  uint64_t delivery_items;   /* in-list items (synapses + spill-over) */
  uint32_t neurons;
  uint32_t max_in_items;     /* longest per-neuron in-list */
  double init_seconds;       /* connectome generation + upload */
  uint64_t device_bytes;     /* device memory held by this network */
  uint32_t window_form;      /* 1: phase 1 runs in the stencil-window form
                                (k_collect + k_sort) where it applies */
  uint32_t reserved;
  uint64_t graph_steps;      /* steps replayed from captured graphs */
} dpsnn_stats;

/* getParameters equivalent: reads the env vars of the calling process and
 * the ini files in `inidir`.  Replaces DPSNN_getParameters.c:23-831. */
int dpsnn_get_parameters(const char *inidir, dpsnn_parameters *out);

/* Perseo simulator .ini files of the network of `inidir`
 * (localNetClass::writeIniFiles, DPSNN_writeIniFiles.c:31-158), at world
 * size 1: PerseoConFile_h0.txt (one line per target/source sub-population
 * pair of each column and stencil cell with a positive connection
 * probability) and PerseoModFile_h0.txt (one line per sub-population of each
 * column) in `outdir`.  The reference compiles its call out
 * (DPSNN_localNet_init.c:247-249, an if(false) region); here it is a call of
 * its own, host only.  env: as dpsnn_options.env (NULL: the process's). */
int dpsnn_write_ini_files(const char *inidir, const char *outdir, const char *env);

/* Builds a network: reads parameters, runs the reference connectome
 * generator on the host (DPSNN_LIFCAcorticalModuleGen.cpp:62-220), builds
 * the delivery layout (DPSNN_localNet_init.c:631-740) and uploads state and
 * per-neuron in-lists to the device.  Replaces localNetClass::init
 * (DPSNN_localNet_init.c:43-307) at world size 1. */
int dpsnn_create(const char *inidir, const dpsnn_options *opt, dpsnn_net **out);
/* dpsnn_create in two stages, so that host work (parameters, generator,
 * layout: no device call) can overlap device work: dpsnn_prepare builds the
 * network on the host, dpsnn_device_bytes_needed tells what its upload will
 * allocate, dpsnn_upload allocates and uploads.  A prepared network can be
 * destroyed without ever being uploaded. */
int dpsnn_prepare(const char *inidir, const dpsnn_options *opt, dpsnn_net **out);
uint64_t dpsnn_device_bytes_needed(const dpsnn_net *net);
/* Host bytes the network's synapse and delivery lists hold (the backward
 * list and, until upload, the in-lists and window layout).  No counterpart
 * in the reference: the server bounds its backlog of built jobs by it. */
uint64_t dpsnn_host_bytes(const dpsnn_net *net);
int dpsnn_upload(dpsnn_net *net);

/* Advances `nsteps` milliseconds (each = one completeTimeStep_ms).  Returns
 * DPSNN_DONE once totalSimTime_ms has been reached (the reference's fire()
 * returning 1, localNetProcess.c:138-157).  Work is queued on the
 * network's stream; files are written at each drain. */
int dpsnn_step(dpsnn_net *net, uint32_t nsteps);
/* dpsnn_step without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory inside
 * their loop (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The step's key scratch comes from a static
 * device array of 128 MiB per device (held from the library's load on that
 * device until the process ends) instead of the network's own buffers, so
 * networks whose step needs more scratch than that are refused
 * (DPSNN_ERR_ARG); so is the replayed drive, whose host-fed buffers grow
 * with the bath.  All allocation-free steps on one device share the array:
 * calls for two networks that overlap in time on one device race.
 * Results are bit-identical to dpsnn_step's. */
int dpsnn_step_noalloc(dpsnn_net *net, uint32_t nsteps);
/* Waits for queued work, drains raster/rates to files. */
int dpsnn_sync(dpsnn_net *net);
uint32_t dpsnn_now_ms(const dpsnn_net *net);
uint32_t dpsnn_total_ms(const dpsnn_net *net);

/* Device state after the last completed step (neuron order). */
int dpsnn_get_state(dpsnn_net *net, double *v, double *c, double *Tr,
                    double *Te, uint8_t *fired);
/* Spikes compacted by the last step (requires keep_spikes). */
int dpsnn_last_spikes(dpsnn_net *net, uint32_t *ids, double *te,
                      uint32_t *count);
/* The delivery-ordered (backward) synapse list the device holds:
 * pre, post, stored delay, int16 weight; plus STDP state. */
uint64_t dpsnn_num_synapses(const dpsnn_net *net);
int dpsnn_get_synapses(dpsnn_net *net, uint32_t *pre, uint32_t *post,
                       uint8_t *delay, int16_t *weight, float *tderiv,
                       int32_t *last_act);
/* Writes the reference's synapse dump text
 * (localNet_h0_N<locN>_SynapticList_BACKWARD_R<rep>.dat format,
 * DPSNN_synapse.h:63-91) of the device's current synapses. */
int dpsnn_write_synapse_dump(dpsnn_net *net, const char *path);
int dpsnn_get_stats(dpsnn_net *net, dpsnn_stats *out);

/* Checkpoint / resume of a network between steps -- the process-level
 * checkpoint hooks the reference's DAL declares but never implements
 * (dal.h:52-55).  dpsnn_checkpoint waits for queued work, writes the
 * pending output, and stores the device-resident state (neuron state, bath
 * cursors, Philox counters, spike history, counters, STDP synapse state)
 * and the host stream state (ran3 position and NormDev cache, replayed
 * bath, open output files and the open rates window) to `path`.
 * dpsnn_restore builds the network of `inidir` with `opt` (the same ini
 * files and env settings as the checkpointed one: the connectome is
 * regenerated, not stored) and loads the state into it; the run then
 * continues as if never interrupted, appending to the second's output
 * files.  A network of a batch must leave it before a checkpoint.
 *
 * File layout (little endian, no padding between items):
 *   header: char magic[8] = "DPSNNCK1"; uint32 N, RS, NW, drive, stdp, t,
 *           nsub, ncols, writer, spikesOpen, ratesOpen, pad; uint64 S, items
 *   double v[N], c[N], Tr[N], Te[N], next external event time[N]
 *   uint64 Philox counter[N]
 *   uint64 fired bitmap[RS][NW]   (slot of ms m: m % RS; bit i % 64 of
 *                                  word i / 64: neuron i fired in step m-1)
 *   double emission time[RS][N]   (same slots; -1: did not fire)
 *   uint64 counters[16][16]       (sum over rows: [0] synaptic events,
 *                                  [1] external events, [2] spikes)
 *   with stdp: int32 last spike ms[N]; then per synapse in target-major
 *           order (each target's incoming synapses in backward-list order):
 *           int16 weight[S], float timeDerivative[S], int32 lastActive[S]
 *   ran3 state: int32 idum, inext, inextp, ma[56], iff, normSet, pad;
 *           double normCached; uint64 draws
 *   replayed drive: double next external event time[N] (host stream)
 *   writer: uint32 per-(column, sub-population) counts of the open rates
 *           window [ncols * nsub] */
int dpsnn_checkpoint(dpsnn_net *net, const char *path);
int dpsnn_restore(const char *inidir, const dpsnn_options *opt, const char *path, dpsnn_net **out);
/* HIP stream the network runs on (hipStream_t as void*). */
void *dpsnn_stream(dpsnn_net *net);
void dpsnn_destroy(dpsnn_net *net);

/* Average device time (ms) of the integration kernel over the launches
 * since the last call, measured with HIP events on the network's stream;
 * *launches receives the launch count.  Requires dpsnn_profile(net, 1). */
int dpsnn_profile(dpsnn_net *net, int enable);
int dpsnn_kernel_time(dpsnn_net *net, int kernel, double *avg_ms,
                      uint64_t *launches);

/* ---- batching -----------------------------------------------------------
 * Many independent networks of uneven size advance in ONE persistent kernel
 * launch per step.  A network joins with dpsnn_batch_add (same device) and
 * is then stepped only through its batch; each batched step advances every
 * member that has steps left by one ms (members may be at different ms).
 * A member's results are bit-identical to stepping it alone.  Files,
 * counters, state and synapse read-back work per member as usual. */
typedef struct dpsnn_batch dpsnn_batch;
typedef struct {
  int chunks;        /* in-list chunks of 64 items held in registers: 8/16/24 */
  int buckets;       /* arrival-time buckets of the LDS sort: 64/256/1024 */
  int waves_per_cu;  /* phase-1 waves per CU (0 = as many as LDS allows) */
  double ms_per_step;/* dpsnn_batch_tune: measured device time of the shape */
  int split;         /* 1: per chunk of tiles, a phase-1 launch (one wave per
                        neuron) then a phase-2 launch; 0: one persistent
                        fused launch per step */
} dpsnn_batch_shape;
int dpsnn_batch_create(int device, dpsnn_batch **out);
int dpsnn_batch_add(dpsnn_batch *b, dpsnn_net *net);
int dpsnn_batch_remove(dpsnn_batch *b, dpsnn_net *net);
uint32_t dpsnn_batch_size(const dpsnn_batch *b);
/* Up to nsteps batched steps; DPSNN_DONE once no member has steps left. */
int dpsnn_batch_step(dpsnn_batch *b, uint32_t nsteps);
int dpsnn_batch_sync(dpsnn_batch *b);
/* Autotuner: runs `trial_steps` batched steps (real steps of the
 * simulation) with each launch shape and keeps the fastest.  Each step's
 * launches are replayed from a captured HIP graph (DPSNN_GRAPH=0: plain
 * launches). */
int dpsnn_batch_tune(dpsnn_batch *b, uint32_t trial_steps, dpsnn_batch_shape *chosen);
int dpsnn_batch_get_shape(const dpsnn_batch *b, dpsnn_batch_shape *s);
int dpsnn_batch_set_shape(dpsnn_batch *b, const dpsnn_batch_shape *s);
/* Average device time (ms) of the batched kernel since the last call. */
int dpsnn_batch_profile(dpsnn_batch *b, int enable);
int dpsnn_batch_kernel_time(dpsnn_batch *b, double *avg_ms, uint64_t *launches);
/* Members are removed (not destroyed). */
void dpsnn_batch_destroy(dpsnn_batch *b);

/* ---- serving -----------------------------------------------------------
 * A request loop for simulation jobs.  Jobs are submitted at any time; host
 * threads build their networks (dpsnn_prepare); a scheduler thread admits a
 * built job as soon as its device footprint fits the HBM budget (first fit
 * in arrival order), uploads it and adds it to the running batch, so it
 * advances in the same batched launches as the jobs already on the GPU
 * (continuous batching); finished jobs leave the batch and free their
 * memory.  State stays on the device from admission to completion.  A job's
 * results (files, counters, state) are those of the network run alone. */
typedef struct dpsnn_server dpsnn_server;
typedef struct {
  int device;
  uint64_t hbm_budget_bytes;  /* device bytes admitted jobs may hold, batch
                                 scratch included (0 = 90% of free memory) */
  uint32_t build_threads;     /* concurrent host builds (0 = 2) */
  uint32_t steps_per_round;   /* batched steps between admissions (0 = 10) */
  int keep_finished;          /* 1 = a finished job keeps its network (and
                                 its memory) until dpsnn_server_take */
  uint32_t max_running;       /* jobs on the device at once (0 = as many as
                                 the budget holds; 1 = one job at a time) */
  uint64_t host_backlog_bytes; /* host bytes (dpsnn_host_bytes) that built
                                 jobs not yet uploaded may hold; builders
                                 start no build beyond it (0 = 32 GiB, at
                                 most a quarter of physical memory) */
} dpsnn_server_options;
typedef enum {
  DPSNN_JOB_QUEUED = 0,       /* waiting for a host builder */
  DPSNN_JOB_BUILDING = 1,     /* network being built on the host */
  DPSNN_JOB_WAITING = 2,      /* built, waiting for device memory */
  DPSNN_JOB_RUNNING = 3,      /* on the device, in the batch */
  DPSNN_JOB_DONE = 4,
  DPSNN_JOB_FAILED = 5
} dpsnn_job_state;
typedef struct {
  uint64_t id;
  int state;                  /* dpsnn_job_state */
  int status;                 /* 0, or the failing call's dpsnn_status */
  double submitted_s, built_s, admitted_s, finished_s;  /* server clock */
  uint64_t device_bytes;
  dpsnn_stats stats;          /* final counters (DONE) */
  char message[160];          /* FAILED: the error */
} dpsnn_job_info;
typedef struct {
  uint64_t submitted, finished, failed, running, waiting, building;
  uint64_t synaptic_events;   /* delivered by all jobs so far */
  uint64_t device_bytes;      /* held by admitted jobs + batch scratch */
  double elapsed_s;           /* server clock */
  double step_s;              /* scheduler time in batched steps */
  double upload_s;            /* uploader time copying admitted jobs (beside the steps) */
  uint64_t queued;            /* submitted, not yet taken by a builder (not in `building`) */
  double admit_s;             /* scheduler time adding uploaded jobs to the batch */
  double account_s;           /* scheduler time reading running jobs' counters */
  double retire_s;            /* scheduler time retiring finished jobs (sync, drain, free) */
} dpsnn_server_stats;
int dpsnn_server_create(const dpsnn_server_options *opt, dpsnn_server **out);
/* Queues a job: the network of `inidir` with options `opt` (copied; its
 * env, output_dir and drive are the job's own). */
int dpsnn_server_submit(dpsnn_server *s, const char *inidir, const dpsnn_options *opt, uint64_t *id);
int dpsnn_server_job(dpsnn_server *s, uint64_t id, dpsnn_job_info *info);
/* Waits until job `id` is DONE or FAILED (timeout_s < 0: no limit);
 * returns DPSNN_DONE when it is, 0 on timeout. */
int dpsnn_server_wait(dpsnn_server *s, uint64_t id, double timeout_s, dpsnn_job_info *info);
/* Waits until every submitted job is DONE or FAILED. */
int dpsnn_server_wait_all(dpsnn_server *s, double timeout_s);
/* keep_finished: hands a DONE job's network to the caller, who destroys it
 * with dpsnn_destroy (its memory leaves the budget then). */
int dpsnn_server_take(dpsnn_server *s, uint64_t id, dpsnn_net **net);
int dpsnn_server_get_stats(dpsnn_server *s, dpsnn_server_stats *st);
/* Stops the loop; jobs not finished are dropped. */
void dpsnn_server_destroy(dpsnn_server *s);

/* ---- host-only access (no device needed) ------------------------------
 * The connectome generator and delivery layout as the device receives
 * them, and the replayed external drive, for inspection and CPU checks. */
typedef struct dpsnn_host_net dpsnn_host_net;
int dpsnn_host_build(const char *inidir, dpsnn_host_net **out);
uint32_t dpsnn_host_num_neurons(const dpsnn_host_net *h);
uint64_t dpsnn_host_num_synapses(const dpsnn_host_net *h);
uint64_t dpsnn_host_num_items(const dpsnn_host_net *h);
/* backward (delivery-ordered) list */
int dpsnn_host_get_synapses(const dpsnn_host_net *h, uint32_t *pre,
                            uint32_t *post, uint8_t *delay, int16_t *weight);
/* per-target in-lists: off[N+1], items (pre | delay << 24), weights and
 * backward-list indices, in delivery insertion order */
int dpsnn_host_get_inlists(const dpsnn_host_net *h, uint64_t *off,
                           uint32_t *pre_delay, int16_t *weight,
                           uint32_t *syn);
/* The in-lists regrouped for the stencil-window form of phase 1 (the
 * device's k_collect input), units of ceil(npc / units_per_col) targets:
 * dims = slots, npc, passes, unit, units per column, units; sizes = the
 * lengths of item, desc, chunk and cols (null arrays: sizes only).
 *   item  window index << 16 | (uint16) weight, padded per target segment
 *         to whole chunks of 8 (padding: window index slots * npc)
 *   desc  per chunk, its target's index within the unit
 *   chunk [units][passes][16] + 1: first chunk of each (unit, pass, wave)
 *         stream; pass p holds run delay passes - 1 - p; wave w streams the
 *         unit's targets w, w + 16, ...
 *   cols  [columns][slots]: the source column of each window slot, or -1 */
int dpsnn_host_window(dpsnn_host_net *h, uint32_t units_per_col, uint32_t dims[6], uint64_t sizes[4],
                      uint32_t *item, uint16_t *desc, uint64_t *chunk, int32_t *cols);
/* per-neuron external bath: 1000/(NuExt*CExt) and the per-ms cap */
int dpsnn_host_get_external(const dpsnn_host_net *h, double *invNuExt,
                            uint32_t *maxCount);
int dpsnn_host_write_synapse_dump(const dpsnn_host_net *h, const char *path);
/* Draws the next step's replayed external events: off[N+1] per neuron,
 * (time, value) pairs into buffers of `cap` entries; *count = total. */
int dpsnn_host_external_step(dpsnn_host_net *h, uint32_t *off, double *time,
                             double *value, uint64_t cap, uint64_t *count);
void dpsnn_host_destroy(dpsnn_host_net *h);

/* Free and total memory of a device. */
int dpsnn_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* Device exp of x[0..n) into y (diagnostic: the device's glibc-exact exp
 * checked against the host libm). */
int dpsnn_device_exp(int device, const double *x, double *y, uint64_t n);

/* a[i] / b[i] into q[0..n) by the device's division by reciprocal, the one
 * the engine uses for its time constants (diagnostic: checked against IEEE
 * division on the host).  b normal, quotients in the normal range. */
int dpsnn_device_div(int device, const double *a, const double *b, double *q, uint64_t n);

const char *dpsnn_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

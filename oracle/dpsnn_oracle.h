/* dpsnn_oracle.h -- CPU restatement of DPSNN's LIFCA per-millisecond step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
 * engine: only the tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it.  The product path never links or calls it.
 *
 * It restates, in plain C and with the reference's own numeric types, the
 * parts of DPSNN (APE-group/201812RealTimeCortSim) that the engine replaces:
 * parameter/ini parsing, the NR ran3 stream, the LIFCA connectome generator,
 * the delay-sorted delivery layout, the per-ms step (delivery, per-neuron
 * event ordering, external Poisson drive, LIFCA dynamics, STDP) and the
 * spike-raster / per-population-rate writers.  Each function cites the
 * reference file:line it follows.
 */
#ifndef DPSNN_ORACLE_H
#define DPSNN_ORACLE_H
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_net ora_net;

/* Reads the env vars and ini files (column.txt, connectivity.txt,
 * bathEfficacy.txt, optional plasticity.txt) found in `dir`, generates the
 * connectome and builds the delivery layout.  Returns NULL and fills `err`
 * on a configuration error. */
ora_net *ora_create(const char *dir, char *err, int errlen);
void ora_destroy(ora_net *n);
/* ora_create with a given backward (delivery-ordered) synapse list instead
 * of the generator's (the engine's host generator is checked against this
 * oracle's, so a large network need not be generated twice). */
ora_net *ora_create_from_backward(const char *dir, uint64_t S, const uint32_t *pre, const uint32_t *post,
                                  const uint8_t *delay, const int16_t *weight, char *err, int errlen);
/* Loads an engine checkpoint file (dpsnn_checkpoint; layout in
 * include/dpsnn_hip.h) of the same network into the oracle's state: the
 * oracle then continues the run from that ms. */
int ora_load_checkpoint(ora_net *n, const char *path, char *err, int errlen);

/* Sizes. */
uint32_t ora_num_neurons(const ora_net *n);
uint64_t ora_num_synapses(const ora_net *n);
uint32_t ora_total_ms(const ora_net *n);
uint32_t ora_now_ms(const ora_net *n);
/* synaptic events delivered since create (each delivered synapse once) */
uint64_t ora_synaptic_events(const ora_net *n);

/* Backward (delivery-ordered) synapse list, as the reference keeps it after
 * inflateIntermToBackwardSynList.  Arrays of length ora_num_synapses(). */
void ora_get_backward(const ora_net *n, uint32_t *pre, uint32_t *post,
                      uint8_t *delay, uint8_t *kind, int16_t *weight,
                      float *tderiv, int32_t *last_act);
/* Same list in generation order (before the delay sort). */
void ora_get_generated(const ora_net *n, uint32_t *pre, uint32_t *post,
                       uint8_t *delay, int16_t *weight);
/* Writes the reference's synapse dump text to `path` (format of
 * DPSNN_synapse.h:63-91). */
int ora_write_synapse_dump(const ora_net *n, const char *path);
/* Per-neuron external drive parameters. */
void ora_get_external(const ora_net *n, double *invNuExt, uint32_t *maxCount);
/* Bath LUT per subPop: out[subPop*256 + i]. */
void ora_get_bath_lut(const ora_net *n, double *out);
uint32_t ora_num_subpops(const ora_net *n);
uint32_t ora_neurons_per_cm(const ora_net *n);

/* Runs one step (one call of completeTimeStep_ms + advanceSimTime_ms).
 * Returns 0 while running, 1 when the programmed time is over (no step). */
int ora_fire(ora_net *n);
/* Spikes compacted in the last step: ids (ascending) and emission times. */
uint32_t ora_last_spikes(const ora_net *n, uint32_t *ids, double *te);
/* Neuron state. */
void ora_get_state(const ora_net *n, double *v, double *c, double *Tr,
                   double *Te, uint8_t *fired);
/* Enables output files (spikes-*.dat / ratesPerPop-*.dat) in `outdir`. */
void ora_set_output_dir(ora_net *n, const char *outdir);
/* When set, the external events drawn during each step are recorded and can
 * be read back: counts per neuron and (time,value) pairs in neuron order. */
void ora_record_external(ora_net *n, int enable);
uint32_t ora_last_external(const ora_net *n, uint32_t *counts, double *times,
                           double *values);
/* glibc exp passthrough, for comparing device exp against the host libm. */
double ora_libm_exp(double x);
float ora_libm_logf(float x);

#ifdef __cplusplus
}
#endif
#endif

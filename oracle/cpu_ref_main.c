/* cpu_ref_main.c -- command-line driver of the CPU restatement (oracle).
 * TEST INFRASTRUCTURE ONLY: used to produce and check fixtures and to time
 * the CPU baseline.  Mirrors the reference main loop
 * (DPSNN_mainUnderMPI.c:64-94): init, then fire() until it returns 1.
 *
 *   dpsnn_cpu_ref <inidir> <outdir> [--dump] [--steps K]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "dpsnn_oracle.h"

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
  char err[512] = {0};
  int dump = 0, i;
  long steps = -1;
  double t0, t1, t2;
  uint64_t nspk = 0;
  ora_net *n;
  if (argc < 3) { fprintf(stderr, "usage: %s inidir outdir [--dump] [--steps K]\n", argv[0]); return 2; }
  for (i = 3; i < argc; i++) {
    if (!strcmp(argv[i], "--dump")) dump = 1;
    else if (!strcmp(argv[i], "--steps") && i + 1 < argc) steps = atol(argv[++i]);
  }
  t0 = now_s();
  n = ora_create(argv[1], err, sizeof err);
  if (!n) { fprintf(stderr, "ERROR: %s\n", err); return 1; }
  t1 = now_s();
  printf("init %.3f s: N=%u S=%llu\n", t1 - t0, ora_num_neurons(n), (unsigned long long)ora_num_synapses(n));
  if (dump) {
    char path[2048];
    snprintf(path, sizeof path, "%s/localNet_h0_N%u_SynapticList_BACKWARD_R-1.dat", argv[2], ora_num_neurons(n));
    ora_write_synapse_dump(n, path);
  }
  ora_set_output_dir(n, argv[2]);
  for (;;) {
    if (steps >= 0 && (long)ora_now_ms(n) >= steps) break;
    if (ora_fire(n)) break;
    nspk += ora_last_spikes(n, NULL, NULL);
  }
  t2 = now_s();
  printf("steps %u in %.3f s (%.6f s/step), spikes %llu, synaptic events %llu\n", ora_now_ms(n), t2 - t1,
         (t2 - t1) / (ora_now_ms(n) ? ora_now_ms(n) : 1), (unsigned long long)nspk,
         (unsigned long long)ora_synaptic_events(n));
  ora_destroy(n);
  return 0;
}

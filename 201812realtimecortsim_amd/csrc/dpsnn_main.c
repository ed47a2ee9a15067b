/* dpsnn_main.c -- the reference main loop (DPSNN_mainUnderMPI.c:33-102)
 * without MPI: init, fire() until it returns 1, finish, with the same
 * CHRONO report lines. */
#include <stdio.h>
#include <time.h>

#include "localNetProcess_hip.h"

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static struct _local_states lnp_ls;
static DALProcess theProcess;

int main(void) {
  DALProcess *p = &theProcess;
  double t0, t1, fireTotal = 0;
  int stop, laps = 0;
  p->local = &lnp_ls;
  t0 = now_s();
  localNetProcess_init(p);
  t1 = now_s();
  printf("CHRONO:MAIN h=0 - localNetProcess_init = %f sec \n", t1 - t0);
  do {
    double a = now_s();
    stop = localNetProcess_fire(p);
    fireTotal += now_s() - a;
    ++laps;
  } while (stop == 0);
  localNetProcess_finish(p);
  printf("CHRONO:MAIN h=0 - mean localNetProcess_fire = %f sec on %d samples total fire time = %f\n",
         fireTotal / laps, laps, fireTotal);
  return 0;
}

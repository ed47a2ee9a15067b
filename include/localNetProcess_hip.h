/* localNetProcess_hip.h -- DAL-shaped process interface of the engine.
 *
 * Same shape and return convention as the reference's process interface
 *   DALProcess            dal.h:25-42
 *   localNetProcess_init  localNetProcess.c:40-100  (localNetProcess.h:72)
 *   localNetProcess_fire  localNetProcess.c:102-158 (localNetProcess.h:73)
 *   localNetProcess_finish localNetProcess.c:160-173 (localNetProcess.h:74)
 * so a main identical in behaviour to DPSNN_mainUnderMPI.c:33-102 (without
 * MPI: one process owns its whole local network on one GPU) drives it.
 * fire() returns 0 to continue and 1 once env_totalSimTime_ms is reached.
 * Parameters come from the environment and the ini files in the current
 * directory (or $DPSNN_INIDIR); output files go to the current directory
 * (or $DPSNN_OUTDIR).  On a configuration or device error init/fire print
 * the message and exit(1) (the reference prints and exits with 0).
 */
#ifndef LOCALNETPROCESS_HIP_H
#define LOCALNETPROCESS_HIP_H

#include "dpsnn_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _local_states *LocalState;
struct _process;
typedef void (*ProcessInit)(struct _process *);
typedef void (*ProcessFinish)(struct _process *);
typedef int (*ProcessFire)(struct _process *);
typedef void *WPTR;

typedef struct _process {
  LocalState local;
  ProcessInit init;
  ProcessFire fire;
  ProcessFinish finish;
  WPTR wptr;
} DALProcess;

/* The process's local state (the reference keeps localNetProcess_localState,
 * localNetProcess.h:49-62; here the network handle replaces the C++ objects). */
struct _local_states {
  dpsnn_parameters parameters;
  dpsnn_net *net;
  int id;
  char name[32];
  int fireCount;
  double fireSeconds;  /* accumulated wall time inside fire() */
};

void localNetProcess_init(DALProcess *p);
int localNetProcess_fire(DALProcess *p);
void localNetProcess_finish(DALProcess *p);

#ifdef __cplusplus
}
#endif
#endif

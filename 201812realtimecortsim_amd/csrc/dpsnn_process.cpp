// dpsnn_process.cpp -- the reference's DAL process interface over the C ABI
// (localNetProcess.c:40-173): init reads parameters and builds the network,
// fire advances one millisecond, finish drains the output files and frees
// the device.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "localNetProcess_hip.h"

static void die(const char *what) {
  std::printf("ERROR %s: %s\n", what, dpsnn_last_error());
  std::fflush(stdout);
  std::exit(1);
}

extern "C" void localNetProcess_init(DALProcess *p) {
  LocalState ls = p->local;
  ls->id = 0;
  std::snprintf(ls->name, sizeof ls->name, "C0- localNetProcess_%d", ls->id);
  const char *inidir = std::getenv("DPSNN_INIDIR");
  const char *outdir = std::getenv("DPSNN_OUTDIR");
  if (!inidir) inidir = ".";
  if (!outdir) outdir = ".";
  if (dpsnn_get_parameters(inidir, &ls->parameters) < 0) die("getParameters");
  dpsnn_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.device = std::getenv("DPSNN_DEVICE") ? std::atoi(std::getenv("DPSNN_DEVICE")) : 0;
  opt.drive = DPSNN_DRIVE_REPLAY;
  if (const char *d = std::getenv("DPSNN_DRIVE")) opt.drive = std::strcmp(d, "philox") == 0 ? DPSNN_DRIVE_PHILOX : DPSNN_DRIVE_REPLAY;
  opt.philox_seed = ls->parameters.globalSeed;
  opt.output_dir = outdir;
  if (dpsnn_create(inidir, &opt, &ls->net) < 0) die("localNet init");
  ls->fireCount = 0;
  ls->fireSeconds = 0;
  std::printf("500- %s: init() - completed - from now it will fire()\n", ls->name);
  std::fflush(stdout);
}

extern "C" int localNetProcess_fire(DALProcess *p) {
  LocalState ls = p->local;
  auto t0 = std::chrono::steady_clock::now();
  const int rc = dpsnn_step(ls->net, 1);
  if (rc < 0) die("fire");
  ls->fireSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  ++ls->fireCount;
  return rc == DPSNN_DONE ? 1 : 0;
}

extern "C" void localNetProcess_finish(DALProcess *p) {
  LocalState ls = p->local;
  if (dpsnn_sync(ls->net) < 0) die("finish");
  dpsnn_destroy(ls->net);
  ls->net = nullptr;
}

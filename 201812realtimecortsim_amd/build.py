"""Builds libdpsnn_hip.so (the C-ABI engine) in-tree for gfx950.

    python -m 201812realtimecortsim_amd.build     (or import and call build())

Host sources are compiled with g++ (no FMA contraction, glibc libm, like the
reference's g++ -O3 build); device sources with hipcc for gfx950 with
-ffp-contract=off so the FP64 LIFCA update rounds exactly where the
reference rounds, and with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OUTDIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUTDIR, "libdpsnn_hip.so")
MAIN = os.path.join(OUTDIR, "dpsnn_hip_main")

ARCH = os.environ.get("DPSNN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-pthread"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (need ROCm)")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    return proc.stdout


def build(verbose=False):
    os.makedirs(OUTDIR, exist_ok=True)
    objdir = os.path.join(OUTDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", INCLUDE, "-I", CSRC]
    hipcc = _hipcc()
    host_obj = os.path.join(objdir, "dpsnn_host.o")
    _run(["g++", *HOST_FLAGS, *inc, "-c", os.path.join(CSRC, "dpsnn_host.cpp"), "-o", host_obj])
    objs = [host_obj]
    for src in ("dpsnn_engine.hip", "dpsnn_process.cpp", "dpsnn_server.cpp"):
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            _run([hipcc, *HIP_FLAGS, *inc, "-c", path, "-o", obj])
        else:
            _run(["g++", *HOST_FLAGS, *inc, "-c", path, "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    main_src = os.path.join(CSRC, "dpsnn_main.c")
    if os.path.exists(main_src):
        _run(["gcc", "-O2", "-Wall", *inc, main_src, "-o", MAIN, f"-L{OUTDIR}", "-ldpsnn_hip",
              f"-Wl,-rpath,{OUTDIR}", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

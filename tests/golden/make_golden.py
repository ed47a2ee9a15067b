"""Writes the golden fixtures of tests/golden/ from runs of the reference.

The reference (APE-group/201812RealTimeCortSim, DPSNN) is built from its own
sources by oracle/ref.mk into oracle/_ref/LIFCA-DPSNN.out and run at MPI
world size 1 on the canonical ini renderings in
201812realtimecortsim_amd/configs/.  Its own output files are reduced to
small fixtures:

  c0_synapses.json  the BACKWARD synapse dump (delivery order): sha256, line
                    count, per-delay counts, head/tail and a strided sample
  c0_raster.json    spikes-sec0 raster of the 2x2-column network (1 s): sha256 of the
                    whole file, line counts and sha256 of each prefix that
                    ends at 100/200/500/1000 ms, per-ms spike counts
  c0_ratesPerPop-sec0.dat   the reference's ratesPerPop file, verbatim
  c2small_*.json / .dat     the same for the STDP network's plasticity at the 2x2-column
                    size (1.2 s), plus sha256 of the synPeriodicProbe A/B
                    dumps written at the end-of-second weight commit

Usage (in the build container, where /root/reference exists):
    python tests/golden/make_golden.py [--rundir-c0 DIR] [--rundir-c2s DIR]
Without --rundir-*, the reference is built and run in a temporary directory
(about 10 and 16 minutes on one core).
"""
import argparse
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CONFIGS = os.path.join(ROOT, "201812realtimecortsim_amd", "configs")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "LIFCA-DPSNN.out")
MPIRUN = "/opt/conda/bin/mpirun"


def sha256_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def read_env(path):
    env = {}
    for line in open(path):
        line = line.strip()
        if line and "=" in line:
            k, v = line.split("=", 1)
            env[k] = v
    return env


def run_reference(cfg, rundir, extra_env=None):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(rundir, exist_ok=True)
    for name in os.listdir(cfg):
        if name.endswith(".txt") and name != "env.txt":
            shutil.copy(os.path.join(cfg, name), rundir)
    env = dict(os.environ)
    env.update(read_env(os.path.join(cfg, "env.txt")))
    env.update(extra_env or {})
    with open(os.path.join(rundir, "stdout.txt"), "w") as out:
        subprocess.run([MPIRUN, "-np", "1", REF_BIN], cwd=rundir, env=env, stdout=out,
                       stderr=subprocess.STDOUT, check=True)


def raster_summary(path, checkpoints):
    """Per-ms counts (a spike written in step t has Te in [t-1, t)) and the
    sha256 of the prefix of lines written by steps < checkpoint."""
    lines = open(path, "rb").read().split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    counts = {}
    steps = []
    for ln in lines:
        te = float(ln.split()[1])
        step = int(te) + 1
        steps.append(step)
        counts[step] = counts.get(step, 0) + 1
    prefix = {}
    for cp in checkpoints:
        n = sum(1 for s in steps if s < cp)
        h = hashlib.sha256(b"".join(ln + b"\n" for ln in lines[:n])).hexdigest()
        prefix[str(cp)] = {"lines": n, "sha256": h}
    last = max(steps) if steps else 0
    return {"sha256": sha256_file(path), "lines": len(lines), "prefix": prefix,
            "per_ms_counts": [counts.get(t, 0) for t in range(last + 1)]}


def synapse_summary(path, n_sample=2000):
    per_delay = {}
    head, tail, sample = [], [], []
    total = 0
    with open(path) as f:
        for i, line in enumerate(f):
            d = int(line.split()[3])
            per_delay[d] = per_delay.get(d, 0) + 1
            if i < 20:
                head.append(line.rstrip("\n"))
            tail.append(line.rstrip("\n"))
            if len(tail) > 20:
                tail.pop(0)
            total += 1
    stride = max(1, total // n_sample)
    with open(path) as f:
        for i, line in enumerate(f):
            if i % stride == 0:
                sample.append([i, line.rstrip("\n")])
    return {"sha256": sha256_file(path), "lines": total, "per_delay": per_delay,
            "head": head, "tail": tail, "stride": stride, "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rundir-c0")
    ap.add_argument("--rundir-c2s")
    args = ap.parse_args()
    tmp = tempfile.mkdtemp(prefix="dpsnn_ref_")
    c0 = args.rundir_c0 or os.path.join(tmp, "c0")
    if not args.rundir_c0:
        run_reference(os.path.join(CONFIGS, "c0"), c0)
    c2s = args.rundir_c2s or os.path.join(tmp, "c2s")
    if not args.rundir_c2s:
        run_reference(os.path.join(CONFIGS, "c2_small"), c2s,
                      {"env_statSynPeriodicProbe_writeDisable": "0",
                       "env_startSynPeriodicProbe_s": "0", "env_stopSynPeriodicProbe_s": "0"})
    dump = os.path.join(c0, "localNet_h0_N5000_SynapticList_BACKWARD_R-1.dat")
    json.dump(synapse_summary(dump), open(os.path.join(HERE, "c0_synapses.json"), "w"), indent=0)
    json.dump(raster_summary(os.path.join(c0, "spikes-sec0-C4-H1-h000.dat"), [30, 100, 200, 500, 1000]),
              open(os.path.join(HERE, "c0_raster.json"), "w"))
    shutil.copy(os.path.join(c0, "ratesPerPop-sec0-C4-H1-h000.dat"),
                os.path.join(HERE, "c0_ratesPerPop-sec0.dat"))
    out = {}
    for sec in (0, 1):
        out[f"spikes-sec{sec}"] = raster_summary(os.path.join(c2s, f"spikes-sec{sec}-C4-H1-h000.dat"),
                                                 [1000 * sec + 100, 1000 * sec + 200])
        shutil.copy(os.path.join(c2s, f"ratesPerPop-sec{sec}-C4-H1-h000.dat"),
                    os.path.join(HERE, f"c2small_ratesPerPop-sec{sec}.dat"))
    for probe in ("A", "B"):
        p = os.path.join(c2s, f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat")
        out[f"synPeriodicProbe_{probe}"] = synapse_summary(p, n_sample=500)
    json.dump(out, open(os.path.join(HERE, "c2small_golden.json"), "w"))
    print("fixtures written to", HERE)
    return 0


if __name__ == "__main__":
    sys.exit(main())

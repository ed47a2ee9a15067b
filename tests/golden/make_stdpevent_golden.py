"""Writes tests/golden/stdpevent_small.json from a run of the reference.

The reference (DPSNN, built from its own sources by oracle/ref.mk into
oracle/_ref/LIFCA-DPSNN.out) runs the STDP network's plasticity on a 2 x 2 grid of
125-neuron columns for 100 ms at MPI world size 1 with the STDPevent file on
(env_statSTDPevent_writeDisable=0) and the reference's debug window at 60 ms
(env_debugPrintEnable_ms=60: contributions of ms 60..80 are written,
statSTDPeventClass::write, DPSNN_stat.c:451-470).  The fixture keeps the
file's sha256, line count, lines per ms, head, tail and a strided sample,
and the sha256 of the run's raster.

Usage (in the build container, where /root/reference exists):
    python tests/golden/make_stdpevent_golden.py
"""
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

COLUMN = ("0 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
          "1 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
          "2 25 0.5 0.0 3 10 20 15 2 0 500 0\n")
ENV = {"env_totalSimTime_ms": "100", "env_statSTDPevent_writeDisable": "0",
       "env_debugPrintEnable_ms": "60", "env_startSTDPevent_s": "0", "env_stopSTDPevent_s": "0"}
FILE = "STDPevent-sec0-C4-H1-h000.dat"


def read_env(path):
    env = {}
    for line in open(path):
        line = line.strip()
        if line and "=" in line:
            k, v = line.split("=", 1)
            env[k] = v
    return env


def write_network(d):
    """The run directory's ini files; returns the env vars of the run."""
    os.makedirs(d, exist_ok=True)
    src = os.path.join(CONFIGS, "c2_small")
    for name in ("connectivity.txt", "bathEfficacy.txt", "plasticity.txt"):
        shutil.copy(os.path.join(src, name), d)
    with open(os.path.join(d, "column.txt"), "w") as f:
        f.write(COLUMN)
    env = read_env(os.path.join(src, "env.txt"))
    env.update(ENV)
    return env


def summary(path, n_sample=400):
    import hashlib
    h = hashlib.sha256()
    per_ms, head, tail, lines = {}, [], [], []
    with open(path, "rb") as f:
        data = f.read()
    h.update(data)
    lines = data.decode().splitlines()
    for ln in lines:
        t = int(ln.split()[0])
        per_ms[t] = per_ms.get(t, 0) + 1
    stride = max(1, len(lines) // n_sample)
    return {"sha256": h.hexdigest(), "lines": len(lines), "per_ms": per_ms, "head": lines[:10],
            "tail": lines[-10:], "stride": stride, "sample": [[i, lines[i]] for i in range(0, len(lines), stride)]}


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    d = tempfile.mkdtemp(prefix="dpsnn_stdpev_")
    env = dict(os.environ)
    env.update(write_network(d))
    with open(os.path.join(d, "stdout.txt"), "w") as out:
        subprocess.run([MPIRUN, "-np", "1", REF_BIN], cwd=d, env=env, stdout=out, stderr=subprocess.STDOUT,
                       check=True)
    res = summary(os.path.join(d, FILE))
    import hashlib
    res["raster_sha256"] = hashlib.sha256(open(os.path.join(d, "spikes-sec0-C4-H1-h000.dat"), "rb").read()).hexdigest()
    with open(os.path.join(HERE, "stdpevent_small.json"), "w") as f:
        json.dump(res, f)
    print("stdpevent_small.json:", res["lines"], "lines")
    return 0


if __name__ == "__main__":
    sys.exit(main())

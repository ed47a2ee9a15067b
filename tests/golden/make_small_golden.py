"""Writes tests/golden/small_golden.json from runs of the reference.

The reference (DPSNN, built from its own sources by oracle/ref.mk into
oracle/_ref/LIFCA-DPSNN.out) runs at MPI world size 1 on two networks that
the other fixtures leave open:

  djext       the 2x2-column network (2 x 2 columns of 1250 neurons, seed 1) with the
              bath efficacy spread DJExt = 0.1 in every sub-population
              (column.txt field 4), 100 ms: the Gaussian bath LUT's
              DJExt > 0 branch (erflib.c:141-169, used by the external drive
              at DPSNN_neuron_sim.c:604-640).  Raster summary (prefixes at
              50 and 100 ms) and the ratesPerPop file verbatim.
              Also the connection-probability matrices the reference
              prints at init (probConnectivityMatrix, DPSNN_LIFCAconnectome
              .cpp:181-192, "%f", rows y, columns x) for the 2x2-column network's
              connectivity: the values writeIniFiles puts in PerseoConFile.
  stdp_small  the STDP network's plasticity on a 2 x 2 grid of 125-neuron columns for
              1100 ms, across the end-of-second weight commit, with the
              synPeriodicProbe dumps on: raster summaries of both seconds,
              both rates files verbatim, sha256 of synPeriodicProbe A/B.

Both networks are written by the functions below, which the tests import
to build the same ini files.

Usage (in the build container, where /root/reference exists):
    python tests/golden/make_small_golden.py
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
FIXTURE = os.path.join(HERE, "small_golden.json")

DJEXT_MS = 100
STDP_SMALL_MS = 1100


def read_env(path):
    env = {}
    for line in open(path):
        line = line.strip()
        if line and "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            env[k.strip()] = v.strip()
    return env


def _copy_ini(src, d, names):
    os.makedirs(d, exist_ok=True)
    for name in names:
        shutil.copy(os.path.join(CONFIGS, src, name), d)


def write_djext(d):
    """the 2x2-column network with DJExt = 0.1; returns the run's env settings."""
    _copy_ini("c0", d, ("connectivity.txt", "bathEfficacy.txt"))
    rows = []
    for line in open(os.path.join(CONFIGS, "c0", "column.txt")):
        f = line.split()
        if f:
            f[3] = "0.1"
            rows.append(" ".join(f) + "\n")
    with open(os.path.join(d, "column.txt"), "w") as out:
        out.write("".join(rows))
    env = read_env(os.path.join(CONFIGS, "c0", "env.txt"))
    env["env_totalSimTime_ms"] = str(DJEXT_MS)
    return env


def write_stdp_small(d):
    """the STDP network's plasticity on 2 x 2 columns of 125 neurons; returns the env."""
    _copy_ini("c2_small", d, ("connectivity.txt", "bathEfficacy.txt", "plasticity.txt"))
    with open(os.path.join(d, "column.txt"), "w") as f:
        f.write("0 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                "1 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                "2 25 0.5 0.0 3 10 20 15 2 0 500 0\n")
    env = read_env(os.path.join(CONFIGS, "c2_small", "env.txt"))
    env.update({"env_totalSimTime_ms": str(STDP_SMALL_MS), "env_statSynPeriodicProbe_writeDisable": "0",
                "env_startSynPeriodicProbe_s": "0", "env_stopSynPeriodicProbe_s": "0"})
    return env


def _run(d, env_over):
    env = dict(os.environ)
    env.update(env_over)
    with open(os.path.join(d, "stdout.txt"), "w") as out:
        subprocess.run([MPIRUN, "-np", "1", REF_BIN], cwd=d, env=env, stdout=out, stderr=subprocess.STDOUT,
                       check=True)


def parse_prob_matrices(stdout_path):
    """{"s,t": rows of the printed matrix} from the reference's stdout."""
    out, key, rows = {}, None, []
    for line in open(stdout_path):
        if "probConnectivityMatrix for sourceSubPop=" in line:
            f = line.replace("=", " ").replace(":", " ").split()
            key = f"{int(f[f.index('sourceSubPop') + 1])},{int(f[f.index('targetSubPop') + 1])}"
            rows = []
            out[key] = rows
        elif key is not None:
            vals = line.split()
            if vals and all(v.replace(".", "", 1).isdigit() for v in vals):
                rows.append(vals)
            elif rows:
                key = None
    return out


def main():
    sys.path.insert(0, HERE)
    from make_golden import raster_summary, sha256_file  # noqa: E402
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    tmp = tempfile.mkdtemp(prefix="dpsnn_small_")
    out = {}
    d = os.path.join(tmp, "djext")
    _run(d, write_djext(d))
    out["djext"] = {"raster": raster_summary(os.path.join(d, "spikes-sec0-C4-H1-h000.dat"), [50, 100]),
                    "rates": open(os.path.join(d, "ratesPerPop-sec0-C4-H1-h000.dat")).read(),
                    "prob_matrices": parse_prob_matrices(os.path.join(d, "stdout.txt"))}
    d = os.path.join(tmp, "stdp_small")
    _run(d, write_stdp_small(d))
    s = {}
    for sec in (0, 1):
        s[f"spikes-sec{sec}"] = raster_summary(os.path.join(d, f"spikes-sec{sec}-C4-H1-h000.dat"),
                                               [1000 * sec + 100])
        s[f"rates-sec{sec}"] = open(os.path.join(d, f"ratesPerPop-sec{sec}-C4-H1-h000.dat")).read()
    for probe in ("A", "B"):
        s[f"synPeriodicProbe_{probe}"] = sha256_file(os.path.join(d, f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat"))
    out["stdp_small"] = s
    with open(FIXTURE, "w") as f:
        json.dump(out, f)
    print("written", FIXTURE, "from", tmp)
    return 0


if __name__ == "__main__":
    sys.exit(main())

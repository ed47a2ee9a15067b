"""Writes c0_philox.json: the raster, rates and counters of the 2x2-column network's
network under the on-device Philox drive (seed 20181201), 500 ms.  The
Philox drive has no counterpart in the reference (it replaces the host's
ran3 bath with a counter-based stream of the same statistics), so this
fixture is a regression vector of the engine itself: it was written by the
engine before the bath draws were moved out of phase 2's merge loop, and the
test holds later builds to it.  Usage (on a machine with the GPU):
    python tests/golden/make_philox_golden.py [library.so]"""
import hashlib
import importlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def run(pkg, out_dir):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=500)
    with pkg.Network(cfg, drive=pkg.DRIVE_PHILOX, philox_seed=20181201, output_dir=out_dir) as net:
        while not net.step(100):
            pass
        net.sync()
        st = net.stats()
        state = net.state()
    files = {}
    for f in sorted(os.listdir(out_dir)):
        with open(os.path.join(out_dir, f), "rb") as fh:
            files[f] = hashlib.sha256(fh.read()).hexdigest()
    h = hashlib.sha256()
    for k in ("v", "c", "Tr", "Te"):
        h.update(state[k].tobytes())
    return {"files": files, "state_sha256": h.hexdigest(),
            "stats": {k: int(st[k]) for k in ("steps", "spikes", "synaptic_events", "external_events")}}


if __name__ == "__main__":
    pkg = importlib.import_module("201812realtimecortsim_amd")
    if len(sys.argv) > 1:
        pkg.LIB_PATH = os.path.abspath(sys.argv[1])
    with tempfile.TemporaryDirectory() as d:
        res = run(pkg, d)
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(HERE, "c0_philox.json")
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res["stats"]))

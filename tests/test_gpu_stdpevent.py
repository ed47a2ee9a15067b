"""The STDPevent file (statSTDPeventClass::write, DPSNN_stat.c:451-470):
every LTP and LTD contribution of the reference's debug window, byte for
byte as the reference wrote it (fixture from oracle/_ref)."""
import importlib.util
import os

import pytest

from helpers import GOLDEN, golden, sha256_file

pytestmark = pytest.mark.gpu


def _fixture_module():
    spec = importlib.util.spec_from_file_location("make_stdpevent_golden",
                                                  os.path.join(GOLDEN, "make_stdpevent_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_stdpevent_file_identical_to_reference(pkg, tmp_path):
    mod = _fixture_module()
    g = golden("stdpevent_small.json")
    d = str(tmp_path / "net")
    env = mod.write_network(d)
    out = tmp_path / "out"
    out.mkdir()
    print("STDPevent: 2x2 x 125 neurons with STDP, 100 ms", flush=True)
    with pkg.Network(d, output_dir=str(out), env=env) as net:
        while not net.step(50):
            pass
        net.sync()
    assert sha256_file(out / "spikes-sec0-C4-H1-h000.dat") == g["raster_sha256"]
    path = out / mod.FILE
    lines = open(path).read().splitlines()
    assert len(lines) == g["lines"]
    assert lines[:10] == g["head"] and lines[-10:] == g["tail"]
    for i, ln in g["sample"]:
        assert lines[i] == ln, i
    assert sha256_file(path) == g["sha256"]

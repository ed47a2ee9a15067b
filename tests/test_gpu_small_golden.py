"""The engine against the reference runs of small_golden.json: the bath
LUT's DJExt > 0 branch on the 2x2-column network and the STDP network's plasticity on a tiny grid
across the end-of-second commit, on both phase-1 forms."""
import os

import pytest

from helpers import golden, sha256_file, small_network

pytestmark = pytest.mark.gpu


def _run(pkg, kind, tmp_path, window, monkeypatch):
    monkeypatch.setenv("DPSNN_WINDOW", window)
    d = tmp_path / kind
    env = small_network(kind, d)
    out = tmp_path / "out"
    out.mkdir()
    print(f"{kind}: window={window}", flush=True)
    with pkg.Network(str(d), output_dir=str(out), env=env) as net:
        while not net.step(100):
            pass
        st = net.stats()
    return out, st


@pytest.mark.parametrize("window", ["0", "1"])
def test_djext_bath_identical_to_reference(pkg, tmp_path, monkeypatch, window):
    g = golden("small_golden.json")["djext"]
    out, st = _run(pkg, "djext", tmp_path, window, monkeypatch)
    assert sha256_file(out / "spikes-sec0-C4-H1-h000.dat") == g["raster"]["sha256"]
    assert open(out / "ratesPerPop-sec0-C4-H1-h000.dat").read() == g["rates"]
    assert st["spikes"] == g["raster"]["lines"]


@pytest.mark.parametrize("window", ["0", "1"])
def test_stdp_small_across_commit_identical_to_reference(pkg, tmp_path, monkeypatch, window):
    g = golden("small_golden.json")["stdp_small"]
    out, _ = _run(pkg, "stdp_small", tmp_path, window, monkeypatch)
    for sec in (0, 1):
        assert sha256_file(out / f"spikes-sec{sec}-C4-H1-h000.dat") == g[f"spikes-sec{sec}"]["sha256"], sec
        assert open(out / f"ratesPerPop-sec{sec}-C4-H1-h000.dat").read() == g[f"rates-sec{sec}"], sec
    for probe in ("A", "B"):
        assert sha256_file(out / f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat") == g[f"synPeriodicProbe_{probe}"]
    assert sorted(os.listdir(out)) == sorted(
        [f"spikes-sec{s}-C4-H1-h000.dat" for s in (0, 1)] + [f"ratesPerPop-sec{s}-C4-H1-h000.dat" for s in (0, 1)] +
        [f"synPeriodicProbe_{p}-sec0-C4-H1-h000.dat" for p in ("A", "B")])

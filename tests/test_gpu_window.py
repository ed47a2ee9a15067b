"""GPU parity of the stencil-window form of phase 1 (k_collect + k_sort,
or the sort fused into k_collect; steps t >= 32 without LTD): forced onto
the 2x2-column network and small networks it must give the reference's raster and the
oracle's state bit for bit, like the gather form."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, OracleNet, golden, sha256_file

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_env(pkg):
    yield
    pkg.apply_env(pkg.config_dir("c0"))


def _columns(tmp, pkg, name, counts, src="c0", nu_ext=3.0):
    base = pkg.config_dir(src)
    d = tmp / name
    d.mkdir(exist_ok=True)
    e, e2, i = counts
    (d / "column.txt").write_text(
        f"0 {e} 0.5 0.0 {nu_ext} 20 20 15 2 1 500 0.05\n"
        f"1 {e2} 0.5 0.0 {nu_ext} 20 20 15 2 1 500 0.05\n"
        f"2 {i} 0.5 0.0 {nu_ext} 10 20 15 2 0 500 0\n")
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(base, f)).read())
    return str(d)


@pytest.mark.parametrize("window,sort_cap,fused", [("1", None, "0"), ("1", "200", "0"), ("1", None, "1"),
                                                  ("1", "200", "1"), ("0", None, "0")])
def test_c0_full_second_identical_to_reference(pkg, tmp_path, monkeypatch, window, sort_cap, fused):
    """sort_cap 200: targets with more than 200 records of a step (nearly all
    of the 2x2-column network's) are sorted by the second sort launch.  fused 1: the sort
    runs inside k_collect, per unit."""
    monkeypatch.setenv("DPSNN_WINDOW", window)
    monkeypatch.setenv("DPSNN_FUSED_SORT", fused)
    if sort_cap:
        monkeypatch.setenv("DPSNN_SORT_CAP", sort_cap)
    g = golden("c0_raster.json")
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=1000)
    print(f"window={window} sort_cap={sort_cap} fused={fused}: the 2x2-column network 1000 ms", flush=True)
    with pkg.Network(cfg, output_dir=str(tmp_path)) as net:
        while not net.step(100):
            pass
        net.sync()
        st = net.stats()
    assert sha256_file(tmp_path / "spikes-sec0-C4-H1-h000.dat") == g["sha256"]
    assert (open(tmp_path / "ratesPerPop-sec0-C4-H1-h000.dat").read() ==
            open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read())
    assert st["spikes"] == g["lines"]


def _trace_vs_oracle(pkg, oracle, cfg, steps, every):
    ora = OracleNet(oracle, cfg)
    try:
        with pkg.Network(cfg, keep_spikes=True) as net:
            for step in range(steps):
                ora.fire()
                net.step(1)
                ids, te = net.last_spikes()
                oids, ote = ora.last_spikes()
                assert np.array_equal(ids, oids) and np.array_equal(te, ote), step
                if step % every == 0 or step == steps - 1:
                    a, b = net.state(), ora.state()
                    for k in ("v", "c", "Tr", "Te"):
                        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), (step, k)
            st = net.stats()
    finally:
        ora.close()
    return st


@pytest.mark.parametrize("fused", ["0", "1"])
def test_window_state_trace_bit_exact_vs_oracle(pkg, oracle, tmp_path, monkeypatch, fused):
    monkeypatch.setenv("DPSNN_WINDOW", "1")
    monkeypatch.setenv("DPSNN_FUSED_SORT", fused)
    cfg = _columns(tmp_path, pkg, "tiny", (50, 50, 25))
    pkg.apply_env(pkg.config_dir("c0"), env_CFX=2, env_CFY=2, env_totalSimTime_ms=200)
    print(f"window=1 fused={fused}: 2x2 x 125 neurons, 200 ms vs oracle", flush=True)
    st = _trace_vs_oracle(pkg, oracle, cfg, 200, 1)
    assert st["spikes"] > 0


@pytest.mark.parametrize("window,fused", [("1", "0"), ("1", "1"), ("0", "0")])
def test_grid24_small_columns_bit_exact_vs_oracle(pkg, oracle, tmp_path, monkeypatch, window, fused):
    """The 24 x 24 grid and stencil (corner, edge and interior
    columns; 576 columns through the work queues) with 40-neuron columns,
    past the packed-scratch step: state and spikes are the oracle's."""
    monkeypatch.setenv("DPSNN_WINDOW", window)
    monkeypatch.setenv("DPSNN_FUSED_SORT", fused)
    cfg = _columns(tmp_path, pkg, "g24", (16, 16, 8), src="c1")
    pkg.apply_env(pkg.config_dir("c1"), env_totalSimTime_ms=80)
    print(f"window={window} fused={fused}: 24x24 x 40 neurons, 80 ms vs oracle", flush=True)
    st = _trace_vs_oracle(pkg, oracle, cfg, 80, 8)
    assert st["neurons"] == 576 * 40 and st["spikes"] > 0 and st["synaptic_events"] > 0

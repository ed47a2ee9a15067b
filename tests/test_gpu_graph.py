"""Graph-replayed plain steps (runs of 10 steps captured once and replayed
with the step counter on the device) against plain launches: two identical
networks, one stepped with graphs and one without (DPSNN_GRAPH=0), must stay
bit-identical; state is read back between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_env(pkg):
    yield
    pkg.apply_env(pkg.config_dir("c0"))


@pytest.mark.parametrize("window", ["0", "1"])
def test_graph_steps_identical_to_plain_launches(pkg, monkeypatch, window):
    monkeypatch.setenv("DPSNN_WINDOW", window)
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=400)
    kw = dict(drive=pkg.DRIVE_PHILOX, philox_seed=5)
    a = pkg.Network(cfg, **kw)
    monkeypatch.setenv("DPSNN_GRAPH", "0")
    b = pkg.Network(cfg, **kw)
    monkeypatch.delenv("DPSNN_GRAPH")
    try:
        a.profile(True)
        for chunk in (50, 37, 100, 13, 199):
            print(f"graph vs plain (window={window}): {chunk} steps from {a.now_ms}", flush=True)
            a.step(chunk)
            b.step(chunk)
            sa, sb = a.state(), b.state()
            for k in ("v", "c", "Tr", "Te"):
                assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64)), (a.now_ms, k)
            assert np.array_equal(sa["fired"], sb["fired"]), a.now_ms
        avg, launches = a.kernel_time()
        assert launches == 399 and avg > 0
        st_a, st_b = a.stats(), b.stats()
        for k in ("steps", "spikes", "synaptic_events", "external_events"):
            assert st_a[k] == st_b[k], k
        assert a.step(10) and b.step(10)  # done at 400 ms
    finally:
        a.close()
        b.close()

"""The allocation-free step (dpsnn_step_noalloc) against the ordinary step
(dpsnn_step), through the C ABI: two identical networks, one advanced by
each entry point, must stay bit-identical (state, spikes, counters).  Two
calls in a row on one device, the second on a smaller network, so that the
shared static scratch holds the first network's keys when the second runs."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _config(tmp, pkg, name, counts, grid):
    src = pkg.config_dir("c0")
    d = tmp / name
    d.mkdir(exist_ok=True)
    e, e2, i = counts
    (d / "column.txt").write_text(
        f"0 {e} 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
        f"1 {e2} 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
        f"2 {i} 0.5 0.0 3 10 20 15 2 0 500 0\n")
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(src, f)).read())
    return str(d)


def _compare(pkg, inidir, steps, env):
    kw = dict(drive=pkg.DRIVE_PHILOX, philox_seed=9, keep_spikes=True, env=env)
    with pkg.Network(inidir, **kw) as a, pkg.Network(inidir, **kw) as b:
        for k in range(0, steps, 10):
            print(f"noalloc compare: {inidir} ms {k}", flush=True)
            a.step(10)
            b.step_noalloc(10)
            sa, sb = a.state(), b.state()
            for key in ("v", "c", "Tr", "Te"):
                assert np.array_equal(sa[key].view(np.uint64), sb[key].view(np.uint64)), (key, k)
            ia, ta = a.last_spikes()
            ib, tb = b.last_spikes()
            assert np.array_equal(ia, ib) and np.array_equal(ta.view(np.uint64), tb.view(np.uint64)), k
        st_a, st_b = a.stats(), b.stats()
        for key in ("steps", "spikes", "synaptic_events", "external_events"):
            assert st_a[key] == st_b[key], key
        assert st_a["spikes"] > 0 and st_a["synaptic_events"] > 0
        return st_a


def test_noalloc_step_matches_step_then_smaller_network(pkg, tmp_path):
    src = pkg.config_dir("c0")
    pkg.apply_env(src)
    # the 2x2-column network (5000 neurons) past the packed-scratch step, then a network
    # a quarter of its size on the same device
    big = _compare(pkg, src, 60, {"env_totalSimTime_ms": 60})
    small_dir = _config(tmp_path, pkg, "small", (250, 250, 125), 2)
    small = _compare(pkg, small_dir, 40, {"env_totalSimTime_ms": 40})
    assert small["neurons"] < big["neurons"]


def test_noalloc_step_refuses_what_it_cannot_hold(pkg, tmp_path):
    src = pkg.config_dir("c0")
    pkg.apply_env(src)
    with pkg.Network(src, drive=pkg.DRIVE_REPLAY, env={"env_totalSimTime_ms": 10}) as net:
        with pytest.raises(pkg.DpsnnError, match="Philox"):
            net.step_noalloc(1)
    # a 3x3 grid of 1250-neuron columns needs more key scratch than 128 MiB
    big_dir = _config(tmp_path, pkg, "big", (500, 500, 250), 3)
    with pkg.Network(big_dir, drive=pkg.DRIVE_PHILOX, env={"env_CFX": 3, "env_CFY": 3,
                                                          "env_totalSimTime_ms": 10}) as net:
        with pytest.raises(pkg.DpsnnError, match="key scratch"):
            net.step_noalloc(1)
        assert net.now_ms == 0
        net.step(2)  # the ordinary step still works
        assert net.now_ms == 2

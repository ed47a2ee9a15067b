"""GPU parity: the HIP engine, through the C ABI, against the reference's
own outputs (golden fixtures) and against the oracle, bit-exact."""
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, OracleNet, golden, prefix_sha, sha256_file

pytestmark = pytest.mark.gpu


def _tiny_config(tmp, pkg, counts=(50, 50, 25), grid=2, nu_ext=3.0):
    """the 2x2-column network's parameters with smaller columns (the stencil keeps the
    same exp(-d) shape and 0.9 total connection probability)."""
    src = pkg.config_dir("c0")
    d = tmp / "tiny"
    d.mkdir(exist_ok=True)
    e, e2, i = counts
    (d / "column.txt").write_text(
        f"0 {e} 0.5 0.0 {nu_ext} 20 20 15 2 1 500 0.05\n"
        f"1 {e2} 0.5 0.0 {nu_ext} 20 20 15 2 1 500 0.05\n"
        f"2 {i} 0.5 0.0 {nu_ext} 10 20 15 2 0 500 0\n")
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (d / f).write_text(open(os.path.join(src, f)).read())
    pkg.apply_env(src, env_CFX=grid, env_CFY=grid)
    return str(d)


def test_device_exp_matches_host_libm(pkg):
    rng = np.random.default_rng(7)
    x = np.concatenate([-rng.random(300000) * 0.01, (rng.random(300000) * 2 - 1) * 700,
                        [0.0, -0.0, 1e-300, -1e-300, -708.5, 709.7]])
    y = pkg.device_exp(x)
    ref = np.array([math.exp(v) for v in x])
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))


def test_device_division_matches_ieee(pkg):
    """The dynamics divide by the time constants through their reciprocals
    (two FMA corrections); the quotients are the IEEE ones.  Random
    magnitudes and signs, event-like intervals, and quotients next to a
    rounding midpoint (built in extended precision), for the time constants
    of the workloads, awkward significands and random divisors."""
    rng = np.random.default_rng(11)
    divisors = [20.0, 10.0, 500.0, 2.0, 3.0, 7.0, 0.1, 0.3, 1e-3, 123.456, 1.5,
                float(np.nextafter(2.0, 0.0)), float(np.nextafter(1.0, 2.0))]
    a_all, b_all = [], []
    for b in divisors:
        n = 400000
        mag = 10.0 ** rng.uniform(-12, 6, n)
        a_all.append(mag * rng.choice([-1.0, 1.0], n))
        dt = rng.random(n) * 2.0
        a_all.append(np.concatenate([dt[: n // 2], dt[n // 2:] / 2]))
        # quotients next to a midpoint: a ~ b * (q + ulp(q) / 2)
        q = 10.0 ** rng.uniform(-8, 3, n)
        half = np.spacing(q).astype(np.longdouble) / 2
        jitter = rng.choice([-1, 0, 1], n).astype(np.longdouble) * np.spacing(q).astype(np.longdouble) * 2.0 ** -20
        a_all.append((np.longdouble(b) * (q.astype(np.longdouble) + half + jitter)).astype(np.float64))
        b_all.append(np.full(3 * n, b))
    n = 2000000
    a_all.append(10.0 ** rng.uniform(-10, 6, n) * rng.choice([-1.0, 1.0], n))
    b_all.append(10.0 ** rng.uniform(-3, 4, n))
    a = np.concatenate(a_all)
    b = np.concatenate(b_all)
    q = pkg.device_div(a, b)
    ref = a / b
    bad = np.flatnonzero(q.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, (bad.size, a[bad[:3]], b[bad[:3]])


def _run_c0(pkg, out, ms, **env):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=ms, **env)
    with pkg.Network(cfg, output_dir=str(out)) as net:
        while not net.step(50):
            pass
        net.sync()
        return net.stats()


def test_c0_raster_first_500ms_identical_to_reference(pkg, tmp_path):
    g = golden("c0_raster.json")
    st = _run_c0(pkg, tmp_path, 500)
    path = tmp_path / "spikes-sec0-C4-H1-h000.dat"
    with open(path, "rb") as f:
        nlines = f.read().count(b"\n")
    assert nlines == g["prefix"]["500"]["lines"]
    for cp in ("30", "100", "200", "500"):
        assert prefix_sha(path, g["prefix"][cp]["lines"]) == g["prefix"][cp]["sha256"], cp
    rates = open(tmp_path / "ratesPerPop-sec0-C4-H1-h000.dat").read().splitlines()
    ref = open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read().splitlines()
    assert rates == ref[:100]
    assert st["spikes"] == g["prefix"]["500"]["lines"]


def test_c0_full_second_identical_to_reference(pkg, tmp_path):
    g = golden("c0_raster.json")
    st = _run_c0(pkg, tmp_path, 1000)
    assert sha256_file(tmp_path / "spikes-sec0-C4-H1-h000.dat") == g["sha256"]
    assert (open(tmp_path / "ratesPerPop-sec0-C4-H1-h000.dat").read() ==
            open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read())
    # synaptic events = sum over spikes of the out-degree delivered
    assert st["steps"] == 1000 and st["spikes"] == g["lines"]


def test_c0_with_minimal_env_D_is_identical(pkg, tmp_path):
    # delays are stored 0..9: env_D = 10 covers them; the result cannot change
    g = golden("c0_raster.json")
    _run_c0(pkg, tmp_path, 100, env_D=10)
    assert prefix_sha(tmp_path / "spikes-sec0-C4-H1-h000.dat", g["prefix"]["100"]["lines"]) == \
        g["prefix"]["100"]["sha256"]


def test_state_trace_bit_exact_vs_oracle(pkg, oracle, tmp_path):
    """v, c, Tr, Te and fired flags of every neuron after every step."""
    cfg = _tiny_config(tmp_path, pkg)
    pkg.apply_env(pkg.config_dir("c0"), env_CFX=2, env_CFY=2, env_totalSimTime_ms=300)
    ora = OracleNet(oracle, cfg)
    with pkg.Network(cfg, keep_spikes=True) as net:
        for step in range(300):
            ora.fire()
            net.step(1)
            a, b = net.state(), ora.state()
            for k in ("v", "c", "Tr", "Te"):
                assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), (step, k)
            assert np.array_equal(a["fired"], b["fired"]), step
            ids, te = net.last_spikes()
            oids, ote = ora.last_spikes()
            assert np.array_equal(ids, oids) and np.array_equal(te, ote), step
    ora.close()


def test_zero_bath_rate_matches_oracle(pkg, oracle, tmp_path):
    """NuExt = 0 makes invNuExt infinite and the per-ms cap 0; the
    reference's do-while still emits one bath event per neuron at t = 0
    (DPSNN_neuron_sim.c:606-622) and never another."""
    cfg = _tiny_config(tmp_path, pkg, nu_ext=0.0)
    pkg.apply_env(pkg.config_dir("c0"), env_CFX=2, env_CFY=2, env_totalSimTime_ms=50)
    ora = OracleNet(oracle, cfg)
    with pkg.Network(cfg) as net:
        net.step(50)
        st = net.stats()
        s = net.state()
    for _ in range(50):
        ora.fire()
    o = ora.state()
    ora.close()
    assert st["external_events"] == st["neurons"] and st["spikes"] == 0
    for k in ("v", "c", "Tr", "Te"):
        assert np.array_equal(s[k].view(np.uint64), o[k].view(np.uint64)), k


def test_philox_drive_statistics(pkg, tmp_path):
    """The on-device Philox drive has the replayed drive's statistics."""
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=400)
    res = {}
    for drive in (pkg.DRIVE_REPLAY, pkg.DRIVE_PHILOX):
        with pkg.Network(cfg, drive=drive, philox_seed=12345) as net:
            net.step(400)
            res[drive] = net.stats()
    a, b = res[pkg.DRIVE_REPLAY], res[pkg.DRIVE_PHILOX]
    # 2.4 kHz per neuron bath over 400 ms, Poisson counts
    assert abs(b["external_events"] / a["external_events"] - 1) < 0.01
    assert abs(b["spikes"] / a["spikes"] - 1) < 0.05


def test_stdp_c2small_identical_to_reference(pkg, tmp_path):
    """the STDP network's plasticity at the 2x2-column size, 1.2 s: raster and rates of
    both seconds and the synapse probes around the t = 999 ms weight commit
    are the reference's, byte for byte."""
    g = golden("c2small_golden.json")
    cfg = pkg.config_dir("c2_small")
    pkg.apply_env(cfg, env_statSynPeriodicProbe_writeDisable=0, env_startSynPeriodicProbe_s=0,
                  env_stopSynPeriodicProbe_s=0)
    try:
        with pkg.Network(cfg, output_dir=str(tmp_path)) as net:
            while not net.step(100):
                pass
            net.sync()
    finally:
        pkg.apply_env(pkg.config_dir("c0"))
    for sec in (0, 1):
        assert sha256_file(tmp_path / f"spikes-sec{sec}-C4-H1-h000.dat") == g[f"spikes-sec{sec}"]["sha256"], sec
        assert (open(tmp_path / f"ratesPerPop-sec{sec}-C4-H1-h000.dat").read() ==
                open(os.path.join(GOLDEN, f"c2small_ratesPerPop-sec{sec}.dat")).read()), sec
    for probe in ("A", "B"):
        assert (sha256_file(tmp_path / f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat") ==
                g[f"synPeriodicProbe_{probe}"]["sha256"]), probe


@pytest.mark.parametrize("dense_max", [0, 400])
def test_phase1_general_path_identical_to_reference(pkg, tmp_path, monkeypatch, dense_max):
    """Phase 1's general path (register batches, gathered twice), for every
    neuron (0) or for those with more than 400 delivered events in a step
    (400, both paths in one step): the 2x2-column network's first second and the STDP network's
    STDP run at the 2x2-column size are still the reference's, byte for byte."""
    monkeypatch.setenv("DPSNN_PHASE1_DENSE_MAX", str(dense_max))
    g = golden("c0_raster.json")
    st = _run_c0(pkg, tmp_path, 1000)
    assert sha256_file(tmp_path / "spikes-sec0-C4-H1-h000.dat") == g["sha256"]
    assert st["spikes"] == g["lines"]
    gs = golden("c2small_golden.json")
    out = tmp_path / "stdp"
    out.mkdir()
    pkg.apply_env(pkg.config_dir("c2_small"), env_statSynPeriodicProbe_writeDisable=0,
                  env_startSynPeriodicProbe_s=0, env_stopSynPeriodicProbe_s=0)
    try:
        with pkg.Network(pkg.config_dir("c2_small"), output_dir=str(out)) as net:
            while not net.step(100):
                pass
            net.sync()
    finally:
        pkg.apply_env(pkg.config_dir("c0"))
    for sec in (0, 1):
        assert sha256_file(out / f"spikes-sec{sec}-C4-H1-h000.dat") == gs[f"spikes-sec{sec}"]["sha256"], sec
    for probe in ("A", "B"):
        assert sha256_file(out / f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat") == \
            gs[f"synPeriodicProbe_{probe}"]["sha256"], probe


def test_stdp_rejects_sub_ms_refractory(pkg, tmp_path):
    cfg = _tiny_config(tmp_path, pkg)
    col = (tmp_path / "tiny" / "column.txt")
    col.write_text(col.read_text().replace(" 2 1 500", " 0.5 1 500"))
    pkg.apply_env(pkg.config_dir("c0"), env_CFX=2, env_CFY=2, env_startPlasticity_ms=0,
                  env_stopPlasticity_ms=100000)
    try:
        with pytest.raises(pkg.DpsnnError):
            pkg.Network(cfg)
    finally:
        pkg.apply_env(pkg.config_dir("c0"))


def test_philox_drive_regression(pkg, tmp_path):
    """The on-device Philox bath reproduces its committed vector (the 2x2-column network,
    seed 20181201, 500 ms): raster, rates, counters and final state."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_philox_golden",
                                                  os.path.join(GOLDEN, "make_philox_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    try:
        got = mod.run(pkg, str(tmp_path))
    finally:
        pkg.apply_env(pkg.config_dir("c0"))
    assert got == golden("c0_philox.json")

"""The on-device Philox bath against the reference's replayed ran3 stream
(whose raster is the reference's, byte for byte): per sub-population rates
and mean ISI CV over configs[0]'s full second and a configs[1] window,
within 1%."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = importlib.import_module("201812realtimecortsim_amd.analysis")


def run_raster(pkg, cfg, drive, seed, ms):
    ids, te = [], []
    with pkg.Network(cfg, drive=drive, philox_seed=seed, keep_spikes=True) as net:
        for _ in range(ms):
            net.step(1)
            i, t = net.last_spikes()
            ids.append(i)
            te.append(t)
        st = net.stats()
    ids = np.concatenate(ids)
    assert len(ids) == st["spikes"]
    return ids, np.concatenate(te)


def _stats(pkg, cfg, drive, seed, ms):
    ids, te = run_raster(pkg, cfg, drive, seed, ms)
    counts = A.column_counts(cfg)
    cols = pkg.get_parameters(cfg)["globCFT"]
    return A.raster_stats(ids, te, A.subpop_of_neurons(counts, cols), float(ms))


def _report(name, ref, got):
    g = A.relative_gaps(ref, got)
    print(f"{name}: replay rate {ref['rate_hz']:.2f} Hz cv {ref['isi_cv']:.4f} | philox rate "
          f"{got['rate_hz']:.2f} Hz cv {got['isi_cv']:.4f} | max gap {g['max']:.4%}", flush=True)
    return g


def test_philox_matches_replay_on_c0_full_second(pkg):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=1000)
    ref = _stats(pkg, cfg, pkg.DRIVE_REPLAY, 1, 1000)
    gaps = []
    for seed in (11, 12, 13):
        got = _stats(pkg, cfg, pkg.DRIVE_PHILOX, seed, 1000)
        gaps.append(_report(f"configs[0] seed {seed}", ref, got)["max"])
    assert max(gaps) < 0.01, gaps


def test_philox_matches_replay_on_c1_window(pkg):
    cfg = pkg.config_dir("c1")
    ms = 200
    pkg.apply_env(cfg, env_totalSimTime_ms=ms)
    try:
        ref = _stats(pkg, cfg, pkg.DRIVE_REPLAY, 1, ms)
        got = _stats(pkg, cfg, pkg.DRIVE_PHILOX, 11, ms)
    finally:
        pkg.apply_env(pkg.config_dir("c0"))
    g = _report("configs[1] 200 ms", ref, got)
    assert g["max"] < 0.01, g

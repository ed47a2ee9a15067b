"""Raster statistics (201812realtimecortsim_amd/analysis.py) against
closed forms and brute force."""
import importlib

import numpy as np

A = importlib.import_module("201812realtimecortsim_amd.analysis")


def test_regular_trains_have_zero_cv_and_exact_rates():
    # neuron k fires every (k + 2) ms from 0.25 ms on, for 1 s
    ids, te = [], []
    for k in range(6):
        t = np.arange(0.25, 1000.0, k + 2.0)
        ids.append(np.full(len(t), k))
        te.append(t)
    ids, te = np.concatenate(ids), np.concatenate(te)
    sub = np.array([0, 0, 0, 1, 1, 1])
    st = A.raster_stats(ids, te, sub, 1000.0)
    assert st["isi_cv"] == 0.0
    counts = [len(np.arange(0.25, 1000.0, k + 2.0)) for k in range(6)]
    assert st["sub"][0]["rate_hz"] == sum(counts[:3]) / 3
    assert st["sub"][1]["rate_hz"] == sum(counts[3:]) / 3


def test_cv_matches_brute_force_and_poisson():
    rng = np.random.default_rng(3)
    n = 50
    ids, te = [], []
    for k in range(n):
        t = np.cumsum(rng.exponential(1.0 / (0.05 + 0.002 * k), size=200 + k))
        ids.append(np.full(len(t), k))
        te.append(t)
    ids, te = np.concatenate(ids), np.concatenate(te)
    perm = rng.permutation(len(ids))  # any order in, same answer
    sub = np.zeros(n, np.int64)
    st = A.raster_stats(ids[perm], te[perm], sub, float(te.max()))
    brute = []
    for k in range(n):
        isi = np.diff(np.sort(te[ids == k]))
        brute.append(np.std(isi, ddof=1) / np.mean(isi))
    assert abs(st["isi_cv"] - np.mean(brute)) < 1e-12
    assert abs(st["isi_cv"] - 1.0) < 0.05  # exponential intervals: CV 1


def test_subpop_layout_and_gaps():
    sub = A.subpop_of_neurons([3, 2, 1], 2)
    assert sub.tolist() == [0, 0, 0, 1, 1, 2, 0, 0, 0, 1, 1, 2]
    a = {"rate_hz": 10.0, "isi_cv": 0.5, "sub": [{"rate_hz": 10.0, "isi_cv": 0.5}]}
    b = {"rate_hz": 10.1, "isi_cv": 0.495, "sub": [{"rate_hz": 9.8, "isi_cv": 0.5}]}
    g = A.relative_gaps(a, b)
    assert abs(g["rate"] - 0.01) < 1e-12 and abs(g["isi_cv"] - 0.01) < 1e-12
    assert abs(g["max"] - 0.02) < 1e-12

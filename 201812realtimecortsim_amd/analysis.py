"""Raster statistics used to hold a drive to the reference's behaviour when
its stream is not the reference's (the on-device Philox bath): per
sub-population firing rates -- what ratesPerPop reports
(statSpikingRatesPerPopClass::write, DPSNN_stat.c:225-273, over the whole
run instead of per sampling window) -- and the coefficient of variation of
each neuron's inter-spike intervals, computed from the raster
(statSpikesClass::write, DPSNN_stat.c:275-296: one (neuron, emission time)
pair per spike).  The reference itself computes no ISI statistic.
"""
import numpy as np


def subpop_of_neurons(counts, columns):
    """Sub-population of every neuron: columns of sum(counts) neurons, each
    laid out sub-population after sub-population (column.txt order)."""
    one = np.repeat(np.arange(len(counts), dtype=np.int64), counts)
    return np.tile(one, columns)


def column_counts(inidir):
    """Neurons per sub-population of one column, from column.txt."""
    counts = []
    with open(f"{inidir}/column.txt") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2:
                counts.append(int(float(parts[1])))
    return counts


def raster_stats(ids, te, subpop, duration_ms, min_spikes=3):
    """Rates (Hz) and mean ISI CV per sub-population and overall.

    ids, te: every spike of the run (neuron, emission time in ms); subpop:
    sub-population of each neuron; duration_ms: simulated time.  A neuron's
    ISI CV is the standard deviation of its inter-spike intervals over their
    mean (sample deviation, n - 1); neurons with fewer than min_spikes spikes
    have none.  The CV of a sub-population is the mean over its neurons."""
    ids = np.asarray(ids, np.int64)
    te = np.asarray(te, np.float64)
    n = len(subpop)
    nsub = int(subpop.max()) + 1 if n else 0
    spikes = np.bincount(ids, minlength=n)
    order = np.lexsort((te, ids))
    ids_s, te_s = ids[order], te[order]
    same = ids_s[1:] == ids_s[:-1]
    isi = np.diff(te_s)[same]
    owner = ids_s[1:][same]
    k = np.bincount(owner, minlength=n).astype(np.float64)
    s1 = np.bincount(owner, weights=isi, minlength=n)
    s2 = np.bincount(owner, weights=isi * isi, minlength=n)
    ok = spikes >= min_spikes
    mean = np.divide(s1, k, out=np.zeros(n), where=k > 0)
    var = np.divide(s2 - k * mean * mean, k - 1, out=np.zeros(n), where=k > 1)
    cv = np.divide(np.sqrt(np.maximum(var, 0.0)), mean, out=np.zeros(n), where=mean > 0)
    out = {"spikes": int(spikes.sum()), "rate_hz": float(spikes.sum()) / n / (duration_ms / 1000.0),
           "isi_cv": float(cv[ok].mean()) if ok.any() else float("nan"), "sub": []}
    for s in range(nsub):
        m = subpop == s
        out["sub"].append({"rate_hz": float(spikes[m].sum()) / m.sum() / (duration_ms / 1000.0),
                           "isi_cv": float(cv[m & ok].mean()) if (m & ok).any() else float("nan")})
    return out


def relative_gaps(a, b):
    """|b / a - 1| of the overall and per-sub-population rates and ISI CVs."""
    def gap(x, y):
        return abs(y / x - 1.0) if x else float("inf")
    g = {"rate": gap(a["rate_hz"], b["rate_hz"]), "isi_cv": gap(a["isi_cv"], b["isi_cv"])}
    g["rate_sub"] = [gap(x["rate_hz"], y["rate_hz"]) for x, y in zip(a["sub"], b["sub"])]
    g["isi_cv_sub"] = [gap(x["isi_cv"], y["isi_cv"]) for x, y in zip(a["sub"], b["sub"])]
    g["max"] = max([g["rate"], g["isi_cv"]] + g["rate_sub"] + g["isi_cv_sub"])
    return g

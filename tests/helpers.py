"""Shared helpers of the test suite (oracle wrappers, raster hashing)."""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def small_network(kind, d):
    """Writes the ini files of one of the networks of small_golden.json
    ("djext" or "stdp_small", tests/golden/make_small_golden.py) into d;
    returns the run's env settings."""
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    import make_small_golden as m
    return {"djext": m.write_djext, "stdp_small": m.write_stdp_small}[kind](str(d))


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleNet:
    """One oracle network (CPU restatement)."""

    def __init__(self, oracle, inidir, backward=None):
        """backward: {"pre", "post", "delay", "weight"} arrays of a backward
        synapse list to take instead of generating the connectome."""
        self.lib = oracle.lib
        err = ctypes.create_string_buffer(512)
        if backward is None:
            self.h = self.lib.ora_create(inidir.encode(), err, 512)
        else:
            b = {k: np.ascontiguousarray(backward[k]) for k in ("pre", "post", "delay", "weight")}
            self.h = self.lib.ora_create_from_backward(inidir.encode(), len(b["pre"]), ptr(b["pre"]), ptr(b["post"]),
                                                       ptr(b["delay"]), ptr(b["weight"]), err, 512)
        if not self.h:
            raise RuntimeError("oracle create failed: " + err.value.decode())
        self.n = self.lib.ora_num_neurons(self.h)
        self.s = self.lib.ora_num_synapses(self.h)

    def load_checkpoint(self, path):
        err = ctypes.create_string_buffer(512)
        if self.lib.ora_load_checkpoint(self.h, str(path).encode(), err, 512) != 0:
            raise RuntimeError("oracle checkpoint load failed: " + err.value.decode())

    def close(self):
        if self.h:
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def backward(self):
        s = self.s
        a = {"pre": np.empty(s, np.uint32), "post": np.empty(s, np.uint32), "delay": np.empty(s, np.uint8),
             "kind": np.empty(s, np.uint8), "weight": np.empty(s, np.int16), "tderiv": np.empty(s, np.float32),
             "last": np.empty(s, np.int32)}
        self.lib.ora_get_backward(self.h, *[ptr(a[k]) for k in ("pre", "post", "delay", "kind", "weight",
                                                                 "tderiv", "last")])
        return a

    def external_params(self):
        inv = np.empty(self.n, np.float64)
        mx = np.empty(self.n, np.uint32)
        self.lib.ora_get_external(self.h, ptr(inv), ptr(mx))
        return inv, mx

    def fire(self):
        return self.lib.ora_fire(self.h)

    def state(self):
        out = {k: np.empty(self.n, np.float64) for k in ("v", "c", "Tr", "Te")}
        out["fired"] = np.empty(self.n, np.uint8)
        self.lib.ora_get_state(self.h, *[ptr(out[k]) for k in ("v", "c", "Tr", "Te", "fired")])
        return out

    def last_spikes(self):
        ids = np.empty(self.n, np.uint32)
        te = np.empty(self.n, np.float64)
        n = self.lib.ora_last_spikes(self.h, ptr(ids), ptr(te))
        return ids[:n].copy(), te[:n].copy()

    def set_output_dir(self, d):
        self.lib.ora_set_output_dir(self.h, d.encode())

    def record_external(self, on=True):
        self.lib.ora_record_external(self.h, int(on))

    def last_external(self):
        cnt = np.empty(self.n, np.uint32)
        cap = self.n * 64
        t, v = np.empty(cap), np.empty(cap)
        n = self.lib.ora_last_external(self.h, ptr(cnt), ptr(t), ptr(v))
        return cnt, t[:n].copy(), v[:n].copy()

    def write_synapse_dump(self, path):
        assert self.lib.ora_write_synapse_dump(self.h, path.encode()) == 0


def sha256_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def prefix_sha(path, nlines):
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    return hashlib.sha256(b"".join(ln + b"\n" for ln in lines[:nlines])).hexdigest()


def raster_lines(ids, te):
    """Raster text of one step, formatted like statSpikesClass::write."""
    return "".join("%d %.7g\n" % (int(i), float(t)) for i, t in zip(ids, te))


def run_lengths(pre, delay, n, D):
    """[n, D] events per delivery of (source, run delay d): the run starts at
    the source's first synapse in delay bucket d of the backward list and
    continues while the source matches, past the bucket's end too
    (findAndGetNextSyn, DPSNN_synapse.c:48-136)."""
    S = len(pre)
    starts = np.flatnonzero(np.r_[True, (pre[1:] != pre[:-1]) | (delay[1:] != delay[:-1])])
    first = np.flatnonzero(np.r_[True, pre[1:] != pre[:-1]])
    end = np.r_[first[1:], S]
    which = np.searchsorted(first, starts, side="right") - 1
    deg = np.zeros((n, D), np.int64)
    deg[pre[starts], delay[starts]] = end[which] - starts
    return deg


def delivered_events(syn, n, spikes):
    """Synaptic events of steps 0..len(spikes)-1 from the backward list and
    the spike ids of each step's scan: step t delivers, for each run delay
    d <= t, the runs of the sources found by the scan of step t - d
    (DPSNN_localNet_sim.c:837-999)."""
    D = int(syn["delay"].max()) + 1 if len(syn["delay"]) else 1
    deg = run_lengths(syn["pre"], syn["delay"], n, D)
    return sum(int(deg[spikes[t - d], d].sum()) for t in range(len(spikes)) for d in range(min(D, t + 1)))


def check_raster(steps, n, tarp_ms=2.0, te_tol=0.0):
    """Invariants of a raster given per step as (ids, te): ids ascend within
    each step, every te lies in [t - 1, t), and no neuron fires twice within
    its refractory period.  te_tol widens both time checks (emission times
    read back from the raster file's 7 significant digits).  Returns the
    spike count."""
    last = np.full(n, -1.0e9)
    total = 0
    for t, (ids, te) in enumerate(steps):
        ids = np.asarray(ids, np.int64)
        assert np.all(np.diff(ids) > 0), t
        if te_tol:
            assert np.all((te >= t - 1 - te_tol) & (te <= t + te_tol)), t
        else:
            assert np.all((te >= t - 1) & (te < t)), t
        assert ids.size == 0 or (ids[0] >= 0 and ids[-1] < n), t
        assert np.all(te - last[ids] >= tarp_ms - 1e-9 - te_tol), t
        last[ids] = te
        total += ids.size
    return total


def read_raster(path, steps):
    """A raster file ("id te" lines, statSpikesClass::write: each step's
    spikes in ascending id order, steps in order) as per-step (ids, te).  A
    spike of step t has te in [t - 1, t), but "%.7g" can print a te just
    below t as t itself: such a line stays in step t while the file's order
    allows it (ids still ascending), else it opens step t + 1."""
    a = np.loadtxt(path, dtype=np.float64, ndmin=2) if os.path.getsize(path) else np.zeros((0, 2))
    ids, te = a[:, 0].astype(np.int64), a[:, 1]
    step = np.floor(te).astype(np.int64) + 1
    cur, last = -1, -1
    for k in range(len(ids)):
        s_ = step[k]
        if te[k] == np.floor(te[k]) and (cur < s_ - 1 or (cur == s_ - 1 and ids[k] > last)):
            s_ -= 1
        assert s_ >= cur, ("raster out of step order", k)
        cur, last = s_, ids[k]
        step[k] = s_
    bounds = np.searchsorted(step, np.arange(steps + 1))
    assert bounds[-1] == len(step), "spikes past the last step"
    return [(ids[bounds[t]:bounds[t + 1]], te[bounds[t]:bounds[t + 1]]) for t in range(steps)]

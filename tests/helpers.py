"""Shared helpers of the test suite (oracle wrappers, raster hashing)."""
import ctypes
import hashlib
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


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

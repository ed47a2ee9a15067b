"""MI355X engine for DPSNN's per-millisecond LIFCA simulation step.

The product is the C-ABI library ``lib/libdpsnn_hip.so`` (declared in
``include/dpsnn_hip.h``); this module is its Python host binding and mirrors
the reference's process interface (``localNetProcess_init`` / ``_fire`` /
``_finish``, localNetProcess.c:40-173; fire returns 0 to continue and 1 when
the programmed time is reached).

Configuration is DPSNN's own: env vars (read from the process environment,
as getParameters does, DPSNN_getParameters.c:23-831) plus column.txt,
connectivity.txt, bathEfficacy.txt and optional plasticity.txt in a
directory.  ``configs/<name>/env.txt`` holds the env vars of each canonical
workload; ``apply_env`` loads one into ``os.environ``.

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdpsnn_hip.so")
CONFIGS = os.path.join(HERE, "configs")

DRIVE_REPLAY = 0
DRIVE_PHILOX = 1
GENERATOR_HOST = 0
GENERATOR_DEVICE = 1
DONE = 1
# dpsnn_status codes
ERR_CONFIG, ERR_DEVICE, ERR_MEMORY, ERR_ARG, ERR_OVERFLOW, ERR_IO = -1, -2, -3, -4, -5, -6

_lib = None


class DpsnnError(RuntimeError):
    """A dpsnn_* entry point returned a negative status."""

    def __init__(self, func, code, message):
        super().__init__(f"{func} failed ({code}): {message}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("drive", ctypes.c_int),
                ("philox_seed", ctypes.c_uint64), ("output_dir", ctypes.c_char_p),
                ("drain_every_ms", ctypes.c_uint32), ("keep_spikes", ctypes.c_int),
                ("env", ctypes.c_char_p), ("drop_host_synapses", ctypes.c_int),
                ("host_threads", ctypes.c_uint32), ("generator", ctypes.c_uint32),
                ("generator_seed", ctypes.c_uint64)]


_PARAM_U32 = ["globH", "locN", "globN", "D", "totalSimTime_ms", "globCFX", "globCFY", "globCFT",
              "neuronsPerCM", "subPopNumber", "tabNeuron", "moduloSec", "ratesSampling",
              "globalSeed", "howManyOutputFiles", "startPlasticity_ms", "stopPlasticity_ms",
              "startStatFiles_s", "stopStatFiles_s", "startRatesPerPop_s", "stopRatesPerPop_s",
              "stencilDim", "overallConnectivity", "delayMin", "delayMax"]
_PARAM_F32 = ["minSynWeight_f", "factorFloat_2_weightType", "factorWeightType_2_Float",
              "STDPmultiplier", "antiCausalCoeff", "tauSTDP_ms", "causalSTDPmaxAbsValue",
              "antiCausalSTDPmaxAbsValue", "tauDerivativeDecay_s", "derivativeDrift"]


class Parameters(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint32) for n in _PARAM_U32] +
                [("spikes_writeDisable", ctypes.c_int32),
                 ("statRatesPerPop_writeDisable", ctypes.c_int32)] +
                [(n, ctypes.c_float) for n in _PARAM_F32] +
                [("statSynPeriodicProbe_writeDisable", ctypes.c_int32),
                 ("startSynPeriodicProbe_s", ctypes.c_uint32), ("stopSynPeriodicProbe_s", ctypes.c_uint32)])

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("spikes", ctypes.c_uint64),
                ("synaptic_events", ctypes.c_uint64), ("external_events", ctypes.c_uint64),
                ("synapses", ctypes.c_uint64), ("delivery_items", ctypes.c_uint64),
                ("neurons", ctypes.c_uint32), ("max_in_items", ctypes.c_uint32),
                ("init_seconds", ctypes.c_double), ("device_bytes", ctypes.c_uint64),
                ("window_form", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("graph_steps", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _p(t):
    return ctypes.POINTER(t)


class ServerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_uint64),
                ("build_threads", ctypes.c_uint32), ("steps_per_round", ctypes.c_uint32),
                ("keep_finished", ctypes.c_int), ("max_running", ctypes.c_uint32),
                ("host_backlog_bytes", ctypes.c_uint64)]


class JobInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("state", ctypes.c_int), ("status", ctypes.c_int),
                ("submitted_s", ctypes.c_double), ("built_s", ctypes.c_double),
                ("admitted_s", ctypes.c_double), ("finished_s", ctypes.c_double),
                ("device_bytes", ctypes.c_uint64), ("stats", Stats), ("message", ctypes.c_char * 160)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("stats", "message")}
        d["state"] = JOB_STATES[self.state]
        d["stats"] = self.stats.as_dict()
        d["message"] = self.message.decode(errors="replace")
        return d


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("finished", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("running", ctypes.c_uint64), ("waiting", ctypes.c_uint64), ("building", ctypes.c_uint64),
                ("synaptic_events", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("elapsed_s", ctypes.c_double), ("step_s", ctypes.c_double), ("upload_s", ctypes.c_double),
                ("queued", ctypes.c_uint64), ("admit_s", ctypes.c_double), ("account_s", ctypes.c_double),
                ("retire_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


JOB_STATES = ("queued", "building", "waiting", "running", "done", "failed")


class BatchShape(ctypes.Structure):
    _fields_ = [("chunks", ctypes.c_int), ("buckets", ctypes.c_int), ("waves_per_cu", ctypes.c_int),
                ("ms_per_step", ctypes.c_double), ("split", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "dpsnn_get_parameters": (ctypes.c_int, [ctypes.c_char_p, _p(Parameters)]),
    "dpsnn_write_ini_files": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "dpsnn_create": (ctypes.c_int, [ctypes.c_char_p, _p(Options), _p(ctypes.c_void_p)]),
    "dpsnn_prepare": (ctypes.c_int, [ctypes.c_char_p, _p(Options), _p(ctypes.c_void_p)]),
    "dpsnn_device_bytes_needed": (ctypes.c_uint64, [ctypes.c_void_p]),
    "dpsnn_host_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
    "dpsnn_upload": (ctypes.c_int, [ctypes.c_void_p]),
    "dpsnn_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "dpsnn_step_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "dpsnn_checkpoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dpsnn_restore": (ctypes.c_int, [ctypes.c_char_p, _p(Options), ctypes.c_char_p, _p(ctypes.c_void_p)]),
    "dpsnn_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "dpsnn_now_ms": (ctypes.c_uint32, [ctypes.c_void_p]),
    "dpsnn_total_ms": (ctypes.c_uint32, [ctypes.c_void_p]),
    "dpsnn_get_state": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5),
    "dpsnn_last_spikes": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3),
    "dpsnn_num_synapses": (ctypes.c_uint64, [ctypes.c_void_p]),
    "dpsnn_get_synapses": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6),
    "dpsnn_write_synapse_dump": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dpsnn_get_stats": (ctypes.c_int, [ctypes.c_void_p, _p(Stats)]),
    "dpsnn_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "dpsnn_destroy": (None, [ctypes.c_void_p]),
    "dpsnn_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dpsnn_kernel_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _p(ctypes.c_double),
                                         _p(ctypes.c_uint64)]),
    "dpsnn_host_build": (ctypes.c_int, [ctypes.c_char_p, _p(ctypes.c_void_p)]),
    "dpsnn_host_num_neurons": (ctypes.c_uint32, [ctypes.c_void_p]),
    "dpsnn_host_num_synapses": (ctypes.c_uint64, [ctypes.c_void_p]),
    "dpsnn_host_num_items": (ctypes.c_uint64, [ctypes.c_void_p]),
    "dpsnn_host_get_synapses": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "dpsnn_host_get_inlists": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "dpsnn_host_window": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dpsnn_host_get_external": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 2),
    "dpsnn_host_write_synapse_dump": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "dpsnn_host_external_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_uint64, _p(ctypes.c_uint64)]),
    "dpsnn_host_destroy": (None, [ctypes.c_void_p]),
    "dpsnn_device_exp": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64]),
    "dpsnn_device_div": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64]),
    "dpsnn_batch_create": (ctypes.c_int, [ctypes.c_int, _p(ctypes.c_void_p)]),
    "dpsnn_batch_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "dpsnn_batch_remove": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "dpsnn_batch_size": (ctypes.c_uint32, [ctypes.c_void_p]),
    "dpsnn_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "dpsnn_batch_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "dpsnn_batch_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _p(BatchShape)]),
    "dpsnn_batch_get_shape": (ctypes.c_int, [ctypes.c_void_p, _p(BatchShape)]),
    "dpsnn_batch_set_shape": (ctypes.c_int, [ctypes.c_void_p, _p(BatchShape)]),
    "dpsnn_batch_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dpsnn_batch_kernel_time": (ctypes.c_int, [ctypes.c_void_p, _p(ctypes.c_double), _p(ctypes.c_uint64)]),
    "dpsnn_batch_destroy": (None, [ctypes.c_void_p]),
    "dpsnn_device_memory": (ctypes.c_int, [ctypes.c_int, _p(ctypes.c_uint64), _p(ctypes.c_uint64)]),
    "dpsnn_server_create": (ctypes.c_int, [_p(ServerOptions), _p(ctypes.c_void_p)]),
    "dpsnn_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _p(Options), _p(ctypes.c_uint64)]),
    "dpsnn_server_job": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _p(JobInfo)]),
    "dpsnn_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, _p(JobInfo)]),
    "dpsnn_server_wait_all": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "dpsnn_server_take": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _p(ctypes.c_void_p)]),
    "dpsnn_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, _p(ServerStats)]),
    "dpsnn_server_destroy": (None, [ctypes.c_void_p]),
    "dpsnn_last_error": (ctypes.c_char_p, []),
}


def library():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first "
                               "(python 201812realtimecortsim_amd/build.py)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _check(func, rc):
    if rc < 0:
        raise DpsnnError(func, rc, library().dpsnn_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def read_env_file(path):
    env = {}
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#") and "=" in line:
                k, v = line.split("=", 1)
                env[k.strip()] = v.strip()
    return env


def apply_env(config_dir, **overrides):
    """Loads <config_dir>/env.txt (and overrides) into os.environ, where the
    engine's getParameters reads them; returns the applied dict."""
    env = read_env_file(os.path.join(config_dir, "env.txt"))
    env.update({k: str(v) for k, v in overrides.items()})
    os.environ.update(env)
    return env


def config_dir(name):
    return os.path.join(CONFIGS, name)


def get_parameters(inidir):
    out = Parameters()
    _check("dpsnn_get_parameters", library().dpsnn_get_parameters(inidir.encode(), ctypes.byref(out)))
    return out.as_dict()


def _options(device=0, drive=DRIVE_REPLAY, philox_seed=1, output_dir=None, drain_every_ms=0, keep_spikes=False,
             env=None, drop_host_synapses=False, generator=GENERATOR_HOST, generator_seed=1):
    """dpsnn_options of the Network keyword arguments (the strings stay
    referenced by the structure)."""
    out = output_dir.encode() if output_dir else None
    env_text = ";".join(f"{k}={v}" for k, v in env.items()).encode() if env else None
    return Options(device, drive, philox_seed, out, drain_every_ms, int(keep_spikes), env_text,
                   int(drop_host_synapses), 0, generator, generator_seed)


def write_ini_files(inidir, outdir, env=None):
    """Perseo .ini files of the network of inidir into outdir
    (dpsnn_write_ini_files; localNetClass::writeIniFiles)."""
    env_text = ";".join(f"{k}={v}" for k, v in env.items()).encode() if env else None
    _check("dpsnn_write_ini_files", library().dpsnn_write_ini_files(inidir.encode(), outdir.encode(), env_text))


class Network:
    """One local network on one GPU (a dpsnn_net)."""

    def __init__(self, inidir, device=0, drive=DRIVE_REPLAY, philox_seed=1, output_dir=None,
                 drain_every_ms=0, keep_spikes=False, env=None, drop_host_synapses=False, checkpoint=None,
                 generator=GENERATOR_HOST, generator_seed=1):
        """env: {"env_name": value} settings that take precedence over the
        process environment for this network only.  checkpoint: a file of
        dpsnn_checkpoint to resume from (dpsnn_restore), taken on the
        network of the same ini files and settings.  generator:
        GENERATOR_HOST (the reference's, bit-exact) or GENERATOR_DEVICE
        (statistically equivalent, drawn on the device with generator_seed)."""
        self._lib = library()
        opt = _options(device, drive, philox_seed, output_dir, drain_every_ms, keep_spikes, env,
                       drop_host_synapses, generator, generator_seed)
        self._out = opt.output_dir
        h = ctypes.c_void_p()
        if checkpoint:
            _check("dpsnn_restore", self._lib.dpsnn_restore(inidir.encode(), ctypes.byref(opt),
                                                            str(checkpoint).encode(), ctypes.byref(h)))
        else:
            _check("dpsnn_create", self._lib.dpsnn_create(inidir.encode(), ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self.n = self.stats()["neurons"]

    def checkpoint(self, path):
        """Writes the network's state to path (dpsnn_checkpoint)."""
        _check("dpsnn_checkpoint", self._lib.dpsnn_checkpoint(self._h, str(path).encode()))

    @classmethod
    def build_many(cls, specs, threads=0, host_threads=0):
        """Builds several networks: their host builds (dpsnn_prepare: ini
        files, connectome, delivery layout; no device call) run on `threads`
        Python threads at once (0: one per host core, at most 16), each
        build on `host_threads` threads of its own (0: the library's
        default); the uploads (dpsnn_upload) follow in order.  specs: a list
        of (inidir, {Network keyword arguments}) pairs.  Returns the
        networks in the order of specs."""
        from concurrent.futures import ThreadPoolExecutor
        lib = library()
        threads = threads or min(16, os.cpu_count() or 1)
        opts = []
        for _, kw in specs:
            o = _options(**kw)
            o.host_threads = host_threads
            opts.append(o)
        handles = [ctypes.c_void_p() for _ in specs]

        def prep(i):
            # ctypes drops the GIL for the call: the builds overlap
            return lib.dpsnn_prepare(specs[i][0].encode(), ctypes.byref(opts[i]), ctypes.byref(handles[i]))

        with ThreadPoolExecutor(max_workers=threads) as ex:
            codes = list(ex.map(prep, range(len(specs))))
        nets = []
        try:
            for i, rc in enumerate(codes):
                _check("dpsnn_prepare", rc)
                _check("dpsnn_upload", lib.dpsnn_upload(handles[i]))
                nets.append(cls._wrap(handles[i]))
                handles[i] = None
        finally:
            for h in handles:
                if h is not None and h.value:
                    lib.dpsnn_destroy(h)
        return nets

    @classmethod
    def _wrap(cls, handle):
        net = cls.__new__(cls)
        net._lib = library()
        net._out = None
        net._h = handle
        net.n = net.stats()["neurons"]
        return net

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dpsnn_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, nsteps=1):
        """Advances nsteps ms; returns True once the programmed time is reached."""
        return _check("dpsnn_step", self._lib.dpsnn_step(self._h, nsteps)) == DONE

    def step_noalloc(self, nsteps=1):
        """step() without device allocations: the key scratch is the
        library's static per-device array (dpsnn_step_noalloc)."""
        return _check("dpsnn_step_noalloc", self._lib.dpsnn_step_noalloc(self._h, nsteps)) == DONE

    def sync(self):
        _check("dpsnn_sync", self._lib.dpsnn_sync(self._h))

    @property
    def now_ms(self):
        return self._lib.dpsnn_now_ms(self._h)

    @property
    def total_ms(self):
        return self._lib.dpsnn_total_ms(self._h)

    def state(self):
        n = self.n
        v, c, tr, te = (np.empty(n, np.float64) for _ in range(4))
        fired = np.empty(n, np.uint8)
        _check("dpsnn_get_state", self._lib.dpsnn_get_state(self._h, _ptr(v), _ptr(c), _ptr(tr), _ptr(te),
                                                            _ptr(fired)))
        return {"v": v, "c": c, "Tr": tr, "Te": te, "fired": fired}

    def last_spikes(self):
        ids = np.empty(self.n, np.uint32)
        te = np.empty(self.n, np.float64)
        cnt = ctypes.c_uint32()
        _check("dpsnn_last_spikes", self._lib.dpsnn_last_spikes(self._h, _ptr(ids), _ptr(te),
                                                                ctypes.byref(cnt)))
        return ids[:cnt.value].copy(), te[:cnt.value].copy()

    def synapses(self):
        s = self._lib.dpsnn_num_synapses(self._h)
        pre, post = np.empty(s, np.uint32), np.empty(s, np.uint32)
        delay, w = np.empty(s, np.uint8), np.empty(s, np.int16)
        td, last = np.empty(s, np.float32), np.empty(s, np.int32)
        _check("dpsnn_get_synapses", self._lib.dpsnn_get_synapses(self._h, _ptr(pre), _ptr(post), _ptr(delay),
                                                                  _ptr(w), _ptr(td), _ptr(last)))
        return {"pre": pre, "post": post, "delay": delay, "weight": w, "tderiv": td, "last": last}

    def write_synapse_dump(self, path):
        _check("dpsnn_write_synapse_dump", self._lib.dpsnn_write_synapse_dump(self._h, path.encode()))

    def stats(self):
        out = Stats()
        _check("dpsnn_get_stats", self._lib.dpsnn_get_stats(self._h, ctypes.byref(out)))
        return out.as_dict()

    def profile(self, enable=True):
        _check("dpsnn_profile", self._lib.dpsnn_profile(self._h, int(enable)))

    def kernel_time(self, kernel=0):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        _check("dpsnn_kernel_time", self._lib.dpsnn_kernel_time(self._h, kernel, ctypes.byref(ms),
                                                                ctypes.byref(n)))
        return ms.value, n.value


class HostConnectome:
    """The host generator and delivery layout, without a device."""

    def __init__(self, inidir):
        self._lib = library()
        h = ctypes.c_void_p()
        _check("dpsnn_host_build", self._lib.dpsnn_host_build(inidir.encode(), ctypes.byref(h)))
        self._h = h
        self.n = self._lib.dpsnn_host_num_neurons(h)
        self.s = self._lib.dpsnn_host_num_synapses(h)
        self.items = self._lib.dpsnn_host_num_items(h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dpsnn_host_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def synapses(self):
        s = self.s
        pre, post = np.empty(s, np.uint32), np.empty(s, np.uint32)
        delay, w = np.empty(s, np.uint8), np.empty(s, np.int16)
        _check("dpsnn_host_get_synapses", self._lib.dpsnn_host_get_synapses(self._h, _ptr(pre), _ptr(post),
                                                                            _ptr(delay), _ptr(w)))
        return {"pre": pre, "post": post, "delay": delay, "weight": w}

    def inlists(self):
        off = np.empty(self.n + 1, np.uint64)
        pd = np.empty(self.items, np.uint32)
        w = np.empty(self.items, np.int16)
        syn = np.empty(self.items, np.uint32)
        _check("dpsnn_host_get_inlists", self._lib.dpsnn_host_get_inlists(self._h, _ptr(off), _ptr(pd),
                                                                          _ptr(w), _ptr(syn)))
        return {"off": off, "pre": pd & 0xFFFFFF, "delay": pd >> 24, "weight": w, "syn": syn}

    def window(self, units_per_col=1):
        """The stencil-window layout of the in-lists (dpsnn_host_window)."""
        dims = np.zeros(6, np.uint32)
        sizes = np.zeros(4, np.uint64)
        null = ctypes.c_void_p()
        _check("dpsnn_host_window", self._lib.dpsnn_host_window(self._h, units_per_col, _ptr(dims), _ptr(sizes),
                                                                null, null, null, null))
        item = np.empty(int(sizes[0]), np.uint32)
        desc = np.empty(int(sizes[1]), np.uint16)
        chunk = np.empty(int(sizes[2]), np.uint64)
        cols = np.empty(int(sizes[3]), np.int32)
        _check("dpsnn_host_window", self._lib.dpsnn_host_window(self._h, units_per_col, _ptr(dims), _ptr(sizes),
                                                                _ptr(item), _ptr(desc), _ptr(chunk), _ptr(cols)))
        keys = ("slots", "npc", "passes", "unit", "units_per_col", "units")
        return dict(zip(keys, (int(v) for v in dims)), item=item, desc=desc, chunk=chunk, cols=cols)

    def external_params(self):
        inv = np.empty(self.n, np.float64)
        mx = np.empty(self.n, np.uint32)
        _check("dpsnn_host_get_external", self._lib.dpsnn_host_get_external(self._h, _ptr(inv), _ptr(mx)))
        return inv, mx

    def external_step(self, cap=None):
        cap = cap or self.n * 64
        off = np.empty(self.n + 1, np.uint32)
        t, v = np.empty(cap, np.float64), np.empty(cap, np.float64)
        cnt = ctypes.c_uint64()
        _check("dpsnn_host_external_step", self._lib.dpsnn_host_external_step(
            self._h, _ptr(off), _ptr(t), _ptr(v), cap, ctypes.byref(cnt)))
        return off, t[:cnt.value].copy(), v[:cnt.value].copy()

    def write_synapse_dump(self, path):
        _check("dpsnn_host_write_synapse_dump", self._lib.dpsnn_host_write_synapse_dump(self._h, path.encode()))


def device_exp(x, device=0):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    _check("dpsnn_device_exp", library().dpsnn_device_exp(device, _ptr(x), _ptr(y), x.size))
    return y


def device_div(a, b, device=0):
    """a / b elementwise by the engine's device division by reciprocal."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(np.broadcast_to(b, a.shape), np.float64)
    q = np.empty_like(a)
    _check("dpsnn_device_div", library().dpsnn_device_div(device, _ptr(a), _ptr(b), _ptr(q), a.size))
    return q


class Batch:
    """Independent networks advancing in one persistent launch per step (a
    dpsnn_batch).  Members are Network objects; while they belong to the
    batch they are stepped only through it."""

    def __init__(self, device=0):
        self._lib = library()
        h = ctypes.c_void_p()
        _check("dpsnn_batch_create", self._lib.dpsnn_batch_create(device, ctypes.byref(h)))
        self._h = h
        self.members = []

    def add(self, net):
        _check("dpsnn_batch_add", self._lib.dpsnn_batch_add(self._h, net._h))
        self.members.append(net)

    def remove(self, net):
        _check("dpsnn_batch_remove", self._lib.dpsnn_batch_remove(self._h, net._h))
        self.members.remove(net)

    def __len__(self):
        return int(self._lib.dpsnn_batch_size(self._h))

    def step(self, nsteps=1):
        """Returns True once no member has steps left."""
        rc = self._lib.dpsnn_batch_step(self._h, nsteps)
        _check("dpsnn_batch_step", rc)
        return rc == 1

    def sync(self):
        _check("dpsnn_batch_sync", self._lib.dpsnn_batch_sync(self._h))

    def tune(self, trial_steps=2):
        s = BatchShape()
        _check("dpsnn_batch_tune", self._lib.dpsnn_batch_tune(self._h, trial_steps, ctypes.byref(s)))
        return s.as_dict()

    def shape(self):
        s = BatchShape()
        _check("dpsnn_batch_get_shape", self._lib.dpsnn_batch_get_shape(self._h, ctypes.byref(s)))
        return s.as_dict()

    def set_shape(self, chunks, buckets, waves_per_cu=0, split=1):
        s = BatchShape(chunks, buckets, waves_per_cu, 0.0, split)
        _check("dpsnn_batch_set_shape", self._lib.dpsnn_batch_set_shape(self._h, ctypes.byref(s)))

    def profile(self, enable=True):
        _check("dpsnn_batch_profile", self._lib.dpsnn_batch_profile(self._h, int(enable)))

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        _check("dpsnn_batch_kernel_time", self._lib.dpsnn_batch_kernel_time(self._h, ctypes.byref(ms),
                                                                            ctypes.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dpsnn_batch_destroy(self._h)
            self._h = None
            self.members = []

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Request loop for simulation jobs with continuous batching (a
    dpsnn_server): submit() queues a network (ini directory + options);
    host threads build it, the scheduler admits it into the running batch
    when its device footprint fits the HBM budget."""

    def __init__(self, device=0, hbm_budget_bytes=0, build_threads=0, steps_per_round=0, keep_finished=False,
                 max_running=0, host_backlog_bytes=0):
        self._lib = library()
        o = ServerOptions(device, hbm_budget_bytes, build_threads, steps_per_round, int(keep_finished), max_running,
                          host_backlog_bytes)
        h = ctypes.c_void_p()
        _check("dpsnn_server_create", self._lib.dpsnn_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def submit(self, inidir, drive=DRIVE_PHILOX, philox_seed=1, output_dir=None, env=None,
               drop_host_synapses=False, keep_spikes=False, generator=GENERATOR_HOST, generator_seed=1):
        env_text = ";".join(f"{k}={v}" for k, v in env.items()).encode() if env else None
        out = output_dir.encode() if output_dir else None
        opt = Options(0, drive, philox_seed, out, 0, int(keep_spikes), env_text, int(drop_host_synapses), 0,
                      generator, generator_seed)
        jid = ctypes.c_uint64()
        _check("dpsnn_server_submit", self._lib.dpsnn_server_submit(self._h, inidir.encode(), ctypes.byref(opt),
                                                                    ctypes.byref(jid)))
        return jid.value

    def job(self, jid):
        info = JobInfo()
        _check("dpsnn_server_job", self._lib.dpsnn_server_job(self._h, jid, ctypes.byref(info)))
        return info.as_dict()

    def wait(self, jid, timeout_s=-1.0):
        info = JobInfo()
        _check("dpsnn_server_wait", self._lib.dpsnn_server_wait(self._h, jid, timeout_s, ctypes.byref(info)))
        return info.as_dict()

    def wait_all(self, timeout_s=-1.0):
        return _check("dpsnn_server_wait_all", self._lib.dpsnn_server_wait_all(self._h, timeout_s)) == DONE

    def take(self, jid):
        """The finished job's network (keep_finished servers) as a Network."""
        h = ctypes.c_void_p()
        _check("dpsnn_server_take", self._lib.dpsnn_server_take(self._h, jid, ctypes.byref(h)))
        return Network._wrap(h)

    def stats(self):
        st = ServerStats()
        _check("dpsnn_server_get_stats", self._lib.dpsnn_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dpsnn_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---- the reference's process interface (DAL shape) ------------------------
class _LocalStates(ctypes.Structure):
    """struct _local_states (include/localNetProcess_hip.h)."""
    _fields_ = [("parameters", Parameters), ("net", ctypes.c_void_p), ("id", ctypes.c_int),
                ("name", ctypes.c_char * 32), ("fireCount", ctypes.c_int), ("fireSeconds", ctypes.c_double)]


class DALProcess(ctypes.Structure):
    """DALProcess (include/localNetProcess_hip.h; the reference's dal.h:25-42)."""
    _fields_ = [("local", ctypes.POINTER(_LocalStates)), ("init", ctypes.c_void_p), ("fire", ctypes.c_void_p),
                ("finish", ctypes.c_void_p), ("wptr", ctypes.c_void_p)]


class LocalNetProcess:
    """localNetProcess_{init,fire,finish} of the library (the reference's
    localNetProcess.c:40-173), called through ctypes on a DALProcess: init
    builds the network from the env and the ini files in ``inidir`` (its
    $DPSNN_INIDIR), fire advances one ms and returns 0, or 1 once
    env_totalSimTime_ms is reached, finish drains the output files and frees
    the device.  As in the reference, a configuration or device error makes
    init/fire print the message and end the process."""

    def __init__(self, inidir=".", output_dir=".", device=0, drive=DRIVE_REPLAY):
        self.inidir, self.output_dir, self.device, self.drive = inidir, output_dir, device, drive
        lib = library()
        for name, res in (("localNetProcess_init", None), ("localNetProcess_fire", ctypes.c_int),
                          ("localNetProcess_finish", None)):
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, [ctypes.POINTER(DALProcess)]
        self._lib = lib
        self._state = _LocalStates()
        self._proc = DALProcess(ctypes.pointer(self._state), None, None, None, None)

    def init(self):
        os.environ.update(DPSNN_INIDIR=self.inidir, DPSNN_OUTDIR=self.output_dir, DPSNN_DEVICE=str(self.device),
                          DPSNN_DRIVE="philox" if self.drive == DRIVE_PHILOX else "replay")
        self._lib.localNetProcess_init(ctypes.byref(self._proc))

    def fire(self):
        return self._lib.localNetProcess_fire(ctypes.byref(self._proc))

    def stats(self):
        st = Stats()
        _check("dpsnn_get_stats", self._lib.dpsnn_get_stats(self._state.net, ctypes.byref(st)))
        return st.as_dict()

    @property
    def parameters(self):
        return self._state.parameters

    @property
    def fire_count(self):
        return self._state.fireCount

    def finish(self):
        self._lib.localNetProcess_finish(ctypes.byref(self._proc))

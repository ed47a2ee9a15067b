import ctypes
import importlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libdpsnn_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("201812realtimecortsim_amd")


def _build_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


class Oracle:
    """ctypes view of the oracle library (oracle/dpsnn_oracle.h)."""

    def __init__(self):
        _build_oracle()
        lib = ctypes.CDLL(ORACLE_SO)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        sigs = {
            "ora_create": (vp, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
            "ora_create_from_backward": (vp, [ctypes.c_char_p, u64, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]),
            "ora_load_checkpoint": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
            "ora_destroy": (None, [vp]),
            "ora_num_neurons": (u32, [vp]), "ora_num_synapses": (u64, [vp]),
            "ora_total_ms": (u32, [vp]), "ora_now_ms": (u32, [vp]), "ora_synaptic_events": (u64, [vp]),
            "ora_get_backward": (None, [vp] + [vp] * 7),
            "ora_get_generated": (None, [vp] + [vp] * 4),
            "ora_write_synapse_dump": (ctypes.c_int, [vp, ctypes.c_char_p]),
            "ora_get_external": (None, [vp, vp, vp]),
            "ora_fire": (ctypes.c_int, [vp]),
            "ora_last_spikes": (u32, [vp, vp, vp]),
            "ora_get_state": (None, [vp] + [vp] * 5),
            "ora_set_output_dir": (None, [vp, ctypes.c_char_p]),
            "ora_record_external": (None, [vp, ctypes.c_int]),
            "ora_last_external": (u32, [vp, vp, vp, vp]),
            "ora_libm_exp": (ctypes.c_double, [ctypes.c_double]),
            "ora_libm_logf": (ctypes.c_float, [ctypes.c_float]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        self.lib = lib


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def pytest_collection_modifyitems(config, items):
    # GPU tests need the built library on the box; CPU tests must never touch the device
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(pytest.mark.timeout(600))

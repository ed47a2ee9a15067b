"""The drop-in boundary itself on the device: the reference's main loop
(DPSNN_mainUnderMPI.c:33-102) as the built program lib/dpsnn_hip_main in a
the 2x2-column network run directory, and localNetProcess_init/fire/finish
(localNetProcess.c:40-173) called through ctypes.  Outputs must be the
reference's files (golden fixtures)."""
import os
import shutil
import subprocess

import pytest

from helpers import GOLDEN, golden, sha256_file

pytestmark = pytest.mark.gpu


def _run_dir(pkg, tmp, ms):
    """A run directory as the reference expects it: the ini files in the
    working directory, the env vars in the environment."""
    src = pkg.config_dir("c0")
    d = tmp / "run"
    d.mkdir()
    for f in ("column.txt", "connectivity.txt", "bathEfficacy.txt"):
        shutil.copy(os.path.join(src, f), d / f)
    env = dict(os.environ)
    env.update({k: str(v) for k, v in pkg.read_env_file(os.path.join(src, "env.txt")).items()})
    env["env_totalSimTime_ms"] = str(ms)
    for k in ("DPSNN_INIDIR", "DPSNN_OUTDIR", "DPSNN_DRIVE"):
        env.pop(k, None)
    return d, env


def test_main_program_reproduces_reference_files(pkg, tmp_path):
    """500 ms through the main program in a fresh process: raster and rates
    are the reference's; fire() was called 501 times (500 steps, then the
    call that returns 1)."""
    main = os.path.join(os.path.dirname(pkg.LIB_PATH), "dpsnn_hip_main")
    assert os.access(main, os.X_OK), main
    d, env = _run_dir(pkg, tmp_path, 500)
    print("dpsnn_hip_main: the 2x2-column network, 500 ms", flush=True)
    proc = subprocess.run([main], cwd=str(d), env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout[-2000:]
    assert "init() - completed" in proc.stdout
    assert "on 501 samples" in proc.stdout, proc.stdout[-500:]
    g = golden("c0_raster.json")
    raster = d / "spikes-sec0-C4-H1-h000.dat"
    assert sha256_file(raster) == g["prefix"]["500"]["sha256"]
    rates = open(d / "ratesPerPop-sec0-C4-H1-h000.dat").read().splitlines()
    ref = open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read().splitlines()
    assert rates == ref[:100]


def test_process_interface_through_ctypes(pkg, tmp_path):
    """init / fire / finish on a DALProcess: fire returns 0 for each of the
    100 steps and 1 on the next call (the reference's convention); the
    files are the reference's."""
    d, env = _run_dir(pkg, tmp_path, 100)
    saved = dict(os.environ)
    try:
        os.environ.clear()
        os.environ.update(env)
        proc = pkg.LocalNetProcess(inidir=str(d), output_dir=str(d))
        print("localNetProcess via ctypes: the 2x2-column network, 100 ms", flush=True)
        proc.init()
        assert proc.parameters.totalSimTime_ms == 100 and proc.parameters.globCFT == 4
        codes = []
        while True:
            codes.append(proc.fire())
            if codes[-1] != 0:
                break
        assert codes == [0] * 100 + [1]
        st = proc.stats()
        assert st["steps"] == 100 and st["neurons"] == 5000
        proc.finish()
    finally:
        os.environ.clear()
        os.environ.update(saved)
    g = golden("c0_raster.json")
    assert sha256_file(d / "spikes-sec0-C4-H1-h000.dat") == g["prefix"]["100"]["sha256"]
    rates = open(d / "ratesPerPop-sec0-C4-H1-h000.dat").read().splitlines()
    ref = open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read().splitlines()
    assert rates == ref[:20]

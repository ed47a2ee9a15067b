"""Checkpoint / resume of device-resident state (the DAL's declared but
unimplemented checkpoint hooks, dal.h:52-55): a run interrupted by a
checkpoint, a destroy and a restore writes the same files as the
uninterrupted reference run, byte for byte."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, golden, sha256_file

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_env(pkg):
    yield
    pkg.apply_env(pkg.config_dir("c0"))


def test_c0_resumed_at_500ms_is_the_reference_second(pkg, tmp_path):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=1000)
    ck = tmp_path / "c0.ckpt"
    print("checkpoint: the 2x2-column network 500 ms, checkpoint, destroy", flush=True)
    with pkg.Network(cfg, output_dir=str(tmp_path)) as net:
        net.step(500)
        net.checkpoint(ck)
        assert net.now_ms == 500
    print("checkpoint: restore, 500 ms more", flush=True)
    with pkg.Network(cfg, output_dir=str(tmp_path), checkpoint=ck) as net:
        assert net.now_ms == 500
        while not net.step(100):
            pass
        net.sync()
        st = net.stats()
    g = golden("c0_raster.json")
    assert sha256_file(tmp_path / "spikes-sec0-C4-H1-h000.dat") == g["sha256"]
    assert (open(tmp_path / "ratesPerPop-sec0-C4-H1-h000.dat").read() ==
            open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read())
    assert st["steps"] == 1000 and st["spikes"] == g["lines"]


def test_stdp_resumed_across_the_commit(pkg, tmp_path):
    """the STDP network's plasticity at the 2x2-column size, checkpointed at 700 ms and
    resumed: both seconds' rasters and the synapse probes around the
    t = 999 ms commit are the reference's."""
    g = golden("c2small_golden.json")
    cfg = pkg.config_dir("c2_small")
    pkg.apply_env(cfg, env_statSynPeriodicProbe_writeDisable=0, env_startSynPeriodicProbe_s=0,
                  env_stopSynPeriodicProbe_s=0)
    ck = tmp_path / "c2.ckpt"
    print("checkpoint: c2_small 700 ms", flush=True)
    with pkg.Network(cfg, output_dir=str(tmp_path)) as net:
        net.step(700)
        net.checkpoint(ck)
    print("checkpoint: c2_small restored, to the end", flush=True)
    with pkg.Network(cfg, output_dir=str(tmp_path), checkpoint=ck) as net:
        while not net.step(100):
            pass
        net.sync()
    for sec in (0, 1):
        assert sha256_file(tmp_path / f"spikes-sec{sec}-C4-H1-h000.dat") == g[f"spikes-sec{sec}"]["sha256"], sec
        assert (open(tmp_path / f"ratesPerPop-sec{sec}-C4-H1-h000.dat").read() ==
                open(os.path.join(GOLDEN, f"c2small_ratesPerPop-sec{sec}.dat")).read()), sec
    for probe in ("A", "B"):
        assert (sha256_file(tmp_path / f"synPeriodicProbe_{probe}-sec0-C4-H1-h000.dat") ==
                g[f"synPeriodicProbe_{probe}"]["sha256"]), probe


def test_philox_resume_matches_uninterrupted_state(pkg, tmp_path):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=300)
    kw = dict(drive=pkg.DRIVE_PHILOX, philox_seed=77)
    ck = tmp_path / "p.ckpt"
    with pkg.Network(cfg, **kw) as a:
        a.step(137)
        a.checkpoint(ck)
        a.step(163)
        sa, st_a = a.state(), a.stats()
    with pkg.Network(cfg, checkpoint=ck, **kw) as b:
        b.step(163)
        sb, st_b = b.state(), b.stats()
    for k in ("v", "c", "Tr", "Te"):
        assert np.array_equal(sa[k].view(np.uint64), sb[k].view(np.uint64)), k
    for k in ("steps", "spikes", "synaptic_events", "external_events"):
        assert st_a[k] == st_b[k], k


def test_restore_refuses_another_network(pkg, tmp_path):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=50)
    ck = tmp_path / "x.ckpt"
    with pkg.Network(cfg, drive=pkg.DRIVE_PHILOX) as a:
        a.step(5)
        a.checkpoint(ck)
    with pytest.raises(pkg.DpsnnError, match="drive"):
        pkg.Network(cfg, drive=pkg.DRIVE_REPLAY, checkpoint=ck)
    small = tmp_path / "small"
    small.mkdir()
    for f in ("connectivity.txt", "bathEfficacy.txt"):
        (small / f).write_text(open(os.path.join(cfg, f)).read())
    (small / "column.txt").write_text("0 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                                      "1 50 0.5 0.0 3 20 20 15 2 1 500 0.05\n"
                                      "2 25 0.5 0.0 3 10 20 15 2 0 500 0\n")
    with pytest.raises(pkg.DpsnnError, match="another network"):
        pkg.Network(str(small), drive=pkg.DRIVE_PHILOX, checkpoint=ck)
    (tmp_path / "junk").write_bytes(b"not a checkpoint, just bytes " * 10)
    with pytest.raises(pkg.DpsnnError, match="not a network checkpoint"):
        pkg.Network(cfg, drive=pkg.DRIVE_PHILOX, checkpoint=tmp_path / "junk")

"""Batching: independent networks of uneven size advancing in one launch per
step give, member by member, the same results as each network stepped alone
(bit-exact state, counters, raster/rates files, STDP synapses), also when
members join at different ms, when the launch shape changes mid-run, after
the autotuner's trial steps, and when members step in the window form
inside the batch."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, golden, prefix_sha, sha256_file

pytestmark = pytest.mark.gpu

# (config, env overrides, drive): grids of 1, 2 and 4 columns, two seeds,
# both drives, one plastic network whose run crosses the t = 999 commit
MEMBERS = [
    ("c0", dict(env_CFX=1, env_CFY=1, env_globalSeed=3, env_totalSimTime_ms=300), "philox"),
    ("c0", dict(env_CFX=2, env_CFY=1, env_globalSeed=5, env_totalSimTime_ms=250), "replay"),
    ("c0", dict(env_CFX=2, env_CFY=2, env_globalSeed=1, env_totalSimTime_ms=320), "replay"),
    ("c2_small", dict(env_totalSimTime_ms=1040), "philox"),
]


def _make(pkg, spec, out):
    name, over, drive = spec
    cfg = pkg.config_dir(name)
    env = pkg.read_env_file(os.path.join(cfg, "env.txt"))
    env.update({k: str(v) for k, v in over.items()})
    return pkg.Network(cfg, drive=pkg.DRIVE_PHILOX if drive == "philox" else pkg.DRIVE_REPLAY,
                       philox_seed=17, output_dir=str(out), env=env)


def _result(net, out, plastic):
    r = {"stats": {k: v for k, v in net.stats().items() if k not in ("init_seconds",)},
         "state": net.state(),
         "files": {f: sha256_file(os.path.join(out, f)) for f in sorted(os.listdir(out))}}
    if plastic:
        r["syn"] = net.synapses()
    return r


def _same(a, b):
    assert a["files"] == b["files"]
    assert a["stats"]["spikes"] == b["stats"]["spikes"]
    assert a["stats"]["synaptic_events"] == b["stats"]["synaptic_events"]
    assert a["stats"]["external_events"] == b["stats"]["external_events"]
    assert a["stats"]["steps"] == b["stats"]["steps"]
    for k in ("v", "c", "Tr", "Te", "fired"):
        assert np.array_equal(a["state"][k].view(np.uint8), b["state"][k].view(np.uint8)), k
    if "syn" in a:
        for k in ("weight", "tderiv", "last"):
            assert np.array_equal(a["syn"][k].view(np.uint8), b["syn"][k].view(np.uint8)), k


@pytest.mark.parametrize("window", ["default", "1"])
def test_batch_members_match_alone(pkg, tmp_path, monkeypatch, window):
    """window 1: every member has a window layout, so from ms 32 on the
    members without LTD in a step run it in the window form on the batch's key
    scratch, beside the batched gather form of the others."""
    if window == "1":
        monkeypatch.setenv("DPSNN_WINDOW", "1")
    alone = []
    for i, spec in enumerate(MEMBERS):
        out = tmp_path / f"alone{i}"
        out.mkdir()
        with _make(pkg, spec, out) as net:
            while not net.step(100):
                pass
            net.sync()
            alone.append(_result(net, str(out), spec[0] == "c2_small"))
        assert alone[-1]["stats"]["spikes"] > 0

    outs = [tmp_path / f"batch{i}" for i in range(len(MEMBERS))]
    for o in outs:
        o.mkdir()
    with pkg.Batch() as b:
        nets = [_make(pkg, MEMBERS[i], outs[i]) for i in (0, 1)]
        for n in nets:
            b.add(n)
        assert not b.step(40)
        # the others join while the first two are at ms 40
        for i in (2, 3):
            nets.append(_make(pkg, MEMBERS[i], outs[i]))
            b.add(nets[-1])
        assert len(b) == 4
        b.set_shape(16, 64)
        b.step(60)
        b.set_shape(8, 1024, 4, split=0)  # the persistent fused launch
        b.step(30)
        b.set_shape(24, 256, 0, split=1)
        b.step(30)
        chosen = b.tune(1)
        assert chosen["chunks"] in (8, 16, 24) and chosen["buckets"] in (64, 256, 1024)
        with pytest.raises(pkg.DpsnnError):
            nets[0].step(1)  # members are stepped through their batch
        while not b.step(200):
            pass
        b.sync()
        got = [_result(n, str(outs[i]), MEMBERS[i][0] == "c2_small") for i, n in enumerate(nets)]
        for n in nets:
            b.remove(n)
            n.close()
    for i, (a, g) in enumerate(zip(alone, got)):
        _same(a, g)
    # member 2 is the 2x2-column network (seed 1, replay drive) cut at 320 ms:
    # its batched raster and rates are the reference's
    g = golden("c0_raster.json")
    assert prefix_sha(outs[2] / "spikes-sec0-C4-H1-h000.dat", g["prefix"]["200"]["lines"]) == \
        g["prefix"]["200"]["sha256"]
    rates = open(outs[2] / "ratesPerPop-sec0-C4-H1-h000.dat").read().splitlines()
    assert rates == open(os.path.join(GOLDEN, "c0_ratesPerPop-sec0.dat")).read().splitlines()[:64]


def test_removed_member_steps_alone_again(pkg, tmp_path):
    spec = ("c0", dict(env_CFX=1, env_CFY=1, env_globalSeed=9, env_totalSimTime_ms=120), "philox")
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    with _make(pkg, spec, tmp_path / "a") as ref:
        while not ref.step(50):
            pass
        ref.sync()
        want = _result(ref, str(tmp_path / "a"), False)
    with pkg.Batch() as b, _make(pkg, spec, tmp_path / "b") as net:
        b.add(net)
        b.step(70)
        b.remove(net)
        while not net.step(50):
            pass
        net.sync()
        _same(want, _result(net, str(tmp_path / "b"), False))


def test_batch_plain_launches_match_graph_replays(pkg, tmp_path, monkeypatch):
    """The same mix stepped with every launch issued directly and with each
    step replayed from captured HIP graphs (several replays per graph)."""
    specs = [("c0", dict(env_CFX=1, env_CFY=1, env_globalSeed=4, env_totalSimTime_ms=150), "philox"),
             ("c0", dict(env_CFX=2, env_CFY=1, env_globalSeed=6, env_totalSimTime_ms=150), "replay")]
    res = {}
    for graph in ("0", "1"):
        monkeypatch.setenv("DPSNN_GRAPH", graph)
        outs = [tmp_path / f"g{graph}_{i}" for i in range(len(specs))]
        for o in outs:
            o.mkdir()
        with pkg.Batch() as b:
            nets = [_make(pkg, sp, o) for sp, o in zip(specs, outs)]
            for n in nets:
                b.add(n)
            while not b.step(25):
                pass
            b.sync()
            res[graph] = [_result(n, str(o), False) for n, o in zip(nets, outs)]
            for n in nets:
                b.remove(n)
                n.close()
    for a, g in zip(res["0"], res["1"]):
        _same(a, g)

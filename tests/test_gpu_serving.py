"""Serving: jobs submitted while others run are admitted into the running
batch (continuous batching); each job's results -- raster/rates files,
counters, final state, STDP synapses -- are bit-identical to the same
network run alone and, for one job, to the oracle's run; the jobs overlap
on the device."""
import os
import time

import numpy as np
import pytest

from helpers import OracleNet, sha256_file

pytestmark = pytest.mark.gpu

JOBS = [  # (side, seed, duration_ms, stdp, drive)
    (1, 11, 400, False, "philox"),
    (2, 12, 300, True, "replay"),
    (3, 13, 250, False, "philox"),
    (1, 14, 1100, True, "philox"),   # crosses the t = 999 STDP commit
    (2, 15, 200, False, "replay"),
    (1, 16, 150, False, "philox"),
]


def _spec(side, seed, dur, stdp):
    return {"side": side, "seed": seed, "nu_ext": 3.0, "duration_ms": dur, "stdp": stdp}


def _drive(pkg, d):
    return pkg.DRIVE_PHILOX if d == "philox" else pkg.DRIVE_REPLAY


def _result(net, out, plastic):
    st = net.stats()
    r = {"stats": {k: st[k] for k in ("steps", "spikes", "synaptic_events", "external_events")},
         "state": net.state(),
         "files": {f: sha256_file(os.path.join(out, f)) for f in sorted(os.listdir(out))}}
    if plastic:
        syn = net.synapses()
        r["syn"] = {k: syn[k] for k in ("weight", "tderiv", "last")}
    return r


def _same(a, b):
    assert a["files"] == b["files"]
    assert a["stats"] == b["stats"]
    for k in ("v", "c", "Tr", "Te", "fired"):
        assert np.array_equal(a["state"][k].view(np.uint8), b["state"][k].view(np.uint8)), k
    for k in a.get("syn", {}):
        assert np.array_equal(a["syn"][k].view(np.uint8), b["syn"][k].view(np.uint8)), k


def test_served_jobs_match_alone_and_overlap(pkg, oracle, tmp_path, monkeypatch):
    W = __import__("importlib").import_module("201812realtimecortsim_amd.workloads")
    dirs, envs = [], []
    for i, (side, seed, dur, stdp, _) in enumerate(JOBS):
        d = tmp_path / f"ini{i}"
        envs.append(W.write_instance(str(d), _spec(side, seed, dur, stdp), raster=True))
        dirs.append(str(d))
    alone = []
    for i, job in enumerate(JOBS):
        out = tmp_path / f"alone{i}"
        out.mkdir()
        with pkg.Network(dirs[i], drive=_drive(pkg, job[4]), philox_seed=100 + i, output_dir=str(out),
                         env=envs[i]) as net:
            while not net.step(100):
                pass
            net.sync()
            alone.append(_result(net, str(out), job[3]))
    outs = [tmp_path / f"served{i}" for i in range(len(JOBS))]
    for o in outs:
        o.mkdir()
    with pkg.Server(build_threads=2, steps_per_round=10, keep_finished=True) as srv:
        ids = []
        for i in range(3):
            ids.append(srv.submit(dirs[i], drive=_drive(pkg, JOBS[i][4]), philox_seed=100 + i,
                                  output_dir=str(outs[i]), env=envs[i]))
        # the rest arrive while the first ones run
        while srv.job(ids[0])["state"] not in ("running", "done"):
            time.sleep(0.01)
        for i in range(3, len(JOBS)):
            ids.append(srv.submit(dirs[i], drive=_drive(pkg, JOBS[i][4]), philox_seed=100 + i,
                                  output_dir=str(outs[i]), env=envs[i]))
        assert srv.wait_all(600.0)
        infos = [srv.job(j) for j in ids]
        for inf in infos:
            assert inf["state"] == "done", inf["message"]
        got = []
        for i, j in enumerate(ids):
            net = srv.take(j)
            got.append(_result(net, str(outs[i]), JOBS[i][3]))
            net.close()
        st = srv.stats()
    for a, g in zip(alone, got):
        _same(a, g)
    # continuous batching: a later job was admitted while an earlier one ran
    overlap = any(infos[k]["admitted_s"] < infos[i]["finished_s"] and infos[k]["admitted_s"] > infos[i]["admitted_s"]
                  for i in range(len(infos)) for k in range(len(infos)) if k != i)
    assert overlap
    assert st["finished"] == len(JOBS) and st["failed"] == 0
    assert st["synaptic_events"] == sum(g["stats"]["synaptic_events"] for g in got)
    # one served job against the oracle: job 4 (2x2 columns, replay drive)
    for k, v in envs[4].items():
        monkeypatch.setenv(k, str(v))
    od = tmp_path / "oracle4"
    od.mkdir()
    ora = OracleNet(oracle, dirs[4])
    try:
        ora.set_output_dir(str(od))
        for _ in range(JOBS[4][2]):
            ora.fire()
        o = ora.state()
    finally:
        ora.close()  # flushes the files
    for k in ("v", "c", "Tr", "Te", "fired"):
        assert np.array_equal(got[4]["state"][k].view(np.uint8), o[k].view(np.uint8)), k
    assert {f: sha256_file(od / f) for f in sorted(os.listdir(od))} == got[4]["files"]


def test_bad_job_fails_alone(pkg, tmp_path):
    """A job whose ini files are missing fails with a configuration error;
    the server keeps serving the others."""
    W = __import__("importlib").import_module("201812realtimecortsim_amd.workloads")
    d = tmp_path / "ok"
    env = W.write_instance(str(d), _spec(1, 21, 50, False))
    with pkg.Server() as srv:
        bad = srv.submit(str(tmp_path / "missing"), env=env)
        good = srv.submit(str(d), env=env)
        assert srv.wait_all(300.0)
        b, g = srv.job(bad), srv.job(good)
    assert b["state"] == "failed" and b["status"] == pkg.ERR_CONFIG and b["message"]
    assert g["state"] == "done" and g["stats"]["steps"] == 50


def test_host_backlog_bound_serves_every_job(pkg, tmp_path):
    """With a host backlog bound of one byte a builder starts a build only
    when no built job waits for its upload: the server does not stall, and
    every job finishes with its counters."""
    W = __import__("importlib").import_module("201812realtimecortsim_amd.workloads")
    dirs, envs = [], []
    for i in range(5):
        d = tmp_path / f"b{i}"
        envs.append(W.write_instance(str(d), _spec(1 + i % 2, 40 + i, 60, False)))
        dirs.append(str(d))
    with pkg.Server(build_threads=2, host_backlog_bytes=1) as srv:
        ids = [srv.submit(dirs[i], drive=pkg.DRIVE_PHILOX, philox_seed=200 + i, env=envs[i]) for i in range(5)]
        assert srv.wait_all(300.0)
        infos = [srv.job(j) for j in ids]
    for inf in infos:
        assert inf["state"] == "done", inf["message"]
        assert inf["stats"]["steps"] == 60 and inf["stats"]["synaptic_events"] > 0

"""The engine at the workloads' own sizes.

* the 2x2-column network and the 24x24-column network (720,000 neurons, 810 M synapses): the oracle,
  built from the engine's backward synapse list, gives the engine's first
  steps; handed an engine checkpoint taken in the steady state (state,
  spike history, ran3 stream), it continues with the engine's next steps,
  state and spikes bit for bit.
* the 24x24-column network for 60 ms: the raster's invariants (ascending ids per step,
  emission times inside their step, refractory period), the engine's
  synaptic-event counter recomputed from the backward list and the spikes,
  and the ratesPerPop file recomputed from the raster.
"""
import os

import numpy as np
import pytest

from helpers import OracleNet, check_raster, delivered_events

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_env(pkg):
    yield
    pkg.apply_env(pkg.config_dir("c0"))


def _same_step(cfg, m, got, ora):
    st, ids, te = got
    o = ora.state()
    for key in ("v", "c", "Tr", "Te"):
        assert np.array_equal(st[key].view(np.uint64), o[key].view(np.uint64)), (m, key)
    assert np.array_equal(st["fired"], o["fired"]), m
    oids, ote = ora.last_spikes()
    assert np.array_equal(ids, oids) and np.array_equal(te.view(np.uint64), ote.view(np.uint64)), m
    print(f"{cfg}: step {m} identical ({len(ids)} spikes)", flush=True)


def _continue_in_oracle(pkg, oracle, cfg, first, settle, steps, tmp_path):
    """The engine's first `first` steps, then, from a checkpoint taken at ms
    `settle`, `steps` more, each compared with the oracle's."""
    ck0, ck = tmp_path / "start.ckpt", tmp_path / "net.ckpt"
    head, tail = [], []
    with pkg.Network(cfg, drive=pkg.DRIVE_REPLAY, keep_spikes=True) as net:
        # the ran3 stream's position at ms 0 (its Gaussian cache included)
        # follows the connectome generation, which the imported oracle skips
        net.checkpoint(ck0)
        for _ in range(first):
            net.step(1)
            head.append((net.state(),) + net.last_spikes())
        print(f"{cfg}: {settle} ms, checkpoint", flush=True)
        net.step(settle - first)
        net.checkpoint(ck)
        for _ in range(steps):
            net.step(1)
            tail.append((net.state(),) + net.last_spikes())
        syn = net.synapses()
    backward = {k: syn[k] for k in ("pre", "post", "delay", "weight")}
    del syn
    print(f"{cfg}: oracle from the backward list", flush=True)
    ora = OracleNet(oracle, cfg, backward=backward)
    try:
        ora.load_checkpoint(ck0)
        for m, got in enumerate(head):
            ora.fire()
            _same_step(cfg, m, got, ora)
        ora.load_checkpoint(ck)
        for k, got in enumerate(tail):
            ora.fire()
            _same_step(cfg, settle + k, got, ora)
        assert sum(len(g[1]) for g in tail) > 0
    finally:
        ora.close()


def test_c0_checkpoint_continues_identically_in_oracle(pkg, oracle, tmp_path):
    cfg = pkg.config_dir("c0")
    pkg.apply_env(cfg, env_totalSimTime_ms=200)
    _continue_in_oracle(pkg, oracle, cfg, 3, 150, 4, tmp_path)


def test_c1_steady_state_bit_exact_vs_oracle(pkg, oracle, tmp_path):
    """the 24x24-column network at full size: its first 3 steps and, past the start-up
    transient, the 2 steps after a checkpoint at ms 80 (stencil-window
    phase 1) are the oracle's."""
    cfg = pkg.config_dir("c1")
    pkg.apply_env(cfg, env_totalSimTime_ms=100)
    _continue_in_oracle(pkg, oracle, cfg, 3, 80, 2, tmp_path)


def test_c1_raster_invariants_and_event_count(pkg, tmp_path):
    cfg = pkg.config_dir("c1")
    T = 60
    pkg.apply_env(cfg, env_totalSimTime_ms=T, env_howManyOutputFiles=1)
    out = tmp_path / "out"
    out.mkdir()
    records = []
    with pkg.Network(cfg, drive=pkg.DRIVE_PHILOX, philox_seed=3, keep_spikes=True, output_dir=str(out)) as net:
        print("the 24x24-column network: 60 ms", flush=True)
        for m in range(T):
            net.step(1)
            records.append(net.last_spikes())
        assert net.step(1)  # past the end: files drained and flushed
        st = net.stats()
        syn = net.synapses()
    par = pkg.get_parameters(cfg)
    N = st["neurons"]
    assert N == 720000 and st["synapses"] == len(syn["pre"]) and st["steps"] == T
    total = check_raster(records, N)
    assert total == st["spikes"] and total > 0
    print(f"the 24x24-column network: {total} spikes; event count from the backward list", flush=True)
    assert delivered_events(syn, N, [r[0] for r in records]) == st["synaptic_events"]
    del syn
    # ratesPerPop (DPSNN_stat.c:225-273): the line of step t (t % sampling
    # == 0) carries t - sampling and the per-(column, sub-population) counts
    # of the scans of steps t - sampling + 1 .. t
    counts = [int(line.split()[1]) for line in open(os.path.join(cfg, "column.txt")) if line.strip()]
    nsub, npc, samp = len(counts), sum(counts), par["ratesSampling"]
    sub_of = np.repeat(np.arange(nsub), counts)
    lines = open(out / "ratesPerPop-sec0-C576-H1-h000.dat").read().splitlines()
    assert len(lines) == (T + samp - 1) // samp
    for k, line in enumerate(lines):
        t = k * samp
        cnt = np.zeros(N // npc * nsub, np.int64)
        for m in range(max(0, t - samp + 1), t + 1):
            ids = records[m][0].astype(np.int64)
            np.add.at(cnt, (ids // npc) * nsub + sub_of[ids % npc], 1)
        den = [float(np.float32(c) * np.float32(samp)) for c in counts]
        want = ["%d" % (t - samp)] + ["%.7g" % (float(np.float32(c)) * 1000.0 / den[i % nsub])
                                      for i, c in enumerate(cnt)]
        assert line.split() == want, k

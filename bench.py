"""Benchmark of the MI355X engine on DPSNN's per-millisecond LIFCA step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1]
    python bench.py --mode batch [--steps K] [--warmup W]   (the batching mix)
    python bench.py --mode serve [--serve-window S]          (the serving stream)
    python bench.py --mode copy                              (counter calibration)

One "step" = one simulated millisecond of the whole local network on one GPU
(the reference's completeTimeStep_ms).  Default workload: the 24x24-column network, a
24x24 grid of 1250-neuron columns (720,000 neurons, ~810 M synapses),
on-device Philox external drive, no raster files; synthetic network from
the reference generator (no external data).  Prints one JSON line:
  value         synaptic events delivered per wall second over the K timed
                steps (inputs resident in HBM; host I/O not in the region)
  ms_per_step   wall ms per simulated ms = real-time factor x 1000
  roofline      one step's device work (phase-1 and phase-2 kernels) against
                HBM bandwidth: algorithmic bytes = 25 B per synaptic event +
                72 B per neuron (+16 B per replayed external event), divided
                by its average time from HIP events on the network's stream;
                traffic = HBM bytes per step from the committed counter
                summary (profiles/step_<config>_traffic.json)
  cpu_baseline  the oracle (scalar C restatement of DPSNN's step) on one host
                core, on the benchmarked network continued from the
                engine's state after the timed steps (a few steps); beside
                it the 2x2-column network on the host: one oracle process per core of
                this process's share, and the reference itself (built from
                its sources into oracle/_ref) under mpirun -np 1
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle
    from helpers import OracleNet
    return Oracle, OracleNet


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_share():
    """Host cores this process may use: its affinity mask, capped by
    OMP_NUM_THREADS where set (the box's CPU share)."""
    n = len(os.sched_getaffinity(0))
    k = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, k) if k > 0 else n


def cpu_baseline_port(pkg, cfg, ckpt, backward, t_ckpt, steps):
    """The oracle (scalar C restatement of the reference's step) on the
    same network, continued from the engine's checkpoint of ms t_ckpt (the
    network's state after the timed region; the oracle's own ran3 drive
    replaces the Philox drive from there), timed over `steps` steps on one
    host core."""
    Oracle, OracleNet = _oracle()
    t0 = time.perf_counter()
    ora = OracleNet(Oracle(), cfg, backward=backward)
    ora.load_checkpoint(ckpt)
    build_s = time.perf_counter() - t0
    e0 = ora.lib.ora_synaptic_events(ora.h)
    t0 = time.perf_counter()
    for _ in range(steps):
        ora.fire()
    el = time.perf_counter() - t0
    ev = ora.lib.ora_synaptic_events(ora.h) - e0
    ora.close()
    return {"value": ev / el, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"oracle (scalar C restatement of DPSNN's step, one host core) on this workload's network "
                      f"(built from the engine's backward list in {build_s:.0f} s, not timed), continued from the "
                      f"engine's state at ms {t_ckpt} with its own ran3 bath: {steps} steps, {ev} synaptic "
                      f"events in {el:.1f} s ({el / steps:.2f} s per simulated ms)",
            "cpu_model": cpu_model()}


def cpu_baseline_c0(pkg, ms):
    """the 2x2-column network on the host: (a) the oracle's driver, one process per host
    core of this process's share, all at once; (b) the reference itself
    (oracle/_ref, built from the reference's sources) at MPI world size 1,
    where that binary and mpirun are present.  Both run `ms` simulated ms;
    the event count of (b) is (a)'s (the same network and raster)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_build", "dpsnn_cpu_ref")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    cfg = pkg.config_dir("c0")
    env = dict(os.environ)
    env.update(pkg.read_env_file(os.path.join(cfg, "env.txt")))
    env["env_spikes_writeDisable"] = "1"
    P = cpu_share()
    tmp = tempfile.mkdtemp(prefix="dpsnn_cpu_")
    procs = []
    for k in range(P):
        out = os.path.join(tmp, f"p{k}")
        os.makedirs(out)
        procs.append(subprocess.Popen([exe, cfg, out, "--steps", str(ms)], env=env, stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    runs = []
    for p in procs:
        text = p.communicate()[0]
        if p.returncode != 0:
            raise RuntimeError(f"dpsnn_cpu_ref failed: {text[-400:]}")
        line = [x for x in text.splitlines() if x.startswith("steps ")][0]
        runs.append((float(line.split(" in ")[1].split(" s")[0]), int(line.rsplit("synaptic events ", 1)[1])))
    ev = runs[0][1]
    assert all(r[1] == ev for r in runs)
    wall = max(r[0] for r in runs)
    res = {"per_core": {"value": P * ev / wall, "unit": "synaptic events/s", "cores": P, "kind": "port",
                        "sample": f"{P} concurrent oracle processes (dpsnn_cpu_ref), each on the 2x2-column network "
                                  f"(5000 neurons) for {ms} ms after init; slowest step loop {wall:.1f} s"}}
    ref = os.path.join(ROOT, "oracle", "_ref", "LIFCA-DPSNN.out")
    mpirun = "/opt/conda/bin/mpirun"
    if os.path.exists(ref) and os.path.exists(mpirun):
        d = os.path.join(tmp, "ref")
        os.makedirs(d)
        for name in ("column.txt", "connectivity.txt", "bathEfficacy.txt"):
            with open(os.path.join(cfg, name)) as f, open(os.path.join(d, name), "w") as g:
                g.write(f.read())
        renv = dict(env, env_totalSimTime_ms=str(ms))
        text = subprocess.run([mpirun, "-np", "1", ref], cwd=d, env=renv, capture_output=True, text=True,
                              timeout=600).stdout
        line = [x for x in text.splitlines() if "mean localNetProcess_fire" in x]
        if line:
            fire_s = float(line[0].rsplit("total fire time =", 1)[1].split()[0])
            res["reference"] = {"value": ev / fire_s, "unit": "synaptic events/s", "cores": 1, "kind": "reference",
                                "sample": f"the reference (oracle/_ref, built from its sources) under mpirun -np 1 on "
                                          f"the 2x2-column network for {ms} ms: {fire_s:.1f} s of localNetProcess_fire "
                                          f"(its own CHRONO line), events as counted by the oracle on the same "
                                          f"raster"}
    else:
        res["reference"] = None
    return res


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batch(pkg, args):
    """the batching mix: 64 independent networks (grid side 2..12, own seeds, bath
    2-6 Hz per input, STDP on half) advancing in one batched step; then the
    same networks stepped one at a time for the same number of steps."""
    import tempfile
    W = importlib.import_module("201812realtimecortsim_amd.workloads")
    mix = W.batch_mix(count=args.batch_count)
    tmp = tempfile.mkdtemp(prefix="dpsnn_batch_")
    specs = []
    for i, spec in enumerate(mix):
        d = os.path.join(tmp, f"n{i:02d}")
        env = W.write_instance(d, spec)
        specs.append((d, dict(drive=pkg.DRIVE_PHILOX, philox_seed=7000 + i, env=env, drop_host_synapses=True)))
    log(f"building {len(mix)} networks (host builds on parallel threads) ...")
    t0 = time.perf_counter()
    nets = pkg.Network.build_many(specs, host_threads=2)
    init_s = time.perf_counter() - t0
    log(f"built {len(nets)} networks in {init_s:.0f} s")
    neurons = sum(n.n for n in nets)
    synapses = sum(n.stats()["synapses"] for n in nets)
    dev_bytes = sum(n.stats()["device_bytes"] for n in nets)
    b = pkg.Batch()
    for n in nets:
        b.add(n)
    b.step(args.warmup)
    b.sync()
    shape = b.tune(1) if not args.no_tune else b.shape()
    log(f"tuned shape {shape}")
    b.step(5)
    b.sync()

    def totals():
        st = [n.stats() for n in nets]
        return sum(x["synaptic_events"] for x in st), sum(x["spikes"] for x in st)

    e0, s0 = totals()
    b.profile(True)
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    kms, launches = b.kernel_time()
    e1, s1 = totals()
    batched = (e1 - e0) / wall
    log(f"batched: {1000 * wall / args.steps:.2f} ms per step, {batched:.3e} events/s")
    for n in list(nets):
        b.remove(n)
    b.close()
    pe, pw = 0, 0.0
    per_net = []
    for n in nets:
        a = n.stats()["synaptic_events"]
        n.step(1)  # stand-alone scratch
        n.sync()
        a = n.stats()["synaptic_events"]
        t0 = time.perf_counter()
        n.step(args.steps)
        n.sync()
        dt = time.perf_counter() - t0
        de = n.stats()["synaptic_events"] - a
        per_net.append((de, dt))
        pw += dt
        pe += de
    plain = pe / pw
    log(f"plain (one network at a time): {plain:.3e} events/s")
    # the same comparison per size class: a batch of only that class's members
    classes = {}
    for lo, hi in ((2, 4), (5, 8), (9, 12)):
        idx = [i for i, m in enumerate(mix) if lo <= m["side"] <= hi]
        if not idx:
            continue
        with pkg.Batch() as bc:
            for i in idx:
                bc.add(nets[i])
            bc.step(2)
            bc.sync()
            ev0 = sum(nets[i].stats()["synaptic_events"] for i in idx)
            t0 = time.perf_counter()
            bc.step(args.steps)
            bc.sync()
            wall_c = time.perf_counter() - t0
            ev1 = sum(nets[i].stats()["synaptic_events"] for i in idx)
            for i in idx:
                bc.remove(nets[i])
        b_rate = (ev1 - ev0) / wall_c
        p_rate = sum(per_net[i][0] for i in idx) / sum(per_net[i][1] for i in idx)
        classes[f"side {lo}-{hi}"] = {"networks": len(idx), "neurons": sum(nets[i].n for i in idx),
                                      "batched_events_per_s": b_rate, "plain_events_per_s": p_rate,
                                      "batched_vs_plain": b_rate / p_rate,
                                      "batched_ms_per_step": 1000.0 * wall_c / args.steps,
                                      "plain_ms_per_step": 1000.0 * sum(per_net[i][1] for i in idx) / args.steps}
        log(f"side {lo}-{hi}: {len(idx)} networks, batched {b_rate:.3e} vs plain {p_rate:.3e} events/s "
            f"({b_rate / p_rate:.2f}x)")
    for n in nets:
        n.close()
    per_launch = (25.0 * (e1 - e0) + 72.0 * neurons * args.steps) / args.steps
    achieved = per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": batched,
        "unit": "synaptic events/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"batching: {len(mix)} independent LIFCA networks, grid side 2-12 "
                               f"columns x 1250 neurons, own seeds, bath 2-6 Hz/input, STDP on "
                               f"{sum(x['stdp'] for x in mix)}; one batched step = 1 ms of every network",
                   "networks": len(mix), "neurons": neurons, "synapses": synapses, "device_bytes": dev_bytes,
                   "drive": "philox", "init_seconds": init_s, "tuned_shape": shape,
                   "mean_rate_hz": (s1 - s0) / args.steps * 1000.0 / neurons},
        "batched_vs_plain": batched / plain,
        "plain_events_per_s": plain,
        "by_size_class": classes,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "batched step (graph of k_batch_gather + k_batch_integrate per chunk)",
                     "kernel_ms": kms, "launches": launches, "algorithmic_bytes_per_launch": per_launch},
        "cpu_baseline": None,
    }


def run_server(pkg, jobs, dirs, envs, args, max_running):
    """Replays the job arrivals against a server for the window; returns
    what it finished and delivered by the window's end."""
    gen = pkg.GENERATOR_DEVICE if args.generator == "device" else pkg.GENERATOR_HOST
    srv = pkg.Server(hbm_budget_bytes=int(args.serve_budget_gb * 1e9), build_threads=args.build_threads,
                     steps_per_round=10, max_running=max_running)
    t0 = time.perf_counter()
    ids, k, last = [], 0, t0
    while True:
        now = time.perf_counter() - t0
        while k < len(jobs) and jobs[k]["arrival_s"] <= now:
            ids.append(srv.submit(dirs[k], drive=pkg.DRIVE_PHILOX, philox_seed=9000 + k, env=envs[k],
                                  drop_host_synapses=True, generator=gen, generator_seed=9000 + k))
            k += 1
        if now >= args.serve_window:
            break
        if time.perf_counter() - last >= 10.0:
            st = srv.stats()
            log(f"{'plain' if max_running == 1 else 'served'} t={now:5.0f} s: submitted {st['submitted']} "
                f"queued {st['queued']} building {st['building']} waiting {st['waiting']} running {st['running']} "
                f"finished {st['finished']} events {st['synaptic_events']:.3e}")
            last = time.perf_counter()
        time.sleep(0.005)
    st = srv.stats()
    infos = [srv.job(j) for j in ids]
    srv.close()
    done = [i for i in infos if i["state"] == "done"]
    lat = np.array([i["finished_s"] - i["submitted_s"] for i in done]) if done else np.zeros(0)
    pct = (lambda q: float(np.percentile(lat, q)) if lat.size else None)
    return {"submitted": len(ids), "finished": len(done), "failed": sum(i["state"] == "failed" for i in infos),
            "running_at_end": st["running"], "waiting_at_end": st["waiting"], "building_at_end": st["building"],
            "queued_at_end": st["queued"],
            "events": st["synaptic_events"], "elapsed_s": st["elapsed_s"],
            "events_per_s": st["synaptic_events"] / st["elapsed_s"],
            "latency_p50_s": pct(50), "latency_p99_s": pct(99),
            "finished_sim_ms": sum(i["stats"]["steps"] for i in done),
            "step_s": st["step_s"], "upload_s": st["upload_s"], "admit_s": st["admit_s"],
            "account_s": st["account_s"], "retire_s": st["retire_s"]}


def bench_serve(pkg, args):
    """the serving stream: Poisson job arrivals (20 jobs/s for 120 s), each job a
    LIFCA network of 1-16 x 1-16 columns, 0.5-5 s simulated, STDP with
    probability 0.3, admitted under a 200 GB HBM budget with continuous
    batching; the same arrival stream served one job at a time is the plain
    comparison."""
    import tempfile
    W = importlib.import_module("201812realtimecortsim_amd.workloads")
    jobs = W.serve_jobs(rate_per_s=args.serve_rate, window_s=args.serve_window)
    tmp = tempfile.mkdtemp(prefix="dpsnn_serve_")
    dirs, envs = [], []
    for i, spec in enumerate(jobs):
        d = os.path.join(tmp, f"j{i:05d}")
        envs.append(W.write_instance(d, spec))
        dirs.append(d)
    # offered work at the saturated rate of these networks (~400 Hz, 1125
    # synapses per neuron)
    offered = sum(j["side"] ** 2 * 1250 * 0.40 * 1125 * j["duration_ms"] for j in jobs) / args.serve_window
    log(f"{len(jobs)} jobs over {args.serve_window:.0f} s; offered ~{offered:.3e} synaptic events/s")
    served = run_server(pkg, jobs, dirs, envs, args, 0)
    log(f"served: {served}")
    plain = run_server(pkg, jobs, dirs, envs, args, 1)
    log(f"plain: {plain}")
    return {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": served["events_per_s"],
        "unit": "synaptic events/s",
        "n_gpus": 1, "steps": None, "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"serving: Poisson arrivals {args.serve_rate:g} jobs/s for "
                               f"{args.serve_window:g} s, grid side 1-16 x 1250 neurons, 0.5-5 s simulated, "
                               f"STDP p=0.3, {args.serve_budget_gb:g} GB HBM admission budget, continuous batching",
                   "jobs_arrived": len(jobs), "offered_events_per_s": offered, "drive": "philox",
                   "generator": args.generator, "build_threads": args.build_threads},
        "served": served,
        "plain": plain,
        "served_vs_plain": served["events_per_s"] / plain["events_per_s"] if plain["events_per_s"] else None,
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_copy(args):
    """A device-to-device copy of known size (256 MiB, five times): the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 256 * 1024 * 1024
    x = torch.empty(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes": n, "copies": 5}


def traffic_from_profile(config):
    """HBM bytes per step of this workload from the committed counter summary
    (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, scaled by the
    same counters read for a plain copy) -- measured by a profiled run of
    the engine, not by this run -- with where and over which steps it was
    measured; or None."""
    path = os.path.join(ROOT, "profiles", f"step_{config}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        j = json.load(f)
    return {"bytes": j.get("hbm_bytes_per_step"),
            "source": f"profiles/step_{config}_traffic.json (counter passes over {j.get('workload', '?')})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--settle", type=int, default=100,
                    help="untimed ms run before the warm-up steps: the start-up transient of the "
                         "network (rates still rising) is not part of the timed regime")
    ap.add_argument("--config", default="c1")
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve", "copy"])
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-window", type=float, default=120.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--build-threads", type=int, default=8)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--drive", default="philox", choices=["philox", "replay"])
    ap.add_argument("--generator", default="host", choices=["host", "device"],
                    help="connectome generator: the reference's (bit-exact) or the device one (statistical)")
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="steps of the oracle on the benchmarked network (from the engine's checkpoint)")
    ap.add_argument("--cpu-c0-ms", type=int, default=40,
                    help="simulated ms of the host runs of the 2x2-column network (one process per core; the reference)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the engine runs one network per GPU and measures one GPU; "
              "--gpus must be 1 (multi-GPU sharding is a design note, not a runnable mode)",
              file=sys.stderr)
        return 2
    pkg = importlib.import_module("201812realtimecortsim_amd")
    if args.mode == "batch":
        print(json.dumps(bench_batch(pkg, args)), flush=True)
        return 0
    if args.mode == "serve":
        print(json.dumps(bench_serve(pkg, args)), flush=True)
        return 0
    if args.mode == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return 0
    cfg = pkg.config_dir(args.config)
    total = args.settle + args.warmup + args.steps
    pkg.apply_env(cfg, env_totalSimTime_ms=max(total, 1), env_spikes_writeDisable=1)
    drive = pkg.DRIVE_PHILOX if args.drive == "philox" else pkg.DRIVE_REPLAY
    print(f"[bench] building {args.config} ...", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    net = pkg.Network(cfg, drive=drive, philox_seed=20181201,
                      generator=pkg.GENERATOR_DEVICE if args.generator == "device" else pkg.GENERATOR_HOST)
    init_s = time.perf_counter() - t0
    st0 = net.stats()
    print(f"[bench] init {init_s:.1f} s: {st0['neurons']} neurons, {st0['synapses']} synapses, "
          f"{st0['device_bytes'] / 1e9:.1f} GB on device", file=sys.stderr, flush=True)
    net.step(args.settle)
    net.step(args.warmup)
    net.sync()
    a = net.stats()
    t_first = args.settle + args.warmup
    net.profile(True)
    t0 = time.perf_counter()
    net.step(args.steps)
    net.sync()
    wall = time.perf_counter() - t0
    b = net.stats()
    kms, launches = net.kernel_time(0)
    if not args.no_cpu_baseline:
        # the network's state after the timed region, for the host baseline
        import tempfile
        ck_dir = tempfile.mkdtemp(prefix="dpsnn_bench_")
        ck = os.path.join(ck_dir, "net.ckpt")
        net.checkpoint(ck)
        t_ck = net.now_ms
        syn = net.synapses()
        backward = {k: syn[k] for k in ("pre", "post", "delay", "weight")}
        del syn
    net.close()
    traffic = traffic_from_profile(args.config)
    par = pkg.get_parameters(cfg)
    grid, npc = f"{par['globCFX']}x{par['globCFY']}", par["neuronsPerCM"]
    plastic = os.path.exists(os.path.join(cfg, "plasticity.txt"))
    events = b["synaptic_events"] - a["synaptic_events"]
    ext = b["external_events"] - a["external_events"]
    spikes = b["spikes"] - a["spikes"]
    neurons = st0["neurons"]
    per_launch = (25.0 * events + 72.0 * neurons * args.steps +
                  (16.0 * ext if drive == pkg.DRIVE_REPLAY else 0.0)) / max(args.steps, 1)
    achieved = per_launch / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    line = {
        "metric": "synaptic events/s + real-time factor on 1 GPU; batched/served vs plain",
        "value": events / wall,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "real_time_factor": (wall / (args.steps / 1000.0)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (reference generator, seeded)",
        "config": {"workload": f"plain LIFCA step, {grid} grid of {npc}-neuron columns "
                               f"({st0['neurons']} neurons, {st0['synapses']} synapses)"
                               f"{', pair-based STDP' if plastic else ''}, {args.drive} external drive "
                               f"(config directory {args.config})",
                   "neurons": neurons, "synapses": st0["synapses"], "drive": args.drive,
                   "generator": args.generator,
                   "spikes_per_step": spikes / args.steps,
                   "mean_rate_hz": spikes / args.steps * 1000.0 / neurons,
                   "init_seconds": init_s,
                   "timed_ms": [t_first, t_first + args.steps],
                   "phase1": ("stencil window (k_collect + k_sort)" if b["window_form"] else
                              "in-list gather (k_gather)") if t_first >= 32 else "mixed (steps before 32 gather)",
                   "graph_replayed_steps": b["graph_steps"] - a["graph_steps"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "one step's device work (phase 1 + k_integrate), HIP events on the "
                               "network's stream",
                     "kernel_ms": kms, "launches": launches,
                     "algorithmic_bytes_per_launch": per_launch},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle on {args.config}, {args.cpu_steps} steps from ms {t_ck} ...")
        pkg.apply_env(cfg, env_totalSimTime_ms=t_ck + args.cpu_steps + 1, env_spikes_writeDisable=1)
        base = cpu_baseline_port(pkg, cfg, ck, backward, t_ck, args.cpu_steps)
        del backward
        os.remove(ck)
        os.rmdir(ck_dir)
        log(f"cpu baseline: {base['sample']}")
        log(f"cpu baseline: the 2x2-column network, {args.cpu_c0_ms} ms, one process per core and the reference ...")
        base.update(cpu_baseline_c0(pkg, args.cpu_c0_ms))
        line["cpu_baseline"] = base
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

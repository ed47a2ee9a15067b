"""DPSNN ini renderings of the workloads, and the instance mixes of the
batching and serving workloads.

A network is written as DPSNN's own input files -- column.txt,
connectivity.txt, bathEfficacy.txt, optional plasticity.txt (formats of
DPSNN_LIFCAconnectome.cpp:302-398, :16-131, :240-300 and
DPSNN_getParameters.c:900-1053) -- plus env.txt, one KEY=VALUE line per
environment variable that getParameters reads (DPSNN_getParameters.c:23-831).

Common neuron / synapse parameters:
  3 sub-populations per column: 500 + 500 excitatory, 250 inhibitory;
  tau_m 20 / 10 ms (exc / inh), theta 20 mV, reset 15 mV, refractory 2 ms,
  calcium tau_c 500 ms, alpha_c 1 and g_c 0.05 on excitatory neurons only;
  external Poisson bath 800 inputs x NuExt Hz at JExt = 0.5 mV (DJExt = 0);
  J ~ N(0.6, 0.06) mV from excitatory, N(-2.4, 0.24) mV from inhibitory
  sources; stored integer delays 0..9 (1-10 ms, uniform) exc, 0 (1 ms) inh;
  connection probability exp(-d) over a 3x3 column stencil (lambda = 1
  column), normalised so that the expected out-degree averaged over the
  grid's columns is 1125: a 1x1 grid keeps it inside the column, in a 2x2
  grid every column is a corner (4 reachable columns), in larger grids
  corner, edge and interior columns reach 4, 6 and 9 cells (the reference's
  generator drops stencil cells past the grid's edges), so an interior
  column's out-degree is somewhat above 1125 and an edge column's below.
"""
import math
import os

import numpy as np

BASE_ENV = {
    "env_overallConnectivity": 0, "env_stencilDim": 2, "env_globH": 1, "env_synGen": 2,
    "env_minSynWeight_f": -6.0, "env_D": 20, "env_startPlasticity_ms": 2000000000,
    "env_stopPlasticity_ms": 0, "env_moduloSec": 1, "env_startStatFiles_s": 0,
    "env_stopStatFiles_s": 1000, "env_spikes_writeDisable": 0,
    "env_statSTDPevent_writeDisable": 1, "env_statSynPeriodicProbe_writeDisable": 1,
    "env_statRatesPerPop_writeDisable": 0, "env_startRatesPerPop_s": 0,
    "env_stopRatesPerPop_s": 1000, "env_startSynPeriodicProbe_s": 0,
    "env_stopSynPeriodicProbe_s": 0, "env_startSTDPevent_s": 0, "env_stopSTDPevent_s": 0,
    "env_ratesSampling": 5, "env_globalSeed": 1, "env_howManyOutputFiles": 1,
    "env_collectivesAlgo": 0, "env_pktLength": 0,
}

# Pair-based STDP (algo 1): antiCausalCoeff -1, STDPmultiplier = Jmax 1.2 mV,
# tau 20 ms, A+ 0.1 and A- 0.12 (fractions of Jmax), derivative reset every
# second (tauDerivativeDecay 1 s), no drift.
STDP = "1 -1.0 1.2 20 0.1 0.12 1.0 0.0\n"
PLASTIC_ENV = {"env_startPlasticity_ms": 0, "env_stopPlasticity_ms": 2000000000}
BATH = "800 800\n800 800\n"


def column_txt(nu_ext=3.0, counts=(500, 500, 250)):
    return (f"0 {counts[0]} 0.5 0.0 {nu_ext:g} 20 20 15 2 1 500 0.05\n"
            f"1 {counts[1]} 0.5 0.0 {nu_ext:g} 20 20 15 2 1 500 0.05\n"
            f"2 {counts[2]} 0.5 0.0 {nu_ext:g} 10 20 15 2 0 500 0\n")


def connectivity_txt(side, npc=1250, out_degree=1125.0):
    """Stencil for a side x side grid (normalisation as in the module doc)."""
    p = {(0, 0): 1.0, (1, 0): math.exp(-1.0), (0, 1): math.exp(-1.0),
         (1, 1): math.exp(-math.sqrt(2.0))}
    if side == 1:
        cells = {(0, 0): 1.0, (1, 0): 0.0, (0, 1): 0.0, (1, 1): 0.0}
        mass = 1.0
    else:
        cells = p
        if side == 2:
            mass = sum(p.values())
        else:
            # the stencil mass a column reaches, averaged over the grid
            reach = 0.0
            for x in range(side):
                for y in range(side):
                    for dx in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            if 0 <= x + dx < side and 0 <= y + dy < side:
                                reach += p[(abs(dx), abs(dy))]
            mass = reach / (side * side)
    scale = out_degree / npc / mass
    vals = [cells[(x, y)] * scale for y in range(2) for x in range(2)]
    lines = []
    for src in range(3):
        for tgt in range(3):
            if src < 2:
                J, DJ, dmin, dmax = 0.6, 0.1, 1, 10
            else:
                J, DJ, dmin, dmax = -2.4, 0.1, 1, 1
            lines.append(f"{src} {tgt} {J} {DJ} {dmin} {dmax} " + " ".join(f"{v:.6f}" for v in vals))
    return "\n".join(lines) + "\n"


def write_network(d, env, column, connectivity, bath=BATH, plasticity=None):
    """Writes one network's ini files and env.txt into directory d."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "column.txt"), "w") as f:
        f.write(column)
    with open(os.path.join(d, "connectivity.txt"), "w") as f:
        f.write(connectivity)
    with open(os.path.join(d, "bathEfficacy.txt"), "w") as f:
        f.write(bath)
    pl = os.path.join(d, "plasticity.txt")
    if plasticity:
        with open(pl, "w") as f:
            f.write(plasticity)
    elif os.path.exists(pl):
        os.remove(pl)
    with open(os.path.join(d, "env.txt"), "w") as f:
        for k, v in env.items():
            f.write(f"{k}={v}\n")


def instance_env(side, seed, duration_ms, stdp, raster=False):
    env = dict(BASE_ENV, env_CFX=side, env_CFY=side, env_globalSeed=seed, env_totalSimTime_ms=duration_ms)
    if stdp:
        env.update(PLASTIC_ENV)
    if not raster:
        env.update(env_spikes_writeDisable=1, env_howManyOutputFiles=0)
    return env


def write_instance(d, spec, raster=False):
    """Writes the network of one mix entry (see batch_mix / serve_jobs) and
    returns its env dict."""
    env = instance_env(spec["side"], spec["seed"], spec["duration_ms"], spec["stdp"], raster)
    write_network(d, env, column_txt(spec["nu_ext"]), connectivity_txt(spec["side"]),
                  plasticity=STDP if spec["stdp"] else None)
    return env


def batch_mix(count=64, seed=20181203):
    """The batching workload: `count` independent networks, grid side drawn
    uniformly from 2..12 columns of 1250 neurons, each with its own seed and
    an external rate of 2-6 Hz per input synapse (800 inputs x 0.5 mV),
    5 s simulated each, STDP on for half of them."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(2, 13, size=count)
    rates = rng.uniform(2.0, 6.0, size=count)
    plastic = np.zeros(count, bool)
    plastic[rng.permutation(count)[: count // 2]] = True
    return [{"side": int(sides[i]), "seed": 1000 + i, "nu_ext": round(float(rates[i]), 4),
             "duration_ms": 5000, "stdp": bool(plastic[i])} for i in range(count)]


def serve_jobs(rate_per_s=20.0, window_s=120.0, seed=20181204):
    """The serving workload: Poisson arrivals at rate_per_s over window_s;
    each job a network of side 1..16 columns (1250 neurons each), 0.5-5 s
    simulated, STDP with probability 0.3, the bath of the 2x2-column network."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t >= window_s:
            break
        jobs.append({"arrival_s": t, "side": int(rng.integers(1, 17)), "seed": 5000 + len(jobs),
                     "nu_ext": 3.0, "duration_ms": int(rng.integers(500, 5001)),
                     "stdp": bool(rng.random() < 0.3)})
    return jobs

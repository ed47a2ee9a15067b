"""Writes the canonical ini renderings of the five workloads.

Each workload directory holds DPSNN's own input files -- column.txt,
connectivity.txt, bathEfficacy.txt, optional plasticity.txt (formats of
DPSNN_LIFCAconnectome.cpp:302-398, :16-131, :240-300 and
DPSNN_getParameters.c:900-1053) -- plus env.txt, one KEY=VALUE line per
environment variable that getParameters reads (DPSNN_getParameters.c:23-831).

Common neuron / synapse parameters (all workloads):
  3 sub-populations per column: 500 + 500 excitatory, 250 inhibitory;
  tau_m 20 / 10 ms (exc / inh), theta 20 mV, reset 15 mV, refractory 2 ms,
  calcium tau_c 500 ms, alpha_c 1 and g_c 0.05 on excitatory neurons only;
  external Poisson bath 800 inputs x NuExt Hz at JExt = 0.5 mV (DJExt = 0);
  J ~ N(0.6, 0.06) mV from excitatory, N(-2.4, 0.24) mV from inhibitory
  sources; stored integer delays 0..9 (1-10 ms, uniform) exc, 0 (1 ms) inh;
  connection probability exp(-d) over a 3x3 column stencil (lambda = 1 column).

The stencil is normalised so that the expected out-degree is 1125: for the
2x2 grid every column is a corner (4 reachable columns); for larger grids
interior columns see all 9 stencil cells.
"""
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

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


def column_txt(nu_ext=3.0, counts=(500, 500, 250)):
    e, i = counts[0], counts[2]
    return (f"0 {counts[0]} 0.5 0.0 {nu_ext:g} 20 20 15 2 1 500 0.05\n"
            f"1 {counts[1]} 0.5 0.0 {nu_ext:g} 20 20 15 2 1 500 0.05\n"
            f"2 {counts[2]} 0.5 0.0 {nu_ext:g} 10 20 15 2 0 500 0\n")


def connectivity_txt(corner_normalised, npc=1250, out_degree=1125.0):
    p = {(0, 0): 1.0, (1, 0): math.exp(-1.0), (0, 1): math.exp(-1.0),
         (1, 1): math.exp(-math.sqrt(2.0))}
    if corner_normalised:   # a corner column reaches (0,0), (1,0), (0,1), (1,1)
        mass = sum(p.values())
    else:                   # an interior column reaches all 9 mirrored cells
        mass = p[(0, 0)] + 2 * p[(1, 0)] + 2 * p[(0, 1)] + 4 * p[(1, 1)]
    scale = out_degree / npc / mass
    cells = [p[(x, y)] * scale for y in range(2) for x in range(2)]
    lines = []
    for src in range(3):
        for tgt in range(3):
            if src < 2:
                J, DJ, dmin, dmax = 0.6, 0.1, 1, 10
            else:
                J, DJ, dmin, dmax = -2.4, 0.1, 1, 1
            lines.append(f"{src} {tgt} {J} {DJ} {dmin} {dmax} " +
                         " ".join(f"{v:.6f}" for v in cells))
    return "\n".join(lines) + "\n"


def write(name, env, column, connectivity, bath="800 800\n800 800\n", plasticity=None):
    d = os.path.join(HERE, name)
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


# Pair-based STDP (algo 1): antiCausalCoeff -1, STDPmultiplier = Jmax 1.2 mV,
# tau 20 ms, A+ 0.1 and A- 0.12 (fractions of Jmax), derivative reset every
# second (tauDerivativeDecay 1 s), no drift.
STDP = "1 -1.0 1.2 20 0.1 0.12 1.0 0.0\n"


def main():
    c0 = dict(BASE_ENV, env_CFX=2, env_CFY=2, env_totalSimTime_ms=1000)
    write("c0", c0, column_txt(), connectivity_txt(True))
    c1 = dict(BASE_ENV, env_CFX=24, env_CFY=24, env_totalSimTime_ms=10000,
              env_spikes_writeDisable=1, env_howManyOutputFiles=0)
    write("c1", c1, column_txt(), connectivity_txt(False))
    c2 = dict(c1, env_startPlasticity_ms=0, env_stopPlasticity_ms=2000000000)
    write("c2", c2, column_txt(), connectivity_txt(False), plasticity=STDP)
    # configs[2]'s STDP at configs[0]'s size: the parity case for plasticity
    c2s = dict(c0, env_totalSimTime_ms=1200, env_startPlasticity_ms=0,
               env_stopPlasticity_ms=2000000000)
    write("c2_small", c2s, column_txt(), connectivity_txt(True), plasticity=STDP)


if __name__ == "__main__":
    sys.exit(main())

"""The percolation stop of the main loop (complab.cpp:3097-3117, 3261): a flow
re-solve that runs out of iterations while nothing flows through the outlet
plane x = nx-2 ends the run after that step; one that runs out while the
outlet still flows only warns.  Checked on the oracle with a channel whose
cross-section an impermeable biofilm (viscosity ratio -99 -> BounceBack,
complab_functions.hh:1386-1388) plugs, or plugs but for one voxel, at the
first step's mask update.  (A precipitated voxel takes the solid code, which
updateNsLatticesDynamics3D leaves alone (part2:779), so in the reference
precipitation never closes the flow lattice and cannot clog it.)"""
import numpy as np

import oracle as O
from complab_amd import driver
from complab_amd import params as P


def sealed_channel(full_seal=True, nxf=10, w=3):
    nx, ny, nz = nxf + 2, w + 2, w + 2
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.c0[0] = 0.3
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0; p.right_btype[0] = P.BC_NEUMANN
    p.n_microbes = 1
    p.bmax = 100.0
    p.thrd_bfilm_frac = 0.1
    p.perm_ratio = 0.0                               # impermeable biofilm: bounce-back
    p.ns_update_interval = 1
    p.ade_update_interval = 1
    p.flow_on = 1
    p.ns_tau = 0.8
    # the re-solves: the tracer needs a full window of 500 samples before it
    # can declare convergence (ValueTracer), so a cap of 450 iterations always
    # runs out -- whether the plugged flow has stopped or not
    p.ns_conv_rule = P.CONV_TRACER
    p.ns_conv_eps = 1e-6
    p.ns_max_iter = 450
    fs = driver.FlowSetup(pe=0.5, char_length=float(nxf), delta_p=1e-4, rule=P.CONV_RELDELTA, eps=1e-10,
                          max_iter=20000, has_biofilm=True)
    ts = driver.TransportSetup(d_pore=[0.1], d_biofilm=[0.1], stabilize_steps=0)
    plug = np.zeros(geom.shape)
    xs = nx // 2
    plug[1:-1, 1:-1, xs] = 50.0                       # above thrd * Bmax = 10, below Bmax: no spill
    if not full_seal:
        plug[2, 2, xs] = 0.0                         # the middle of the cross-section stays open
    return p, geom, fs, ts, plug


def run_case(case, p, fs, ts, seal, steps=12):
    rep = driver.setup(case, p, fs, ts)
    case.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    case.write(P.FIELD_BIOMASS, 0, seal)
    return case.run_steps(steps)


def test_sealed_outlet_stops_the_run():
    p, geom, fs, ts, seal = sealed_channel(True)
    o = O.OracleCase(p, geom)
    rc, done = run_case(o, p, fs, ts, seal)
    assert rc == P.CLOGGED and done == 1                 # the first step's mask update plugs and clogs
    c = o.counters()
    assert c[4] == 1 and c[5] == 1 and c[7] == 1
    assert o.run_steps(5) == (P.CLOGGED, 0)              # a clogged case refuses further steps


def test_partly_sealed_outlet_keeps_running():
    p, geom, fs, ts, seal = sealed_channel(False)
    o = O.OracleCase(p, geom)
    rc, done = run_case(o, p, fs, ts, seal)
    assert rc == 0 and done == 12
    c = o.counters()
    assert c[4] == 1                                     # the re-solve ran to its cap, the outlet still flows
    ux = o.read(P.FIELD_VEL, 0)
    assert ux[2, 2, p.nx - 2] > 1e-8

"""Mineral dissolution on the GPU against the oracle: a calcite grain leached
by acid in a closed box (inventory, water, counters of shares parked and
collected), the reopening of dissolved grain voxels, and a reopening with the
flow re-solved around the new pore space."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import params as P
import oracle as O

from test_oracle_dissolution import calcite_case, CA, CAL, H

pytestmark = pytest.mark.gpu

REL = 1e-10


def relerr(a, b, mask=None):
    if mask is not None:
        a = a[mask]; b = b[mask]
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def compare(g, o, n_species=3):
    mg, mo = g.read(P.FIELD_MASK), o.read(P.FIELD_MASK)
    assert np.array_equal(mg, mo), "mask differs"
    assert np.array_equal(g.read(P.FIELD_PHASE), o.read(P.FIELD_PHASE)), "phase differs"
    # open voxels of the reaction box (a closed x plane is a bounce-back cell:
    # the oracle reports the populations it holds in transit, the GPU 0)
    om = mo >= 2
    om[:, :, 0] = False
    om[:, :, -1] = False
    for s in range(n_species):
        mask = om if s != CAL else None
        e = relerr(g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s), mask)
        print("species", s, "rel err", e)
        assert e <= REL
    cg, co = g.counters(), o.counters()
    print("counters gpu", cg, "oracle", co)
    assert np.array_equal(cg[8:12], co[8:12])
    assert cg[0] == co[0] and cg[5] == co[5]


def run_both(ctx, p, geom, steps):
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    for c in (g, o):
        c.set_omegas([1.2, 1.2, 1.0], [1.2, 1.2, 1.0], 1.0)
    rg, ng = g.run_steps(steps)
    ro, no = o.run_steps(steps)
    assert rg == ro == 0 and ng == no == steps
    return g, o


def test_calcite_leaching_matches_oracle(ctx):
    p, geom = calcite_case(k=5e-2, block=3, n=8, nx=12)
    g, o = run_both(ctx, p, geom, 30)
    compare(g, o)
    c = g.counters()
    assert c[8] > 0 and c[8] == c[9]
    g.close(); o.close()


def test_reopening_matches_oracle(ctx):
    p, geom = calcite_case(k=100.0, h0=1.0, interval=2, block=2)
    p.dissol_reopen_frac = 0.6
    g, o = run_both(ctx, p, geom, 9)
    compare(g, o)
    assert g.counters()[11] > 0
    g.close(); o.close()


def test_reopening_with_flow_resolve(ctx):
    """A grain in a duct dissolves and reopens; the flow is re-solved over the
    new pore space (complab.cpp:3209-3236) on both sides."""
    nx, n = 14, 8
    p, geom = calcite_case(k=60.0, h0=1.0, interval=3, block=2, nx=nx, n=n)
    p.dissol_reopen_frac = 0.6
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL
    geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    for s in (H, CA):
        p.left_btype[s] = P.BC_DIRICHLET; p.right_btype[s] = P.BC_NEUMANN
    p.left_bc[H] = 1.0; p.left_bc[CA] = 0.0
    p.flow_on = 1
    p.ns_tau = 0.8
    p.ns_max_iter = 4000
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-9
    p.perm_ratio = 0.1
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    its = []
    for c in (g, o):
        c.ns_init(1e-3)
        its.append(c.ns_solve(p.ns_max_iter, p.ns_conv_rule, p.ns_conv_eps, True))
        c.ns_couple()
        c.set_omegas([1.2, 1.2, 1.0], [1.2, 1.2, 1.0], 1.0)
    assert its[0] == its[1]
    rg, ng = g.run_steps(8)
    ro, no = o.run_steps(8)
    assert rg == ro == 0 and ng == no == 8
    compare(g, o)
    cg = g.counters()
    assert cg[11] > 0 and cg[4] >= 1, "no reopening / no re-solve happened"
    for k in range(3):
        assert relerr(g.read(P.FIELD_VEL, k), o.read(P.FIELD_VEL, k)) <= REL
    g.close(); o.close()

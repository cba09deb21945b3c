"""Mineral dissolution on the GPU against the oracle: a calcite grain leached
by acid in a closed box (inventory, water, counters of shares parked and
collected), the reopening of dissolved grain voxels, and a reopening with the
flow re-solved around the new pore space."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import params as P
import oracle as O

from test_oracle_dissolution import calcite_case, two_phase_case, CA, CAL, DOL, H, MG

pytestmark = pytest.mark.gpu

REL = 1e-10


def relerr(a, b, mask=None):
    if mask is not None:
        a = a[mask]; b = b[mask]
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def compare(g, o, n_species=3, immobile=(CAL,)):
    mg, mo = g.read(P.FIELD_MASK), o.read(P.FIELD_MASK)
    assert np.array_equal(mg, mo), "mask differs"
    assert np.array_equal(g.read(P.FIELD_PHASE), o.read(P.FIELD_PHASE)), "phase differs"
    # open voxels of the reaction box (a closed x plane is a bounce-back cell:
    # the oracle reports the populations it holds in transit, the GPU 0)
    om = mo >= 2
    om[:, :, 0] = False
    om[:, :, -1] = False
    for s in range(n_species):
        mask = om if s not in immobile else None
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


def test_two_grain_phases_match_oracle(ctx):
    """Two declared grain phases with their own inventories and rate laws
    (first order in acid and mineral; second order in acid, zero order in
    mineral), reopening at their own thresholds."""
    p, geom = two_phase_case(k1=0.4, k2=0.02, h0=1.0)
    p.dissol_reopen_frac = 0.6
    p.dissol_interval = 3
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    for c in (g, o):
        c.set_omegas([1.2, 1.2, 1.0, 1.2, 1.0], [1.2, 1.2, 1.0, 1.2, 1.0], 1.0)
    rg, ng = g.run_steps(12)
    ro, no = o.run_steps(12)
    assert rg == ro == 0 and ng == no == 12
    compare(g, o, n_species=5, immobile=(CAL, DOL))
    c = g.counters()
    assert c[8] > 0 and c[8] == c[9] and c[11] > 0
    g.close(); o.close()


def test_precipitate_phase_beside_a_grain_phase_matches_oracle(ctx):
    """A grain phase (calcite, id 1) and a precipitate phase (id 2, a mineral
    made from Ca and Mg on the grain surface, sealing voxels that then
    dissolve back in acid): each voxel's gather skips the mineral of the
    phase it collects from."""
    p, geom = two_phase_case(k1=0.05, k2=0.0, h0=1.0)
    geom[geom == P.WALL] = P.PORE
    p.dissol_n_phases = 1
    p.c0[MG] = 0.5
    p.n_rxn = 1
    p.rxn_r1[0] = CA; p.rxn_r2[0] = MG; p.rxn_k[0] = 5.0
    p.rxn_stoich[0][CA] = -1.0; p.rxn_stoich[0][MG] = -1.0; p.rxn_stoich[0][DOL] = 1.0
    p.surface_only = 1
    P.set_phase(p, 1, DOL, 0.02, precipitate=True, k_rate=3.0, acid=H, stoich={CA: 1.0, MG: 1.0, H: -2.0})
    p.precip_on = 1; p.precip_species = DOL; p.precip_max_rho = 0.02; p.precip_interval = 2
    p.dissol_reopen_frac = 0.9
    p.dissol_interval = 2
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    for c in (g, o):
        c.set_omegas([1.2, 1.2, 1.0, 1.2, 1.0], [1.2, 1.2, 1.0, 1.2, 1.0], 1.0)
    rg, ng = g.run_steps(20)
    ro, no = o.run_steps(20)
    assert rg == ro == 0 and ng == no == 20
    compare(g, o, n_species=5, immobile=(CAL, DOL))
    c = o.counters()
    print("oracle counters", c)
    assert c[8] > 0 and c[5] > 0, "no dissolution / no seal happened"
    g.close(); o.close()

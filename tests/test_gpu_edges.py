"""Edge cases at the boundary.  Inputs the path refuses with a status code
(no open voxel, sizes and counts out of range) and degenerate or limit-sized
inputs it runs exactly as the oracle does (FP64, 1e-10 relative): a channel
one voxel wide (every y/z neighbour beyond the grid), eight mobile species
(the largest unroll), two CA microbes, x planes without an open voxel."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs
from complab_amd import params as P
import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-10


def relerr(a, b, mask=None):
    if mask is not None:
        a = a[mask]; b = b[mask]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def refused(ctx, p, geo, code):
    with pytest.raises(clb.ClbError) as e:
        clb.Case(ctx, p, geo)
    assert e.value.code == code, str(e.value)


def test_refusals(ctx):
    p = P.default_params(6, 4, 4, 1)
    refused(ctx, p, np.zeros((4, 4, 6), np.int8), P.E_GEOMETRY)          # no open voxel
    geo = np.full((4, 4, 6), P.PORE, np.int8)
    q = P.default_params(2, 4, 4, 1)
    refused(ctx, q, np.full((4, 4, 2), P.PORE, np.int8), P.E_ARG)          # nx below 3
    q = P.default_params(6, 4, 4, P.MAX_SPECIES + 1)
    refused(ctx, q, geo, P.E_ARG)
    q = P.default_params(6, 4, 4, 1); q.n_microbes = P.MAX_MICROBES + 1
    refused(ctx, q, geo, P.E_ARG)
    q = P.default_params(6, 4, 4, 1); q.left_btype[0] = 7
    refused(ctx, q, geo, P.E_ARG)
    bad = geo.copy(); bad[0, 0, 0] = 42
    refused(ctx, p, bad, P.E_GEOMETRY)


def run_both(ctx, p, geo, steps, omp, omb):
    g = clb.Case(ctx, p, geo)
    o = O.OracleCase(p, geo)
    for c in (g, o):
        c.set_omegas(omp, omb, 1.0)
    assert g.run_steps(steps) == o.run_steps(steps) == (0, steps)
    return g, o


def test_one_voxel_channel(ctx):
    nx = 17
    p = P.default_params(nx, 1, 1, 2)
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0; p.right_btype[0] = P.BC_NEUMANN
    p.left_btype[1] = P.BC_CLOSED; p.right_btype[1] = P.BC_DIRICHLET; p.right_bc[1] = 0.3
    p.n_rxn = 1; p.rxn_r1[0] = 0; p.rxn_k[0] = 2e-3; p.rxn_stoich[0][0] = -1; p.rxn_stoich[0][1] = 1
    geo = np.full((1, 1, nx), P.PORE, np.int8)
    g, o = run_both(ctx, p, geo, 300, [1.3, 1.1], [1.3, 1.1])
    # the reaction box for both species; species 0's planes are open (Dirichlet,
    # Neumann) and compare too -- a closed plane reads 0 on the device, where
    # the oracle's bounce-back cell holds what is in flight
    for s in range(2):
        assert relerr(g.read(P.FIELD_CONC, s)[..., 1:-1], o.read(P.FIELD_CONC, s)[..., 1:-1]) <= REL
    assert relerr(g.read(P.FIELD_CONC, 0), o.read(P.FIELD_CONC, 0)) <= REL
    assert np.allclose(g.species_totals(), o.species_totals(), rtol=1e-12, atol=0)


def test_eight_mobile_species(ctx):
    w = configs.config1(scale=0.25)
    p = w.params
    S = P.MAX_SPECIES
    p.n_species = S
    for s in range(S):
        p.mobile[s] = 1
        p.c0[s] = 0.1 * (s + 1)
        p.left_btype[s] = (P.BC_DIRICHLET, P.BC_NEUMANN, P.BC_CLOSED)[s % 3]
        p.left_bc[s] = 0.5
        p.right_btype[s] = (P.BC_NEUMANN, P.BC_CLOSED, P.BC_DIRICHLET)[s % 3]
        p.right_bc[s] = 0.05 * s
    p.n_rxn = P.MAX_RXN
    for r in range(P.MAX_RXN):
        p.rxn_r1[r] = 2 * r; p.rxn_r2[r] = 2 * r + 1; p.rxn_k[r] = 0.01 * (r + 1)
        p.rxn_stoich[r][2 * r] = -1; p.rxn_stoich[r][2 * r + 1] = -1; p.rxn_stoich[r][(2 * r + 2) % S] = 1
    p.flow_on = 0
    g, o = run_both(ctx, p, w.geometry, 120, [1.0 + 0.05 * s for s in range(S)], [1.0] * S)
    om = o.read(P.FIELD_MASK) >= 2
    om[:, :, 0] = False; om[:, :, -1] = False            # a closed plane reads 0 on the device
    for s in range(S):
        assert relerr(g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s), om) <= REL, s


def test_two_microbes(ctx):
    w = configs.config2(scale=0.1)
    p = w.params
    geo, _ = configs.seed_biofilm(w.geometry, 0.02, 11, code=P.BIOFILM + 1)
    p.n_microbes = 2
    p.mu_max[0] = 2e-2; p.mu_max[1] = 1.5e-2
    p.ks[1] = p.ks[0]; p.ke[1] = p.ke[0]; p.kdecay[1] = 0.0
    for s in range(p.n_species):
        p.bio_stoich[1][s] = p.bio_stoich[0][s]
    p.b0[0] = 95.0; p.b0[1] = 90.0
    p.flow_on = 0
    g, o = run_both(ctx, p, geo, 60, [1.2] * p.n_species, [1.1] * p.n_species)
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    assert np.array_equal(g.read(P.FIELD_AGE), o.read(P.FIELD_AGE))
    for m in range(2):
        assert relerr(g.read(P.FIELD_BIOMASS, m), o.read(P.FIELD_BIOMASS, m)) <= REL
    cg, co = g.counters(), o.counters()
    assert cg[6] == co[6] and cg[6] > 0                                     # CA sweeps ran
    assert np.allclose(g.field_stats(), o.field_stats(), rtol=1e-10, atol=1e-300)


def test_closed_x_planes(ctx):
    w = configs.config1(scale=0.25)
    geo = w.geometry.copy()
    geo[:, :, 0] = P.WALL; geo[:, :, -1] = P.WALL                         # no open plane voxel
    p = w.params
    for s in range(2):
        p.left_btype[s] = P.BC_CLOSED; p.right_btype[s] = P.BC_CLOSED
    p.c0[0] = 1.0
    p.flow_on = 0
    g, o = run_both(ctx, p, geo, 100, [1.5, 1.5], [1.5, 1.5])
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(2):
        assert relerr(g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s), om) <= REL
    st = g.field_stats()
    tot = st[:2, 0] + st[:2, 3]
    assert np.allclose(tot, o.field_stats()[:2, 0] + o.field_stats()[:2, 3], rtol=1e-12)


def test_flow_kernel_timing_counts_working_iterations(ctx):
    """clb_flow_kernel_time: with every iteration timed, the timed launches of
    a solve that stops on its rule equal the iterations it reports (launches
    queued after the stop return at once and are not counted), and the mean
    is a positive device time."""
    w = configs.config1(scale=0.25)
    c = clb.Case(ctx, w.params, w.geometry)
    c.set_timing_stride(1)
    c.flow_kernel_time(reset=True)
    c.ns_init(w.flow.delta_p)
    it, cv = c.ns_solve(5000, P.CONV_RELDELTA, 1e-7, reset=True)
    ms, n = c.flow_kernel_time()
    assert cv and 0 < it < 5000
    assert n == it and ms > 0
    c.flow_kernel_time(reset=True)
    assert c.flow_kernel_time() == (0.0, 0)
    c.close()

"""The diagnostics row on the device against the oracle's (FP64, 1e-10
relative per column), and the closed-box mass balance on the device: total +
held conserved while the total alone moves."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver, diagnostics as D
from complab_amd import params as P
import oracle as O

pytestmark = pytest.mark.gpu


def close_rows(a, b, tol=1e-10):
    for col in range(4):
        scale = max(np.max(np.abs(b[:, col])), 1e-300)
        e = float(np.max(np.abs(a[:, col] - b[:, col])) / scale)
        assert e <= tol, (col, e, a[:, col], b[:, col])


def closed_box(make):
    w = configs.config1(scale=0.25)
    p = w.params
    for s in range(2):
        p.left_btype[s] = P.BC_CLOSED; p.right_btype[s] = P.BC_CLOSED
    p.n_rxn = 0
    c = make(p, w.geometry)
    rng = np.random.default_rng(0)
    closed = w.geometry < 2
    for a in range(3):
        u = rng.uniform(-0.03, 0.03, w.geometry.shape)
        u[closed] = 0
        c.write(P.FIELD_VEL, a, u)
    v = rng.uniform(0, 2, w.geometry.shape)
    v[closed] = p.c0[0]
    c.write(P.FIELD_CONC, 0, v)
    c.set_omegas([1.5, 1.5], [1.5, 1.5], 1.0)
    return c


def test_closed_box_stats_and_balance(ctx):
    g = closed_box(lambda p, geo: clb.Case(ctx, p, geo))
    o = closed_box(O.OracleCase)
    close_rows(g.field_stats(), o.field_stats())          # before the first launch
    first = None
    for k in range(5):
        g.ade_steps(41); o.ade_steps(41)
        sg, so = g.field_stats(), o.field_stats()
        close_rows(sg, so)
        cons = sg[:, 0] + sg[:, 3]
        if first is None:
            first = cons
            t0 = sg[0, 0]
        assert np.all(np.abs(cons - first) <= 1e-12 * np.abs(first)), (cons, first)
    assert abs(sg[0, 0] - t0) > 1e-6 * t0
    g.close(); o.close()


def test_biotic_row_matches_oracle(ctx):
    w = configs.config2(scale=0.12)
    w.transport.stabilize_steps = 30
    w.params.ns_update_interval = 5
    w.params.ns_max_iter = 3000
    w.params.b0[0] = 95.0
    w.params.mu_max[0] = 2e-2
    g = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    driver.setup(g, w.params, w.flow, w.transport)
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    close_rows(g.field_stats(), o.field_stats())          # after the reset
    assert g.run_steps(40) == o.run_steps(40) == (0, 40)
    sg, so = g.field_stats(), o.field_stats()
    close_rows(sg, so)
    names = ["S", "E", "P1", "P2"]
    rg, rr = D.row_of(g, 40, names, ["bio"]), D.row_of(o, 40, names, ["bio"])
    assert rg.open_voxels == rr.open_voxels and rg.porosity == rr.porosity
    g.close(); o.close()


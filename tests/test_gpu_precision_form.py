"""The precision form's physics (carbonate-calcium equilibrium by FP64 Newton,
calcite precipitation sealing voxels and acid dissolution reopening them, the
flow re-solved at every change) on the GPU against the oracle in FP64, and the
reduced precisions against FP64 on the same seed."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver, study
from complab_amd import params as P
import oracle as O

pytestmark = pytest.mark.gpu


def small(prec):
    w = configs.config3(scale=1 / 16, precision=prec)
    w.flow.eps = 1e-6
    w.params.ns_conv_eps = 1e-6
    return w


def test_precision_form_fp64_matches_oracle(ctx):
    w = small(P.FP64)
    g = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    rg = driver.setup(g, w.params, w.flow, w.transport)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    assert rg.iters == ro.iters
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    for chunk in range(10):
        a, b = g.run_steps(100), o.run_steps(100)
        assert a == b == (0, 100)
    cg, co = g.counters(), o.counters()
    print("counters gpu", cg, "oracle", co)
    assert cg[4] == co[4] and cg[4] > 0                 # flow re-solves
    assert cg[2] == co[2] and cg[5] == co[5]            # clamp binds, mask changes
    assert np.array_equal(cg[8:12], co[8:12]) and cg[8] > 0
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    assert np.array_equal(g.read(P.FIELD_PHASE), o.read(P.FIELD_PHASE))
    sg, so = g.field_stats(), o.field_stats()            # diagnostics row: total, min, max, held
    for col in range(4):
        scale = max(np.max(np.abs(so[:, col])), 1e-300)
        assert np.max(np.abs(sg[:, col] - so[:, col])) <= 1e-10 * scale, (col, sg[:, col], so[:, col])
    om = o.read(P.FIELD_MASK) >= 2
    om[:, :, 0] = False; om[:, :, -1] = False
    for s in range(6):
        a, b = g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s)
        sel = om if s != configs.CALCITE_ else np.ones_like(om)
        e = float(np.max(np.abs(a[sel] - b[sel])) / max(np.max(np.abs(b[sel])), 1e-300))
        print("species", s, "rel err", e)
        assert e <= 1e-10
    g.close(); o.close()


def test_reduced_precision_against_fp64(ctx):
    # field bounds about four times, drift bounds about ten times the measured
    # values of this seeded case (FP32 transport: 5.4e-6 field, 1.0e-5 drift;
    # full FP32 with its FP64-arithmetic flow collision: 5.9e-5 field, 1.1e-3
    # drift; no mask voxel differs)
    steps = 600
    hi = study.run(ctx, small(P.FP64), steps)
    bounds = {P.FP32_ADE: (2e-5, 1e-4), P.FP32: (2e-4, 1e-2)}
    for prec in (P.FP32_ADE, P.FP32):
        lo = study.run(ctx, small(prec), steps)
        r = study.compare(lo, hi, steps)
        print(prec, r)
        field, drift = bounds[prec]
        assert r["mask_voxels_differ"] == 0
        assert max(r["field_max_rel_err"]) <= field
        assert max(r["total_drift_per_1e4_steps"]) <= drift

"""The plain-method configurations at their stated sizes, device against the
oracle (FP64: fields within 1e-10 relative, masks, ages and counters exact).

Config 1: 64x32x32 (+2 ghost columns) sphere pack, the whole set-up (flow
solve, Peclet rescale, 10 000 stabilization steps) and 2000 coupled steps on
both sides.
Config 2: 256x128x128 (+2) biotic pack.  The device first runs the config from
its own initial state (biofilm seeds at 50) until the CA spill is under way;
its fields (concentrations, biomass, ages, live mask) then seed a fresh case on
each side, which run 100 coupled steps in the spill regime.  The oracle takes
the device's flow field (a flow solve of this size is hours of CPU), and no
re-solve falls in the window, so the comparison is of transport, kinetics,
the CA spill, the mask/age updates and the counters."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver
from complab_amd import params as P
import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-10


def relerr(a, b, mask=None):
    if mask is not None:
        a = a[mask]; b = b[mask]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_config1_full_size(ctx):
    w = configs.config1(scale=1.0)
    assert (w.params.nx, w.params.ny, w.params.nz) == (66, 32, 32) and w.steps == 2000
    O.set_threads(0)
    g = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    rg = driver.setup(g, w.params, w.flow, w.transport)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    print("flow iterations gpu", rg.iters, "oracle", ro.iters, "dP", rg.delta_p, ro.delta_p, flush=True)
    assert rg.iters == ro.iters
    assert abs(rg.delta_p - ro.delta_p) <= 1e-10 * abs(ro.delta_p)
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    assert g.run_steps(w.steps) == (0, w.steps)
    assert o.run_steps(w.steps) == (0, w.steps)
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(2):
        e = relerr(g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s), om)
        print("species", s, "max rel err", e, flush=True)
        assert e <= REL
    for k in range(3):
        assert relerr(g.read(P.FIELD_VEL, k), o.read(P.FIELD_VEL, k)) <= REL
    np.testing.assert_allclose(g.species_totals(), o.species_totals(), rtol=REL, atol=0)
    np.testing.assert_allclose(g.field_stats(), o.field_stats(), rtol=REL, atol=1e-300)
    assert list(g.counters()) == list(o.counters())
    g.close(); o.close()


def test_config2_full_size_in_the_spill_regime(ctx):
    w = configs.config2(scale=1.0)
    p = w.params
    assert (p.nx, p.ny, p.nz) == (258, 128, 128)
    O.set_threads(0)
    # 1. the config from its own initial state until the spill runs
    w.transport.stabilize_steps = 0
    pre = clb.Case(ctx, p, w.geometry)
    driver.setup(pre, p, w.flow, w.transport)
    done = 0
    while pre.counters()[6] < 20 and done < 4000:
        rc, n = pre.run_steps(200)
        assert rc == 0
        done += n
    cnt = pre.counters()
    print("pre-run", done, "steps: CA sweeps", cnt[6], "mask changes", cnt[5], flush=True)
    assert cnt[6] >= 20
    mask = pre.read(P.FIELD_MASK).astype(np.int8)
    conc = [pre.read(P.FIELD_CONC, s) for s in range(p.n_species)]
    bio = pre.read(P.FIELD_BIOMASS, 0)
    age = pre.read(P.FIELD_AGE)
    pre.close()
    assert np.max(bio) > 0.9 * p.bmax
    # 2. that state on a fresh case per side; no re-solve in the window
    p.ns_update_interval = 1000000
    g = clb.Case(ctx, p, mask)
    o = O.OracleCase(p, mask)
    rg = driver.setup(g, p, w.flow, w.transport)
    for c in (g, o):
        for s in range(p.n_species):
            c.write(P.FIELD_CONC, s, conc[s])
        c.write(P.FIELD_BIOMASS, 0, bio)
        c.write(P.FIELD_AGE, 0, age)
        c.set_omegas(rg.omega_pore, rg.omega_biofilm, 1.0)
    for k in range(3):
        o.write(P.FIELD_VEL, k, g.read(P.FIELD_VEL, k))
    steps = 100
    assert g.run_steps(steps) == (0, steps)
    assert o.run_steps(steps) == (0, steps)
    cg, co = g.counters(), o.counters()
    print("window counters gpu", list(cg), "oracle", list(co), flush=True)
    assert co[6] > 0                                        # the spill ran in the window
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    assert np.array_equal(g.read(P.FIELD_AGE), o.read(P.FIELD_AGE))
    assert relerr(g.read(P.FIELD_BIOMASS, 0), o.read(P.FIELD_BIOMASS, 0)) <= REL
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(p.n_species):
        e = relerr(g.read(P.FIELD_CONC, s), o.read(P.FIELD_CONC, s), om)
        print("species", s, "max rel err", e, flush=True)
        assert e <= REL
    for k in (2, 4, 5, 6, 7):
        assert cg[k] == co[k], (k, cg[k], co[k])
    g.close(); o.close()


def config3_capped(prec, closed=False, cap=3000):
    """Config 3 at its stated 512x256x256 (+2) size with every flow solve capped
    at `cap` iterations: the precisions compared below share that (FP64) flow,
    so its convergence does not enter the comparison."""
    w = configs.config3(scale=1.0, precision=prec)
    w.flow.max_iter = cap
    w.params.ns_max_iter = cap
    if closed:
        for s in range(w.params.n_species):
            w.params.left_btype[s] = P.BC_CLOSED
            w.params.right_btype[s] = P.BC_CLOSED
    return w


def element_sums(stats):
    """Ca, C and H of the carbonate-calcium system from a diagnostics row
    (total + held per species): every reaction, the speciation and the
    precipitation / dissolution conserve them."""
    H, CO3, CA, HCO3, CAAQ, CAL = range(6)
    th = stats[:, 0] + stats[:, 3]
    return np.array([th[CA] + th[CAAQ] + th[CAL], th[CO3] + th[HCO3] + th[CAAQ] + th[CAL], th[H] + th[HCO3]])


def test_config3_full_size_closed_box_conserves_elements(ctx):
    """FP64 at 512x256x256 with every x plane closed and the speciation step
    off: with nothing entering or leaving, calcium, carbon and the proton
    balance stay constant through transport and the surface kinetics of
    calcite growth (total + held per species, to 1e-12 relative), checked
    every 50 steps for 300 steps while the clamp never binds.  (The
    speciation step is left out because the reference's limiter caps each
    species' change at 10 % on its own, part4:1461-1475, which moves the
    element totals whenever it binds.)"""
    w = config3_capped(P.FP64, closed=True)
    assert (w.params.nx, w.params.ny, w.params.nz) == (514, 256, 256)
    w.params.eq_on = 0
    g = clb.Case(ctx, w.params, w.geometry)
    rep = driver.setup(g, w.params, w.flow, w.transport)
    g.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    e0 = element_sums(g.field_stats())
    for k in range(6):
        assert g.run_steps(50) == (0, 50)
        e = element_sums(g.field_stats())
        c = g.counters()
        print("step", 50 * (k + 1), "element drift", (e - e0) / e0, "clamp binds", c[2], "mask changes", c[5], flush=True)
        assert c[2] == 0 and c[5] == 0
        assert np.all(np.abs(e - e0) <= 1e-12 * np.abs(e0))
    st = g.field_stats()
    assert st[configs.CALCITE_, 0] > 0                  # calcite grew on the grain surface
    g.close()


def test_config3_full_size_fp32_transport_against_fp64(ctx):
    """FP32 transport lattices (FP64 flow, speciation and reductions) against
    FP64 at 512x256x256 after 300 coupled steps of the workload as stated
    (acid inflow, calcite precipitation and dissolution, FP64 Newton): per
    species field error, and the seal / reopen topology."""
    out = {}
    for prec in (P.FP64, P.FP32_ADE):
        w = config3_capped(prec)
        g = clb.Case(ctx, w.params, w.geometry)
        rep = driver.setup(g, w.params, w.flow, w.transport)
        g.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        assert g.run_steps(300) == (0, 300)
        out[prec] = ([g.read(P.FIELD_CONC, s) for s in range(6)], g.read(P.FIELD_MASK), g.counters(),
                     list(rep.iters))
        g.close()
    (fh, mh, ch, ih), (fl, ml, cl, il) = out[P.FP64], out[P.FP32_ADE]
    assert ih == il                                    # the same FP64 flow
    print("counters fp64", list(ch), "fp32_ade", list(cl), flush=True)
    om = mh >= 2
    om[:, :, 0] = False; om[:, :, -1] = False
    errs = []
    for s in range(6):
        sel = om if s != configs.CALCITE_ else np.ones_like(om)
        errs.append(relerr(fl[s], fh[s], sel))
    print("fp32_ade max rel err per species", errs, "mask voxels differing", int(np.sum(ml != mh)), flush=True)
    assert max(errs) <= 1e-4
    assert int(np.sum(ml != mh)) <= 100


def test_config4_all_64_cases_in_one_batch():
    """The batched form at its stated size: the 64 seeded cases (edges 32 to
    160, aspect up to 4, spheres / cylinders / fracture, 2-6 species, tau
    0.6-1.2, Monod parameters over a decade) set up and stepped together in
    FP32 transport with an FP64 flow; each of eight cases spread over the
    sizes ends bit-identical to its own one-at-a-time run, and every case
    runs its steps."""
    from complab_amd import sweep
    steps = 200
    ws = [configs.sweep_case(s, P.FP32_ADE, steps=steps, stabilize=50) for s in range(64)]
    vox = [w.params.nx * w.params.ny * w.params.nz for w in ws]
    grp, wall = sweep.run_group(ws)
    print("64 cases, %.1f M voxels, %.1f s" % (sum(vox) / 1e6, wall), flush=True)
    assert all(r.status == 0 and r.steps == steps for r in grp)
    pick = list(np.argsort(vox)[np.linspace(0, 63, 8).astype(int)])
    alone, _ = sweep.run_sequential([configs.sweep_case(int(s), P.FP32_ADE, steps=steps, stabilize=50) for s in pick])
    for s, a in zip(pick, alone):
        b = grp[s]
        assert a.steps == b.steps and a.flow_iters == b.flow_iters, a.name
        for fa, fb in zip(a.fields, b.fields):
            assert np.array_equal(fa, fb), a.name


def test_config5_serving_up_to_256_cubed():
    """The serving form at its stated sizes: 12 ABC-style requests with edges
    drawn in 48..256 (up to 255 and down to 63), arriving as a Poisson stream and
    admitted into the running batch between runs of 100 steps; each ends with
    the steps, flow iterations and fields of its one-at-a-time run, bit for
    bit, whatever ran beside it."""
    from complab_amd import sweep
    reqs = [configs.serving_case(100 + s, P.FP32_ADE, size=(48, 256), max_steps=400, stabilize=50) for s in range(12)]
    edges = [e for w in reqs for e in (w.params.nx - 2, w.params.ny, w.params.nz)]
    assert max(edges) > 250 and min(edges) < 70
    served, wall = sweep.serve_group(reqs, rate_per_s=4.0, steady_tol=1e-6, seed=5)
    alone, _ = sweep.run_sequential([configs.serving_case(100 + s, P.FP32_ADE, size=(48, 256), max_steps=400, stabilize=50)
                                     for s in range(12)], steady_tol=1e-6)
    print("12 requests, wall %.1f s, p50 %.2f s" % (wall, sweep.percentile([r.latency for r in served], 50)), flush=True)
    for a, b in zip(alone, served):
        assert a.status == b.status and a.status in (0, P.CLOGGED), a.name
        assert a.steps == b.steps and a.flow_iters == b.flow_iters, a.name
        for fa, fb in zip(a.fields, b.fields):
            assert np.array_equal(fa, fb), a.name

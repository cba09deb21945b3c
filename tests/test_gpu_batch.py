"""The batched form (clb_batch_*) on the GPU: every member of a batch ends
bit-identical to the same case run alone -- flow iterations, fields, mask,
counters, status -- whatever the mix (sizes and species counts, a CA member
that spills and changes its mask with flow re-solves, a precipitation member
with host steps, an equilibrium member on its own path), the transport
workgroup size, and members joining and leaving between calls."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver
from complab_amd import params as P

pytestmark = pytest.mark.gpu

STEPS = 120


def ca_case(prec):
    w = configs.config2(scale=0.12, precision=prec)
    w.transport.stabilize_steps = 30
    w.params.ns_update_interval = 5
    w.params.ns_max_iter = 3000
    w.params.b0[0] = 95.0
    w.params.mu_max[0] = 2e-2
    return w


def precip_case(prec):
    w = configs.serving_case(101, prec, size=(24, 32), max_steps=STEPS, stabilize=20)
    w.params.precip_interval = 25
    w.params.rxn_k[0] = 0.5                 # fast enough to seal voxels within the run
    return w


def eq_case(prec):
    w = configs.config3(scale=1 / 16, precision=prec)
    w.flow.eps = 1e-6
    w.params.ns_conv_eps = 1e-6
    return w


def mix(prec, with_eq=True):
    ws = [configs.sweep_case(s, prec, size=(16, 40), steps=STEPS, stabilize=25) for s in range(4)]
    ws += [ca_case(prec), precip_case(prec)]
    if with_eq:
        ws.append(eq_case(prec))
    return ws


def state(c, w):
    out = {"mask": c.read(P.FIELD_MASK), "counters": c.counters()}
    for s in range(w.params.n_species):
        out["c%d" % s] = c.read(P.FIELD_CONC, s)
    for m in range(w.params.n_microbes):
        out["b%d" % m] = c.read(P.FIELD_BIOMASS, m)
    if w.params.flow_on:
        for k in range(3):
            out["u%d" % k] = c.read(P.FIELD_VEL, k)
    return out


def alone(ctx, w, steps):
    c = clb.Case(ctx, w.params, w.geometry)
    rep = driver.setup(c, w.params, w.flow, w.transport)
    c.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    res = c.run_steps(steps)
    s = state(c, w)
    c.close()
    return rep.iters, res, s


def compare(name, a, b):
    ia, ra, sa = a
    ib, rb, sb = b
    assert ia == ib, (name, ia, ib)
    assert ra == rb, (name, ra, rb)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), (name, k)


@pytest.mark.parametrize("prec", [P.FP64, P.FP32, P.FP32_ADE])
def test_members_end_as_they_would_alone(ctx, prec):
    ws = mix(prec)
    print("batch of %d: %s" % (len(ws), [w.name for w in ws]), flush=True)
    ref = [alone(ctx, w, STEPS) for w in ws]
    cases = [clb.Case(ctx, w.params, w.geometry) for w in ws]
    reps = driver.setup_batch(ctx, cases, ws)
    for c, r in zip(cases, reps):
        c.set_omegas(r.omega_pore, r.omega_biofilm, 1.0)
    b = clb.Batch(ctx, cases)
    out = b.run_steps(STEPS)
    tile, us = b.tuning(P.BATCH_STEPS)
    print("tuned workgroup", tile, us, flush=True)
    assert tile in (64, 128, 256)
    got = [(r.iters, out[k], state(cases[k], ws[k])) for k, r in enumerate(reps)]
    b.close()
    for c in cases:
        c.close()
    for w, a, g in zip(ws, ref, got):
        compare(w.name, a, g)
    ca = got[4][2]["counters"]
    print("CA member counters", ca, "precipitation member counters", got[5][2]["counters"], flush=True)
    assert ca[6] > 0 and ca[5] > 0 and ca[4] > 0       # spill sweeps, mask changes, flow re-solves in the run


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_every_workgroup_size_gives_the_same_steps(ctx, tile):
    ws = mix(P.FP32_ADE, with_eq=False)
    ref = [alone(ctx, w, STEPS) for w in ws]
    cases = [clb.Case(ctx, w.params, w.geometry) for w in ws]
    reps = driver.setup_batch(ctx, cases, ws)
    for c, r in zip(cases, reps):
        c.set_omegas(r.omega_pore, r.omega_biofilm, 1.0)
    b = clb.Batch(ctx, cases)
    b.set_tile(P.BATCH_STEPS, tile)
    out = b.run_steps(STEPS)
    assert b.tuning(P.BATCH_STEPS)[0] == tile
    got = [(r.iters, out[k], state(cases[k], ws[k])) for k, r in enumerate(reps)]
    b.close()
    for c in cases:
        c.close()
    for w, a, g in zip(ws, ref, got):
        compare(w.name, a, g)


def test_members_join_and_leave_between_calls(ctx):
    ws = [configs.sweep_case(s, P.FP32, size=(16, 40), steps=STEPS, stabilize=25) for s in range(3)]
    ws.append(ca_case(P.FP32))
    ref = [alone(ctx, w, STEPS) for w in ws]
    cases = [clb.Case(ctx, w.params, w.geometry) for w in ws]
    reps = driver.setup_batch(ctx, cases, ws)
    for c, r in zip(cases, reps):
        c.set_omegas(r.omega_pore, r.omega_biofilm, 1.0)
    done = [0] * len(ws)

    def run(b, n):
        order = [cases.index(c) for c in b.cases]
        for k, (rc, d) in zip(order, b.run_steps(n)):
            assert rc == 0
            done[k] += d

    b = clb.Batch(ctx, cases[:2])                    # 0 and 1 start, 2 and 3 join later
    run(b, 40)
    b.add(cases[2])
    b.add(cases[3])
    b.remove(cases[0])
    assert cases[0].run_steps(STEPS - done[0]) == (0, STEPS - 40)   # a member that left goes on alone
    done[0] = STEPS
    assert clb.lib().clb_case_destroy(ctx.h, cases[2].id) < 0       # a member cannot be destroyed
    run(b, 50)
    b.remove(cases[1])
    assert cases[1].run_steps(STEPS - done[1]) == (0, STEPS - 90)
    done[1] = STEPS
    run(b, STEPS - 50)
    assert done == [STEPS] * len(ws)
    b.close()
    got = [(r.iters, (0, STEPS), state(c, w)) for r, c, w in zip(reps, cases, ws)]
    for c in cases:
        c.close()
    for w, a, g in zip(ws, ref, got):
        compare(w.name, a, g)

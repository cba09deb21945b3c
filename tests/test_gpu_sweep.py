"""The batched and serving forms: a case gives the same fields alone and in
the mix, FP32 transport with an FP64 flow stays within 1e-5 of the FP64 path,
and the serving loop admits every request and reports its latency."""
import numpy as np
import pytest

from complab_amd import configs, sweep
from complab_amd import params as P

pytestmark = pytest.mark.gpu


def small_cases(prec, n=4, steps=200):
    return [configs.sweep_case(s, prec, size=(32, 48), steps=steps, stabilize=50) for s in range(n)]


def test_same_result_alone_and_in_the_mix():
    alone, _ = sweep.run_sequential(small_cases(P.FP32_ADE))
    mix, _ = sweep.run_batched(small_cases(P.FP32_ADE), workers=4)
    for a, b in zip(alone, mix):
        assert a.status == 0 and b.status == 0
        assert a.steps == b.steps and a.flow_iters == b.flow_iters
        for fa, fb in zip(a.fields, b.fields):
            assert np.array_equal(fa, fb), a.name


def test_mixed_precision_within_1e5_of_fp64():
    r64, _ = sweep.run_batched(small_cases(P.FP64), workers=4)
    r32, _ = sweep.run_batched(small_cases(P.FP32_ADE), workers=4)
    for a, b in zip(r32, r64):
        e = sweep.field_error(a, b)
        print(a.name, "rel err", e)
        assert e <= 1e-5, a.name


def test_serving_admits_every_request():
    reqs = [configs.serving_case(100 + s, P.FP32_ADE, size=(40, 48), max_steps=400, stabilize=50) for s in range(4)]
    res, wall = sweep.serve(reqs, rate_per_s=8.0, workers=2, steady_tol=1e-6, seed=1)
    assert all(r.status == 0 for r in res)
    assert all(0 < r.latency <= wall + 1e-6 for r in res)
    assert all(0 < r.steps <= 400 for r in res)


def test_group_run_equals_one_case_at_a_time():
    """sweep.run_group (members stepped by the batched entry points) against
    each case's own run: bit-identical fields, steps and flow iterations."""
    ws = small_cases(P.FP32_ADE, n=5)
    alone, _ = sweep.run_sequential(ws)
    grp, _ = sweep.run_group(small_cases(P.FP32_ADE, n=5))
    for a, b in zip(alone, grp):
        assert a.status == 0 and b.status == 0
        assert a.steps == b.steps and a.flow_iters == b.flow_iters, a.name
        for fa, fb in zip(a.fields, b.fields):
            assert np.array_equal(fa, fb), a.name


def test_continuous_batching_equals_one_request_at_a_time():
    """sweep.serve_group: requests arriving while others run join the batch
    and leave it at steady state or their step limit; each ends exactly as its
    one-at-a-time run (same steps, bit-identical fields), whatever it ran
    beside."""
    def reqs():
        return [configs.serving_case(100 + s, P.FP32_ADE, size=(40, 56), max_steps=700, stabilize=50)
                for s in range(5)]
    alone, _ = sweep.run_sequential(reqs(), steady_tol=1e-4)
    served, wall = sweep.serve_group(reqs(), rate_per_s=6.0, steady_tol=1e-4, seed=3)
    for a, b in zip(alone, served):
        assert a.status == b.status and a.status in (0, P.CLOGGED), a.name
        assert a.steps == b.steps and a.flow_iters == b.flow_iters, a.name
        assert 0 < b.latency <= wall + 1e-6
        for fa, fb in zip(a.fields, b.fields):
            assert np.array_equal(fa, fb), a.name

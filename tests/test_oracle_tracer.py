"""The flow solve's convergence rule in the oracle (ValueTracer: a window of
the last 500 mean energies, converged when the population standard deviation
over the window divided by |mean| is below epsilon, complab.cpp:1097) checked
against that statement computed directly with numpy, on sequences where a
window one sample shorter or longer, or the sample (n - 1) deviation, stops at
another iteration.  The device rule is held to the oracle's by the flow tests'
equal iteration counts."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle as O  # noqa: E402


def oracle_stop(vals, eps):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    out = C.c_int(0)
    assert O.lib().orc_tracer_stop(v.ctypes.data, len(v), eps, C.byref(out)) == 0
    return out.value


def direct_stop(vals, eps, window=500, ddof=0):
    for k in range(window, len(vals) + 1):
        w = np.asarray(vals[k - window:k])
        r = np.std(w, ddof=ddof) / abs(np.mean(w))
        if np.isnan(r) or r < eps:
            return k
    return 0


def decaying(n, tau, amp=0.5, seed=0, noise=0.0):
    t = np.arange(n, dtype=np.float64)
    rng = np.random.default_rng(seed)
    return 1.0 + amp * np.exp(-t / tau) + noise * rng.standard_normal(n)


@pytest.mark.parametrize("tau,eps", [(800.0, 1e-3), (2000.0, 1e-4), (300.0, 1e-6)])
def test_stop_equals_the_stated_rule(tau, eps):
    vals = decaying(20000, tau)
    k = oracle_stop(vals, eps)
    assert k > 0
    assert k == direct_stop(vals, eps)


def test_the_rule_is_discriminated_from_near_misses():
    # a slowly settling sequence: the stop moves with the window length and
    # with the deviation's divisor
    vals = decaying(40000, 3000.0)
    eps = 1e-4
    k = oracle_stop(vals, eps)
    assert k == direct_stop(vals, eps)
    assert k != direct_stop(vals, eps, window=499)
    assert k != direct_stop(vals, eps, window=501)
    assert k != direct_stop(vals, eps, ddof=1)


def test_no_decision_before_a_full_window_and_on_divergence():
    flat = np.full(499, 2.0)
    assert oracle_stop(flat, 1e-12) == 0                 # 499 samples: never
    assert oracle_stop(np.full(500, 2.0), 1e-12) == 500  # a constant full window: converged at once
    noisy = decaying(600, 50.0, noise=1e-3, seed=3)
    assert oracle_stop(noisy, 1e-9) == 0 == direct_stop(noisy, 1e-9)
    bad = decaying(600, 50.0, noise=1e-3, seed=3)
    bad[550] = np.nan
    assert oracle_stop(bad, 1e-9) == 551                 # a NaN in the window stops the solve (diverged)

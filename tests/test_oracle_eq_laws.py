"""The oracle's speciation pinned by the laws its answer must satisfy, for
both solvers (Anderson-accelerated PCF, part4_eqsolver.hh:960-1180, and the
Newton solve) and three tableaux: the carbonate system of
examples/04_equilibrium (2 components), the carbonate-calcium system of the
precision form (3 components, log K drawn within +-2), and a 6-component
carbonate-calcium tableau (H, CO3, Ca, Mg, Na, Cl; HCO3 and CaCO3(aq)).

A converged answer conserves every component total of its input (mass
balance) and holds every complex at 10^logK times the product of its
components' concentrations (mass action).  A dropped term, a wrong sign or a
transposed stoichiometry fails one of the two."""
import numpy as np
import pytest

import oracle as O
from complab_amd import configs
from complab_amd import params as P

import sys, os
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import carbonate, calcium_carbonate       # noqa: E402  (the tableaux of the fixtures)


def tableau(name, solver, seed=0):
    if name == "carbonate2":
        p = P.default_params(3, 1, 1, 4)
        carbonate(p)
    elif name == "precision3":
        p = configs.config3(scale=1 / 32, seed=3 + seed).params
    else:
        p = P.default_params(3, 1, 1, 8)
        calcium_carbonate(p, np.random.default_rng(seed))
    p.eq_solver = solver
    p.eq_tol = 1e-12 if solver == P.EQ_NEWTON else 1e-10
    p.eq_max_iter = 400
    p.eq_anderson_depth = 4
    return p


def laws(p, cin, out):
    ns, nc = p.n_species, p.eq_nc
    S = np.array([[p.eq_stoich[i][j] for j in range(nc)] for i in range(ns)])
    eq = np.array([np.any(np.abs(S[i]) > 1e-10) for i in range(ns)])
    logk = np.array([p.eq_logk[i] for i in range(ns)])
    comp = [p.eq_comp_species[j] for j in range(nc)]
    T_in = S[eq].T @ cin[eq]
    T_out = S[eq].T @ out[eq]
    scale = np.abs(S[eq]).T @ cin[eq]
    mb = np.max(np.abs(T_out - T_in) / scale)
    lx = np.log10(out[comp])
    ma = 0.0
    for i in range(ns):
        if eq[i]:
            ma = max(ma, abs(np.log10(out[i]) - logk[i] - S[i] @ lx))
    return mb, ma


@pytest.mark.parametrize("name", ["carbonate2", "precision3", "calcium6"])
@pytest.mark.parametrize("solver", [P.EQ_NEWTON, P.EQ_AAPCF])
def test_speciation_meets_mass_balance_and_mass_action(name, solver):
    rng = np.random.default_rng(11)
    checked = 0
    worst_mb = worst_ma = 0.0
    for k in range(40):
        p = tableau(name, solver, seed=k % 4)
        cin = 10 ** rng.uniform(-6, -2, size=p.n_species)
        S = np.array([[p.eq_stoich[i][j] for j in range(p.eq_nc)] for i in range(p.n_species)])
        if np.any(S.T @ cin <= 0):
            continue                                    # a non-positive total: no positive answer exists
        out, it, res, cv = O.eq_solve(p, cin)
        if not cv:
            continue
        assert np.all(out > 0)
        if np.any(out >= 10.0) or np.any(out <= 1e-30):
            continue                                    # a clamp of the solver bound: laws not expected
        mb, ma = laws(p, cin, out)
        worst_mb, worst_ma = max(worst_mb, mb), max(worst_ma, ma)
        checked += 1
    print(name, solver, "checked", checked, "mass balance", worst_mb, "mass action", worst_ma)
    assert checked >= 20
    assert worst_ma <= 1e-14                            # complexes from their components, to rounding
    assert worst_mb <= (1e-11 if solver == P.EQ_NEWTON else 1e-9)   # tolerance 1e-12 / 1e-10


def test_laws_catch_a_wrong_answer():
    """The checks fail a plausible mistake: the carbonate answer with one
    complex's stoichiometry transposed, or its log K sign flipped."""
    p = tableau("carbonate2", P.EQ_NEWTON)
    cin = np.array([1e-3, 1e-5, 2e-4, 3e-4])
    out, _, _, cv = O.eq_solve(p, cin)
    assert cv
    mb, ma = laws(p, cin, out)
    assert mb <= 1e-10 and ma <= 1e-12
    q = tableau("carbonate2", P.EQ_NEWTON)
    q.eq_stoich[2][0], q.eq_stoich[2][1] = q.eq_stoich[2][1], q.eq_stoich[2][0]
    assert max(laws(q, cin, out)) > 1e-3
    q = tableau("carbonate2", P.EQ_NEWTON)
    q.eq_logk[3] = -q.eq_logk[3]
    assert max(laws(q, cin, out)) > 1e-3

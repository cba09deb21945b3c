"""A biotic CA case whose microbe takes its rates from a .sym law (ranges,
per-hour units, chained rates, functions, the shared budget, a gate on the
organism) on the device against the oracle: FP64 fields 1e-10, masks exact."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver
from complab_amd import params as P
from complab_amd import symbolic as S
import oracle as O

pytestmark = pytest.mark.gpu

LAW = """# a Monod-like law with a product term and an inhibition
units per_hour
vars S0 S1 S2 Bug
rate growth = 72 * S0 / (0.1 + S0) * max(S1, 0) / (0.1 + S1) * exp(-S2 / 2) + 0 * tanh(S2)
rate S0 = -1.8 * growth * Bug
rate S1 = -0.9 * growth * Bug
rate S2 = 0.9 * growth * Bug + 1e-3 * sqrt(S0)
range S0 0 0.45
"""


def test_symbolic_law_matches_oracle(ctx, tmp_path):
    print("case: biotic 0.12-scale pack, symbolic law with the gate on microbe 0", flush=True)
    w = configs.config2(scale=0.12)
    w.transport.stabilize_steps = 30
    w.params.ns_max_iter = 3000
    p = w.params
    p.thm_on = 1
    p.thm_microbe = 0
    p.thm_rt = 8.314462618e-3 * 298.15
    p.thm_threshold = 12.5
    for s, v in enumerate([-1.0, -1.0, 1.0, 1.0]):
        p.thm_nu[s] = v
    p.thm_dg0 = -12.5 - 2.0 * p.thm_rt - p.thm_rt * (2 * np.log(p.thm_cmin) - 2 * np.log(0.5))
    f = tmp_path / "law.sym"
    f.write_text(LAW)
    prog = S.compile_law(S.read_sym(str(f)), ["S%d" % s for s in range(p.n_species)], "Bug")
    g = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    g.set_symbolic(0, prog)
    o.set_symbolic(0, prog)
    rg = driver.setup(g, w.params, w.flow, w.transport)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    assert rg.iters == ro.iters
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    assert g.run_steps(60) == o.run_steps(60) == (0, 60)
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    om = o.read(P.FIELD_MASK) >= 2
    b = o.read(P.FIELD_BIOMASS, 0)
    assert np.max(np.abs(g.read(P.FIELD_BIOMASS, 0) - b)) <= 1e-10 * np.max(np.abs(b))
    for s in range(4):
        a, c = g.read(P.FIELD_CONC, s)[om], o.read(P.FIELD_CONC, s)[om]
        assert np.max(np.abs(a - c)) <= 1e-10 * max(np.max(np.abs(c)), 1e-300), s
    cg, co = g.counters(), o.counters()
    assert cg[5] == co[5] and cg[6] == co[6]


def test_bad_programs_are_refused(ctx, tmp_path):
    print("case: a program that would underflow its stack is refused before any run", flush=True)
    w = configs.config2(scale=0.1)
    w.transport.stabilize_steps = 0
    w.flow.max_iter = 200
    p = w.params
    f = tmp_path / "law.sym"
    f.write_text("vars S0 Bug\nrate growth = 1e-3 * S0\nrate S0 = -growth * Bug\n")
    prog = S.compile_law(S.read_sym(str(f)), ["S%d" % s for s in range(p.n_species)], "Bug")
    g = clb.Case(ctx, p, w.geometry)
    bad = P.SymProgram()
    bad.n_rates, bad.n_ops, bad.rate_end[0], bad.op[0], bad.unit_scale = 1, 1, 1, P.SYM_ADD, 1.0   # pops two from nothing
    with pytest.raises(clb.ClbError):
        g.set_symbolic(0, bad)
    g.set_symbolic(0, prog)
    g.set_symbolic(0, None)
    g.close()

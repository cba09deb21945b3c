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
    # the energy snapshot of the gated organism at the standing concentrations
    dg, do = g.read(P.FIELD_DG, 0), o.read(P.FIELD_DG, 0)
    assert np.max(np.abs(dg - do)) <= 1e-10 * np.max(np.abs(do)) and np.max(np.abs(do)) > 0
    fg, fo = g.read(P.FIELD_FT, 0), o.read(P.FIELD_FT, 0)
    assert np.max(np.abs(fg - fo)) <= 1e-10 and 0.0 <= fo.min() and fo.max() <= 1.0


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


ABIOTIC = """# A -> B through one extent line, fast enough that the budget holds A in places
units per_hour
vars A B
reaction A -1 B 1
rate extent = 1080 * A + 7200 * A^3 + 36 * sqrt(A) * exp(-B)
range A 0 0.8
"""


def _pair(ctx, w, prog):
    g = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    for c in (g, o):
        if prog is not None:
            c.set_symbolic(P.SYM_ABIOTIC, prog)
    return g, o


def _close(a, c, what):
    assert np.max(np.abs(a - c)) <= 1e-10 * max(np.max(np.abs(c)), 1e-300), what


def test_abiotic_symbolic_law_matches_oracle(ctx, tmp_path):
    print("case: A->B pack at 0.3 scale, abiotic .sym law only, then a checkpoint round trip", flush=True)
    w = configs.config1(scale=0.3)
    w.transport.stabilize_steps = 10
    w.params.n_rxn = 0
    f = tmp_path / "ab.sym"
    f.write_text(ABIOTIC)
    prog = S.compile_law(S.read_sym(str(f)), ["A", "B"], None)
    g, o = _pair(ctx, w, prog)
    rg = driver.setup(g, w.params, w.flow, w.transport)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    assert rg.iters == ro.iters
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    assert g.run_steps(40) == o.run_steps(40) == (0, 40)
    om = o.read(P.FIELD_MASK) >= 2                              # open voxels (closed ones hold no populations here)
    for s in range(2):
        _close(g.read(P.FIELD_CONC, s)[om], o.read(P.FIELD_CONC, s)[om], s)
    assert o.read(P.FIELD_CONC, 1)[om].max() > 0.1             # the law ran
    blob = g.checkpoint()
    g.run_steps(20)
    want = [g.read(P.FIELD_CONC, s) for s in range(2)]
    g.restore(blob)
    g.run_steps(20)
    for s in range(2):
        assert np.array_equal(g.read(P.FIELD_CONC, s), want[s]), s
    g.close()


def test_abiotic_symbolic_beside_compiled_biotic_law(ctx, tmp_path):
    print("case: biotic 0.12-scale pack, compiled dual-Monod plus an abiotic .sym law", flush=True)
    w = configs.config2(scale=0.12)
    w.transport.stabilize_steps = 30
    w.params.ns_max_iter = 3000
    names = ["S%d" % s for s in range(w.params.n_species)]
    f = tmp_path / "ab.sym"
    f.write_text("vars S0 S2\nreaction S2 -1 S3 1\nrate extent = 0.3 * S2 * S0 / (1e-4 + S0)\n")
    prog = S.compile_law(S.read_sym(str(f)), names, None)
    g, o = _pair(ctx, w, prog)
    rg = driver.setup(g, w.params, w.flow, w.transport)
    ro = driver.setup(o, w.params, w.flow, w.transport)
    assert rg.iters == ro.iters
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
    assert g.run_steps(60) == o.run_steps(60) == (0, 60)
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    om = o.read(P.FIELD_MASK) >= 2
    _close(g.read(P.FIELD_BIOMASS, 0), o.read(P.FIELD_BIOMASS, 0), "biomass")
    for s in range(4):
        _close(g.read(P.FIELD_CONC, s)[om], o.read(P.FIELD_CONC, s)[om], s)
    assert o.read(P.FIELD_CONC, 3)[om].max() > 0
    g.close()


def test_compiled_and_symbolic_organisms_match_oracle(ctx, tmp_path):
    print("case: two organisms at 0.1 scale, microbe 0 compiled, microbe 1 symbolic (gate on microbe 1)", flush=True)
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
    p.thm_on = 1
    p.thm_microbe = 1
    p.thm_threshold = 12.5
    for s, v in enumerate([-1.0, -1.0, 1.0, 1.0]):
        p.thm_nu[s] = v
    p.thm_dg0 = -12.5 - 2.0 * p.thm_rt - p.thm_rt * (2 * np.log(p.thm_cmin) - 2 * np.log(0.5))
    f = tmp_path / "law.sym"
    f.write_text(LAW)
    prog = S.compile_law(S.read_sym(str(f)), ["S%d" % s for s in range(p.n_species)], "Bug")
    g = clb.Case(ctx, p, geo)
    o = O.OracleCase(p, geo)
    for c in (g, o):
        c.set_symbolic(1, prog)
        c.set_omegas([1.2] * p.n_species, [1.1] * p.n_species, 1.0)
    assert g.run_steps(60) == o.run_steps(60) == (0, 60)
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    om = o.read(P.FIELD_MASK) >= 2
    for m in range(2):
        _close(g.read(P.FIELD_BIOMASS, m), o.read(P.FIELD_BIOMASS, m), "biomass %d" % m)
    for s in range(4):
        _close(g.read(P.FIELD_CONC, s)[om], o.read(P.FIELD_CONC, s)[om], s)
    cg, co = g.counters(), o.counters()
    assert cg[5] == co[5] and cg[6] == co[6]
    g.close()

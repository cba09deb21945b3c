"""Symbolic (.sym) rate laws: the expression language against values the
mathematics fixes (through the oracle's evaluator of the compiled postfix
code), the reader's refusals, derived lines equal to written ones, and a
.sym restatement of the dual-Monod law running the oracle's biotic case like
the compiled law does."""
import math
import os

import numpy as np
import pytest

from complab_amd import configs, driver
from complab_amd import params as P
from complab_amd import symbolic as S
import oracle as O

REF = "/root/reference"


def law_of(tmp_path, text, species=("A", "B"), microbe="Bug"):
    f = tmp_path / "law.sym"
    f.write_text(text)
    return S.compile_law(S.read_sym(str(f)), list(species), microbe)


def value(tmp_path, expr, A=2.0, B=3.0, bug=5.0):
    prog = law_of(tmp_path, "vars A B Bug\nrate growth = %s\n" % expr)
    return O.sym_evaluate(prog, [A, B], bug)[0]


@pytest.mark.parametrize("expr,want", [
    ("1 + 2 * 3", 7.0), ("(1 + 2) * 3", 9.0), ("8 / 2 / 2", 2.0), ("2 - 3 - 4", -5.0),
    ("-2^2", -4.0), ("2^-1", 0.5), ("2^3^2", 512.0), ("-A", -2.0), ("+A", 2.0), ("--A", 2.0),
    ("A * B + Bug", 11.0), ("A^B", 8.0), ("pow(A, B)", 8.0), ("pow(-2, 3)", -8.0),
    ("pow(-8, 1/3)", 0.0), ("(-8)^0.5", 0.0),                 # negative base, fractional power: 0
    ("1 / 0", 0.0), ("A / (B - 3)", 0.0),                      # division by |q| < 1e-300: 0
    ("log(-1)", 0.0), ("log(0)", 0.0), ("log10(1000)", 3.0), ("sqrt(-4)", 0.0), ("sqrt(16)", 4.0),
    ("exp(0)", 1.0), ("abs(-2.5)", 2.5), ("min(A, B)", 2.0), ("max(A, B)", 3.0), ("tanh(0)", 0.0),
    ("exp(1000)", 1e300), ("-exp(1000)", -1e300),              # held within +-1e300
    ("1.5e-3 * 2", 3e-3), (".5 + 5.", 5.5), ("2.e1", 20.0),
])
def test_expression_values(tmp_path, expr, want):
    assert value(tmp_path, expr) == pytest.approx(want, rel=1e-15, abs=0)


def test_constants_and_transcendentals(tmp_path):
    assert value(tmp_path, "pi") == math.pi and value(tmp_path, "e") == math.e
    assert value(tmp_path, "log(e)") == pytest.approx(1.0, rel=1e-15)
    assert value(tmp_path, "tanh(1)") == pytest.approx(math.tanh(1), rel=1e-15)


def test_chained_rates_units_and_ranges(tmp_path):
    prog = law_of(tmp_path, "units per_hour\nvars A B Bug\nrate growth = 0.36 * A\n"
                            "rate A = -2 * growth * Bug\nrate B = growth\nrange A 0.1 1.0\n")
    out = O.sym_evaluate(prog, [2.0, 0.0], 5.0)                # A held at 1.0
    growth = 0.36 * 1.0
    assert out[0] == pytest.approx(growth / 3600, rel=1e-15)
    assert out[1] == pytest.approx(-2 * growth * 5.0 / 3600, rel=1e-15)   # chained value unscaled
    assert out[2] == pytest.approx(growth / 3600, rel=1e-15)
    out = O.sym_evaluate(prog, [0.01, 0.0], 5.0)                # held at 0.1 from below
    assert out[0] == pytest.approx(0.036 / 3600, rel=1e-15)
    # a name that is both a variable and an earlier rate reads the variable
    prog = law_of(tmp_path, "vars A B Bug\nrate A = 7\nrate growth = A\n")
    assert O.sym_evaluate(prog, [2.0, 0.0], 1.0)[1] == 2.0


def test_reader_refusals(tmp_path):
    bad = {"unknown function": "vars A\nrate growth = foo(A)\n",
           "forward reference": "vars A\nrate growth = B2\nrate B2 = A\n",
           "missing paren": "vars A\nrate growth = (A + 1\n",
           "trailing text": "vars A\nrate growth = A A\n",
           "duplicate rate": "vars A\nrate growth = A\nrate growth = 2\n",
           "unknown keyword": "vars A\nratee growth = A\n",
           "bad units": "units per_day\nvars A\nrate growth = A\n",
           "rate and reaction": "vars A B Bug\nreaction A -1\nyield A 0.5\nbiomass Bug\nrate growth = A\nrate A = 1\n"}
    for what, text in bad.items():
        f = tmp_path / "bad.sym"
        f.write_text(text)
        with pytest.raises(S.SymError):
            S.read_sym(str(f))
    f = tmp_path / "ok.sym"
    f.write_text("vars A X\nrate growth = A\n")
    with pytest.raises(S.SymError):                                   # X binds nothing in this run
        S.compile_law(S.read_sym(str(f)), ["A", "B"], "Bug")


def test_derived_lines_equal_written_ones(tmp_path):
    written = law_of(tmp_path, "units per_hour\nvars A B Bug\nrate growth = 0.35 * A / (5e-4 + A) * B / (1e-4 + B)\n"
                               "rate A = -2.5 * growth * Bug\nrate B = -5.0 * growth * Bug\n")
    derived = law_of(tmp_path, "units per_hour\nvars A B Bug\nreaction A -1 B -2\nyield A 0.4\nbiomass Bug\n"
                               "rate growth = 0.35 * A / (5e-4 + A) * B / (1e-4 + B)\n")
    for c in ([1e-3, 1e-3], [2e-4, 5e-5], [0.0, 1e-3]):
        assert np.array_equal(O.sym_evaluate(written, c, 2.0), O.sym_evaluate(derived, c, 2.0))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_reference_example_17_laws():
    d = os.path.join(REF, "examples", "17_symbolic_law", "input")
    a = S.compile_law(S.read_sym(os.path.join(d, "growth.sym")), ["acetate", "o2"], "Bug")
    b = S.compile_law(S.read_sym(os.path.join(d, "growth_stoich.sym")), ["acetate", "o2"], "Bug")
    for c in ([1e-3, 1e-3], [3e-3, 1.5e-3], [1e-7, 1e-7], [9e-3, 9e-3]):      # inside, at and past the ranges
        oa, ob = O.sym_evaluate(a, c, 2.0), O.sym_evaluate(b, c, 2.0)
        assert np.array_equal(oa, ob)
        ac, o2 = min(max(c[0], 1e-6), 5e-3), min(max(c[1], 1e-6), 2e-3)
        g = 0.35 * ac / (5.0e-4 + ac) * o2 / (1.0e-4 + o2)
        assert oa == pytest.approx([g / 3600, -2.5 * g * 2.0 / 3600, -5.0 * g * 2.0 / 3600], rel=1e-14)


def monod_sym(p):
    """The dual-Monod law of microbe 0 written as a .sym file."""
    names = ["S%d" % s for s in range(p.n_species)]
    lines = ["units per_second", "vars %s Bug" % " ".join(names),
             "rate growth = %.17g * S%d / (%.17g + S%d) * S%d / (%.17g + S%d)"
             % (p.mu_max[0], p.donor, p.ks[0], p.donor, p.acceptor, p.ke[0], p.acceptor)]
    for s in range(p.n_species):
        if p.bio_stoich[0][s] != 0:
            lines.append("rate S%d = %.17g * growth * Bug" % (s, p.bio_stoich[0][s]))
    return "\n".join(lines) + "\n", names


def test_symbolic_monod_runs_like_the_compiled_law(tmp_path):
    w = configs.config2(scale=0.1)
    w.transport.stabilize_steps = 20
    w.params.ns_max_iter = 1500
    w.params.mu_max[0] = 2e-2
    w.params.kdecay[0] = 0.0
    text, names = monod_sym(w.params)
    prog = law_of(tmp_path, text, species=names)
    runs = []
    for sym in (False, True):
        o = O.OracleCase(w.params, w.geometry)
        if sym:
            o.set_symbolic(0, prog)
        rep = driver.setup(o, w.params, w.flow, w.transport)
        o.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        b0 = o.read(P.FIELD_BIOMASS, 0).sum()
        assert o.run_steps(30) == (0, 30)
        assert o.read(P.FIELD_BIOMASS, 0).sum() > b0                         # the biofilm grew
        runs.append(o)
    a, b = runs
    assert np.array_equal(a.read(P.FIELD_MASK), b.read(P.FIELD_MASK))
    ba, bb = a.read(P.FIELD_BIOMASS, 0), b.read(P.FIELD_BIOMASS, 0)
    assert np.max(np.abs(ba - bb)) <= 1e-9 * np.max(np.abs(ba))
    for s in range(w.params.n_species):
        ca, cb = a.read(P.FIELD_CONC, s), b.read(P.FIELD_CONC, s)
        assert np.max(np.abs(ca - cb)) <= 1e-9 * max(np.max(np.abs(ca)), 1e-300), s


# ---- the abiotic form: one extent line (tests/test_sym_stoich.cpp:138-216)

def test_abiotic_extent_matches_hand_written(tmp_path):
    hand = tmp_path / "ahand.sym"
    hand.write_text("units per_second\nvars Fe HS\nrate Fe = -1.6e2 * Fe * HS\nrate HS = -1.6e2 * Fe * HS\n")
    decl = tmp_path / "adecl.sym"
    decl.write_text("units per_second\nvars Fe HS FeS\nreaction Fe -1  HS -1  FeS 1\nrate extent = 1.6e2 * Fe * HS\n")
    lh, ld = S.read_sym(str(hand)), S.read_sym(str(decl))
    assert ld.derived == 3 and S.is_abiotic(ld) and not S.is_abiotic(lh)
    sp = ["Fe", "HS", "FeS"]
    ah, ad = S.compile_law(lh, sp, None), S.compile_law(ld, sp, None)
    assert list(ad.rate_species[:4]) == [-1, 0, 1, 2]
    for i in range(1, 41):
        c = [1e-3 * i / 40.0, 1e-3 * (41 - i) / 40.0, 0.0]
        rh, rd = O.sym_evaluate(ah, c, 0.0), O.sym_evaluate(ad, c, 0.0)
        assert np.array_equal(rh, rd[1:3])                     # derived Fe and HS equal the written pair
    rd = O.sym_evaluate(ad, [4e-4, 7e-4, 0.0], 0.0)
    assert rd[1] == rd[2] and rd[3] == -rd[1] and rd[3] > 0 > rd[1]
    assert rd[0] == pytest.approx(1.6e2 * 4e-4 * 7e-4, rel=1e-15)


def test_abiotic_refusals_and_binding(tmp_path):
    bad = {"extent with biomass": "vars Fe HS Bug\nreaction Fe -1 HS -1\nbiomass Bug\nrate extent = Fe\n",
           "extent with yield": "vars Fe HS\nreaction Fe -1 HS -1\nyield Fe 0.4\nrate extent = Fe\n",
           "growth and extent": "vars Fe HS Bug\nreaction Fe -1 HS -1\nyield Fe 0.4\nbiomass Bug\n"
                                "rate growth = Fe\nrate extent = Fe\n",
           "extent, no reaction": "vars Fe HS\nrate extent = Fe * HS\n",
           "reaction, no rate of it": "vars Fe HS\nreaction Fe -1 HS -1\nrate Fe = -1 * Fe\n",
           "rate line and reaction place": "vars Fe HS\nreaction Fe -1 HS -1\nrate extent = Fe\nrate Fe = -1\n",
           "yield without reaction": "vars A Bug\nyield A 0.4\nrate growth = A\n",
           "no vars": "rate growth = 1\n"}
    for what, text in bad.items():
        f = tmp_path / "bad.sym"
        f.write_text(text)
        with pytest.raises(S.SymError):
            S.read_sym(str(f))
    f = tmp_path / "ab.sym"
    f.write_text("vars Fe HS FeS\nreaction Fe -1  HS -1  FeS 1\nrate extent = 1.6e2 * Fe * HS\n")
    ab = S.read_sym(str(f))
    with pytest.raises(S.SymError):                            # an abiotic law handed to an organism
        S.compile_law(ab, ["Fe", "HS", "FeS"], "Bug")
    S.compile_law(ab, ["Fe", "HS", "FeS"], None)
    f.write_text("vars acetate o2 Bug\nreaction acetate -1 o2 -2\nyield acetate 0.4\nbiomass Bug\n"
                 "rate growth = 0.35 * acetate\n")
    bio = S.read_sym(str(f))
    with pytest.raises(S.SymError):                            # a biotic law through <abiotic_file>
        S.compile_law(bio, ["acetate", "o2"], None)
    S.compile_law(bio, ["acetate", "o2"], "Bug")
    f.write_text("vars Fe X\nrate Fe = -Fe\n")
    with pytest.raises(S.SymError):                            # X binds no substrate
        S.compile_law(S.read_sym(str(f)), ["Fe", "HS"], None)


def test_abiotic_symbolic_runs_like_mass_action(tmp_path):
    """A -> B at k = 1e-3 written as a .sym extent law: the oracle's run equals
    the mass-action run bit for bit (the same products, the same order)."""
    w = configs.config1(scale=0.15)
    w.transport.stabilize_steps = 10
    w.params.ns_max_iter = 2000
    text = "units per_second\nvars A B\nreaction A -1 B 1\nrate extent = 1e-3 * A\n"
    prog = law_of(tmp_path, text, species=("A", "B"), microbe=None)
    runs = []
    for sym in (False, True):
        p = P.copy_params(w.params)
        if sym:
            p.n_rxn = 0
        o = O.OracleCase(p, w.geometry)
        if sym:
            o.set_symbolic(P.SYM_ABIOTIC, prog)
        rep = driver.setup(o, p, w.flow, w.transport)
        o.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        o.run_steps(40)
        runs.append(o)
    a, b = runs
    for s in range(2):
        assert np.array_equal(a.read(P.FIELD_CONC, s), b.read(P.FIELD_CONC, s)), s
    assert a.read(P.FIELD_CONC, 1).max() > 0                   # the product appeared


def two_microbe_case():
    """config2 at 0.1 scale with a second organism seeded beside the first,
    both on the dual-Monod law with their own mu_max."""
    w = configs.config2(scale=0.1)
    p = w.params
    geo, _ = configs.seed_biofilm(w.geometry, 0.02, 11, code=P.BIOFILM + 1)
    p.n_microbes = 2
    p.mu_max[0] = 2e-2; p.mu_max[1] = 1.5e-2
    p.ks[1] = p.ks[0]; p.ke[1] = p.ke[0]; p.kdecay[0] = 0.0; p.kdecay[1] = 0.0
    for s in range(p.n_species):
        p.bio_stoich[1][s] = p.bio_stoich[0][s]
    p.b0[0] = 95.0; p.b0[1] = 90.0
    p.flow_on = 0
    return w, geo


def monod_sym_of(p, m):
    names = ["S%d" % s for s in range(p.n_species)]
    lines = ["vars %s Bug" % " ".join(names),
             "rate growth = %.17g * S%d / (%.17g + S%d) * S%d / (%.17g + S%d)"
             % (p.mu_max[m], p.donor, p.ks[m], p.donor, p.acceptor, p.ke[m], p.acceptor)]
    for s in range(p.n_species):
        if p.bio_stoich[m][s] != 0:
            lines.append("rate S%d = %.17g * growth * Bug" % (s, p.bio_stoich[m][s]))
    return "\n".join(lines) + "\n", names


def test_compiled_and_symbolic_organisms_in_one_run(tmp_path):
    """Microbe 0 on the compiled law, microbe 1 on a .sym restatement of its
    own compiled law: the run equals the all-compiled run within 1e-9, and
    each organism grows by its own law (a symbolic law left out, or applied
    to the wrong organism, moves microbe 1's biomass)."""
    w, geo = two_microbe_case()
    text, names = monod_sym_of(w.params, 1)
    prog = law_of(tmp_path, text, species=names)
    runs = []
    for mixed in (False, True):
        o = O.OracleCase(w.params, geo)
        if mixed:
            o.set_symbolic(1, prog)
        o.set_omegas([1.2] * w.params.n_species, [1.1] * w.params.n_species, 1.0)
        b0 = [o.read(P.FIELD_BIOMASS, m).sum() for m in range(2)]
        assert o.run_steps(40) == (0, 40)
        assert all(o.read(P.FIELD_BIOMASS, m).sum() > b0[m] for m in range(2))
        runs.append(o)
    a, b = runs
    assert np.array_equal(a.read(P.FIELD_MASK), b.read(P.FIELD_MASK))
    for m in range(2):
        ba, bb = a.read(P.FIELD_BIOMASS, m), b.read(P.FIELD_BIOMASS, m)
        assert np.max(np.abs(ba - bb)) <= 1e-9 * np.max(np.abs(ba)), m
    for s in range(w.params.n_species):
        ca, cb = a.read(P.FIELD_CONC, s), b.read(P.FIELD_CONC, s)
        assert np.max(np.abs(ca - cb)) <= 1e-9 * max(np.max(np.abs(ca)), 1e-300), s
    # the same law given to microbe 0 instead runs microbe 0 at microbe 1's rate
    o = O.OracleCase(w.params, geo)
    o.set_symbolic(0, prog)
    o.set_omegas([1.2] * w.params.n_species, [1.1] * w.params.n_species, 1.0)
    o.run_steps(40)
    assert not np.allclose(o.read(P.FIELD_BIOMASS, 0), a.read(P.FIELD_BIOMASS, 0), rtol=1e-6, atol=0)

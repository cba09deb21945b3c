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
           "abiotic extent": "vars A B\nreaction A -1 B 1\nrate extent = A\n",
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

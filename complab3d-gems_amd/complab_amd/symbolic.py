"""Biotic rate laws read from a .sym file (complab3d_symbolic.hh), compiled to
the postfix code of clb_sym_program that the kinetics kernel runs.

The file: `#` comments, then lines

    version 1
    units   per_second | per_hour           (per_hour divides every rate by 3600)
    vars    acetate o2 Bug                  (substrates and the organism's own biomass)
    rate    growth  = 0.35 * acetate / (5e-4 + acetate) * o2 / (1e-4 + o2)
    rate    acetate = -2.5 * growth * Bug
    range   acetate 1e-6 5e-3               (the variable is held in [lo, hi])
    reaction acetate -1  o2 -2              (with yield and biomass: the substrate
    yield    acetate 0.4                     lines are derived, coefficient
    biomass  Bug                             (nu_i / |nu_key|) / Y)

An abiotic law (complab3d_symbolic.hh:84-115, 653-684) has no organism: one
`rate extent = ...` line and a reaction line, from which the species lines
`nu_i * extent` are derived.  A law is biotic with a "growth" line and
abiotic with an "extent" line, never both.

Expressions (the reference's grammar, complab3d_symbolic.hh:185-305): + and -
left-associative, then * and /, then unary - and +, then ^ (right-associative,
its exponent a unary expression: -2^2 = -4, 2^-1 = 0.5); numbers, the
declared variables and the rates defined earlier (a name matches the
variables first), pi, e, exp log log10 sqrt abs tanh (one argument), min max
pow (two).  A "growth" rate is the specific growth rate, an "extent" rate the
rate of the abiotic reaction; a substrate rate is that substrate's rate of
change.
"""
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import params as P


class SymError(ValueError):
    pass


_UNARY = {"exp": P.SYM_EXP, "log": P.SYM_LOG, "log10": P.SYM_LOG10, "sqrt": P.SYM_SQRT,
          "abs": P.SYM_ABS, "tanh": P.SYM_TANH}
_BINARY = {"min": P.SYM_MIN, "max": P.SYM_MAX, "pow": P.SYM_POW}
_NUMBER = re.compile(r"(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_NAME = re.compile(r"[A-Za-z_][A-Za-z0-9_]*")


class _Parser:
    """Recursive descent into postfix code: [(op, arg)], constants by value."""

    def __init__(self, text: str, names: Sequence[str]):
        self.s, self.p, self.names = text, 0, list(names)
        self.code: List[Tuple[int, object]] = []

    def fail(self, msg):
        raise SymError("%s at character %d of '%s'" % (msg, self.p + 1, self.s))

    def skip(self):
        while self.p < len(self.s) and self.s[self.p] in " \t":
            self.p += 1

    def peek(self):
        self.skip()
        return self.s[self.p] if self.p < len(self.s) else ""

    def parse(self):
        self.expr()
        if self.peek():
            self.fail("unexpected text after the expression")
        return self.code

    def expr(self):
        self.term()
        while True:
            c = self.peek()
            if c == "+":
                self.p += 1; self.term(); self.code.append((P.SYM_ADD, 0))
            elif c == "-":
                self.p += 1; self.term(); self.code.append((P.SYM_SUB, 0))
            else:
                return

    def term(self):
        self.unary()
        while True:
            c = self.peek()
            if c == "*":
                self.p += 1; self.unary(); self.code.append((P.SYM_MUL, 0))
            elif c == "/":
                self.p += 1; self.unary(); self.code.append((P.SYM_DIV, 0))
            else:
                return

    def unary(self):
        c = self.peek()
        if c == "-":
            self.p += 1; self.unary(); self.code.append((P.SYM_NEG, 0))
        elif c == "+":
            self.p += 1; self.unary()
        else:
            self.power()

    def power(self):
        self.atom()
        if self.peek() == "^":
            self.p += 1; self.unary(); self.code.append((P.SYM_POW, 0))

    def atom(self):
        c = self.peek()
        if not c:
            self.fail("expression ended early")
        if c == "(":
            self.p += 1; self.expr()
            if self.peek() != ")":
                self.fail("missing ')'")
            self.p += 1
            return
        if c.isdigit() or c == ".":
            m = _NUMBER.match(self.s, self.p)
            if not m:
                self.fail("could not read a number")
            self.p = m.end()
            self.code.append((P.SYM_CONST, float(m.group(0))))
            return
        m = _NAME.match(self.s, self.p)
        if not m:
            self.fail("unexpected character")
        name = m.group(0)
        self.p = m.end()
        if self.peek() == "(":
            self.p += 1
            self.expr()
            if name in _BINARY:
                if self.peek() != ",":
                    self.fail("'%s' takes two arguments" % name)
                self.p += 1
                self.expr()
            if self.peek() != ")":
                self.fail("missing ')' after '%s'" % name)
            self.p += 1
            if name in _UNARY:
                self.code.append((_UNARY[name], 0))
            elif name in _BINARY:
                self.code.append((_BINARY[name], 0))
            else:
                self.fail("unknown function '%s'" % name)
            return
        if name == "pi":
            self.code.append((P.SYM_CONST, 3.14159265358979323846))
        elif name == "e":
            self.code.append((P.SYM_CONST, 2.71828182845904523536))
        elif name in self.names:
            self.code.append((P.SYM_VAR, self.names.index(name)))
        else:
            self.fail("'%s' is not one of the declared variables" % name)


@dataclass
class SymLaw:
    vars: List[str] = field(default_factory=list)
    rates: List[Tuple[str, str]] = field(default_factory=list)      # (output name, expression)
    ranges: Dict[str, Tuple[float, float]] = field(default_factory=dict)
    unit_scale: float = 1.0
    units: str = "per_second"
    reaction: List[Tuple[str, float]] = field(default_factory=list)
    yield_of: str = ""
    yield_val: float = 0.0
    biomass: str = ""
    derived: int = 0
    code: List[List[Tuple[int, object]]] = field(default_factory=list)   # per rate


def parse_expression(text: str, names: Sequence[str]):
    """Postfix code of one expression over `names` (value slots in order)."""
    return _Parser(text, names).parse()


def read_sym(path) -> SymLaw:
    law = SymLaw()
    with open(path) as fh:
        lines = fh.readlines()
    for ln, raw in enumerate(lines, 1):
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        at = " (%s line %d)" % (path, ln)
        key, _, rest = line.partition(" ")
        rest = rest.strip()
        t = rest.split()
        if key == "version":
            continue
        if key == "units":
            u = rest
            if u in ("per_second", "per_sec", "s"):
                law.unit_scale, law.units = 1.0, "per_second"
            elif u in ("per_hour", "per_hr", "h"):
                law.unit_scale, law.units = 1.0 / 3600.0, "per_hour"
            else:
                raise SymError("'units' must be per_second or per_hour, not '%s'%s" % (u, at))
        elif key == "vars":
            if not t:
                raise SymError("'vars' names no variable%s" % at)
            for v in t:
                if v in law.vars:
                    raise SymError("variable '%s' declared twice%s" % (v, at))
                law.vars.append(v)
        elif key == "rate":
            name, eq, expr = rest.partition("=")
            name = name.strip()
            if not eq or not name:
                raise SymError("a rate line is 'rate NAME = EXPRESSION'%s" % at)
            if any(r[0] == name for r in law.rates):
                raise SymError("rate '%s' defined twice%s" % (name, at))
            law.code.append(parse_expression(expr.strip(), law.vars + [r[0] for r in law.rates]))
            law.rates.append((name, expr.strip()))
        elif key == "range":
            if len(t) != 3 or t[0] not in law.vars:
                raise SymError("a range line is 'range VARIABLE LO HI' for a declared variable%s" % at)
            law.ranges[t[0]] = (float(t[1]), float(t[2]))
        elif key == "reaction":
            if len(t) < 2 or len(t) % 2:
                raise SymError("'reaction' takes name/coefficient pairs%s" % at)
            for n, v in zip(t[0::2], t[1::2]):
                if float(v) == 0.0:
                    raise SymError("'%s' has coefficient 0 in the reaction%s" % (n, at))
                law.reaction.append((n, float(v)))
        elif key == "yield":
            if len(t) != 2:
                raise SymError("'yield' takes a species and a value%s" % at)
            law.yield_of, law.yield_val = t[0], float(t[1])
        elif key == "biomass":
            if len(t) != 1 or t[0] not in law.vars:
                raise SymError("'biomass' names one declared variable%s" % at)
            law.biomass = t[0]
        else:
            raise SymError("'%s' is not a keyword of a .sym file%s" % (key, at))
    if not law.vars:
        raise SymError("%s has no 'vars' line" % path)
    names = [r[0] for r in law.rates]
    growth, extent = "growth" in names, "extent" in names
    if growth and extent:
        raise SymError("%s: a law has a 'growth' line (an organism's) or an 'extent' line (no organism's), "
                       "not both" % path)
    if extent and not law.reaction:
        raise SymError("%s: an 'extent' line needs the 'reaction' it is the rate of" % path)
    if law.reaction and extent:
        # abiotic: nu_i * extent per species (complab3d_symbolic.hh:653-684)
        if law.yield_of or law.biomass:
            raise SymError("%s: an 'extent' law belongs to no organism; 'yield' and 'biomass' do not apply" % path)
        for n, v in law.reaction:
            if n in names:
                raise SymError("%s: '%s' has a rate line and a place in the reaction; keep one" % (path, n))
            text = "%s * extent" % _g10(v)
            law.code.append(parse_expression(text, law.vars + [r[0] for r in law.rates]))
            law.rates.append((n, text))
            law.derived += 1
    elif law.reaction:
        # derived substrate lines (complab3d_symbolic.hh:686-745)
        if not growth:
            raise SymError("%s: a 'reaction' line needs a 'growth' line (an organism's law) or an "
                           "'extent' line (no organism's)" % path)
        if not law.yield_of or not law.yield_val > 0 or not law.biomass:
            raise SymError("%s: a biotic 'reaction' needs 'yield' (positive) and 'biomass'" % path)
        key = [v for n, v in law.reaction if n == law.yield_of]
        if not key:
            raise SymError("%s: the yield species '%s' is not in the reaction" % (path, law.yield_of))
        for n, v in law.reaction:
            if n in names:
                raise SymError("%s: '%s' has a rate line and a place in the reaction; keep one" % (path, n))
            coef = (v / abs(key[0])) / law.yield_val
            text = "%s * growth * %s" % (_g10(coef), law.biomass)
            law.code.append(parse_expression(text, law.vars + [r[0] for r in law.rates]))
            law.rates.append((n, text))
            law.derived += 1
    elif law.yield_of or law.biomass:
        raise SymError("%s: 'yield' or 'biomass' without a 'reaction' line derives nothing" % path)
    if not law.rates:
        raise SymError("%s declares no rate" % path)
    return law


def _g10(x: float) -> str:
    """printf("%.10g", x)."""
    return "%.10g" % x


def is_abiotic(law: SymLaw) -> bool:
    return any(r[0] == "extent" for r in law.rates)


def compile_law(law: SymLaw, species: Sequence[str], microbe: Optional[str]) -> P.SymProgram:
    """The law bound to a run (bindToSubstrates, complab3d_symbolic.hh:812-883):
    each variable to a substrate or to this organism's biomass (its name),
    each rate to 'growth' or a substrate.  microbe None binds the abiotic law
    (<symbolic><abiotic_file>): substrates only, 'extent' instead of
    'growth'."""
    abiotic = microbe is None
    if abiotic and any(r[0] == "growth" for r in law.rates):
        raise SymError("the abiotic law declares a 'growth' rate; an abiotic law belongs to no organism")
    if not abiotic and is_abiotic(law):
        raise SymError("an 'extent' law is abiotic and belongs in <abiotic_file>, not to microbe '%s'" % microbe)
    prog = P.SymProgram()
    if len(law.vars) > P.SYM_MAXVARS or len(law.rates) > P.SYM_MAXRATES:
        raise SymError("the law has more variables or rates than the kernel holds")
    prog.n_vars, prog.n_rates = len(law.vars), len(law.rates)
    for i, v in enumerate(law.vars):
        if v in species:
            prog.var_species[i] = list(species).index(v)
        elif not abiotic and v == microbe:
            prog.var_species[i] = -1
        elif abiotic:
            raise SymError("variable '%s' of the abiotic law is not a substrate of this run" % v)
        else:
            raise SymError("variable '%s' is neither a substrate nor the microbe '%s'" % (v, microbe))
        lo, hi = law.ranges.get(v, (0.0, 0.0))
        prog.lo[i], prog.hi[i] = lo, hi
    consts: List[float] = []
    ops: List[Tuple[int, int]] = []
    for r, ((name, _), code) in enumerate(zip(law.rates, law.code)):
        if name in ("growth", "extent"):
            prog.rate_species[r] = -1
        elif name in species:
            prog.rate_species[r] = list(species).index(name)
        else:
            raise SymError("rate '%s' is neither '%s' nor a substrate of this run"
                           % (name, "extent" if abiotic else "growth"))
        depth = 0
        for op, arg in code:
            if op == P.SYM_CONST:
                consts.append(float(arg))
                ops.append((op, len(consts) - 1))
                depth += 1
            elif op == P.SYM_VAR:
                ops.append((op, int(arg)))
                depth += 1
            else:
                ops.append((op, 0))
                if op in (P.SYM_ADD, P.SYM_SUB, P.SYM_MUL, P.SYM_DIV, P.SYM_POW, P.SYM_MIN, P.SYM_MAX):
                    depth -= 1
            if depth > P.SYM_STACK:
                raise SymError("rate '%s' needs a deeper stack than the kernel holds" % name)
        prog.rate_end[r] = len(ops)
    if len(ops) > P.SYM_MAXOPS or len(consts) > P.SYM_MAXCONST:
        raise SymError("the law is longer than the kernel holds")
    prog.n_ops, prog.n_const = len(ops), len(consts)
    for k, (op, arg) in enumerate(ops):
        prog.op[k], prog.arg[k] = op, arg
    for k, c in enumerate(consts):
        prog.konst[k] = c
    prog.unit_scale = law.unit_scale
    return prog

"""Run diagnostics: one row of field statistics per interval and the
mass-balance checks over it (complab3d_diagnostics.hh:57-250, rows built at
complab.cpp:2680-2790).

A row holds, per species and per microbe, the total over the reaction box, the
mean over the open voxels, the minimum, the maximum and `held` -- the mass the
total cannot see (the x-plane voxels, populations in flight into closed
voxels).  A conservation check names a sum of fields, `"A + B"`, and compares
the sum of their conserved amounts (total + held) against the first row it
saw, relative to the larger of the two magnitudes.  The statistics come from
one device call per row (`clb_field_stats`); nothing here touches a field.

The CSV file the reference writes is left to the caller: `Diagnostics.rows`
keeps every row recorded.
"""
from dataclasses import dataclass, field
from typing import List, Sequence


@dataclass
class FieldStat:
    name: str
    total: float = 0.0
    mean: float = 0.0
    minv: float = 0.0
    maxv: float = 0.0
    held: float = 0.0

    def conserved(self) -> float:
        """Everything the lattice holds of the field, wherever it sits."""
        return self.total + self.held


@dataclass
class Row:
    iteration: int = 0
    porosity: float = 0.0
    open_voxels: int = 0
    fields: List[FieldStat] = field(default_factory=list)


def split_plus(expr: str) -> List[str]:
    """The terms of a '+'-separated sum, blanks removed, empty terms dropped."""
    out = []
    for part in expr.split("+"):
        t = "".join(ch for ch in part if ch not in " \t")
        if t:
            out.append(t)
    return out


@dataclass
class ConserveCheck:
    expr: str
    terms: List[str]
    first: float = 0.0
    have_first: bool = False
    comparisons: int = 0
    worst_drift: float = 0.0
    warned: bool = False
    skipped: bool = False


class Diagnostics:
    """complab_diag::Diagnostics without the file output."""

    def __init__(self, tol=1e-6, enabled=True):
        self.enabled = enabled
        self.tol = tol
        self.rows: List[Row] = []
        self.checks: List[ConserveCheck] = []
        self.any_fail = False
        self._neg_warned = set()

    def add_conserve(self, expr: str):
        terms = split_plus(expr)
        if terms:
            self.checks.append(ConserveCheck(expr, terms))

    def failed(self) -> bool:
        return self.any_fail

    def record(self, row: Row) -> str:
        """Keeps the row, runs every check on it; returns what to print
        (empty when all is well)."""
        if not self.enabled:
            return ""
        self.rows.append(row)
        by_name = {f.name: f for f in row.fields}
        msg = []
        for k in self.checks:
            if any(t not in by_name for t in k.terms):
                if not k.warned:
                    k.warned = True
                    k.skipped = True
                    msg.append("  [DIAG] <conserve>%s</conserve> names a field that does not exist; "
                               "the check is skipped.\n" % k.expr)
                continue
            total = sum(by_name[t].conserved() for t in k.terms)
            if not k.have_first:
                k.first, k.have_first = total, True
                continue
            scale = max(abs(k.first), abs(total), 1e-300)
            drift = abs(total - k.first) / scale
            k.comparisons += 1
            k.worst_drift = max(k.worst_drift, drift)
            if drift > self.tol and not k.warned:
                k.warned = True
                self.any_fail = True
                msg.append("  [DIAG] MASS BALANCE: %s drifted %.3e (tolerance %.1e) by step %d.\n"
                           "  [DIAG]   started %.10g, now %.10g.\n"
                           "  [DIAG]   An open boundary, an unbalanced stoichiometry or a time step too "
                           "long for the rate are the usual causes.\n"
                           % (k.expr, drift, self.tol, row.iteration, k.first, total))
        for f in row.fields:
            if f.minv < -1e-9 and f.name not in self._neg_warned:
                self._neg_warned.add(f.name)
                self.any_fail = True
                msg.append("  [DIAG] %s went negative (%.3e) at step %d.\n" % (f.name, f.minv, row.iteration))
        return "".join(msg)

    def verdicts(self):
        """Per check: (expr, worst drift, 'PASS' | 'FAIL' | 'SKIPPED ...'); a
        check that never compared two rows is SKIPPED, not PASS."""
        out = []
        for k in self.checks:
            if k.skipped:
                v = "SKIPPED (unknown field)"
            elif k.comparisons == 0:
                v = "SKIPPED (never compared: fewer than two rows)"
            else:
                v = "PASS" if k.worst_drift <= self.tol else "FAIL"
            out.append((k.expr, k.worst_drift, v))
        return out

    def verdict(self) -> str:
        ran = [k for k in self.checks if not k.skipped and k.comparisons > 0]
        if not self.checks:
            return ""
        if not ran:
            return "NOT CHECKED"
        return "FAIL" if self.any_fail else "PASS"


def row_of(case, iteration: int, species_names: Sequence[str], microbe_names: Sequence[str] = ()) -> Row:
    """One diagnostics row from a case (the device case or the oracle's): the
    open-voxel count of the reaction box on the live mask, porosity over
    (nx-2)*ny*nz, and the field statistics."""
    p = case.p
    st = case.field_stats()
    open_voxels = int(case.counters()[0])
    box = (p.nx - 2) * p.ny * p.nz
    r = Row(iteration=iteration, porosity=open_voxels / box if box > 0 else 0.0, open_voxels=open_voxels)
    names = list(species_names) + list(microbe_names)
    if len(names) != st.shape[0]:
        raise ValueError("%d names for %d fields" % (len(names), st.shape[0]))
    for name, (tot, lo, hi, held) in zip(names, st):
        r.fields.append(FieldStat(name, float(tot), float(tot) / open_voxels if open_voxels > 0 else 0.0,
                                  float(lo), float(hi), float(held)))
    return r

"""The diagnostics row on the oracle and the mass-balance checks over rows.

`held` is pinned by what it must add up to: the whole lattice content (every
cell's populations, bounce-back cells included) minus the total minus the rest
content closed cells swap among themselves from set-up, which no open cell ever
meets and which therefore stays constant.  So whole - (total + held) must not
move, whatever the boundaries, the flow and the reactions do; and in a closed
box without reactions total + held itself must not move while total alone does.
"""
import numpy as np
import pytest

import oracle as O
from complab_amd import configs, diagnostics as D
from complab_amd import params as P


def box_case(btype, rxn):
    w = configs.config1(scale=0.25, packer=O.geom_pack)
    p = w.params
    for s in range(2):
        p.left_btype[s] = btype; p.right_btype[s] = btype
    if btype == P.BC_DIRICHLET:
        p.right_btype[0] = P.BC_NEUMANN; p.right_btype[1] = P.BC_NEUMANN
    if not rxn:
        p.n_rxn = 0
    o = O.OracleCase(p, w.geometry)
    rng = np.random.default_rng(0)
    closed = w.geometry < 2
    for a in range(3):
        u = rng.uniform(-0.03, 0.03, w.geometry.shape)
        u[closed] = 0
        o.write(P.FIELD_VEL, a, u)
    c = rng.uniform(0, 2, w.geometry.shape)
    c[closed] = p.c0[0]
    o.write(P.FIELD_CONC, 0, c)
    o.set_omegas([1.5, 1.5], [1.5, 1.5], 1.0)
    return o


def whole(o, s):
    return float(o.read(P.FIELD_CONC, s).sum())


def test_closed_box_total_plus_held_is_conserved():
    o = box_case(P.BC_CLOSED, rxn=False)
    o.ade_steps(3)
    st0 = o.field_stats()
    seen = []
    for _ in range(8):
        o.ade_steps(37)
        st = o.field_stats()
        for s in range(2):
            c0, c1 = st0[s, 0] + st0[s, 3], st[s, 0] + st[s, 3]
            assert abs(c1 - c0) <= 1e-12 * abs(c0), (s, c0, c1)
        seen.append(abs(st[0, 0] - st0[0, 0]) / st0[0, 0])
    assert max(seen) > 1e-6            # the total alone moves: held is doing work
    assert st[0, 3] > 0


@pytest.mark.parametrize("btype,rxn", [(P.BC_CLOSED, True), (P.BC_DIRICHLET, True), (P.BC_DIRICHLET, False)])
def test_held_is_the_lattice_content_the_total_misses(btype, rxn):
    o = box_case(btype, rxn)
    o.ade_steps(2)
    base = None
    for _ in range(6):
        o.ade_steps(23)
        st = o.field_stats()
        parked = np.array([whole(o, s) - st[s, 0] - st[s, 3] for s in range(2)])
        if base is None:
            base = parked
            continue
        scale = max(abs(whole(o, 0)), 1.0)
        assert np.all(np.abs(parked - base) <= 1e-12 * scale), (parked, base)


def test_minimum_and_maximum_over_the_open_voxels():
    o = box_case(P.BC_CLOSED, rxn=False)
    o.ade_steps(5)
    st = o.field_stats()
    c = o.read(P.FIELD_CONC, 0)
    m = o.read(P.FIELD_MASK)
    sel = m >= P.PORE
    sel[:, :, 0] = False; sel[:, :, -1] = False
    assert st[0, 1] == c[sel].min() and st[0, 2] == c[sel].max()
    assert st[0, 0] == pytest.approx(c[sel].sum(), rel=1e-13)


def test_record_checks_drift_against_the_larger_magnitude():
    d = D.Diagnostics(tol=1e-6)
    d.add_conserve("A + B")
    d.add_conserve(" C ")
    d.add_conserve("A+Z")
    rows = []
    for it, (a, b, ha) in enumerate([(1.0, 2.0, 0.5), (1.2, 2.0, 0.3), (1.2, 2.0, 0.3 + 7e-6)]):
        r = D.Row(iteration=it, porosity=0.4, open_voxels=10,
                  fields=[D.FieldStat("A", a, held=ha), D.FieldStat("B", b), D.FieldStat("C", 0.0)])
        rows.append(d.record(r))
    assert "names a field that does not exist" in rows[0]
    assert rows[1] == ""                                   # 3.5 == 3.5: no drift
    assert "MASS BALANCE: A + B drifted" in rows[2]
    v = {e: (drift, verdict) for e, drift, verdict in d.verdicts()}
    assert v["A + B"][1] == "FAIL" and v["A + B"][0] == pytest.approx(7e-6 / (3.5 + 7e-6))
    assert v[" C "] == (0.0, "PASS")                        # 0 -> 0 compares with scale 1e-300
    assert v["A+Z"][1].startswith("SKIPPED")
    assert d.failed() and d.verdict() == "FAIL"


def test_empty_start_drifts_by_at_most_one_and_single_row_is_not_a_pass():
    d = D.Diagnostics(tol=1e-6)
    d.add_conserve("P")
    d.record(D.Row(0, fields=[D.FieldStat("P", 0.0)]))
    assert d.verdict() == "NOT CHECKED"
    assert d.verdicts()[0][2].startswith("SKIPPED (never compared")
    d.record(D.Row(1, fields=[D.FieldStat("P", 5.0)]))
    assert d.verdicts()[0][1] == 1.0


def test_negative_concentration_is_reported_once():
    d = D.Diagnostics()
    m1 = d.record(D.Row(3, fields=[D.FieldStat("H", 1.0, minv=-2e-9)]))
    m2 = d.record(D.Row(4, fields=[D.FieldStat("H", 1.0, minv=-3e-9)]))
    assert "H went negative" in m1 and m2 == "" and d.failed()


def test_split_plus():
    assert D.split_plus(" Ca + HCO3+\tCaCO3 ++ ") == ["Ca", "HCO3", "CaCO3"]


def test_row_of_the_oracle_case():
    o = box_case(P.BC_CLOSED, rxn=False)
    o.ade_steps(4)
    r = D.row_of(o, 4, ["A", "B"])
    n = int(o.counters()[0])
    assert r.open_voxels == n and r.porosity == pytest.approx(n / ((o.p.nx - 2) * o.p.ny * o.p.nz))
    assert r.fields[0].mean == pytest.approx(r.fields[0].total / n)
    with pytest.raises(ValueError):
        D.row_of(o, 4, ["A"])

"""The oracle's restatement of mineral dissolution (dissolutionVOP.hh:378-812)
pinned by closed forms and invariants: the first-step inventory of a wetted
grain voxel, the half-per-step limiter, the element balance between the
mineral removed and the products released (every share parked is collected),
the enclosed-voxel rule, and the reopening threshold."""
import numpy as np

import oracle as O
from complab_amd import params as P

H, CA, CAL = 0, 1, 2            # proton, calcium (mobile), calcite inventory (immobile)


def calcite_case(k=1e-2, block=2, nx=10, n=6, surface_only=1, interval=1, h0=0.1):
    """A closed box of pore water with a cube of calcite grain (material 0)
    in its middle: example 14's rate law, rate = k [H+] mineral, releasing one
    Ca2+ and consuming one H+ per CaCO3."""
    p = P.default_params(nx, n, n, 3)
    p.mobile[CAL] = 0
    for s in (H, CA):
        p.left_btype[s] = P.BC_CLOSED; p.right_btype[s] = P.BC_CLOSED
        p.omega_pore[s] = 1.2; p.omega_biofilm[s] = 1.2
    p.c0[H] = h0
    p.clamp = 1
    P.set_phase(p, 0, CAL, 27.1, material=P.SOLID, initial=27.1, k_rate=k, acid=H, stoich={H: -1.0, CA: 1.0})
    p.dissol_reopen_frac = 0.9
    p.dissol_surface_only = surface_only
    p.dissol_interval = interval
    geom = np.full((n, n, nx), P.PORE, np.int8)
    lo = (n - block) // 2
    x0 = (nx - block) // 2
    geom[lo:lo + block, lo:lo + block, x0:x0 + block] = P.SOLID
    return p, geom


def grain_voxels(geom):
    return geom == P.SOLID


def box_total(o, s):
    """A mobile species over every voxel: the open water plus what the
    bounce-back cells hold in transit (a Palabos BounceBack cell stores the
    populations that streamed into it until it sends them back)."""
    return float(o.read(P.FIELD_CONC, s).sum())


def test_first_step_inventory_closed_form():
    """A wetted grain voxel loses k [H+] M dt in its first step: 27.1 (1 - 1e-3)."""
    p, geom = calcite_case()
    o = O.OracleCase(p, geom)
    o.run_steps(1)
    m = o.read(P.FIELD_CONC, CAL)[grain_voxels(geom)]
    np.testing.assert_allclose(m, 27.1 * (1.0 - 1e-2 * 0.1), rtol=1e-12)
    o.close()


def test_limiter_removes_half_per_step():
    """With k [H+] dt > 1/2 a step removes exactly half the inventory
    (MAX_DISSOL_FRACTION, dissolutionVOP.hh:485) and scales the products by
    the same factor."""
    p, geom = calcite_case(k=100.0, h0=1.0)
    p.dissol_reopen_frac = 0.01            # keep the voxels mineral for this check
    o = O.OracleCase(p, geom)
    o.run_steps(1)
    g = grain_voxels(geom)
    np.testing.assert_allclose(o.read(P.FIELD_CONC, CAL)[g], 27.1 * 0.5, rtol=1e-15)
    released = box_total(o, CA)
    np.testing.assert_allclose(released, 27.1 * 0.5 * g.sum(), rtol=1e-12)
    o.close()


def test_element_balance_and_every_share_collected():
    """Over many steps: Ca in the water equals the calcite removed, H + Ca is
    conserved (one proton per CaCO3), and the shares parked equal the shares
    collected (the invariant of dissolutionVOP.hh:243-265)."""
    p, geom = calcite_case(k=5e-2)
    o = O.OracleCase(p, geom)
    g = grain_voxels(geom)
    m0 = o.read(P.FIELD_CONC, CAL)[g].sum()
    h0 = box_total(o, H)
    o.run_steps(40)
    m1 = o.read(P.FIELD_CONC, CAL)[g].sum()
    assert m1 < m0
    np.testing.assert_allclose(box_total(o, CA), m0 - m1, rtol=1e-11)
    np.testing.assert_allclose(box_total(o, H) + box_total(o, CA), h0, rtol=1e-12)
    c = o.counters()
    assert c[8] > 0 and c[8] == c[9]
    assert c[10] == 0                      # nothing clamped at this rate
    o.close()


def test_enclosed_voxel_does_not_dissolve():
    """The centre of a 3x3x3 grain has no open face: with surface_only it keeps
    its inventory.  Without surface_only it reacts with the water it sees -- none,
    averaged over no face, so an acid-driven rate is zero -- and with a rate that
    needs no acid it dissolves while parking nothing (no face to park on)."""
    p, geom = calcite_case(block=3, n=7, nx=11)
    p.dissol_reopen_frac = 0.01
    o = O.OracleCase(p, geom)
    o.run_steps(3)
    m = o.read(P.FIELD_CONC, CAL)
    assert m[3, 3, 5] == 27.1
    assert m[3, 3, 4] < 27.1
    o.close()
    p.dissol_surface_only = 0
    o = O.OracleCase(p, geom)
    o.run_steps(1)
    assert o.read(P.FIELD_CONC, CAL)[3, 3, 5] == 27.1
    o.close()
    p.dissol_acid[0] = -1
    o = O.OracleCase(p, geom)
    o.run_steps(1)
    m = o.read(P.FIELD_CONC, CAL)
    np.testing.assert_allclose(m[3, 3, 5], 27.1 * (1 - 1e-2), rtol=1e-12)
    c = o.counters()
    assert c[8] == c[9] and c[8] > 0      # the wetted voxels park, the centre does not
    o.close()


def test_reopening_below_the_threshold():
    """dissolNodeConversion3D: a phase voxel whose inventory falls below
    reopen_fraction x full_density turns pore and loses its phase at the next
    check; it then carries solute like any pore voxel."""
    p, geom = calcite_case(k=100.0, h0=1.0, interval=2)
    o = O.OracleCase(p, geom)
    g = grain_voxels(geom)
    o.run_steps(1)                         # iT = 0 is a check step: 13.55 < 24.39 reopens
    mask = o.read(P.FIELD_MASK)
    assert np.all(mask[g] == P.PORE)
    assert np.all(o.read(P.FIELD_PHASE)[g] == 0)
    assert o.counters()[11] == g.sum()
    o.run_steps(4)
    ca = o.read(P.FIELD_CONC, CA)
    assert np.all(ca[g] > 0)               # calcium diffuses into the reopened voxels
    o.close()


MG, DOL = 3, 4                  # magnesium (mobile), a second mineral's inventory (immobile)


def two_phase_case(k1=1e-2, k2=2e-3, h0=0.1, acid_order2=2.0, mineral_order2=0.0):
    """Two declared grain phases in one closed box: phase 1 (id 1) a calcite
    cube on the no-dynamics material with example 14's law, phase 2 (id 2) a
    cube on the bounce-back material whose law is second order in acid and
    zero order in its mineral, releasing Mg2+ and consuming two protons."""
    nx, n = 14, 8
    p = P.default_params(nx, n, n, 5)
    p.mobile[CAL] = 0
    p.mobile[DOL] = 0
    for s in (H, CA, MG):
        p.left_btype[s] = P.BC_CLOSED; p.right_btype[s] = P.BC_CLOSED
        p.omega_pore[s] = 1.2; p.omega_biofilm[s] = 1.2
    p.c0[H] = h0
    P.set_phase(p, 0, CAL, 27.1, material=P.SOLID, initial=27.1, k_rate=k1, acid=H, stoich={H: -1.0, CA: 1.0})
    P.set_phase(p, 1, DOL, 15.0, material=P.WALL, initial=15.0, k_rate=k2, acid=H, acid_order=acid_order2,
                mineral_order=mineral_order2, stoich={H: -2.0, MG: 1.0})
    p.dissol_reopen_frac = 0.1
    p.dissol_interval = 1
    geom = np.full((n, n, nx), P.PORE, np.int8)
    geom[3:5, 3:5, 3:5] = P.SOLID
    geom[3:5, 3:5, 9:11] = P.WALL
    return p, geom


def test_two_grain_phases_first_step_and_ids():
    """Each phase takes the voxels of its material (ids 1 and 2) and its own
    inventory; in the first step a wetted voxel of phase 1 loses k1 [H] M and
    one of phase 2 loses k2 [H]^2 (zero order in its mineral)."""
    p, geom = two_phase_case()
    o = O.OracleCase(p, geom)
    ph = o.read(P.FIELD_PHASE)
    assert np.all(ph[geom == P.SOLID] == 1) and np.all(ph[geom == P.WALL] == 2)
    assert np.all(ph[geom == P.PORE] == 0)
    o.run_steps(1)
    np.testing.assert_allclose(o.read(P.FIELD_CONC, CAL)[geom == P.SOLID], 27.1 * (1 - 1e-2 * 0.1), rtol=1e-12)
    np.testing.assert_allclose(o.read(P.FIELD_CONC, DOL)[geom == P.WALL], 15.0 - 2e-3 * 0.1 ** 2, rtol=1e-12)
    # each mineral stays in its own phase's voxels
    assert np.all(o.read(P.FIELD_CONC, CAL)[geom == P.WALL] == 0.0)
    assert np.all(o.read(P.FIELD_CONC, DOL)[geom == P.SOLID] == 0.0)
    o.close()


def test_two_grain_phases_element_balance():
    """Over many steps: Ca in the water equals the calcite removed, Mg the
    second mineral removed, H + Ca + 2 Mg is conserved, and every share
    parked is collected."""
    p, geom = two_phase_case(k1=5e-2, k2=5e-2)
    o = O.OracleCase(p, geom)
    s1, s2 = geom == P.SOLID, geom == P.WALL
    m1, m2 = o.read(P.FIELD_CONC, CAL)[s1].sum(), o.read(P.FIELD_CONC, DOL)[s2].sum()
    h0 = box_total(o, H)
    o.run_steps(30)
    d1 = m1 - o.read(P.FIELD_CONC, CAL)[s1].sum()
    d2 = m2 - o.read(P.FIELD_CONC, DOL)[s2].sum()
    assert d1 > 0 and d2 > 0
    np.testing.assert_allclose(box_total(o, CA), d1, rtol=1e-11)
    np.testing.assert_allclose(box_total(o, MG), d2, rtol=1e-11)
    np.testing.assert_allclose(box_total(o, H) + box_total(o, CA) + 2 * box_total(o, MG), h0, rtol=1e-12)
    c = o.counters()
    assert c[8] > 0 and c[8] == c[9]
    o.close()


def test_each_phase_reopens_at_its_own_threshold():
    """dissolNodeConversion3D with two phases: each voxel compares its own
    phase's inventory with reopen_fraction x that phase's full density.  With
    the second mineral's threshold set two first-step losses below its
    inventory, phase 2 stays at the check after one step (iT 0) and reopens at
    the check after three (iT 2), while phase 1 stays."""
    p, geom = two_phase_case(k1=1e-4, k2=1e-4, h0=0.5)
    p.dissol_interval = 2
    loss = 1e-4 * 0.25                                # phase 2's first-step loss, k2 [H]^2
    p.dissol_full[1] = (15.0 - 2 * loss) / p.dissol_reopen_frac
    o = O.OracleCase(p, geom)
    o.run_steps(1)                                    # iT 0: nothing below its threshold yet
    assert o.counters()[11] == 0
    o.run_steps(2)                                    # iT 2: phase 2 below, phase 1 not
    mask = o.read(P.FIELD_MASK)
    wetted2 = (geom == P.WALL)
    assert np.all(mask[wetted2] == P.PORE)
    assert np.all(mask[geom == P.SOLID] == P.SOLID)
    assert np.all(o.read(P.FIELD_PHASE)[geom == P.SOLID] == 1)
    assert o.counters()[11] == wetted2.sum()
    o.close()


def test_phases_on_one_material_and_a_precipitate_phase():
    """Two grain phases naming the same material: the first takes the voxels'
    phase id, both seed their inventories there (complab.cpp:1672-1698).  A
    precipitate phase declared beside them starts nowhere."""
    p, geom = two_phase_case()
    p.dissol_material[1] = P.SOLID
    geom[geom == P.WALL] = P.PORE
    P.set_phase(p, 2, DOL, 5.0, precipitate=True, k_rate=0.0)
    p.precip_on = 1; p.precip_species = DOL; p.precip_max_rho = 1e30
    o = O.OracleCase(p, geom)
    ph = o.read(P.FIELD_PHASE)
    assert np.all(ph[geom == P.SOLID] == 1) and np.all(ph[geom != P.SOLID] == 0)
    assert np.all(o.read(P.FIELD_CONC, CAL)[geom == P.SOLID] == 27.1)
    assert np.all(o.read(P.FIELD_CONC, DOL)[geom == P.SOLID] == 15.0)
    o.close()

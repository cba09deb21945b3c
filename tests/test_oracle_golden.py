"""The oracle pinned against the reference's own code (golden vectors written by
tests/golden/make_golden.py from the reference's equilibrium solver and
geometry generator, compiled against its own Palabos stub) and against
closed forms of the equilibrium problem."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
from complab_amd import params as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def carbonate():
    p = P.default_params(3, 1, 1, 4)
    p.eq_on = 1; p.eq_nc = 2
    S = [[1, 0], [0, 1], [1, -1], [1, 1]]
    lk = [0.0, 0.0, -10.33, 6.35]
    for i in range(4):
        for j in range(2):
            p.eq_stoich[i][j] = S[i][j]
        p.eq_logk[i] = lk[i]
    p.eq_comp_species[0] = 0; p.eq_comp_species[1] = 1
    p.eq_tol = 1e-8; p.eq_max_iter = 200; p.eq_anderson_depth = 4
    return p


def calcium_carbonate(logk):
    p = P.default_params(3, 1, 1, 8)
    p.eq_on = 1; p.eq_nc = 6
    rows = [[1, 0, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0],
            [0, 0, 0, 0, 1, 0], [0, 0, 0, 0, 0, 1], [1, 1, 0, 0, 0, 0], [0, 1, 1, 0, 0, 0]]
    for i in range(8):
        for j in range(6):
            p.eq_stoich[i][j] = rows[i][j]
        p.eq_logk[i] = logk[i]
    for j in range(6):
        p.eq_comp_species[j] = j
    p.eq_tol = 1e-8; p.eq_max_iter = 200; p.eq_anderson_depth = 4
    return p


@pytest.mark.parametrize("fixture", ["eq_carbonate.npz", "eq_calcium_carbonate.npz"])
def test_aapcf_matches_reference_vectors(fixture):
    d = np.load(os.path.join(GOLD, fixture))
    p = carbonate() if "logk" not in d else calcium_carbonate(d["logk"])
    for k in range(d["cin"].shape[0]):
        out, it, res, cv = O.eq_solve(p, d["cin"][k])
        assert it == d["iters"][k]
        assert cv == bool(d["converged"][k])
        np.testing.assert_allclose(out, d["out"][k], rtol=1e-14, atol=0)
        vox = O.eq_voxel(p, d["cin"][k])
        np.testing.assert_allclose(vox, d["voxel"][k], rtol=1e-12, atol=1e-300)


def test_aapcf_matches_live_reference():
    if O.ref() is None:
        pytest.skip("reference tree not built here (oracle/_ref)")
    p = carbonate()
    rng = np.random.default_rng(101)
    for _ in range(40):
        c = 10 ** rng.uniform(-8, 0, size=4)
        a = O.eq_solve(p, c)
        b = O.ref_eq_solve(p, c)
        assert a[1] == b[1] and a[3] == b[3]
        np.testing.assert_allclose(a[0], b[0], rtol=1e-14, atol=0)


def _carbonate_bisection(T_c, T_h):
    """Brute-force root of the carbonate system: with h = [H+], the species are
    HCO3 = x, CO3 = x 10^-10.33 / h, H2CO3 = x h 10^6.35, and the two mass
    balances T_c = x (1 + a/h + b h), T_h = h - x a/h + x b h fix h by
    bisection on log h."""
    a, b = 10 ** -10.33, 10 ** 6.35

    def g(lh):
        h = 10 ** lh
        x = T_c / (1 + a / h + b * h)
        return h - x * a / h + x * b * h - T_h

    lo, hi = -30.0, 1.0
    glo = g(lo)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        gm = g(mid)
        if (gm > 0) == (glo > 0):
            lo, glo = mid, gm
        else:
            hi = mid
    h = 10 ** (0.5 * (lo + hi))
    x = T_c / (1 + a / h + b * h)
    return np.array([x, h, x * a / h, x * b * h])


def test_equilibrium_closed_form():
    p = carbonate()
    p.eq_solver = P.EQ_NEWTON
    p.eq_tol = 1e-13
    rng = np.random.default_rng(5)
    for _ in range(20):
        c = 10 ** rng.uniform(-6, -2, size=4)
        T_c = c[0] + c[2] + c[3]
        T_h = c[1] - c[2] + c[3]
        if T_h <= 0:
            continue
        exact = _carbonate_bisection(T_c, T_h)
        out, it, res, cv = O.eq_solve(p, c)
        assert cv
        np.testing.assert_allclose(out, exact, rtol=1e-9)
        # the AA-PCF solver reaches the same root to its own tolerance
        p2 = carbonate()
        out2, _, _, cv2 = O.eq_solve(p2, c)
        if cv2:
            np.testing.assert_allclose(out2, exact, rtol=1e-5)


def test_equilibrium_limiter_rules():
    """run_equilibrium_biotic: a change is capped at max(1e-9, 0.1|c|) and a
    result is floored at 1e-20 (part4_eqsolver.hh:1472-1485)."""
    p = carbonate()
    c = np.array([1e-3, 1e-9, 5e-3, 1e-6])
    eq, _, _, _ = O.eq_solve(p, c)
    v = O.eq_voxel(p, c)
    d = v - c
    cap = np.maximum(1e-9, 0.1 * np.abs(c))
    assert np.all(np.abs(d) <= cap * (1 + 1e-12))
    full = np.abs(eq - c) <= cap
    np.testing.assert_allclose(v[full], eq[full], rtol=1e-12)


def test_rng_and_generator_match_reference():
    d = np.load(os.path.join(GOLD, "geometry_generate.npz"))
    assert np.array_equal(O.rng_draws(int(d["rng_seed"]), 32), d["rng"])
    sph = O.geom_spheres(24, 16, 12, 3.0, 14, 77, 1)
    assert np.array_equal(sph, d["spheres"])


def _clusters(open_mask):
    from scipy import ndimage
    lab, n = ndimage.label(open_mask, structure=ndimage.generate_binary_structure(3, 1))
    return lab, n


def test_abc_geometry_fixture_inspection():
    """ABC_sweep_package/geometries/geom_0000 (32^3, index.csv porosity_block
    0.317749): counts, porosity and the single spanning cluster."""
    meta = json.load(open(os.path.join(GOLD, "abc_geometries.json")))["geom_0000"]
    codes = np.load(os.path.join(GOLD, "abc_geom_0000.npz"))["codes"]
    for k, v in meta["counts"].items():
        assert int((codes == int(k)).sum()) == v
    assert abs((codes >= 2).mean() - 0.317749) < 5e-7
    lab, n = _clusters(codes >= 2)
    assert n == meta["clusters"]
    span = set(np.unique(lab[:, :, 0])) & set(np.unique(lab[:, :, -1])) - {0}
    assert len(span) == meta["spanning"] == 1
    # no no-dynamics voxel touches pore through a face (make_geometries.py convention)
    solid = codes == 0
    pore = codes == 2
    for ax in range(3):
        for sh in (1, -1):
            rolled = np.roll(pore, sh, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = 0 if sh == 1 else -1
            rolled[tuple(edge)] = False
            assert not np.any(solid & rolled)


def test_pack_builder_conventions():
    rc, lat, phi = O.geom_pack(40, 20, 20, 3.0, 0.4, 9)
    assert rc == 0
    nz, ny, nx = lat.shape
    assert nx == 42
    assert np.array_equal(lat[:, :, 0], lat[:, :, 1]) and np.array_equal(lat[:, :, -1], lat[:, :, -2])
    assert np.all(lat[:, 0, :] == P.WALL) and np.all(lat[0, :, :] == P.WALL)
    core = lat[:, :, 1:-1]
    assert abs((core == P.PORE).mean() - phi) < 1e-12
    lab, n = _clusters(core == P.PORE)
    assert n == 1                                   # pockets walled off
    span = set(np.unique(lab[:, :, 0])) & set(np.unique(lab[:, :, -1])) - {0}
    assert len(span) == 1
    solid = core == P.SOLID
    for ax in range(3):
        for sh in (1, -1):
            r = np.roll(core == P.PORE, sh, axis=ax)
            e = [slice(None)] * 3; e[ax] = 0 if sh == 1 else -1; r[tuple(e)] = False
            assert not np.any(solid & r)


def test_params_layout_matches_c():
    out = np.zeros(8, dtype=np.int64)
    O.lib().orc_params_layout.argtypes = [C.c_void_p]
    O.lib().orc_params_layout(out.ctypes.data_as(C.c_void_p))
    CP = P.CaseParams
    assert list(out) == [C.sizeof(CP), CP.dt.offset, CP.bio_stoich.offset, CP.ca_seed.offset,
                         CP.eq_stoich.offset, CP.precip_on.offset, CP.dissol_stoich.offset, CP.thm_nu.offset]

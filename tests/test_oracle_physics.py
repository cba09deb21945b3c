"""The oracle's restatement of the Palabos pieces (BGK advection-diffusion with
cs^2 = 1/4, bounce-back, regularized planes, incompressible BGK flow, the
pressure planes) and of the reference processors, pinned by closed forms and
invariants.  Each check is chosen so a plausible mistake fails it: cs^2 = 1/3
instead of 1/4 misses the diffusion decay by 33 %, a pressure plane without
the 1/cs^2 factor misses Poiseuille by 3x, a lost deposit weight breaks the
exact first-order decay, a skipped clamp drives a concentration negative."""
import numpy as np
import pytest

import oracle as O
from complab_amd import configs, driver
from complab_amd import params as P


def channel_params(nx, species=1):
    p = P.default_params(nx, 1, 1, species)
    for s in range(species):
        p.left_btype[s] = P.BC_CLOSED
        p.right_btype[s] = P.BC_CLOSED
        p.omega_pore[s] = 1.0
        p.omega_biofilm[s] = 1.0
    return p


def test_diffusion_mode_decay_cs2_quarter():
    nx = 66                                     # fluid x = 1..64, closed planes at 0 and 65
    p = channel_params(nx)
    p.c0[0] = 1.0                               # the closed planes hold the background
    geom = np.full((1, 1, nx), P.PORE, np.int8)
    o = O.OracleCase(p, geom)
    L = nx - 2
    x = np.arange(nx)
    k = np.pi / L
    c = 1.0 + 0.1 * np.cos(k * (x - 0.5))
    o.write(P.FIELD_CONC, 0, c.reshape(1, 1, nx))
    T = 2000
    o.ade_steps(T)
    r = o.read(P.FIELD_CONC, 0).reshape(-1)[1:-1]
    xs = x[1:-1]
    amp = 2.0 / L * np.sum((r - r.mean()) * np.cos(k * (xs - 0.5)))
    D = 0.25 * (1.0 / 1.0 - 0.5)              # cs^2 (tau - 1/2), cs^2 = 1/4
    expect = 0.1 * np.exp(-D * k * k * T)
    assert abs(amp / expect - 1) < 5e-3
    total = o.read(P.FIELD_CONC, 0).sum()       # fluid and the two bounce-back planes
    assert abs(total - (c[1:-1].sum() + 2.0)) < 1e-12


def test_closed_box_conserves_mass_with_flow():
    w = configs.config1(scale=0.25, packer=O.geom_pack)
    p = w.params
    for s in range(2):
        p.left_btype[s] = P.BC_CLOSED; p.right_btype[s] = P.BC_CLOSED
    p.n_rxn = 0
    o = O.OracleCase(p, w.geometry)
    rng = np.random.default_rng(0)
    closed = w.geometry < 2
    for a in range(3):
        u = rng.uniform(-0.03, 0.03, w.geometry.shape)
        u[closed] = 0
        o.write(P.FIELD_VEL, a, u)
    c = rng.uniform(0, 2, w.geometry.shape)
    o.write(P.FIELD_CONC, 0, c)
    o.set_omegas([1.5, 1.5], [1.5, 1.5], 1.0)
    tot0 = o.read(P.FIELD_CONC, 0).sum()        # every cell, bounce-back links included
    o.ade_steps(300)
    tot1 = o.read(P.FIELD_CONC, 0).sum()
    assert abs(tot1 - tot0) <= 1e-12 * tot0


def test_dirichlet_planes_give_linear_steady_profile():
    nx = 34
    p = channel_params(nx)
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0
    p.right_btype[0] = P.BC_DIRICHLET; p.right_bc[0] = 0.0
    geom = np.full((1, 1, nx), P.PORE, np.int8)
    o = O.OracleCase(p, geom)
    o.ade_steps(30000)
    r = o.read(P.FIELD_CONC, 0).reshape(-1)
    xs = np.arange(1, nx - 1)
    a, b = np.polyfit(xs, r[1:-1], 1)
    assert np.max(np.abs(np.polyval([a, b], xs) - r[1:-1])) < 1e-9
    # node-based planes: the line passes through 1 at x = 0 and 0 at x = nx-1
    assert abs(b - 1.0) < 1e-6 and abs(a * (nx - 1) + b) < 1e-6


def test_neumann_outlet_copies_inner_density():
    nx = 20
    p = channel_params(nx)
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0
    p.right_btype[0] = P.BC_NEUMANN
    geom = np.full((1, 1, nx), P.PORE, np.int8)
    o = O.OracleCase(p, geom)
    o.ade_steps(40000)
    r = o.read(P.FIELD_CONC, 0).reshape(-1)
    # no flow, zero-gradient outlet, constant inlet: the whole channel fills to 1
    assert np.max(np.abs(r[1:] - 1.0)) < 1e-6


def first_order_box(k, clamp=1):
    nx = 6
    p = channel_params(nx, species=2)
    p.ny = p.nz = 4
    p.c0[0] = 1.0
    p.n_rxn = 1
    p.rxn_r1[0] = 0; p.rxn_k[0] = k
    p.rxn_stoich[0][0] = -1; p.rxn_stoich[0][1] = 1
    p.clamp = clamp
    geom = np.full((4, 4, nx), P.PORE, np.int8)
    return p, O.OracleCase(p, geom)


def test_first_order_decay_exact():
    """One step of A -> B from a uniform field: every voxel of the reaction box
    that is not next to a plane holds exactly 1 - k (the deposit weights sum to
    the increment); over many steps A + B stays exactly 1 in every voxel (both
    lattices relax alike and the reaction moves mass between them)."""
    k = 1e-2
    p, o = first_order_box(k)
    p.nx = 10
    o = O.OracleCase(p, np.full((4, 4, 10), P.PORE, np.int8))
    o.ade_steps(1)
    A = o.read(P.FIELD_CONC, 0)
    np.testing.assert_allclose(A[:, :, 2:-2], 1 - k, rtol=1e-14)
    o.ade_steps(49)
    A = o.read(P.FIELD_CONC, 0)[:, :, 1:-1]
    B = o.read(P.FIELD_CONC, 1)[:, :, 1:-1]
    np.testing.assert_allclose(A + B, 1.0, rtol=1e-13)
    assert np.all(A < 1 - k)


def test_clamp_caps_consumption_at_availability():
    """test_abiotic.cpp:119-125 / test_symbolic_processor.cpp:125-148: a debit
    larger than the voxel holds takes exactly what is there; production is
    never clamped."""
    p, o = first_order_box(3.0, clamp=1)
    p.nx = 10
    o = O.OracleCase(p, np.full((4, 4, 10), P.PORE, np.int8))
    o.ade_steps(1)
    A = o.read(P.FIELD_CONC, 0)[:, :, 2:-2]             # away from the closed planes
    B = o.read(P.FIELD_CONC, 1)[:, :, 2:-2]
    assert np.all(A == 0.0)
    np.testing.assert_allclose(B, 3.0, rtol=1e-12)        # the product term is R dt, unclamped
    assert o.counters()[2] > 0
    p2, o2 = first_order_box(3.0, clamp=0)
    p2.nx = 10
    o2 = O.OracleCase(p2, np.full((4, 4, 10), P.PORE, np.int8))
    o2.ade_steps(1)
    np.testing.assert_allclose(o2.read(P.FIELD_CONC, 0)[:, :, 2:-2], -2.0, rtol=1e-12)


def square_duct_mean_velocity(G, mu, W):
    s = 0.0
    for n in range(1, 200, 2):
        s += np.tanh(n * np.pi / 2) / n ** 5
    Q = G * W ** 4 / (12 * mu) * (1 - 192 / np.pi ** 5 * s)
    return Q / W ** 2


def test_poiseuille_square_duct():
    nxf, W = 30, 12
    nx, ny, nz = nxf + 2, W + 2, W + 2
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.flow_on = 1
    p.ns_tau = 0.8
    o = O.OracleCase(p, geom)
    dP = 1e-4
    o.ns_init(dP)
    it, cv = o.ns_solve(60000, P.CONV_RELDELTA, 1e-12, reset=True)
    assert cv
    o.ns_couple()
    ux = o.read(P.FIELD_VEL, 0)
    mid = ux[1:-1, 1:-1, nx // 2]
    G = dP / (nx - 1)                      # p = rho/3: the planes differ by dP over nx-1
    mu = (0.8 - 0.5) / 3.0
    expect = square_duct_mean_velocity(G, mu, W)
    assert abs(mid.mean() / expect - 1) < 0.03
    # the profile is symmetric in y and z
    np.testing.assert_allclose(mid, mid[::-1, :], rtol=1e-9)
    np.testing.assert_allclose(mid, mid.T, rtol=1e-9)


def test_peclet_rescale_hits_target():
    w = configs.config1(scale=0.25, packer=O.geom_pack)
    w.transport.stabilize_steps = 0
    o = O.OracleCase(w.params, w.geometry)
    rep = driver.setup(o, w.params, w.flow, w.transport)
    assert abs(rep.pe_achieved - w.flow.pe) / w.flow.pe < 0.02
    assert 0.5 < rep.ref_tau <= 2


def test_ca_spill_invariants():
    w = configs.config2(scale=0.1, packer=O.geom_pack)
    p = w.params
    p.flow_on = 0
    o = O.OracleCase(p, w.geometry)
    rng = np.random.default_rng(1)
    openv = w.geometry >= 2
    openv[:, :, 0] = False; openv[:, :, -1] = False
    b = np.zeros(w.geometry.shape)
    b[openv] = rng.uniform(0, 60, openv.sum())
    hot = np.flatnonzero(openv.reshape(-1))[::5]
    b.reshape(-1)[hot] = rng.uniform(120, 300, hot.size)
    o.write(P.FIELD_BIOMASS, 0, b)
    rc, sweeps = o.ca_spill()
    assert rc == 0 and sweeps > 0
    a = o.read(P.FIELD_BIOMASS, 0)
    assert abs(a.sum() - b.sum()) <= 1e-11 * b.sum()
    assert a.max() <= p.bmax + 1e-9
    assert np.all(a[w.geometry < 2] == 0)
    m = o.read(P.FIELD_MASK)
    thr = p.thrd_bfilm_frac * p.bmax
    assert np.all(m[(a >= thr) & openv] == P.BIOFILM)
    assert np.all(m[(a < thr) & openv] == P.PORE)


def gauss_seidel_age(age, dist, btot, bmax, once=False):
    """updateAgeDistance3D as the reference runs it: in place, x outermost,
    then y, then z (part2:868-955); repeated until the age sum stops changing,
    or one sweep (`once`, the pass after a mask change, complab.cpp:3079)."""
    nz, ny, nx = age.shape
    age = age.copy()
    T = 1e-14
    while True:
        s0 = age.sum()
        for x in range(nx):
            for y in range(ny):
                for z in range(nz):
                    B = btot[z, y, x]
                    if not B > T:
                        continue
                    nb = []
                    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                        X, Y, Z = x + dx, y + dy, z + dz
                        if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz and dist[Z, Y, X] > 0:
                            nb.append((Z, Y, X))
                    a = age[z, y, x]
                    if a == 0:
                        age[z, y, x] = 1
                    elif a == 1 and (B - bmax) > -T:
                        if not any(age[q] == 0 for q in nb):
                            age[z, y, x] = 2
                    else:
                        if all(age[q] >= a and (btot[q] - bmax) > -T for q in nb):
                            age[z, y, x] = a + 1
        if once or age.sum() == s0:
            return age


def jacobi_age_once(age, dist, btot, bmax):
    """The same rule against a snapshot of the ages: what a parallel sweep
    without ordering would give."""
    nz, ny, nx = age.shape
    old = age.copy()
    new = age.copy()
    T = 1e-14
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                B = btot[z, y, x]
                if not B > T:
                    continue
                nb = [(z + dz, y + dy, x + dx) for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))
                      if 0 <= x + dx < nx and 0 <= y + dy < ny and 0 <= z + dz < nz and dist[z + dz, y + dy, x + dx] > 0]
                a = old[z, y, x]
                if a == 0:
                    new[z, y, x] = 1
                elif a == 1 and (B - bmax) > -T:
                    if not any(old[q] == 0 for q in nb):
                        new[z, y, x] = 2
                elif all(old[q] >= a and (btot[q] - bmax) > -T for q in nb):
                    new[z, y, x] = a + 1
    return new


def test_single_age_sweep_is_the_in_place_pass():
    """The sweep after a mask change is ONE in-place pass in the processor's
    loop order (x outer, z inner): a voxel reads the ages its -x/-y/-z
    neighbours already took.  Checked against a pure-Python pass of the
    reference loop on a seeded 10x7x6 box whose outcome differs from a sweep
    against a snapshot (so an unordered sweep would fail here), and from
    several starting age fields."""
    nx, ny, nz = 10, 7, 6
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.n_microbes = 1
    p.bmax = 10.0
    dist = np.zeros(geom.shape, int)
    inner = np.zeros(geom.shape, bool); inner[:-1, :-1, :-1] = True
    dist[(geom >= 2) & inner] = 1
    differs = 0
    for seed in range(6):
        rng = np.random.default_rng(40 + seed)
        b = np.where(geom >= 2, rng.choice([0.0, 5.0, 10.0, 10.0, 10.0], size=geom.shape), 0.0)
        age0 = np.where(geom >= 2, rng.integers(0, 3, size=geom.shape), -1)
        o = O.OracleCase(p, geom)
        o.write(P.FIELD_BIOMASS, 0, b)
        o.write(P.FIELD_AGE, 0, age0)
        o.age_relax(False)
        got = o.read(P.FIELD_AGE).astype(int)
        o.close()
        want = gauss_seidel_age(age0, dist, b, p.bmax, once=True)
        assert np.array_equal(got, want), seed
        differs += int(not np.array_equal(want, jacobi_age_once(age0, dist, b, p.bmax)))
    assert differs >= 3


def test_age_relaxation_reaches_the_reference_fixed_point():
    """To the fixed point (the CA loop's relaxation, complab.cpp:3025-3033)
    the restatement matches the reference's in-place sweeps repeated until
    the age sum stops changing."""
    nx, ny, nz = 10, 7, 6
    rng = np.random.default_rng(4)
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.n_microbes = 1
    p.bmax = 10.0
    o = O.OracleCase(p, geom)
    b = np.where(geom >= 2, rng.choice([0.0, 5.0, 10.0, 10.0], size=geom.shape), 0.0)
    o.write(P.FIELD_BIOMASS, 0, b)
    age0 = np.where(geom >= 2, 0, -1)
    o.write(P.FIELD_AGE, 0, age0)
    o.age_relax(True)
    jac = o.read(P.FIELD_AGE).astype(int)
    dist = np.zeros_like(age0)
    inner = np.zeros(geom.shape, bool); inner[:-1, :-1, :-1] = True
    dist[(geom >= 2) & inner] = 1
    gs = gauss_seidel_age(age0, dist, b, p.bmax)
    assert np.array_equal(jac, gs)


def monod_params(mu, ks, kd, ke=0.1, acceptor=-1):
    p = P.default_params(3, 1, 1, 3)
    p.bio_law = P.BIO_DUAL_MONOD
    p.n_microbes = 1
    p.donor = 0
    p.acceptor = acceptor
    p.mu_max[0] = mu; p.ks[0] = ks; p.ke[0] = ke; p.kdecay[0] = kd
    p.bio_stoich[0][0] = -2.0
    p.bio_stoich[0][1] = -0.5
    p.bio_stoich[0][2] = 0.25
    return p


def test_monod_rates_at_closed_form_points():
    """defineKinetics.biotic.hh:90-126 (dual Monod).  At S = Ks the Monod factor
    is exactly 1/2; with S >> Ks growth is zero order (mu_max B); with the
    acceptor at E = Ke a second exact 1/2; decay enters with a minus sign and
    the stoichiometry scales growth only; a closed voxel reacts not at all."""
    mu, ks, kd, b = 0.3, 0.2, 0.01, 5.0
    p = monod_params(mu, ks, kd)
    sr, br, _ = O.rates(p, [ks, 1.0, 0.0], [b], 3)
    assert br[0] == pytest.approx((0.5 * mu - kd) * b, rel=1e-15)
    np.testing.assert_allclose(sr, np.array([-2.0, -0.5, 0.25]) * 0.5 * mu * b, rtol=1e-15)
    sr, br, _ = O.rates(p, [1e12 * ks, 1.0, 0.0], [b], 3)
    assert br[0] == pytest.approx((mu - kd) * b, rel=1e-11)
    p2 = monod_params(mu, ks, kd, ke=0.4, acceptor=1)
    sr, br, _ = O.rates(p2, [ks, 0.4, 0.0], [b], 3)
    assert br[0] == pytest.approx((0.25 * mu - kd) * b, rel=1e-15)
    assert sr[0] == pytest.approx(-2.0 * 0.25 * mu * b, rel=1e-15)
    sr, br, _ = O.rates(p, [ks, 1.0, 0.0], [b], P.SOLID)
    assert np.all(sr == 0) and np.all(br == 0)
    sr, br, _ = O.rates(p, [ks, 1.0, 0.0], [0.0], 3)
    assert np.all(sr == 0) and np.all(br == 0)


def test_precipitation_converts_at_threshold():
    """precipNodeConversion3D (precipitationVOP.hh:166-227): an open voxel of the
    reaction box whose mineral holds at least max_precipRho turns solid and
    takes the phase id; below the threshold, on the ghost columns and in
    closed voxels nothing changes."""
    nx, ny, nz = 8, 3, 3
    p = P.default_params(nx, ny, nz, 2)
    p.mobile[1] = 0
    p.precip_on = 1; p.precip_species = 1; p.precip_max_rho = 2.0
    p.precip_interval = 1; p.precip_phase_id = 5
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[1, 1, 4] = P.WALL
    o = O.OracleCase(p, geom)
    m = np.zeros((nz, ny, nx))
    m[0, 0, 2] = 2.0                       # exactly at the threshold: converts
    m[0, 0, 3] = 2.0 - 1e-12               # just below: stays open
    m[2, 2, 5] = 7.0                       # well above: converts
    m[1, 1, 0] = 9.0                       # ghost column: outside the reaction box
    m[1, 1, 4] = 9.0                       # already closed
    o.write(P.FIELD_CONC, 1, m)
    assert o.precip_convert()
    mask = o.read(P.FIELD_MASK)
    phase = o.read(P.FIELD_PHASE)
    conv = np.zeros_like(mask, bool)
    conv[0, 0, 2] = conv[2, 2, 5] = True
    assert np.all(mask[conv] == P.SOLID) and np.all(phase[conv] == 5)
    assert mask[0, 0, 3] == P.PORE and mask[1, 1, 0] == P.PORE and mask[1, 1, 4] == P.WALL
    assert np.all(phase[~conv] == 0)
    assert not o.precip_convert()          # nothing left above the threshold
    o.close()


def test_threads_do_not_change_results():
    """The oracle's per-cell loops run on every thread: one thread and all of
    them give bit-identical fields, totals and counters (biotic with CA and
    flow re-solves, so every threaded loop and the CA run)."""
    w = configs.config2(scale=0.08, packer=O.geom_pack)
    w.transport.stabilize_steps = 5
    w.params.ns_update_interval = 5
    w.params.ns_max_iter = 500
    w.params.b0[0] = 95.0
    w.params.mu_max[0] = 2e-2
    w.flow.max_iter = 500
    out = []
    n_all = O.set_threads(0)
    for n in (1, n_all):
        O.set_threads(n)
        o = O.OracleCase(w.params, w.geometry)
        rep = driver.setup(o, w.params, w.flow, w.transport)
        o.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        assert o.run_steps(30) == (0, 30)
        out.append((rep.iters, [o.read(P.FIELD_CONC, s) for s in range(4)], o.read(P.FIELD_BIOMASS, 0),
                    o.read(P.FIELD_MASK), o.counters(), o.read(P.FIELD_VEL, 0)))
    O.set_threads(n_all)
    a, b = out
    assert a[4][6] > 0 and a[4][4] > 0                    # the CA and flow re-solves ran
    assert a[0] == b[0]
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    for k in (2, 3, 4, 5):
        assert np.array_equal(a[k], b[k])


def ca_two_microbe_case(method, seed):
    """A 9x7x7 box (walls on y and z), two CA microbes, random biomass with
    some voxels above the cap; ages random so the age fallback varies."""
    nx, ny, nz = 9, 7, 7
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    rng = np.random.default_rng(seed)
    geom[1:-1, 1:-1, 1:-1][rng.uniform(size=(nz - 2, ny - 2, nx - 2)) < 0.15] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.n_microbes = 2
    p.bmax = 10.0
    p.thrd_bfilm_frac = 0.1
    p.ca_method = method
    p.ca_seed = 77 + seed
    openv = geom >= 2
    inner = openv.copy(); inner[:, :, 0] = False; inner[:, :, -1] = False
    b = [np.where(inner, rng.uniform(0, 6, geom.shape), 0.0) for _ in range(2)]
    hot = inner & (rng.uniform(size=geom.shape) < 0.3)
    b[0][hot] += rng.uniform(3, 12, hot.sum())
    b[1][hot] += rng.uniform(3, 12, hot.sum())
    geom[hot & (rng.uniform(size=geom.shape) < 0.7)] = P.BIOFILM     # most of the full voxels are biofilm
    o = O.OracleCase(p, geom)
    for m in range(2):
        o.write(P.FIELD_BIOMASS, m, b[m])
    o.write(P.FIELD_AGE, 0, np.where(openv, rng.integers(0, 4, geom.shape), -1).astype(float))
    return p, geom, o, b


@pytest.mark.parametrize("seed", range(4))
def test_ca_fraction_sweep_moves_the_excess_in_species_shares(seed):
    """One sweep of the fraction rule (pushExcessBiomass3D, part1:500-732):
    every over-full voxel with an open face neighbour inside the CA box ends
    at exactly the cap; the amount that leaves is sum (B - Bmax)+ over those
    voxels; each microbe leaves in its share B_m / B_tot of the donor (so
    B_m ends at B_m Bmax / B_tot); what leaves arrives in open face
    neighbours only, and the biomass of every microbe is conserved."""
    p, geom, o, b = ca_two_microbe_case(P.CA_FRACTION, seed)
    tot = b[0] + b[1]
    o.ca_sweep()
    a = [o.read(P.FIELD_BIOMASS, m) for m in range(2)]
    for m in range(2):
        assert abs(a[m].sum() - b[m].sum()) <= 1e-12 * b[m].sum()
        assert np.all(a[m][geom < 2] == 0)
    nz, ny, nx = geom.shape
    donors = np.zeros(geom.shape, bool)
    for z in range(nz):
        for y in range(ny):
            for x in range(1, nx - 1):
                if not (tot[z, y, x] > p.bmax and tot[z, y, x] - p.bmax > 1e-12):
                    continue
                nbs = [(x + 1, y, z), (x - 1, y, z), (x, y + 1, z), (x, y - 1, z), (x, y, z + 1), (x, y, z - 1)]
                gaps = [x == nx - 2, x == 1, y == ny - 1, y == 0, z == nz - 1, z == 0]
                if any(not g and geom[Z, Y, X] >= 2 for (X, Y, Z), g in zip(nbs, gaps)):
                    donors[z, y, x] = True
    assert donors.sum() > 5
    moved = np.sum(tot[donors] - p.bmax)
    # a donor receives from over-full neighbours too: its outflow is B - Bmax
    # exactly, so its own share ratio fixes what it kept before the gather
    kept = [b[m] * p.bmax / np.where(tot > 0, tot, 1.0) for m in range(2)]
    gained = [a[m] - np.where(donors, kept[m], b[m]) for m in range(2)]
    assert np.all(gained[0] >= -1e-12) and np.all(gained[1] >= -1e-12)
    assert abs(gained[0].sum() + gained[1].sum() - moved) <= 1e-10 * moved
    for m in range(2):
        out_m = np.sum(b[m][donors] - kept[m][donors])
        assert abs(out_m - np.sum(gained[m])) <= 1e-10 * moved


def test_ca_half_sweep_keeps_half_in_biofilm_donors():
    """The half rule (halfPushExcessBiomass3D, part1:735-970): an over-full
    biofilm voxel sends half of its total away; pore voxels never push."""
    p, geom, o, b = ca_two_microbe_case(P.CA_HALF, 9)
    m0 = o.read(P.FIELD_MASK)
    tot = b[0] + b[1]
    o.ca_sweep()
    a = [o.read(P.FIELD_BIOMASS, m) for m in range(2)]
    for m in range(2):
        assert abs(a[m].sum() - b[m].sum()) <= 1e-12 * b[m].sum()
    donors = (tot > p.bmax) & (m0 >= P.BIOFILM)
    pore_over = (tot > p.bmax) & (m0 == P.PORE)
    after = a[0] + a[1]
    # a donor keeps half unless a neighbour sent it something
    assert np.sum(donors) > 0
    assert np.all(after[donors] >= 0.5 * tot[donors] - 1e-12)
    assert np.all(after[pore_over] >= tot[pore_over] - 1e-12)

"""The reference's processors applied one at a time through the C-ABI (the
form a caller that keeps the reference's main loop uses), against the fused
whole-step path and against the CPU oracle."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs
from complab_amd import params as P
import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-10


def relerr(a, b, mask=None):
    if mask is not None:
        a = a[mask]; b = b[mask]
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def reactive_case(scale=0.12):
    """config 2's dual-Monod biofilm plus a first-order abiotic reaction, a
    seeded velocity field and random initial concentrations."""
    w = configs.config2(scale=scale)
    p = w.params
    p.flow_on = 0
    p.n_rxn = 1
    p.rxn_r1[0] = 2; p.rxn_r2[0] = -1; p.rxn_k[0] = 5e-3
    p.rxn_stoich[0][2] = -1.0; p.rxn_stoich[0][3] = 1.0
    p.mu_max[0] = 2e-2
    rng = np.random.default_rng(11)
    u = rng.uniform(-0.02, 0.02, size=(3,) + w.geometry.shape)
    closed = w.geometry < P.PORE
    for k in range(3):
        u[k][closed] = 0.0
    conc = [rng.uniform(0.05, 1.0, size=w.geometry.shape) for _ in range(p.n_species)]
    return w, u, conc


def prepare(c, u, conc):
    for k in range(3):
        c.write(P.FIELD_VEL, k, u[k])
    for s, f in enumerate(conc):
        c.write(P.FIELD_CONC, s, f)
    c.set_omegas([1.2, 1.25, 1.3, 1.1], [1.6, 1.65, 1.7, 1.5], 1.0)


def test_split_kinetics_matches_fused_and_oracle(ctx):
    """collide+stream then run_kinetics/update_rxnLattices as separate calls
    equals the fused kernel bit for bit, and both equal the oracle."""
    w, u, conc = reactive_case()
    fused = clb.Case(ctx, w.params, w.geometry)
    split = clb.Case(ctx, w.params, w.geometry)
    o = O.OracleCase(w.params, w.geometry)
    for c in (fused, split, o):
        prepare(c, u, conc)
    steps = 25
    fused.ade_steps(steps)
    for _ in range(steps):
        split.ade_collide_stream(1)
        split.kinetics(P.KIN_BIOTIC | P.KIN_ABIOTIC)
    o.ade_steps(steps)
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(w.params.n_species):
        a, b = fused.read(P.FIELD_CONC, s), split.read(P.FIELD_CONC, s)
        assert np.array_equal(a, b), "species %d: split differs from fused" % s
        e = relerr(a, o.read(P.FIELD_CONC, s), om)
        print("species", s, "rel err vs oracle", e)
        assert e <= REL
    assert np.array_equal(fused.read(P.FIELD_BIOMASS, 0), split.read(P.FIELD_BIOMASS, 0))
    assert relerr(fused.read(P.FIELD_BIOMASS, 0), o.read(P.FIELD_BIOMASS, 0)) <= REL
    assert fused.counters()[2] == split.counters()[2]
    o.close(); fused.close(); split.close()


def test_kinetics_selector_applies_one_family(ctx):
    """Biotic then abiotic as two calls equals both in one call (the main loop
    applies them in that order); the abiotic family leaves biomass alone and
    the biotic family moves it."""
    w, u, conc = reactive_case()
    one = clb.Case(ctx, w.params, w.geometry)
    two = clb.Case(ctx, w.params, w.geometry)
    for c in (one, two):
        prepare(c, u, conc)
        c.ade_collide_stream(1)
    b0 = two.read(P.FIELD_BIOMASS, 0)
    two.kinetics(P.KIN_BIOTIC)
    b1 = two.read(P.FIELD_BIOMASS, 0)
    assert not np.array_equal(b1, b0)
    two.kinetics(P.KIN_ABIOTIC)
    assert np.array_equal(two.read(P.FIELD_BIOMASS, 0), b1)
    one.kinetics(P.KIN_BIOTIC | P.KIN_ABIOTIC)
    for s in range(w.params.n_species):
        assert np.array_equal(one.read(P.FIELD_CONC, s), two.read(P.FIELD_CONC, s))
    assert np.array_equal(one.read(P.FIELD_BIOMASS, 0), two.read(P.FIELD_BIOMASS, 0))
    one.close(); two.close()


def spill_inputs(w):
    rng = np.random.default_rng(5)
    b = np.zeros(w.geometry.shape)
    openv = w.geometry >= 2
    openv[:, :, 0] = False; openv[:, :, -1] = False
    b[openv] = rng.uniform(0, 60, size=openv.sum())
    hot = np.flatnonzero(openv.reshape(-1))[:: 7]
    b.reshape(-1)[hot] = rng.uniform(150, 400, size=hot.size)
    return b


def test_granular_ca_loop_matches_block(ctx):
    """The CA block of complab.cpp:2793-3055 written as the reference writes it,
    with updateLocalMaskNtotalLattices3D, pushExcessBiomass3D and
    updateAgeDistance3D called one at a time, equals clb_ca_spill."""
    w = configs.config2(scale=0.1)
    p = w.params
    p.flow_on = 0
    b = spill_inputs(w)
    block = clb.Case(ctx, p, w.geometry)
    loop = clb.Case(ctx, p, w.geometry)
    for c in (block, loop):
        c.write(P.FIELD_BIOMASS, 0, b)
    rc, sweeps_block = block.ca_spill()
    assert rc == 0
    gmax, _ = loop.mask_total()
    prev, stall, wc, sweeps = gmax, 0, 0, 0
    while gmax - p.bmax > 1e-12:
        loop.ca_sweep()
        gmax, _ = loop.mask_total()
        sweeps += 1
        stall = stall + 1 if gmax > prev - 1e-12 else 0
        assert stall < 100
        prev = gmax
        if wc % 50 == 0:
            loop.age_update(True)
        wc += 1
    assert sweeps == sweeps_block
    for f in (P.FIELD_BIOMASS, P.FIELD_BTOTAL, P.FIELD_MASK, P.FIELD_AGE):
        assert np.array_equal(block.read(f), loop.read(f))
    block.close(); loop.close()


def test_precip_convert_and_solute_solid_match_oracle(ctx):
    """precipNodeConversion3D then updateSoluteSolidDynamics3D, one call each,
    against the oracle's conversion; the converted voxels stop carrying solute."""
    nx, ny, nz = 18, 8, 8
    p = P.default_params(nx, ny, nz, 2)
    p.mobile[1] = 0
    p.precip_on = 1; p.precip_species = 1; p.precip_max_rho = 1.0
    p.precip_interval = 1; p.precip_phase_id = 3
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0
    p.right_btype[0] = P.BC_NEUMANN
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL
    geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    rng = np.random.default_rng(2)
    m = rng.uniform(0.0, 1.3, size=geom.shape)
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    for c in (g, o):
        c.set_omegas([1.2, 1.0], [1.2, 1.0], 1.0)
        c.write(P.FIELD_CONC, 1, m)
    assert g.precip_convert() == o.precip_convert() == True
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    assert np.array_equal(g.read(P.FIELD_PHASE), o.read(P.FIELD_PHASE))
    assert g.update_solute_solid() == 0
    g.ade_collide_stream(3)
    conc = g.read(P.FIELD_CONC, 0)
    solid = g.read(P.FIELD_MASK) == P.SOLID
    assert np.all(conc[solid] == 0.0)
    assert not g.precip_convert()
    g.close(); o.close()


def test_cpp_processor_form_matches_whole_steps_and_oracle(ctx, tmp_path):
    """The reference's loop in processor form, driven from C++ through
    include/clb_processors.hh alone (complab3d-gems_amd/clb_run_processors,
    built with the library): collide and stream, run_kinetics +
    update_rxnLattices, run_abiotic_kinetics + update_abiotic_rxnLattices, and
    the CA block with updateLocalMaskNtotalLattices3D, pushExcessBiomass3D and
    updateAgeDistance3D, on a biotic case with an abiotic reaction whose
    biomass starts above Bmax in places.  Equals the same steps through the
    fused transport launch and the CA block from Python bit for bit, and the
    oracle at 1e-10 (masks and ages exact)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(clb.LIB_PATH), "clb_run_processors")
    if not os.path.exists(exe):
        pytest.fail("clb_run_processors is not built (run the build)")
    w, u, conc = reactive_case()
    p = w.params
    b = spill_inputs(w)
    omp, omb = [1.2, 1.25, 1.3, 1.1], [1.6, 1.65, 1.7, 1.5]
    steps = 12
    (tmp_path / "p.bin").write_bytes(bytes(p))
    (tmp_path / "g.bin").write_bytes(np.ascontiguousarray(w.geometry, dtype=np.int8).tobytes())
    init = np.concatenate([np.ravel(x) for x in list(u) + conc + [b]] + [np.array(omp + omb)]).astype(np.float64)
    init.tofile(str(tmp_path / "init.bin"))
    r = subprocess.run([exe, str(tmp_path / "p.bin"), str(tmp_path / "g.bin"), str(tmp_path / "init.bin"), str(steps),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sweeps_c = int(r.stdout.split()[1])
    out = np.fromfile(str(tmp_path / "out.bin"), dtype=np.float64).reshape(-1, *w.geometry.shape)
    S = p.n_species
    cpp = {"conc": out[:S], "bio": out[S], "mask": out[S + 1], "age": out[S + 2]}

    g = clb.Case(ctx, p, w.geometry)
    o = O.OracleCase(p, w.geometry)
    for c in (g, o):
        prepare(c, u, conc)
        c.write(P.FIELD_BIOMASS, 0, b)
    sweeps_g = sweeps_o = 0
    for _ in range(steps):
        g.ade_steps(1)
        o.ade_steps(1)
        rc, n = g.ca_spill()
        assert rc == 0
        sweeps_g += n
        rc, n = o.ca_spill()
        assert rc == 0
        sweeps_o += n
    print("CA sweeps: C++", sweeps_c, "python", sweeps_g, "oracle", sweeps_o)
    assert sweeps_c == sweeps_g == sweeps_o and sweeps_c > 0
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(S):
        assert np.array_equal(cpp["conc"][s], g.read(P.FIELD_CONC, s)), s
        assert relerr(cpp["conc"][s], o.read(P.FIELD_CONC, s), om) <= REL, s
    assert np.array_equal(cpp["bio"], g.read(P.FIELD_BIOMASS, 0))
    assert relerr(cpp["bio"], o.read(P.FIELD_BIOMASS, 0)) <= REL
    for f, key in ((P.FIELD_MASK, "mask"), (P.FIELD_AGE, "age")):
        assert np.array_equal(cpp[key], g.read(f)) and np.array_equal(cpp[key], o.read(f)), key
    g.close(); o.close()

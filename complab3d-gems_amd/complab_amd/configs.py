"""The workloads: parameter blocks + seeded geometry for each configuration.

Rates given "per step" are passed with a kinetic time step of 1, so the
increment of a step is rate * 1 exactly as the reference's R * ade_dt with
R = k_step / ade_dt.
"""
from dataclasses import dataclass

import numpy as np

from . import params as P
from .driver import FlowSetup, TransportSetup


@dataclass
class Workload:
    name: str
    params: P.CaseParams
    geometry: np.ndarray          # (nz, ny, nx) int8 lattice incl. ghost columns
    porosity: float
    flow: FlowSetup
    transport: TransportSetup
    steps: int
    rates_per_step: bool = True


def _pack(nx_file, ny, nz, radius, porosity, seed, kind=P.GEOM_SPHERES, packer=None):
    if packer is None:
        from . import geom_pack as packer
    for k in range(64):                          # a seed that does not percolate is skipped
        rc, lat, phi = packer(nx_file, ny, nz, radius, porosity, seed + k, kind)
        if rc == 0:
            return lat, phi
    raise RuntimeError("no percolating packing found")


def config1(scale=1.0, precision=P.FP64, packer=None, seed=1):
    """Plain method, FP64: 64x32x32, spheres r=4 to porosity 0.45, walls on y/z,
    tau 0.8, Pe 1, two D3Q7 species A->B (k = 1e-3 per step), D 0.1 in pore,
    A Dirichlet 1 at the inlet, Neumann outlets, 2000 steps."""
    nxf, ny, nz = max(8, int(64 * scale)), max(6, int(32 * scale)), max(6, int(32 * scale))
    radius = max(1.5, 4.0 * scale)
    lat, phi = _pack(nxf, ny, nz, radius, 0.45, seed, packer=packer)
    p = P.default_params(nxf + 2, ny, nz, 2)
    p.precision = precision
    p.c0[0] = 0.0; p.c0[1] = 0.0
    p.left_btype[0] = P.BC_DIRICHLET; p.left_bc[0] = 1.0; p.right_btype[0] = P.BC_NEUMANN
    p.left_btype[1] = P.BC_DIRICHLET; p.left_bc[1] = 0.0; p.right_btype[1] = P.BC_NEUMANN
    p.n_rxn = 1
    p.rxn_r1[0] = 0; p.rxn_r2[0] = -1
    p.rxn_k[0] = 1e-3
    p.rxn_stoich[0][0] = -1.0; p.rxn_stoich[0][1] = 1.0
    p.ns_tau = 0.8
    p.flow_on = 1
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-8
    p.ns_max_iter = 100000
    # characteristic length = the domain length, the convention of the reference's
    # ABC sweep (ABC_sweep_package/tools/make_campaign.py:236, CHAR_LEN = NX*DX)
    fs = FlowSetup(pe=1.0, char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=0.8,
                   rule=P.CONV_RELDELTA, eps=1e-8, max_iter=100000)
    ts = TransportSetup(d_pore=[0.1, 0.1], d_biofilm=[0.1, 0.1])
    return Workload("plain-fp64 %dx%dx%d spheres r%g phi0.45 A->B" % (nxf, ny, nz, radius), p, lat, phi, fs, ts, 2000)


def seed_biofilm(lat, fraction, seed, code=P.BIOFILM):
    """Marks `fraction` of the grain-surface pore voxels (pore with a bounce-back
    face neighbour, x = 1..nx-2) as biofilm seeds; deterministic for a seed."""
    nz, ny, nx = lat.shape
    pore = lat == P.PORE
    wall = lat == P.WALL
    touch = np.zeros_like(pore)
    for ax, sh in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
        touch |= np.roll(wall, sh, axis=ax)
    surf = pore & touch
    surf[:, :, 0] = False; surf[:, :, nx - 1] = False
    surf[:, 0, :] = False; surf[:, ny - 1, :] = False; surf[0, :, :] = False; surf[nz - 1, :, :] = False
    idx = np.flatnonzero(surf.reshape(-1))
    rng = np.random.default_rng(seed)
    k = max(1, int(round(fraction * idx.size)))
    pick = rng.choice(idx, size=k, replace=False)
    out = lat.copy().reshape(-1)
    out[pick] = code
    return out.reshape(lat.shape), k


def config2(scale=1.0, precision=P.FP64, packer=None, seed=2):
    """Biotic plain method on GPU: 256x128x128, spheres r=6 to porosity 0.35,
    flow re-solved every 50 steps, four species (substrate, acceptor, two
    products), D 0.1 pore / 0.02 biofilm, dual Monod mu_max 1e-3, K 0.1, one
    biofilm seeded on 2 % of grain-surface voxels at 50, CA fraction rule,
    Bmax 100, viscosity ratio 10, 10k steps."""
    nxf, ny, nz = max(12, int(256 * scale)), max(8, int(128 * scale)), max(8, int(128 * scale))
    radius = max(1.5, 6.0 * scale)
    lat, phi = _pack(nxf, ny, nz, radius, 0.35, seed, packer=packer)
    lat, nseed = seed_biofilm(lat, 0.02, seed)
    p = P.default_params(nxf + 2, ny, nz, 4)
    p.precision = precision
    c0 = [0.5, 0.5, 0.0, 0.0]
    inlet = [1.0, 1.0, 0.0, 0.0]
    for s in range(4):
        p.c0[s] = c0[s]
        p.left_btype[s] = P.BC_DIRICHLET; p.left_bc[s] = inlet[s]
        p.right_btype[s] = P.BC_NEUMANN
    p.bio_law = P.BIO_DUAL_MONOD
    p.donor = 0; p.acceptor = 1
    p.n_microbes = 1
    p.mu_max[0] = 1e-3; p.ks[0] = 0.1; p.ke[0] = 0.1; p.kdecay[0] = 0.0
    for s, nu in enumerate((-0.02, -0.01, 0.01, 0.005)):
        p.bio_stoich[0][s] = nu
    p.b0[0] = 50.0
    p.bmax = 100.0
    p.thrd_bfilm_frac = 0.1
    p.ca_method = P.CA_FRACTION
    p.ca_seed = 1234 + seed
    p.perm_ratio = 0.1                 # viscosity ratio 10
    p.ns_tau = 0.8
    p.flow_on = 1
    p.ns_update_interval = 50
    p.ade_update_interval = 1
    p.ns_max_iter = 20000
    p.ns_conv_rule = P.CONV_TRACER
    p.ns_conv_eps = 1e-6
    # a reduced domain keeps the full Peclet number only at full scale: tiny grains
    # would need a pressure drop beyond the incompressible limit
    fs = FlowSetup(pe=min(1.0, scale), char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=0.8,
                   rule=P.CONV_TRACER, eps=1e-8, max_iter=100000, has_biofilm=True)
    ts = TransportSetup(d_pore=[0.1] * 4, d_biofilm=[0.02] * 4)
    return Workload("biotic-fp64 %dx%dx%d spheres r%g phi0.35 4 species CA" % (nxf, ny, nz, radius),
                    p, lat, phi, fs, ts, 10000)


def _draw_box(rng, lo, hi, max_aspect=4.0):
    """Edge lengths uniform in [lo, hi] with the longest at most max_aspect
    times the shortest."""
    while True:
        e = rng.integers(lo, hi + 1, size=3)
        if e.max() <= max_aspect * e.min():
            return int(e[0]), int(e[1]), int(e[2])


def sweep_case(seed, precision=P.FP32, size=(32, 160), steps=2000, stabilize=200):
    """Batched form, one case: a seeded medium (spheres / cylinders / fracture,
    porosity 0.25-0.5) of a cuboid drawn in size^3 with aspect up to 4:1,
    2-6 D3Q7 species, flow tau in [0.6, 1.2], Monod growth of a biofilm seeded
    on 2 % of the grain surface with mu_max and K log-uniform over one decade.
    The biofilm never spills (no CA threshold is reached), so the case's
    topology is fixed and its fields measure transport and kinetics alone."""
    rng = np.random.default_rng(seed)
    nxf, ny, nz = _draw_box(rng, size[0], size[1])
    kind = (P.GEOM_SPHERES, P.GEOM_CYLINDERS, P.GEOM_FRACTURE)[int(rng.integers(0, 3))]
    phi = float(rng.uniform(0.25, 0.5))
    radius = float(rng.uniform(3.0, 6.0))
    gseed = int(rng.integers(1, 1 << 30))
    try:
        lat, phi_got = _pack(nxf, ny, nz, radius, phi, gseed, kind=kind)
    except RuntimeError:
        # a drawn cylinder/fracture medium that never percolates in this box:
        # the same box and porosity as a sphere pack
        kind = P.GEOM_SPHERES
        lat, phi_got = _pack(nxf, ny, nz, radius, phi, gseed, kind=kind)
    lat, _ = seed_biofilm(lat, 0.02, seed)
    ns = int(rng.integers(2, 7))
    p = P.default_params(nxf + 2, ny, nz, ns)
    p.precision = precision
    for s in range(ns):
        p.c0[s] = 0.5 if s < 2 else 0.0
        p.left_btype[s] = P.BC_DIRICHLET; p.left_bc[s] = 1.0 if s < 2 else 0.0
        p.right_btype[s] = P.BC_NEUMANN
    p.bio_law = P.BIO_DUAL_MONOD
    p.donor = 0
    p.acceptor = 1 if ns >= 3 else -1
    p.n_microbes = 1
    p.mu_max[0] = float(10 ** rng.uniform(-3, -2))
    p.ks[0] = float(10 ** rng.uniform(-1.3, -0.3))
    p.ke[0] = float(10 ** rng.uniform(-1.3, -0.3))
    p.bio_stoich[0][0] = -0.02
    if ns >= 3:
        p.bio_stoich[0][1] = -0.01
    for s in range(2 if ns >= 3 else 1, ns):
        p.bio_stoich[0][s] = 0.01 / (s + 1)
    p.b0[0] = 10.0
    p.bmax = 1e30
    p.thrd_bfilm_frac = 0.0
    tau = float(rng.uniform(0.6, 1.2))
    p.ns_tau = tau
    p.flow_on = 1
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-6
    p.ns_max_iter = 20000
    fs = FlowSetup(pe=1.0, char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=tau,
                   rule=P.CONV_RELDELTA, eps=1e-6, max_iter=20000, has_biofilm=True)
    ts = TransportSetup(d_pore=[0.1] * ns, d_biofilm=[0.05] * ns, stabilize_steps=stabilize)
    names = {P.GEOM_SPHERES: "spheres", P.GEOM_CYLINDERS: "cylinders", P.GEOM_FRACTURE: "fracture"}
    return Workload("sweep-%d %dx%dx%d %s phi%.2f %d species tau%.2f" % (seed, nxf, ny, nz, names[kind], phi_got, ns, tau),
                    p, lat, phi_got, fs, ts, steps)


def serving_case(seed, precision=P.FP32, size=(48, 256), max_steps=20000, stabilize=200):
    """Serving form, one request: an ABC-style case (A + B -> P on the grain
    surface, P an immobile mineral that seals a voxel at 0.5 mol/L, the flow
    re-solved at every precipitation check, 100 steps apart) in a seeded sphere
    pack of a cuboid drawn in size^3, 3-5 species (A, B, P and up to two inert
    tracers), run to steady state or max_steps."""
    rng = np.random.default_rng(seed)
    nxf, ny, nz = _draw_box(rng, size[0], size[1])
    lat, phi_got = _pack(nxf, ny, nz, float(rng.uniform(3.0, 5.0)), float(rng.uniform(0.3, 0.45)),
                         int(rng.integers(1, 1 << 30)))
    ns = int(rng.integers(3, 6))
    A, B, PM = 0, 1, 2
    p = P.default_params(nxf + 2, ny, nz, ns)
    p.precision = precision
    p.mobile[PM] = 0
    for s in range(ns):
        p.c0[s] = 0.0
        p.left_btype[s] = P.BC_DIRICHLET; p.right_btype[s] = P.BC_NEUMANN
        p.left_bc[s] = 0.0
    p.left_bc[A] = 1.0
    p.c0[B] = 1.0
    for s in range(3, ns):
        p.left_bc[s] = 1.0                 # inert tracers fed at the inlet
    p.n_rxn = 1
    p.rxn_r1[0] = A; p.rxn_r2[0] = B
    p.rxn_k[0] = float(10 ** rng.uniform(-2.5, -1.5))
    p.rxn_stoich[0][A] = -1.0; p.rxn_stoich[0][B] = -1.0; p.rxn_stoich[0][PM] = 1.0
    p.surface_only = 1
    p.precip_on = 1; p.precip_species = PM; p.precip_max_rho = 0.5
    p.precip_interval = 100; p.precip_phase_id = 1
    p.precip_perm_ratio = 0.0
    p.ns_tau = 0.8
    p.flow_on = 1
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-6
    p.ns_max_iter = 20000
    fs = FlowSetup(pe=1.0, char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=0.8,
                   rule=P.CONV_RELDELTA, eps=1e-6, max_iter=20000)
    ts = TransportSetup(d_pore=[0.1] * ns, d_biofilm=[0.1] * ns, stabilize_steps=stabilize)
    return Workload("serve-%d %dx%dx%d spheres phi%.2f %d species" % (seed, nxf, ny, nz, phi_got, ns),
                    p, lat, phi_got, fs, ts, max_steps)


H_, CO3_, CA_, HCO3_, CACO3AQ_, CALCITE_ = range(6)


def config3_six_components(scale=1.0, precision=P.FP64, seed=3, steps=5000):
    """The precision form's grain field and inflow with a six-component
    carbonate-calcium tableau: components H+, CO3--, Ca++, Mg++, Na+, Cl-,
    complexes HCO3- (H + CO3) and CaCO3(aq) (Ca + CO3), log K drawn within +-2
    of 10.33 and 3.22 from the seed; eight mobile D3Q7 lattices (the species
    limit leaves no slot for a mineral, so no precipitation or dissolution);
    per-voxel FP64 Newton to 1e-12."""
    rng = np.random.default_rng(seed)
    nxf, ny, nz = max(12, int(512 * scale)), max(8, int(256 * scale)), max(8, int(256 * scale))
    lat, phi = _pack(nxf, ny, nz, 3.0, 0.3, seed, kind=P.GEOM_RANDOM)
    p = P.default_params(nxf + 2, ny, nz, 8)
    p.precision = precision
    MG, NA, CL, HCO3, CAAQ = 3, 4, 5, 6, 7
    c0 = [1e-7, 1e-5, 1e-3, 5e-4, 1e-2, 1.2e-2, 2.1e-2, 1.7e-5]
    inlet = [1e-5, 1e-7, 1e-3, 5e-4, 1e-2, 1.2e-2, 2.1e-2, 1.7e-7]
    for s in range(8):
        p.c0[s] = c0[s]
        p.left_btype[s] = P.BC_DIRICHLET; p.left_bc[s] = inlet[s]
        p.right_btype[s] = P.BC_NEUMANN
    p.eq_on = 1
    p.eq_solver = P.EQ_NEWTON
    p.eq_nc = 6
    for j in range(6):
        p.eq_stoich[j][j] = 1.0
        p.eq_comp_species[j] = j
    p.eq_stoich[HCO3][H_] = 1.0; p.eq_stoich[HCO3][CO3_] = 1.0
    p.eq_stoich[CAAQ][CA_] = 1.0; p.eq_stoich[CAAQ][CO3_] = 1.0
    p.eq_logk[HCO3] = 10.33 + float(rng.uniform(-2, 2))
    p.eq_logk[CAAQ] = 3.22 + float(rng.uniform(-2, 2))
    p.eq_tol = 1e-12
    p.eq_max_iter = 200
    p.ns_tau = 0.8
    p.flow_on = 1
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-6
    p.ns_max_iter = 100000
    fs = FlowSetup(pe=1.0, char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=0.8,
                   rule=P.CONV_RELDELTA, eps=1e-8, max_iter=100000)
    ts = TransportSetup(d_pore=[0.1] * 8, d_biofilm=[0.1] * 8, stabilize_steps=0)
    return Workload("precision6-%s %dx%dx%d random phi%.2f six-component carbonate-calcium eq"
                    % ({P.FP64: "fp64", P.FP32: "fp32", P.FP32_ADE: "fp32ade"}[precision], nxf, ny, nz, phi),
                    p, lat, phi, fs, ts, steps)


def config3(scale=1.0, precision=P.FP64, seed=3, steps=5000):
    """Precision form: 512x256x256 seeded random grain field at porosity 0.3;
    the carbonate-calcium system as six species -- H+, CO3--, Ca++ (the
    equilibrium components), HCO3- and CaCO3(aq) (complexes, log K 10.33 and
    3.22 drawn within +-2 from the seed), all five mobile D3Q7 lattices, and an
    immobile calcite inventory.  Ca + HCO3 precipitates calcite on the grain
    surface releasing H (sealing a voxel at 0.05 mol/L), acid dissolves it back
    (CaCO3 + H -> Ca + HCO3, reopening below 0.9 of full); speciation by
    per-voxel FP64 Newton to 1e-12.  Acidic inflow at x = 0."""
    rng = np.random.default_rng(seed)
    nxf, ny, nz = max(12, int(512 * scale)), max(8, int(256 * scale)), max(8, int(256 * scale))
    lat, phi = _pack(nxf, ny, nz, 3.0, 0.3, seed, kind=P.GEOM_RANDOM)
    p = P.default_params(nxf + 2, ny, nz, 6)
    p.precision = precision
    p.mobile[CALCITE_] = 0
    c0 = {H_: 1e-7, CO3_: 1e-5, CA_: 1e-3, HCO3_: 2.1e-2, CACO3AQ_: 1.7e-5, CALCITE_: 0.0}
    inlet = {H_: 1e-5, CO3_: 1e-7, CA_: 1e-3, HCO3_: 2.1e-2, CACO3AQ_: 1.7e-7}
    for s in range(6):
        p.c0[s] = c0[s]
        if s != CALCITE_:
            p.left_btype[s] = P.BC_DIRICHLET; p.left_bc[s] = inlet[s]
            p.right_btype[s] = P.BC_NEUMANN
    # equilibrium: components H, CO3, Ca; complexes HCO3 = H + CO3, CaCO3(aq) = Ca + CO3
    p.eq_on = 1
    p.eq_solver = P.EQ_NEWTON
    p.eq_nc = 3
    rows = {H_: (1, 0, 0), CO3_: (0, 1, 0), CA_: (0, 0, 1), HCO3_: (1, 1, 0), CACO3AQ_: (0, 1, 1)}
    logk = {H_: 0.0, CO3_: 0.0, CA_: 0.0,
            HCO3_: 10.33 + float(rng.uniform(-2, 2)), CACO3AQ_: 3.22 + float(rng.uniform(-2, 2))}
    for s, row in rows.items():
        for j in range(3):
            p.eq_stoich[s][j] = row[j]
        p.eq_logk[s] = logk[s]
    p.eq_comp_species[0] = H_; p.eq_comp_species[1] = CO3_; p.eq_comp_species[2] = CA_
    p.eq_tol = 1e-12
    p.eq_max_iter = 200
    # Ca + HCO3 -> calcite + H on the grain surface
    p.n_rxn = 1
    p.rxn_r1[0] = CA_; p.rxn_r2[0] = HCO3_; p.rxn_k[0] = 5.0
    p.rxn_stoich[0][CA_] = -1.0; p.rxn_stoich[0][HCO3_] = -1.0; p.rxn_stoich[0][CALCITE_] = 1.0
    p.rxn_stoich[0][H_] = 1.0
    p.surface_only = 1
    p.precip_on = 1; p.precip_species = CALCITE_; p.precip_max_rho = 0.05
    p.precip_interval = 100; p.precip_phase_id = 1; p.precip_perm_ratio = 0.0
    # CaCO3 + H -> Ca + HCO3 on the wetted faces of the sealed voxels
    P.set_phase(p, 0, CALCITE_, 0.05, precipitate=True, k_rate=20.0, acid=H_,
                stoich={H_: -1.0, CA_: 1.0, HCO3_: 1.0})
    p.dissol_reopen_frac = 0.9
    p.dissol_surface_only = 1
    p.dissol_interval = 100
    p.ns_tau = 0.8
    p.flow_on = 1
    # the set-up solve to 1e-8 and the in-step re-solves to 1e-6: the
    # reference's defaults of ns_converge_iT1 / ns_converge_iT2, with its caps
    # ns_max_iT1 = ns_max_iT2 = 100 000 (complab_functions.hh:1258-1265)
    p.ns_conv_rule = P.CONV_RELDELTA
    p.ns_conv_eps = 1e-6
    p.ns_max_iter = 100000
    fs = FlowSetup(pe=1.0, char_length=float(nxf), delta_p=1e-5 * (nxf + 1), ns_tau=0.8,
                   rule=P.CONV_RELDELTA, eps=1e-8, max_iter=100000)
    ts = TransportSetup(d_pore=[0.1] * 6, d_biofilm=[0.1] * 6, stabilize_steps=0)
    return Workload("precision-%s %dx%dx%d random phi%.2f carbonate-calcium eq + calcite precip/dissol"
                    % ({0: "fp64", 1: "fp32", 2: "fp32_ade"}[precision], nxf, ny, nz, phi), p, lat, phi, fs, ts, steps)

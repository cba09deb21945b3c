"""ctypes mirror of ``clb_case_params`` (include/clb.h) and the constants.

The field order and types here must match the C struct exactly; a layout
check against ``sizeof(clb_case_params)`` as reported by the library runs at
load time.
"""
import ctypes as C

ABI_VERSION = 2       # CLB_ABI_VERSION of include/clb.h
MAX_SPECIES = 8
MAX_MICROBES = 2
MAX_RXN = 4
MAX_COMP = 8
MAX_PHASES = 4

OK = 0
CLOGGED = 1           # not an error: the run stopped, the pore space no longer percolates
E_ARG, E_HIP, E_NOMEM, E_STATE, E_NONFINITE, E_CA_STALL, E_CA_SWEEPS, E_AGE_SWEEPS, E_GEOMETRY = (
    -1, -2, -3, -4, -5, -6, -7, -8, -9)

SOLID, WALL, PORE, BIOFILM = 0, 1, 2, 3
FP64, FP32, FP32_ADE = 0, 1, 2
BC_DIRICHLET, BC_NEUMANN, BC_CLOSED = 0, 1, 2
CONV_TRACER, CONV_RELDELTA = 0, 1
TUNE_NS_SHAPE = 0     # clb_set_tuning knobs
TUNE_ADE_BB = 1       # the transport bounce-back store: 1 on, 0 off (before the first transport launch)
BATCH_STEPS, BATCH_STABILIZE = 0, 1   # clb_batch_set_tile / clb_batch_tuning modes
BIO_NONE, BIO_DUAL_MONOD = 0, 1
CA_FRACTION, CA_HALF = 0, 1
KIN_BIOTIC, KIN_ABIOTIC = 1, 2
EQ_AAPCF, EQ_NEWTON = 0, 1

(FIELD_CONC, FIELD_BIOMASS, FIELD_BTOTAL, FIELD_MASK, FIELD_VEL, FIELD_AGE, FIELD_PHASE, FIELD_NSRHO, FIELD_RATE,
 FIELD_DG, FIELD_FT) = range(11)
GEOM_SPHERES, GEOM_RANDOM, GEOM_CYLINDERS, GEOM_FRACTURE, GEOM_CHANNEL = range(5)

_d = C.c_double
_i = C.c_int32


class CaseParams(C.Structure):
    _fields_ = [
        ("nx", _i), ("ny", _i), ("nz", _i),
        ("precision", _i),
        ("n_species", _i),
        ("mobile", _i * MAX_SPECIES),
        ("c0", _d * MAX_SPECIES),
        ("omega_pore", _d * MAX_SPECIES),
        ("omega_biofilm", _d * MAX_SPECIES),
        ("left_btype", _i * MAX_SPECIES), ("right_btype", _i * MAX_SPECIES),
        ("left_bc", _d * MAX_SPECIES), ("right_bc", _d * MAX_SPECIES),
        ("dt", _d),
        ("clamp", _i),
        ("bio_law", _i),
        ("donor", _i), ("acceptor", _i),
        ("mu_max", _d * MAX_MICROBES), ("ks", _d * MAX_MICROBES), ("ke", _d * MAX_MICROBES),
        ("kdecay", _d * MAX_MICROBES),
        ("bio_stoich", (_d * MAX_SPECIES) * MAX_MICROBES),
        ("n_rxn", _i),
        ("rxn_r1", _i * MAX_RXN), ("rxn_r2", _i * MAX_RXN),
        ("rxn_k", _d * MAX_RXN),
        ("rxn_stoich", (_d * MAX_SPECIES) * MAX_RXN),
        ("rxn_cap_frac", _d), ("rxn_cap_dt", _d), ("rxn_min_conc", _d),
        ("abiotic_rate_scale", _d),
        ("surface_only", _i),
        ("n_microbes", _i),
        ("b0", _d * MAX_MICROBES),
        ("bmax", _d),
        ("thrd_bfilm_frac", _d),
        ("ca_method", _i),
        ("ca_seed", C.c_uint64),
        ("ade_update_interval", _i),
        ("ns_tau", _d),
        ("delta_p", _d),
        ("perm_ratio", _d),
        ("precip_perm_ratio", _d),
        ("ns_update_interval", _i),
        ("ns_max_iter", _i),
        ("ns_conv_rule", _i),
        ("ns_conv_eps", _d),
        ("flow_on", _i),
        ("eq_on", _i),
        ("eq_solver", _i),
        ("eq_nc", _i),
        ("eq_comp_species", _i * MAX_COMP),
        ("eq_stoich", (_d * MAX_COMP) * MAX_SPECIES),
        ("eq_logk", _d * MAX_SPECIES),
        ("eq_tol", _d),
        ("eq_max_iter", _i),
        ("eq_anderson_depth", _i),
        ("precip_on", _i),
        ("precip_species", _i),
        ("precip_max_rho", _d),
        ("precip_interval", _i),
        ("precip_phase_id", _i),
        ("dissol_on", _i),
        ("dissol_n_phases", _i),
        ("dissol_surface_only", _i),
        ("dissol_interval", _i),
        ("dissol_reopen_frac", _d),
        ("dissol_species", _i * MAX_PHASES),
        ("dissol_material", _i * MAX_PHASES),
        ("dissol_is_precipitate", _i * MAX_PHASES),
        ("dissol_acid", _i * MAX_PHASES),
        ("dissol_full", _d * MAX_PHASES),
        ("dissol_initial", _d * MAX_PHASES),
        ("dissol_k", _d * MAX_PHASES),
        ("dissol_acid_order", _d * MAX_PHASES),
        ("dissol_mineral_order", _d * MAX_PHASES),
        ("dissol_stoich", (_d * MAX_SPECIES) * MAX_PHASES),
        ("thm_on", _i),
        ("thm_microbe", _i),
        ("thm_rt", _d),
        ("thm_dg0", _d),
        ("thm_dgloss", _d),
        ("thm_threshold", _d),
        ("thm_chi", _d),
        ("thm_cmin", _d),
        ("thm_conc_scale", _d),
        ("thm_nu", _d * MAX_SPECIES),
    ]


def default_params(nx, ny, nz, n_species):
    """A parameter block with the reference's defaults
    (complab_functions.hh:1254-1265, part4_eqsolver.hh:204-215)."""
    p = CaseParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.precision = FP64
    p.n_species = n_species
    for s in range(MAX_SPECIES):
        p.mobile[s] = 1
        p.omega_pore[s] = 1.0
        p.omega_biofilm[s] = 1.0
        p.left_btype[s] = BC_NEUMANN
        p.right_btype[s] = BC_NEUMANN
    p.dt = 1.0
    p.clamp = 1
    p.bio_law = BIO_NONE
    p.donor = -1
    p.acceptor = -1
    for r in range(MAX_RXN):
        p.rxn_r1[r] = -1
        p.rxn_r2[r] = -1
    p.abiotic_rate_scale = 1.0
    p.bmax = 999999999.0
    p.thrd_bfilm_frac = 0.0
    p.ca_method = CA_FRACTION
    p.ca_seed = 12345
    p.ade_update_interval = 1
    p.ns_tau = 1.0
    p.perm_ratio = 1.0
    p.precip_perm_ratio = 1.0
    p.ns_update_interval = 1
    p.ns_max_iter = 100000
    p.ns_conv_rule = CONV_TRACER
    p.ns_conv_eps = 1e-6
    p.flow_on = 0
    p.eq_solver = EQ_AAPCF
    for j in range(MAX_COMP):
        p.eq_comp_species[j] = -1
    p.eq_tol = 1e-8
    p.eq_max_iter = 200
    p.eq_anderson_depth = 4
    p.precip_species = -1
    p.precip_interval = 1
    for k in range(MAX_PHASES):
        p.dissol_species[k] = -1
        p.dissol_material[k] = -1
        p.dissol_acid[k] = -1
        p.dissol_acid_order[k] = 1.0
        p.dissol_mineral_order[k] = 1.0
    p.dissol_surface_only = 1
    p.dissol_interval = 1
    p.dissol_reopen_frac = 0.9
    # the .thm defaults (complab3d_thermo.hh:174-197): 298.15 K, R in kJ/mol/K
    p.thm_rt = 8.314462618e-3 * 298.15
    p.thm_chi = 1.0
    p.thm_cmin = 1e-9
    p.thm_conc_scale = 1.0
    p.thm_microbe = -1
    return p


SYM_MAXOPS, SYM_MAXCONST, SYM_MAXVARS, SYM_MAXRATES, SYM_STACK = 256, 64, 16, 16, 16
SYM_ABIOTIC = -1                        # clb_set_symbolic: the abiotic law
(SYM_CONST, SYM_VAR, SYM_NEG, SYM_ADD, SYM_SUB, SYM_MUL, SYM_DIV, SYM_POW, SYM_EXP, SYM_LOG, SYM_LOG10,
 SYM_SQRT, SYM_ABS, SYM_TANH, SYM_MIN, SYM_MAX) = range(1, 17)


class SymProgram(C.Structure):
    """clb_sym_program of include/clb.h."""
    _fields_ = [
        ("n_ops", _i), ("n_vars", _i), ("n_rates", _i), ("n_const", _i),
        ("op", C.c_uint8 * SYM_MAXOPS),
        ("arg", C.c_int16 * SYM_MAXOPS),
        ("rate_end", C.c_int16 * SYM_MAXRATES),
        ("var_species", C.c_int16 * SYM_MAXVARS),
        ("rate_species", C.c_int16 * SYM_MAXRATES),
        ("lo", _d * SYM_MAXVARS), ("hi", _d * SYM_MAXVARS),
        ("konst", _d * SYM_MAXCONST),
        ("unit_scale", _d),
    ]


def copy_params(p):
    q = CaseParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(CaseParams))
    return q


def set_phase(p, k, species, full, material=-1, initial=0.0, precipitate=False, k_rate=0.0, acid=-1,
              acid_order=1.0, mineral_order=1.0, stoich=None):
    """Declares dissolvable phase k (id k + 1, <phase k> of <dissolution>) in
    the parameter block and turns dissolution on; the rate law is
    rate = k_rate max(C[acid],0)^acid_order max(mineral,0)^mineral_order with
    products stoich[s] * rate (defineDissolutionRate of the reference's
    examples: both orders 1).  A precipitate phase is also made the one
    precipitation stamps."""
    if not 0 <= k < MAX_PHASES:
        raise ValueError("phase index %d outside 0..%d" % (k, MAX_PHASES - 1))
    p.dissol_on = 1
    p.dissol_n_phases = max(p.dissol_n_phases, k + 1)
    p.dissol_species[k] = species
    p.dissol_full[k] = full
    p.dissol_material[k] = -1 if precipitate else material
    p.dissol_initial[k] = 0.0 if precipitate else initial
    p.dissol_is_precipitate[k] = 1 if precipitate else 0
    p.dissol_k[k] = k_rate
    p.dissol_acid[k] = acid
    p.dissol_acid_order[k] = acid_order
    p.dissol_mineral_order[k] = mineral_order
    for s in range(MAX_SPECIES):
        p.dissol_stoich[k][s] = 0.0
    for s, v in (stoich or {}).items():
        p.dissol_stoich[k][s] = v
    if precipitate:
        p.precip_phase_id = k + 1

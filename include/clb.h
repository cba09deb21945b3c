/*
 * clb.h -- C-ABI of the MI355X-native coupled reaction step.
 *
 * One library (libclb.so, HIP for gfx950) owns the device memory of every case
 * and runs the coupled step of the CompLaB3D main loop on it:
 *   D3Q19 incompressible BGK flow with bounce-back on the live mask,
 *   one D3Q7 advection-diffusion lattice per mobile species,
 *   per-voxel kinetics with the clamp dC = max(R*dt, -C),
 *   per-voxel equilibrium speciation,
 *   the biofilm cellular-automaton spill,
 *   precipitation / dissolution mask updates.
 *
 * Plain C types only: ints, doubles, pointers, sizes.  Every entry point
 * returns 0 on success and a negative CLB_E_* code on failure; the message of
 * the last failure on a context is clb_last_error(ctx).  Nothing aborts the
 * process.  Calls on one context are not thread-safe; use one context per
 * host thread.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the CompLaB3D-GEMS source tree).
 */
#ifndef CLB_H
#define CLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLB_ABI_VERSION 2

#define CLB_MAX_SPECIES   8   /* substrates (mobile + immobile)              */
#define CLB_MAX_MICROBES  2   /* CA biofilm populations                      */
#define CLB_MAX_RXN       4   /* abiotic mass-action reactions               */
#define CLB_MAX_COMP      8   /* equilibrium components                      */
#define CLB_MAX_PHASES    4   /* declared dissolvable solid phases           */

/* status codes */
#define CLB_OK              0
#define CLB_CLOGGED         1   /* not an error: clb_run_steps stopped after a step whose flow
                                   re-solve ran out of iterations with no outlet flow, or
                                   diverged (percolationFlag, complab.cpp:3113, 3119, 3176,
                                   3261); the case refuses further steps                     */
#define CLB_E_ARG          -1   /* bad argument / inconsistent parameters    */
#define CLB_E_HIP          -2   /* HIP runtime failure                       */
#define CLB_E_NOMEM        -3   /* device allocation failed                  */
#define CLB_E_STATE        -4   /* call not valid in the case's state        */
#define CLB_E_NONFINITE    -5   /* NaN/Inf detected (biomass or flow energy) */
#define CLB_E_CA_STALL     -6   /* CA spill made no progress in 100 sweeps   */
#define CLB_E_CA_SWEEPS    -7   /* CA spill did not settle in 1000 sweeps    */
#define CLB_E_AGE_SWEEPS   -8   /* age relaxation did not settle             */
#define CLB_E_GEOMETRY     -9   /* geometry does not percolate / invalid     */

/* material codes, as in complab3d_geometry.hh:61 and ABC make_geometries.py */
#define CLB_SOLID   0   /* no-dynamics grain interior                         */
#define CLB_WALL    1   /* bounce-back                                        */
#define CLB_PORE    2   /* open pore water                                    */
#define CLB_BIOFILM 3   /* first CA microbe's material number (3 + m)         */

/* precisions */
#define CLB_FP64     0
#define CLB_FP32     1   /* FP32 storage and collide arithmetic, flow included       */
#define CLB_FP32_ADE 2   /* FP32 transport lattices, FP64 flow lattice               */

/* boundary types of the x planes, complab_functions.hh:382-392 */
#define CLB_BC_DIRICHLET 0
#define CLB_BC_NEUMANN   1
#define CLB_BC_CLOSED    2

/* flow convergence rules */
#define CLB_CONV_TRACER   0  /* Palabos ValueTracer(1,1000,eps) on the mean energy */
#define CLB_CONV_RELDELTA 1  /* |E_t - E_{t-1}| / E_t < eps                          */

/* biotic rate laws (defineRxnKinetics, config/kinetics/defineKinetics.biotic.hh:90-126) */
#define CLB_BIO_NONE       0
#define CLB_BIO_DUAL_MONOD 1  /* mu = mu_max S/(Ks+S) * E/(Ke+E); E term off if acceptor<0 */

/* CA rules, complab_functions.hh:1414-1439 */
#define CLB_CA_FRACTION 0   /* pushExcessBiomass3D      */
#define CLB_CA_HALF     1   /* halfPushExcessBiomass3D  */

/* equilibrium solvers */
#define CLB_EQ_AAPCF  0   /* Anderson-accelerated PCF, part4_eqsolver.hh:960-1180 */
#define CLB_EQ_NEWTON 1   /* Newton on log-components (NR_reserve.hh:512)         */

typedef struct clb_case_params {
    /* lattice, including the two ghost x columns (complab_functions.hh:999) */
    int32_t nx, ny, nz;
    int32_t precision;                 /* CLB_FP64 / CLB_FP32 / CLB_FP32_ADE  */

    /* ---- species (substrates) -------------------------------------------- */
    int32_t n_species;
    int32_t mobile[CLB_MAX_SPECIES];   /* 0 = immobile scalar (never streams) */
    double  c0[CLB_MAX_SPECIES];       /* initial concentration               */
    double  omega_pore[CLB_MAX_SPECIES];
    double  omega_biofilm[CLB_MAX_SPECIES];
    int32_t left_btype[CLB_MAX_SPECIES], right_btype[CLB_MAX_SPECIES];
    double  left_bc[CLB_MAX_SPECIES],    right_bc[CLB_MAX_SPECIES];
    double  dt;                        /* ade_dt, seconds per step            */
    int32_t clamp;                     /* 1: dC = max(R dt, -C) on every path */

    /* ---- biotic kinetics ------------------------------------------------- */
    int32_t bio_law;                   /* CLB_BIO_*                           */
    int32_t donor, acceptor;           /* species indices (-1 = unused)       */
    double  mu_max[CLB_MAX_MICROBES], ks[CLB_MAX_MICROBES], ke[CLB_MAX_MICROBES];
    double  kdecay[CLB_MAX_MICROBES];
    double  bio_stoich[CLB_MAX_MICROBES][CLB_MAX_SPECIES]; /* dC_s = nu * growth */

    /* ---- abiotic mass-action kinetics (defineAbioticRxnKinetics) ---------- */
    int32_t n_rxn;
    int32_t rxn_r1[CLB_MAX_RXN], rxn_r2[CLB_MAX_RXN];      /* reactants (-1 none) */
    double  rxn_k[CLB_MAX_RXN];                            /* 1/s or L/(mol s)    */
    double  rxn_stoich[CLB_MAX_RXN][CLB_MAX_SPECIES];
    double  rxn_cap_frac, rxn_cap_dt, rxn_min_conc;        /* ABC per-step cap    */
    double  abiotic_rate_scale;                            /* g_abioticRateScale  */
    int32_t surface_only;                                  /* precipitationVOP.hh:53 */

    /* ---- CA biofilm ------------------------------------------------------ */
    int32_t n_microbes;                /* CA biofilm populations              */
    double  b0[CLB_MAX_MICROBES];      /* density written on seed voxels      */
    double  bmax;                      /* maximum_biomass_density             */
    double  thrd_bfilm_frac;           /* thrd_biofilm_fraction               */
    int32_t ca_method;                 /* CLB_CA_*                            */
    uint64_t ca_seed;                  /* counter-based RNG for the spill     */
    int32_t ade_update_interval;       /* solute omega switch interval        */

    /* ---- flow ------------------------------------------------------------ */
    double  ns_tau;                    /* tau of the D3Q19 lattice            */
    double  delta_p;                   /* pressure drop (lattice)             */
    double  perm_ratio;                /* 1/viscosity_ratio in biofilm; <=1e-14 -> BB */
    double  precip_perm_ratio;
    int32_t ns_update_interval;        /* re-solve interval (steps)           */
    int32_t ns_max_iter;               /* ns_max_iT2 cap per re-solve         */
    int32_t ns_conv_rule;              /* CLB_CONV_*                          */
    double  ns_conv_eps;
    int32_t flow_on;                   /* 0: Pe = 0, velocity stays zero      */

    /* ---- equilibrium ----------------------------------------------------- */
    int32_t eq_on;
    int32_t eq_solver;                 /* CLB_EQ_*                            */
    int32_t eq_nc;                     /* components                          */
    int32_t eq_comp_species[CLB_MAX_COMP];  /* species holding each component's guess (-1 none) */
    double  eq_stoich[CLB_MAX_SPECIES][CLB_MAX_COMP];
    double  eq_logk[CLB_MAX_SPECIES];
    double  eq_tol;
    int32_t eq_max_iter;
    int32_t eq_anderson_depth;

    /* ---- precipitation (precipitationVOP.hh) ---------------------------- */
    int32_t precip_on;
    int32_t precip_species;            /* immobile mineral species            */
    double  precip_max_rho;
    int32_t precip_interval;
    int32_t precip_phase_id;

    /* ---- dissolution (dissolutionVOP.hh): declared solid phases ----------
     * Phase k (<phase k>) has id k + 1 (complab.cpp:516-531).  A grain phase
     * is seeded on every voxel of its material at creation, the first phase
     * naming a material taking its voxels (initPhaseLattice3D,
     * complab.cpp:1672-1698); a precipitate phase starts nowhere and
     * precipitation stamps precip_phase_id, which must name it.  Rate law of
     * phase k (defineDissolutionRate for phase id k + 1):
     *   rate = k * max(C[acid],0)^acid_order * max(mineral,0)^mineral_order,
     *   mineralR = -rate, subsR[s] = stoich[s] * rate
     * (acid < 0: factor 1; the examples' law is both orders 1). */
    int32_t dissol_on;
    int32_t dissol_n_phases;           /* 1..CLB_MAX_PHASES                                      */
    int32_t dissol_surface_only;       /* <surface_only>: enclosed voxels do not react          */
    int32_t dissol_interval;           /* <update_interval> of the reopen check                 */
    double  dissol_reopen_frac;        /* <reopen_fraction>                                     */
    int32_t dissol_species[CLB_MAX_PHASES];        /* <substrate>: immobile inventory species */
    int32_t dissol_material[CLB_MAX_PHASES];       /* grain phase: CLB_SOLID or CLB_WALL; -1 none */
    int32_t dissol_is_precipitate[CLB_MAX_PHASES]; /* <is_precipitate>                          */
    int32_t dissol_acid[CLB_MAX_PHASES];           /* rate law: species C[acid]; -1 none        */
    double  dissol_full[CLB_MAX_PHASES];           /* <full_density>                            */
    double  dissol_initial[CLB_MAX_PHASES];        /* <initial_fill> of a grain phase           */
    double  dissol_k[CLB_MAX_PHASES];
    double  dissol_acid_order[CLB_MAX_PHASES];
    double  dissol_mineral_order[CLB_MAX_PHASES];
    double  dissol_stoich[CLB_MAX_PHASES][CLB_MAX_SPECIES];
    /* thermodynamic gate of the biotic rates, one .thm reaction block
     * (complab3d_thermo.hh:215-255, applied at part1:127-140): the voxel's
     * whole biotic rate vector times F_T = max(0, 1 - exp(f / (chi RT))),
     * f = dG0 + RT ln Q + dGloss + atp dGatp,
     * ln Q = sum_s nu_s ln max(conc_scale C_s, cmin); energies in kJ/mol */
    int32_t thm_on;
    int32_t thm_microbe;               /* the organism gated on the symbolic path; -1: every one */
    double  thm_rt;                    /* gas constant x temperature                          */
    double  thm_dg0;
    double  thm_dgloss;                /* >= 0                                                 */
    double  thm_threshold;             /* atp x dGatp                                          */
    double  thm_chi;                   /* > 0                                                  */
    double  thm_cmin;                  /* activity floor, > 0                                  */
    double  thm_conc_scale;            /* lattice concentration -> mol/L                       */
    double  thm_nu[CLB_MAX_SPECIES];   /* stoichiometry: negative consumed, positive produced */
} clb_case_params;

typedef struct clb_ctx clb_ctx;

/* ---- context ------------------------------------------------------------- */
/* plbInit (complab.cpp:261) */
int  clb_ctx_create(int device, clb_ctx **out);
int  clb_ctx_destroy(clb_ctx *ctx);
const char *clb_last_error(const clb_ctx *ctx);
int  clb_abi_version(void);
/* fills dst with a short text naming the device; for reports */
int  clb_device_name(clb_ctx *ctx, char *dst, size_t n);
int  clb_sync(clb_ctx *ctx);

/* ---- cases --------------------------------------------------------------- */
/* the lattices built in complab.cpp:1380-1711 (soluteDomainSetup,
 * bmassDomainSetup, NSdomainSetup, CopyGeometryScalar2maskLattice3D).
 * geometry: nx*ny*nz material codes, x fastest (ghost columns included).
 * biomass0: n_microbes*nx*ny*nz initial biofilm densities or NULL. */
int  clb_case_create(clb_ctx *ctx, const clb_case_params *p, const int8_t *geometry,
                     const double *biomass0, int *case_id);
int  clb_case_destroy(clb_ctx *ctx, int case_id);
/* device bytes held by a case built from these parameters (admission control) */
size_t clb_case_footprint(const clb_case_params *p);
/* the parameters a case runs with (after clb_set_ade_relaxation); the values
 * the reference's processors receive as constructor arguments */
int  clb_case_get_params(clb_ctx *ctx, int case_id, clb_case_params *out);

/* ---- flow ---------------------------------------------------------------- */
/* NSdomainSetup(...) + initializeAtEquilibrium(PressureGradient)
 * (complab_functions.hh:82-94, 296-345) with the current delta_p */
int  clb_ns_init(clb_ctx *ctx, int case_id, double delta_p);
/* collideAndStream loop with a ValueTracer (complab.cpp:1111-1124); the
 * tracer keeps its window across calls unless reset != 0 (ns_convg2 at
 * complab.cpp:3090-3093 is never reset).  *iters = iterations run,
 * *converged = 1 if the rule was met before max_iter. */
int  clb_ns_solve(clb_ctx *ctx, int case_id, int max_iter, int rule, double eps,
                  int reset_tracer, int *iters, int *converged);
/* latticeToPassiveAdvDiff (complab.cpp:2012-2032): velocity of the current
 * flow state copied into the field every ADE lattice reads */
int  clb_ns_couple(clb_ctx *ctx, int case_id);
/* reductions of complab.cpp:1136-1157: box-mean |u| over x=1..nx-2, max u_x,
 * outlet-plane mean u_x at x=nx-2, and the pore+biofilm mean |u| */
int  clb_ns_stats(clb_ctx *ctx, int case_id, double out[4]);
/* updateNsLatticesDynamics3D (part2:748-818) from the live mask */
int  clb_ns_update_dynamics(clb_ctx *ctx, int case_id, double perm_ratio);

/* ---- advection-diffusion + reactions -------------------------------------- */
/* the relaxation rates and time step of the ADE lattices after the Peclet
 * mapping of complab.cpp:1220-1283 (omega per species in pore / biofilm) */
int  clb_set_ade_relaxation(clb_ctx *ctx, int case_id, const double *omega_pore,
                            const double *omega_biofilm, double dt);
/* the 10 000-step stabilization pre-loop and stabilizeADElattice3D reset of
 * pore/biofilm voxels to c0 (complab.cpp:2033-2041; part3:261-310) */
int  clb_ade_stabilize(clb_ctx *ctx, int case_id, int steps);
/* nsteps coupled reaction steps of the main loop (complab.cpp:2178-3262):
 * collide, biotic kinetics + apply, abiotic kinetics (+ dissolution) + apply,
 * equilibrium, CA spill, mask-driven dynamics updates and flow re-solves,
 * precipitation / dissolution node conversion, stream.  *done = steps run
 * (fewer when a stop condition is met; the status says which). */
int  clb_run_steps(clb_ctx *ctx, int case_id, int nsteps, int *done);
/* only the fused transport + kinetics kernel, nsteps times (no CA, no
 * equilibrium, no conversions); the timed unit of the roofline */
int  clb_ade_steps(clb_ctx *ctx, int case_id, int nsteps);
/* run_equilibrium_biotic (part4:1359-1560) once over the reaction box */
int  clb_equilibrium(clb_ctx *ctx, int case_id);
/* the CA block of complab.cpp:2793-3055 once; *sweeps = push/pull sweeps */
int  clb_ca_spill(clb_ctx *ctx, int case_id, int *sweeps);

/* ---- the processors one at a time -------------------------------------------
 * For callers that keep the reference's loop and apply its processors one by
 * one (include/clb_processors.hh).  Between calls a case holds post-collision,
 * pre-stream populations: the state the reference's processors see between
 * collide() and stream().  A transport launch streams that state and collides. */
/* ADE collide + stream without kinetics, nsteps times (complab.cpp:2307-2315,
 * 3251-3259) */
int  clb_ade_collide_stream(clb_ctx *ctx, int case_id, int nsteps);
#define CLB_KIN_BIOTIC  1   /* run_kinetics + update_rxnLattices (part1:85-187, 406-497)            */
#define CLB_KIN_ABIOTIC 2   /* run_abiotic_kinetics + update_abiotic_rxnLattices (part1:199-404);
                               with dissolution also surfaceDissolutionKinetics3D and
                               dissolutionGather3D (dissolutionVOP.hh:378-720), one increment */
/* rates, clamp and D3Q7 deposit on the current post-collision state; biotic
 * first, then abiotic, as the main loop applies them */
int  clb_kinetics(clb_ctx *ctx, int case_id, int which);
/* updateLocalMaskNtotalLattices3D (part2:71-259) + the max of the total biomass
 * and the mask sum the main loop reduces (complab.cpp:1983, 2793-2800) */
int  clb_mask_total(clb_ctx *ctx, int case_id, double *max_total, int64_t *mask_sum);
/* one sweep of pushExcessBiomass3D / halfPushExcessBiomass3D (part1:500-970),
 * the rule chosen by ca_method */
int  clb_ca_sweep(clb_ctx *ctx, int case_id);
/* updateAgeDistance3D (part2:840-965): one sweep, or sweeps to the fixed point */
int  clb_age_update(clb_ctx *ctx, int case_id, int to_fixed_point);
/* updateSoluteDynamics3D (part2:485-553): pore / biofilm relaxation per voxel */
int  clb_update_solute_dynamics(clb_ctx *ctx, int case_id);
/* updateSoluteSolidDynamics3D (part2:585-647): bounce-back on solid voxels,
 * reopened voxels rebuilt at rest; *reopened = voxels reopened */
int  clb_update_solute_solid(clb_ctx *ctx, int case_id, int *reopened);
/* precipNodeConversion3D (precipitationVOP.hh:166-227) over the reaction box;
 * *changed = 1 if any voxel turned solid */
int  clb_precip_convert(clb_ctx *ctx, int case_id, int *changed);
/* dissolNodeConversion3D (dissolutionVOP.hh:744-812) over the reaction box:
 * phase voxels below reopen_fraction x full_density turn pore; *reopened =
 * voxels converted (follow with clb_ns_update_dynamics / clb_ns_solve and
 * clb_update_solute_solid, as complab.cpp:3209-3246 does) */
int  clb_dissol_convert(clb_ctx *ctx, int case_id, int *reopened);

/* ---- fields ---------------------------------------------------------------- */
#define CLB_FIELD_CONC     0  /* species concentration (density after stream) */
#define CLB_FIELD_BIOMASS  1  /* CA biomass of microbe `index`                 */
#define CLB_FIELD_BTOTAL   2  /* total biofilm biomass                          */
#define CLB_FIELD_MASK     3  /* live material code (as double)                 */
#define CLB_FIELD_VEL      4  /* velocity component `index`                     */
#define CLB_FIELD_AGE      5  /* CA age                                         */
#define CLB_FIELD_PHASE    6  /* solid phase id                                 */
#define CLB_FIELD_NSRHO    7  /* flow density (1 at closed voxels)              */
#define CLB_FIELD_RATE     8  /* species `index`'s rate of the last captured step (clb_rate_capture) */
#define CLB_FIELD_DG       9  /* the gate's dG (kJ/mol) of gated microbe `index` at the standing
                                 concentrations, 0 off the open interior (dG_<microbe>_*.vti) */
#define CLB_FIELD_FT      10  /* the gate's F_T of gated microbe `index`, likewise (FT_<microbe>_*.vti) */
/* copies nx*ny*nz doubles (x fastest) of one field into dst */
int  clb_read_field(clb_ctx *ctx, int case_id, int field, int index, double *dst);
/* overwrites a species concentration field with equilibrium populations at
 * rest, a biomass field (the total follows), a velocity component, or the CA
 * age field (rounded to integers); for tests and restarts */
int  clb_write_field(clb_ctx *ctx, int case_id, int field, int index, const double *src);

/* ---- reductions (PopulationSum3D, MaskedScalarCounts3D, computeSum) --------- */
/* per species, over the reaction box x = 1..nx-2: a mobile species summed
 * over the voxels open on the live mask (mask not solid/wall), an immobile one
 * over every voxel (a mineral keeps its inventory in the voxels it sealed,
 * complab3d_diagnostics.hh:255-261); out has n_species entries */
int  clb_species_totals(clb_ctx *ctx, int case_id, double *out);
/* clb_species_totals, allocating nothing: the block partials and sums live in
 * a static device array of CLB_NOALLOC_SCRATCH_BYTES that the library holds on
 * every GPU it runs on, not in the case's buffers.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  CLB_E_ARG when the case's reductions
 * need more than the array holds (an immobile species needs 32 bytes per 256
 * voxels of the lattice: about 41.9 M voxels at most).  Calls on one GPU share
 * the array: calls that overlap on one GPU (several host threads or streams)
 * race; nothing in it needs clearing between calls. */
#define CLB_NOALLOC_SCRATCH_BYTES (5u << 20)
int  clb_species_totals_noalloc(clb_ctx *ctx, int case_id, double *out);
/* diagnostics row (complab.cpp:2736-2785, complab3d_diagnostics.hh:57-69): for
 * each species, then each microbe, four values -- total, minimum, maximum and
 * held.  total is clb_species_totals' (a microbe: its biomass over every voxel
 * of the reaction box); minimum and maximum are taken over the voxels the total
 * counts.  held is the mass the total does not see, so that total + held is
 * what the lattice holds (PopulationSum3D over the whole box, part3:505-536,
 * minus the total).  A mobile species: the content of the x = 0 and x = nx-1
 * voxels open to it plus the populations in flight into closed voxels; the rest
 * content that closed voxels keep among themselves from set-up is not part of
 * it.  An immobile species: 0 (nothing streams).  A microbe: its biomass in the
 * x = 0 and x = nx-1 planes.  out has 4 * (n_species + n_microbes) entries. */
int  clb_field_stats(clb_ctx *ctx, int case_id, double *out);
/* counters: [0] open voxels, [1] biofilm voxels, [2] clamp binds (debits the
 * clamp capped that exceeded the availability by more than 1e-12 of it),
 * [3] equilibrium non-converged voxels, [4] flow re-solves, [5] mask changes,
 * [6] CA sweeps, [7] steps run, [8] dissolution product shares parked,
 * [9] shares collected (equal to [8]: the deposit invariant of
 * dissolutionVOP.hh:243-265), [10] collected shares the positivity rule
 * clamped, [11] voxels reopened by dissolution */
#define CLB_NCOUNTERS 12
int  clb_counters(clb_ctx *ctx, int case_id, int64_t out[CLB_NCOUNTERS]);

/* ---- symbolic rate laws (complab3d_symbolic.hh; run_symbolic3D,
 * complab3d_processors_symbolic.hh:64-200) ------------------------------------
 * One organism's .sym law compiled to postfix code over a value vector that
 * holds the declared variables (slots 0..n_vars-1) and then the rate outputs
 * already computed (slot n_vars + r).  Rate r's code is op[rate_end[r-1] ..
 * rate_end[r]); its value, NaN -> 0 and held within +-1e300, is stored
 * unscaled in its slot and leaves the law times unit_scale.  Division by
 * |q| < 1e-300, a negative base to a fractional power, and the logarithms and
 * square root of a non-positive argument give 0.  Variables with hi > lo are
 * held within [lo, hi].  A variable binds a species' concentration or, with
 * var_species -1, the organism's own biomass; a rate with rate_species s is
 * species s's rate, with -1 the specific growth rate (times the biomass).
 * The abiotic law (microbe CLB_SYM_ABIOTIC, run_symbolic_abiotic3D,
 * complab3d_processors_abiotic_learned.hh:49-150) binds species only; its
 * rate_species -1 is the extent line, which feeds the species lines derived
 * from it and moves nothing itself. */
#define CLB_SYM_MAXOPS    256
#define CLB_SYM_MAXCONST  64
#define CLB_SYM_MAXVARS   16
#define CLB_SYM_MAXRATES  16
#define CLB_SYM_STACK     16
#define CLB_SYM_ABIOTIC   (-1)                  /* clb_set_symbolic: the abiotic law */
enum {
    CLB_SYM_CONST = 1, CLB_SYM_VAR, CLB_SYM_NEG, CLB_SYM_ADD, CLB_SYM_SUB, CLB_SYM_MUL, CLB_SYM_DIV,
    CLB_SYM_POW, CLB_SYM_EXP, CLB_SYM_LOG, CLB_SYM_LOG10, CLB_SYM_SQRT, CLB_SYM_ABS, CLB_SYM_TANH,
    CLB_SYM_MIN, CLB_SYM_MAX
};
typedef struct clb_sym_program {
    int32_t n_ops, n_vars, n_rates, n_const;
    uint8_t op[CLB_SYM_MAXOPS];
    int16_t arg[CLB_SYM_MAXOPS];                /* CONST: constant index; VAR: value slot */
    int16_t rate_end[CLB_SYM_MAXRATES];
    int16_t var_species[CLB_SYM_MAXVARS];
    int16_t rate_species[CLB_SYM_MAXRATES];
    double  lo[CLB_SYM_MAXVARS], hi[CLB_SYM_MAXVARS];
    double  konst[CLB_SYM_MAXCONST];
    double  unit_scale;                         /* 1 per second, 1/3600 per hour */
} clb_sym_program;
/* The biotic rates of microbe `microbe` come from `prog` instead of the
 * compiled law (set before the first step; the compiled law keeps the
 * microbes without one, and both add into the step's increments as
 * run_kinetics and run_symbolic3D do, complab.cpp:2348-2393);
 * the program is checked (op codes, slots, stack depth) and CLB_E_ARG
 * returned when it does not run; NULL switches the microbe back.  With
 * microbe = CLB_SYM_ABIOTIC, `prog` is the abiotic law: it runs at every open
 * voxel of the reaction box after the biotic update, with its own shared
 * budget, and its increments join the mass-action ones (complab.cpp:2455-2493;
 * with or without n_rxn > 0).  Symbolic laws are refused with dissolution. */
int  clb_set_symbolic(clb_ctx *ctx, int case_id, int microbe, const clb_sym_program *prog);

/* ---- checkpoint / restart (saveBinaryBlock / loadBinaryBlock of every lattice,
 * complab.cpp:2293-2302 and 3274-3310, read back by <IO><read_ADE_file>) -------
 * One self-describing block of this library's own format holding everything a
 * bit-identical continuation needs: the persistent device fields (masks,
 * phases, both population buffers of every lattice, immobile species, biomass,
 * ages, velocities, flow state, classes), the step state and the parameter
 * block (relaxation and dt included).  A checkpoint loads into a case created
 * with the same parameters and geometry (CLB_E_ARG otherwise); the library's
 * precision and ABI version must match. */
int  clb_checkpoint_size(clb_ctx *ctx, int case_id, size_t *bytes);
int  clb_checkpoint_save(clb_ctx *ctx, int case_id, void *dst, size_t bytes);
int  clb_checkpoint_load(clb_ctx *ctx, int case_id, const void *src, size_t bytes);

/* ---- rate snapshots (rateField + accumulateRates, complab.cpp:2078-2105,
 * written as rate_<species>_<iT>.vti at 2548-2561) -----------------------------
 * With capture on, every following step records each species' requested
 * reaction increment of the step -- the biotic and abiotic increments as the
 * rate laws asked for them, before the clamp, equilibrium not counted -- over
 * dt, readable as CLB_FIELD_RATE (0 where nothing reacted).  The steps run
 * the split kinetics kernel while capture is on.  Refused with dissolution. */
int  clb_rate_capture(clb_ctx *ctx, int case_id, int on);

/* ---- profiling ------------------------------------------------------------- */
/* average device time (ms) of the fused transport kernel of the interior
 * voxels (stream, collide, kinetics; the x planes run after it in a kernel of
 * their own) over the timed launches since the last reset, measured with HIP
 * events on the case's stream */
int  clb_kernel_time(clb_ctx *ctx, int case_id, int reset, double *avg_ms, int64_t *launches);
/* the same for the flow kernel (one D3Q19 iteration, k_ns_step, its
 * convergence reduction left out) over the timed iterations of every flow
 * solve since the last reset that did work (iterations queued after the stop
 * return at once and are not counted); timed at the transport stride.  No
 * counterpart in the reference. */
int  clb_flow_kernel_time(clb_ctx *ctx, int case_id, int reset, double *avg_ms, int64_t *launches);
/* time every stride-th transport launch only (default 1: every launch); each
 * timed launch puts two event markers on the stream, which the device waits
 * on.  No counterpart in the reference. */
int  clb_set_timing_stride(clb_ctx *ctx, int case_id, int stride);

/* launch shapes of the hot kernels, the knobs of the autotuner
 * (complab_amd.tune); results do not depend on them.  No counterpart in the
 * reference.  CLB_TUNE_NS_SHAPE: the flow interior launch, 0 one workgroup per
 * 256 voxels, n > 0 n workgroups each looping over a contiguous run of them and
 * loading the next one's neighbour records ahead.  CLB_TUNE_ADE_BB: 1 the
 * transport pull reads the populations returning from closed neighbours from
 * a dense store beside each population buffer, 0 from the voxel's own slots;
 * chosen at creation (the store where closed links average at most one per
 * interior voxel and no equilibrium runs), settable before the first
 * transport launch only, and 1 is refused with equilibrium. */
#define CLB_TUNE_NS_SHAPE 0
#define CLB_TUNE_ADE_BB   1
int  clb_set_tuning(clb_ctx *ctx, int case_id, int knob, int value);

/* ---- the batched form ------------------------------------------------------
 * No counterpart in the reference, which runs one case per process
 * (complab.cpp main); a parameter sweep there is many processes
 * (ABC_sweep_package/README_1_SCENARIO.md:552-557).  A batch steps many cases
 * of one precision together: each kernel of a flow iteration or of a transport
 * step covers every member (a table of member descriptors and a map of the
 * launch's workgroups to member workgroups), runs of steps and of flow
 * iterations replay HIP graphs, and the transport workgroup size is picked by
 * timing the candidates (64, 128, 256) on the batch's first steps.  Every
 * member ends exactly where its own calls (clb_ns_solve, clb_ade_stabilize,
 * clb_run_steps) would take it: the same voxel arithmetic, the same block
 * partials summed in the same order, the same step decisions.  Members the
 * group launches do not cover (equilibrium, dissolution, symbolic laws, rate
 * capture, no mobile species) run on their own path inside the same call.
 * Per-member outputs are in member order (creation order, then additions).
 * A member's own failure (clog, non-finite state, ...) is its status entry,
 * not the call's return code.  A case in a batch cannot be destroyed. */
#define CLB_BATCH_STEPS     0
#define CLB_BATCH_STABILIZE 1
int  clb_batch_create(clb_ctx *ctx, int precision, const int *case_ids, int n, int *batch_id);
int  clb_batch_destroy(clb_ctx *ctx, int batch_id);
/* continuous batching: members join and leave between calls */
int  clb_batch_add(clb_ctx *ctx, int batch_id, int case_id);
int  clb_batch_remove(clb_ctx *ctx, int batch_id, int case_id);
int  clb_batch_members(clb_ctx *ctx, int batch_id, int *n, int *case_ids, int cap);
/* clb_ns_solve of every member; iters / converged / status per member */
int  clb_batch_ns_solve(clb_ctx *ctx, int batch_id, int max_iter, int rule, double eps, int reset_tracer,
                        int *iters, int *converged, int *status);
/* clb_ade_stabilize of every member */
int  clb_batch_ade_stabilize(clb_ctx *ctx, int batch_id, int steps, int *status);
/* clb_run_steps of every member: done / status per member */
int  clb_batch_run_steps(clb_ctx *ctx, int batch_id, int nsteps, int *done, int *status);
/* the transport workgroup size of a mode: 64, 128, 256, or 0 to tune again */
int  clb_batch_set_tile(clb_ctx *ctx, int batch_id, int mode, int tile);
/* the pick (0: not tuned yet) and each candidate's measured microseconds per step */
int  clb_batch_tuning(clb_ctx *ctx, int batch_id, int mode, int *tile, double us_per_step[3]);

/* device copy bandwidth of a plain vectorised copy kernel over `bytes` bytes,
 * best of `reps` launches (HIP events); the measured ceiling the roofline
 * fractions are read against and the calibration of the HBM byte counters.
 * No counterpart in the reference. */
int  clb_stream_copy(clb_ctx *ctx, size_t bytes, int reps, double *gbps, double *ms);

/* ---- geometry (complab3d_geometry.hh) -------------------------------------- */
/* Rng of complab3d_geometry.hh:77-86: next `n` raw 64-bit draws from seed */
int  clb_rng_draws(uint64_t seed, int n, uint64_t *out);
/* complab_geom::generate "spheres" with `count` spheres (x fastest, no ghosts) */
int  clb_geom_spheres(int nx, int ny, int nz, double radius, int count, uint64_t seed,
                      int walls_yz, int8_t *out);
/* seeded medium (spheres / cylinders inserted until porosity <= target, or a
 * random field / fracture of that porosity), grain shell -> wall, y/z walls,
 * isolated pockets -> wall, percolation checked; writes the lattice
 * with ghost columns (nx_file+2) * ny * nz.  *porosity = open fraction of the
 * file box.  Returns CLB_E_GEOMETRY when the pore space does not span x. */
int  clb_geom_pack(int nx_file, int ny, int nz, double radius, double porosity,
                   uint64_t seed, int kind, int8_t *out_lattice, double *porosity_out);
#define CLB_GEOM_SPHERES   0
#define CLB_GEOM_RANDOM    1  /* smoothed random field thresholded at the quantile */
#define CLB_GEOM_CYLINDERS 2
#define CLB_GEOM_FRACTURE  3  /* rough aperture around y = ny/2 */
#define CLB_GEOM_CHANNEL   4  /* nothing carved */
/* complab_geom::generate + addWalls (complab3d_geometry.hh:92-202): the
 * reference's seeded media, x fastest, no ghost columns; walls: bit 1 x, 2 y,
 * 4 z */
int clb_geom_generate(int nx, int ny, int nz, int kind, double radius, int count, double porosity,
                      double aperture, double roughness, uint64_t seed, int walls, int8_t *out);

#ifdef __cplusplus
}
#endif
#endif /* CLB_H */

/*
 * clb_processors.hh -- host facade of the reference's data-processor API over
 * the C-ABI of clb.h (header only, C++11).
 *
 * CompLaB3D applies its hot-path physics as Palabos data processors:
 *     applyProcessingFunctional(new run_kinetics<T,Desc>(nx, S, M, dt, Kc, mu, solid, bb),
 *                               reactionBox, lattices);          (complab.cpp:2348-2351)
 * The classes below carry the same names and constructor signatures
 * (complab3d_processors_part1.hh, _part2.hh, _part3.hh, _part4_eqsolver.hh,
 * precipitationVOP.hh), and their process() forwards to the clb_* call that
 * runs the same operation on the device-resident fields of a case.  A caller
 * keeps its loop and swaps the lattice vector for a Target (context + case).
 *
 * The physical parameters live in the case (clb_case_create); a processor's
 * constructor arguments are checked against them in process(), which returns
 * CLB_E_ARG on a mismatch instead of running with different numbers.  Box
 * arguments are implicit: reactions run over the reference's reactionBox
 * (x = 1..nx-2, complab.cpp:2337), the CA and conversions over the whole case.
 *
 * The rate evaluation and its application are one device pass: run_kinetics
 * does both and update_rxnLattices is a no-op kept for call-site
 * compatibility (likewise the abiotic and equilibrium pairs).
 */
#ifndef CLB_PROCESSORS_HH
#define CLB_PROCESSORS_HH

#include <cmath>
#include <cstdint>
#include <vector>

#include "clb.h"

namespace clb_proc {

typedef long plint;

/* what the reference passes as the vector of lattices */
struct Target {
    clb_ctx *ctx;
    int case_id;
};

namespace detail {
inline int params(const Target &t, clb_case_params &p) { return clb_case_get_params(t.ctx, t.case_id, &p); }
inline bool same(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fmax(std::fabs(a), std::fabs(b)); }
inline int mismatch(const Target &) { return CLB_E_ARG; }
}  // namespace detail

/* complab3d_processors_part1.hh:85-187 */
template <typename T>
class run_kinetics {
public:
    run_kinetics(plint nx_, plint subsNum_, plint bioNum_, T dt_, std::vector<std::vector<T> > vec2_Kc_kns_,
                 std::vector<T> vec1_mu_kns_, plint solid_, plint bb_)
        : nx(nx_), subsNum(subsNum_), bioNum(bioNum_), dt(dt_), Kc(vec2_Kc_kns_), mu(vec1_mu_kns_), solid(solid_), bb(bb_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || subsNum != p.n_species || bioNum > CLB_MAX_MICROBES || !detail::same(dt, p.dt) ||
            solid != CLB_SOLID || bb != CLB_WALL)
            return detail::mismatch(t);
        for (size_t m = 0; m < mu.size() && m < (size_t) CLB_MAX_MICROBES; ++m)
            if (!detail::same(mu[m], p.mu_max[m])) return detail::mismatch(t);
        // the half-saturation constants, Kc[microbe][substrate], of the donor and the acceptor
        for (size_t m = 0; m < Kc.size() && m < (size_t) CLB_MAX_MICROBES; ++m) {
            if (p.donor >= 0 && (size_t) p.donor < Kc[m].size() && !detail::same(Kc[m][(size_t) p.donor], p.ks[m]))
                return detail::mismatch(t);
            if (p.acceptor >= 0 && (size_t) p.acceptor < Kc[m].size() && !detail::same(Kc[m][(size_t) p.acceptor], p.ke[m]))
                return detail::mismatch(t);
        }
        return clb_kinetics(t.ctx, t.case_id, CLB_KIN_BIOTIC);
    }
private:
    plint nx, subsNum, bioNum;
    T dt;
    std::vector<std::vector<T> > Kc;
    std::vector<T> mu;
    plint solid, bb;
};

/* complab3d_processors_part1.hh:406-497 (applied by run_kinetics) */
template <typename T>
class update_rxnLattices {
public:
    update_rxnLattices(plint nx_, plint subsNum_, plint bioNum_, plint solid_, plint bb_)
        : nx(nx_), subsNum(subsNum_), bioNum(bioNum_), solid(solid_), bb(bb_) {}
    int process(const Target &) { return CLB_OK; }
private:
    plint nx, subsNum, bioNum, solid, bb;
};

/* complab3d_processors_part1.hh:199-297 */
template <typename T>
class run_abiotic_kinetics {
public:
    run_abiotic_kinetics(plint nx_, plint subsNum_, T dt_, plint solid_, plint bb_)
        : nx(nx_), subsNum(subsNum_), dt(dt_), solid(solid_), bb(bb_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || subsNum != p.n_species || !detail::same(dt, p.dt)) return detail::mismatch(t);
        return clb_kinetics(t.ctx, t.case_id, CLB_KIN_ABIOTIC);
    }
private:
    plint nx, subsNum;
    T dt;
    plint solid, bb;
};

/* complab3d_processors_part1.hh:306-404 (applied by run_abiotic_kinetics) */
template <typename T>
class update_abiotic_rxnLattices {
public:
    update_abiotic_rxnLattices(plint nx_, plint subsNum_, plint solid_, plint bb_,
                               const std::vector<bool> &immobile_ = std::vector<bool>())
        : nx(nx_), subsNum(subsNum_), solid(solid_), bb(bb_), immobile(immobile_) {}
    int process(const Target &) { return CLB_OK; }
private:
    plint nx, subsNum, solid, bb;
    std::vector<bool> immobile;
};

/* precipitationVOP.hh:53-153: the abiotic kinetics gated to surface voxels
 * (the gate is the case's surface_only) */
template <typename T>
class surfaceAbioticKinetics3D {
public:
    surfaceAbioticKinetics3D(plint nx_, plint ny_, plint nz_, plint subsNum_, T dt_, plint solid_, plint bb_, int surfaceOnly_)
        : nx(nx_), ny(ny_), nz(nz_), subsNum(subsNum_), dt(dt_), solid(solid_), bb(bb_), surfaceOnly(surfaceOnly_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || ny != p.ny || nz != p.nz || subsNum != p.n_species || !detail::same(dt, p.dt) ||
            (surfaceOnly != 0) != (p.surface_only != 0))
            return detail::mismatch(t);
        return clb_kinetics(t.ctx, t.case_id, CLB_KIN_ABIOTIC);
    }
private:
    plint nx, ny, nz, subsNum;
    T dt;
    plint solid, bb;
    int surfaceOnly;
};

/* complab3d_processors_part4_eqsolver.hh:1359-1560; eqChem's tableau is the
 * case's eq_* parameters */
template <typename T, typename EqChem>
class run_equilibrium_biotic {
public:
    run_equilibrium_biotic(plint nx_, plint subsNum_, EqChem &eqChem_, plint solid_, plint bb_)
        : nx(nx_), subsNum(subsNum_), eqChem(&eqChem_), solid(solid_), bb(bb_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || subsNum != p.n_species || !p.eq_on) return detail::mismatch(t);
        return clb_equilibrium(t.ctx, t.case_id);
    }
private:
    plint nx, subsNum;
    EqChem *eqChem;
    plint solid, bb;
};

/* complab3d_processors_part4_eqsolver.hh:1695-1790 (applied by run_equilibrium_biotic) */
template <typename T>
class update_equilibriumLattices {
public:
    update_equilibriumLattices(plint nx_, plint subsNum_, plint solid_, plint bb_)
        : nx(nx_), subsNum(subsNum_), solid(solid_), bb(bb_) {}
    int process(const Target &) { return CLB_OK; }
private:
    plint nx, subsNum, solid, bb;
};

/* complab3d_processors_part2.hh:71-259; the reduced max total and mask sum
 * are kept for the caller (the reference reads them with computeMax /
 * computeSum on the lattices right after) */
template <typename T>
class updateLocalMaskNtotalLattices3D {
public:
    updateLocalMaskNtotalLattices3D(plint nx_, plint ny_, plint nz_, plint length_, plint bb_, plint solid_,
                                    std::vector<std::vector<plint> > bio_, std::vector<plint> pore_, T bMassFrac_,
                                    T maxbMassRho_)
        : nx(nx_), ny(ny_), nz(nz_), length(length_), bb(bb_), solid(solid_), bio(bio_), pore(pore_),
          thrd(bMassFrac_ * maxbMassRho_), maxTotal(0), maskSum(0) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || ny != p.ny || nz != p.nz || !detail::same(thrd, p.thrd_bfilm_frac * p.bmax)) return detail::mismatch(t);
        return clb_mask_total(t.ctx, t.case_id, &maxTotal, &maskSum);
    }
    double getMaxTotal() const { return maxTotal; }
    int64_t getMaskSum() const { return maskSum; }
private:
    plint nx, ny, nz, length, bb, solid;
    std::vector<std::vector<plint> > bio;
    std::vector<plint> pore;
    T thrd;
    double maxTotal;
    int64_t maskSum;
};

/* complab3d_processors_part1.hh:500-732 (the case's ca_method selects the rule) */
template <typename T>
class pushExcessBiomass3D {
public:
    pushExcessBiomass3D(T Bmax_, plint nx_, plint ny_, plint nz_, plint bdryGap_, plint length_, plint solid_, plint bb_,
                        std::vector<plint> pore_)
        : Bmax(Bmax_), nx(nx_), ny(ny_), nz(nz_), bdryGap(bdryGap_), length(length_), solid(solid_), bb(bb_), pore(pore_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (!detail::same(Bmax, p.bmax) || nx != p.nx || ny != p.ny || nz != p.nz || bdryGap != 1 ||
            p.ca_method != CLB_CA_FRACTION)
            return detail::mismatch(t);
        return clb_ca_sweep(t.ctx, t.case_id);
    }
private:
    T Bmax;
    plint nx, ny, nz, bdryGap, length, solid, bb;
    std::vector<plint> pore;
};

/* complab3d_processors_part1.hh:735-970 */
template <typename T>
class halfPushExcessBiomass3D {
public:
    halfPushExcessBiomass3D(T Bmax_, plint nx_, plint ny_, plint nz_, plint bdryGap_, plint length_, plint solid_,
                            plint bb_, std::vector<plint> pore_)
        : Bmax(Bmax_), nx(nx_), ny(ny_), nz(nz_), bdryGap(bdryGap_), length(length_), solid(solid_), bb(bb_), pore(pore_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (!detail::same(Bmax, p.bmax) || nx != p.nx || ny != p.ny || nz != p.nz || bdryGap != 1 ||
            p.ca_method != CLB_CA_HALF)
            return detail::mismatch(t);
        return clb_ca_sweep(t.ctx, t.case_id);
    }
private:
    T Bmax;
    plint nx, ny, nz, bdryGap, length, solid, bb;
    std::vector<plint> pore;
};

/* complab3d_processors_part2.hh:840-965: one sweep per process(); the main
 * loop repeats it until the age sum stops changing (complab.cpp:3025-3033),
 * which relaxToFixedPoint() does on the device */
template <typename T>
class updateAgeDistance3D {
public:
    updateAgeDistance3D(T Bmax_, plint nx_, plint ny_, plint nz_) : Bmax(Bmax_), nx(nx_), ny(ny_), nz(nz_) {}
    int process(const Target &t) { return run(t, 0); }
    int relaxToFixedPoint(const Target &t) { return run(t, 1); }
private:
    int run(const Target &t, int fixed) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (!detail::same(Bmax, p.bmax) || nx != p.nx || ny != p.ny || nz != p.nz) return detail::mismatch(t);
        return clb_age_update(t.ctx, t.case_id, fixed);
    }
    T Bmax;
    plint nx, ny, nz;
};

/* complab3d_processors_part2.hh:485-553 */
template <typename T>
class updateSoluteDynamics3D {
public:
    updateSoluteDynamics3D(plint subsNum_, plint bb_, plint solid_, std::vector<plint> pore_,
                           std::vector<T> substrOMEGAinbMass_, std::vector<T> substrOMEGAinPore_)
        : subsNum(subsNum_), bb(bb_), solid(solid_), pore(pore_), omB(substrOMEGAinbMass_), omP(substrOMEGAinPore_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (subsNum != p.n_species) return detail::mismatch(t);
        for (size_t s = 0; s < omP.size() && s < (size_t) p.n_species; ++s)
            if (!detail::same(omP[s], p.omega_pore[s]) || (s < omB.size() && !detail::same(omB[s], p.omega_biofilm[s])))
                return detail::mismatch(t);
        return clb_update_solute_dynamics(t.ctx, t.case_id);
    }
private:
    plint subsNum, bb, solid;
    std::vector<plint> pore;
    std::vector<T> omB, omP;
};

/* complab3d_processors_part2.hh:585-647; reopened() = voxels reopened */
template <typename T>
class updateSoluteSolidDynamics3D {
public:
    updateSoluteSolidDynamics3D(plint subsNum_, plint bb_, plint solid_, std::vector<plint> pore_,
                                std::vector<T> substrOMEGAinbMass_, std::vector<T> substrOMEGAinPore_,
                                std::vector<bool> immobile_)
        : subsNum(subsNum_), bb(bb_), solid(solid_), pore(pore_), omB(substrOMEGAinbMass_), omP(substrOMEGAinPore_),
          immobile(immobile_), nReopened(0) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (subsNum != p.n_species) return detail::mismatch(t);
        return clb_update_solute_solid(t.ctx, t.case_id, &nReopened);
    }
    int reopened() const { return nReopened; }
private:
    plint subsNum, bb, solid;
    std::vector<plint> pore;
    std::vector<T> omB, omP;
    std::vector<bool> immobile;
    int nReopened;
};

/* complab3d_processors_part2.hh:748-818: bioX is the permeability ratio
 * (1/viscosity_ratio; <= 1e-14 makes biofilm bounce-back) */
template <typename T>
class updateNsLatticesDynamics3D {
public:
    updateNsLatticesDynamics3D(T nsOmega_, T bioX_, std::vector<plint> pore_, plint solid_, plint bb_)
        : nsOmega(nsOmega_), bioX(bioX_), pore(pore_), solid(solid_), bb(bb_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (!detail::same(nsOmega, 1.0 / p.ns_tau)) return detail::mismatch(t);
        return clb_ns_update_dynamics(t.ctx, t.case_id, (double) bioX);
    }
private:
    T nsOmega, bioX;
    std::vector<plint> pore;
    plint solid, bb;
};

/* precipitationVOP.hh:166-227; changed() = a voxel turned solid */
template <typename T>
class precipNodeConversion3D {
public:
    precipNodeConversion3D(plint nx_, plint ny_, plint nz_, T max_precipRho_, plint solid_, plint bb_, plint precipCode_,
                           plint precipPhaseId_ = 0)
        : nx(nx_), ny(ny_), nz(nz_), maxRho(max_precipRho_), solid(solid_), bb(bb_), code(precipCode_),
          phaseId(precipPhaseId_), anyChange(0) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || ny != p.ny || nz != p.nz || !detail::same(maxRho, p.precip_max_rho) || code != CLB_SOLID ||
            phaseId != p.precip_phase_id)
            return detail::mismatch(t);
        return clb_precip_convert(t.ctx, t.case_id, &anyChange);
    }
    bool changed() const { return anyChange != 0; }
private:
    plint nx, ny, nz;
    T maxRho;
    plint solid, bb, code, phaseId;
    int anyChange;
};

/* dissolutionVOP.hh:309-342: the phase ids of the grain phases are seeded from
 * the geometry when the case is created; this checks that the (material, id)
 * pairs the caller passes (complab.cpp:1672-1683, grain phases only) are the
 * case's: phase id k + 1 on dissol_material[k] */
template <typename T>
class initPhaseLattice3D {
public:
    initPhaseLattice3D(const std::vector<plint> &materials_, const std::vector<plint> &ids_)
        : materials(materials_), ids(ids_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (!p.dissol_on || materials.size() != ids.size()) return detail::mismatch(t);
        for (size_t k = 0; k < materials.size(); ++k) {
            const plint q = ids[k] - 1;
            if (q < 0 || q >= p.dissol_n_phases || p.dissol_is_precipitate[q] || materials[k] != p.dissol_material[q])
                return detail::mismatch(t);
        }
        return CLB_OK;
    }
private:
    std::vector<plint> materials, ids;
};

/* dissolutionVOP.hh:378-569 and 607-720: the park and gather passes run inside
 * the abiotic increment (clb_kinetics with CLB_KIN_ABIOTIC), so these two only
 * check that the case's dissolution matches */
template <typename T, typename Config>
class surfaceDissolutionKinetics3D {
public:
    surfaceDissolutionKinetics3D(plint nx_, plint ny_, plint nz_, plint subsNum_, T dt_, plint solid_, plint bb_,
                                 const std::vector<plint> &pore_, const Config *cfg_)
        : nx(nx_), ny(ny_), nz(nz_), subsNum(subsNum_), dt(dt_), solid(solid_), bb(bb_), pore(pore_), cfg(cfg_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || ny != p.ny || nz != p.nz || subsNum != p.n_species || !detail::same(dt, p.dt) || !p.dissol_on)
            return detail::mismatch(t);
        return CLB_OK;
    }
private:
    plint nx, ny, nz, subsNum;
    T dt;
    plint solid, bb;
    std::vector<plint> pore;
    const Config *cfg;
};
template <typename T, typename Config>
class dissolutionGather3D {
public:
    dissolutionGather3D(plint nx_, plint ny_, plint nz_, plint subsNum_, plint solid_, plint bb_, const Config *cfg_)
        : nx(nx_), ny(ny_), nz(nz_), subsNum(subsNum_), solid(solid_), bb(bb_), cfg(cfg_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (nx != p.nx || ny != p.ny || nz != p.nz || subsNum != p.n_species || !p.dissol_on) return detail::mismatch(t);
        return CLB_OK;
    }
private:
    plint nx, ny, nz, subsNum, solid, bb;
    const Config *cfg;
};

/* dissolutionVOP.hh:744-812; reopened() = voxels that turned pore */
template <typename T, typename Config>
class dissolNodeConversion3D {
public:
    dissolNodeConversion3D(plint subsNum_, plint poreCode_, plint solid_, plint bb_, const Config *cfg_)
        : subsNum(subsNum_), poreCode(poreCode_), solid(solid_), bb(bb_), cfg(cfg_), nReopened(0) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (subsNum != p.n_species || poreCode != CLB_PORE || !p.dissol_on) return detail::mismatch(t);
        return clb_dissol_convert(t.ctx, t.case_id, &nReopened);
    }
    int reopened() const { return nReopened; }
private:
    plint subsNum, poreCode, solid, bb;
    const Config *cfg;
    int nReopened;
};

/* complab3d_processors_part3.hh:261-310: the reset of the stabilization
 * pre-loop; the steps themselves are clb_ade_stabilize's */
template <typename T>
class stabilizeADElattice3D {
public:
    stabilizeADElattice3D(T c0_, std::vector<plint> pore_, std::vector<std::vector<plint> > bio_)
        : c0(c0_), pore(pore_), bio(bio_) {}
    int process(const Target &t) { return clb_ade_stabilize(t.ctx, t.case_id, 0); }
private:
    T c0;
    std::vector<plint> pore;
    std::vector<std::vector<plint> > bio;
};

/* complab3d_processors_part3.hh:460-536: sums of every species over the open
 * reaction box; getSum(s) after process() */
template <typename T>
class MaskedPopulationSumFunctional3D {
public:
    MaskedPopulationSumFunctional3D(std::vector<plint> materials_, T restingValue_)
        : materials(materials_), resting(restingValue_) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        sums.assign((size_t) p.n_species, 0.0);
        return clb_species_totals(t.ctx, t.case_id, sums.data());
    }
    double getSum(int species) const { return sums.at((size_t) species); }
private:
    std::vector<plint> materials;
    T resting;
    std::vector<double> sums;
};

/* complab3d_processors_part3.hh:505-536: everything a species' lattice holds
 * (PopulationSum3D over the whole box); here the diagnostics' total + held of
 * clb_field_stats, so getSum(s) - the reaction-box total is the held column of
 * complab.cpp:2758-2764 */
class PopulationSumFunctional3D {
public:
    explicit PopulationSumFunctional3D(int species_) : species(species_), sum(0.0) {}
    int process(const Target &t) {
        clb_case_params p;
        int rc = detail::params(t, p);
        if (rc) return rc;
        if (species < 0 || species >= p.n_species) return CLB_E_ARG;
        std::vector<double> st((size_t) 4 * (p.n_species + p.n_microbes));
        rc = clb_field_stats(t.ctx, t.case_id, st.data());
        if (!rc) sum = st[(size_t) 4 * species] + st[(size_t) 4 * species + 3];
        return rc;
    }
    double getSum() const { return sum; }
private:
    int species;
    double sum;
};
inline int PopulationSum3D(const Target &t, int species, double *sum) {
    PopulationSumFunctional3D f(species);
    const int rc = f.process(t);
    if (!rc && sum) *sum = f.getSum();
    return rc;
}

/* complab3d_processors_part3.hh:386-427 over the reaction box on the live
 * mask: MaskedScalarCounts3D for the open materials (pore and biofilm) is
 * clb_counters' [0], for the biofilm materials alone its [1] */
inline int MaskedScalarCounts3D(const Target &t, bool biofilmOnly, int64_t *count) {
    int64_t c[CLB_NCOUNTERS];
    const int rc = clb_counters(t.ctx, t.case_id, c);
    if (!rc && count) *count = biofilmOnly ? c[1] : c[0];
    return rc;
}

/* Palabos' applyProcessingFunctional: the callee owns the new-ed processor */
template <class Proc>
int applyProcessingFunctional(Proc *proc, const Target &t) {
    const int rc = proc->process(t);
    delete proc;
    return rc;
}
/* the reductive form: the caller keeps the processor and reads it after */
template <class Proc>
int applyProcessingFunctional(Proc &proc, const Target &t) {
    return proc.process(t);
}

}  // namespace clb_proc

#endif /* CLB_PROCESSORS_HH */

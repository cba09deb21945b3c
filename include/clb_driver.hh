/*
 * clb_driver.hh -- the set-up sequence of a CompLaB3D run over the C-ABI, for a
 * C++ host program (header-only, C++11).
 *
 * The order and formulas are those of complab.cpp: the flow solve and the
 * pressure-drop correction for the target Peclet number (complab.cpp:1098-1183),
 * the relaxation and time step of every advection-diffusion lattice
 * (complab.cpp:1220-1252), the velocity coupling and the stabilization pre-loop
 * with its reset (complab.cpp:2009-2042).  Every call returns the library's
 * status code; a negative one stops the sequence where it failed.
 */
#ifndef CLB_DRIVER_HH
#define CLB_DRIVER_HH

#include "clb.h"

#include <cmath>
#include <vector>

namespace clb_driver {

const double CS2_ADE = 0.25;          // D3Q7: rest 1/4, faces 1/8
const double CS2_NS = 1.0 / 3.0;

struct FlowSetup {
    double pe = 0.0;                  // target Peclet number (0: no flow)
    double char_length = 0.0;         // lattice units
    double delta_p = 0.0;             // first pressure drop
    double ns_tau = 0.8;
    int max_iter = 100000;            // ns_max_iT1
    int rule = CLB_CONV_TRACER;
    double eps = 1e-8;                // ns_converge_iT1
    double tau_ade_fixed = 0.8;       // complab.cpp:332
    bool has_biofilm = false;         // the pore + biofilm mean velocity
};

struct TransportSetup {
    std::vector<double> d_pore, d_biofilm;   // physical diffusivities per species
    double dx = 1.0;
    double tau_if_no_flow = 0.8;
    int stabilize_steps = 10000;
};

struct Report {
    double delta_p = 0.0, pore_mean_u = 0.0, max_ux = 0.0, outlet_ux = 0.0;
    double permeability = 0.0, pe_achieved = 0.0, ref_tau = 0.0, dt = 0.0;
    int tau_marginal = 0, tau_high = 0;      /* the tau screen's warnings */
    std::vector<int> iters;
    std::vector<int> converged;
    std::vector<double> omega_pore, omega_biofilm;
};

/* the flow solve and the Peclet rescale (complab.cpp:1098-1183) */
inline int solveFlow(clb_ctx *ctx, int id, const FlowSetup &fs, Report &rep, double *poreMean) {
    const double nu = CS2_NS * (fs.ns_tau - 0.5);
    double dp = fs.delta_p;
    int rc, it = 0, cv = 0;
    if ((rc = clb_ns_init(ctx, id, dp))) return rc;
    if ((rc = clb_ns_solve(ctx, id, fs.max_iter, fs.rule, fs.eps, 1, &it, &cv))) return rc;
    rep.iters.push_back(it); rep.converged.push_back(cv);
    double st[4];
    if ((rc = clb_ns_stats(ctx, id, st))) return rc;
    double pm = fs.has_biofilm ? st[3] : st[0];
    double outlet = st[2], maxux = st[1];
    const double dLat = CS2_ADE * (fs.tau_ade_fixed - 0.5);
    const double perm = outlet * nu * fs.char_length / dp;
    const double ut = fs.pe * dLat / fs.char_length;
    const double dpNew = perm > 0 ? ut * nu * fs.char_length / perm : dp;
    if (std::fabs(dpNew - dp) / dp > 0.01) {
        if ((rc = clb_ns_init(ctx, id, dpNew))) return rc;
        if ((rc = clb_ns_solve(ctx, id, fs.max_iter, fs.rule, fs.eps, 1, &it, &cv))) return rc;
        rep.iters.push_back(it); rep.converged.push_back(cv);
        if ((rc = clb_ns_stats(ctx, id, st))) return rc;
        pm = st[0];                                   // complab.cpp:1175 averages the box
        maxux = st[1];
        outlet = st[2];
        dp = dpNew;
    }
    rep.delta_p = dp; rep.pore_mean_u = pm; rep.max_ux = maxux; rep.outlet_ux = outlet;
    rep.permeability = perm;
    rep.pe_achieved = outlet * fs.char_length / dLat;
    *poreMean = pm;
    return CLB_OK;
}

/* the tau screen of complab.cpp:1315-1378 over the relaxation rates of the
 * streaming (mobile) species, pore and biofilm: CLB_E_ARG below 0.51 (the
 * reference terminates); warnings for [0.51, 0.55) and above 2 counted in
 * rep.tau_marginal / rep.tau_high */
inline int tauScreen(int nSpecies, const int32_t *mobile, Report &rep) {
    rep.tau_marginal = rep.tau_high = 0;
    int low = 0;
    for (int s = 0; s < nSpecies; ++s) {
        if (mobile && !mobile[s]) continue;
        const double taus[2] = {1.0 / rep.omega_pore[s], 1.0 / rep.omega_biofilm[s]};
        for (double tau : taus) {
            if (tau < 0.51) ++low;
            else if (tau < 0.55) ++rep.tau_marginal;
            else if (tau > 2.0) ++rep.tau_high;
        }
    }
    return low ? CLB_E_ARG : CLB_OK;
}

/* lattice viscosity, relaxation and time step per species
 * (complab.cpp:1220-1252) and the tau screen; CLB_E_ARG where the reference
 * stops.  mobile: n_species flags (NULL: every species streams) */
inline int adeRelaxation(int nSpecies, double poreMeanU, const FlowSetup &fs, const TransportSetup &ts, Report &rep,
                         const int32_t *mobile = nullptr) {
    double refTau, refNu;
    if (fs.pe > 1e-12) {
        refNu = poreMeanU * fs.char_length / fs.pe;
        refTau = refNu / CS2_ADE + 0.5;
        if (refTau > 2 || refTau <= 0.5) return CLB_E_ARG;               // complab.cpp:1224
    } else {
        refTau = ts.tau_if_no_flow;
        refNu = CS2_ADE * (refTau - 0.5);
    }
    if ((int) ts.d_pore.size() < nSpecies || (int) ts.d_biofilm.size() < nSpecies || !(ts.d_pore[0] > 0))
        return CLB_E_ARG;                                                 // complab.cpp:1239
    rep.ref_tau = refTau;
    rep.dt = refNu * ts.dx * ts.dx / ts.d_pore[0];
    rep.omega_pore.assign(nSpecies, 0.0);
    rep.omega_biofilm.assign(nSpecies, 0.0);
    for (int s = 0; s < nSpecies; ++s) {
        const double nuS = refNu * ts.d_pore[s] / ts.d_pore[0];
        rep.omega_pore[s] = s > 0 ? 1.0 / (nuS / CS2_ADE + 0.5) : 1.0 / refTau;
        rep.omega_biofilm[s] = 1.0 / (refNu * ts.d_biofilm[s] / ts.d_pore[0] / CS2_ADE + 0.5);
    }
    return tauScreen(nSpecies, mobile, rep);
}

/* the whole set-up on a created case */
inline int setup(clb_ctx *ctx, int id, const clb_case_params &p, const FlowSetup &fs, const TransportSetup &ts, Report &rep) {
    double poreMean = 0.0;
    int rc;
    const bool flow = fs.pe > 0 && p.flow_on;
    if (flow && (rc = solveFlow(ctx, id, fs, rep, &poreMean))) return rc;
    if ((rc = adeRelaxation(p.n_species, poreMean, fs, ts, rep, p.mobile))) return rc;
    if ((rc = clb_set_ade_relaxation(ctx, id, rep.omega_pore.data(), rep.omega_biofilm.data(), rep.dt))) return rc;
    if (flow) {
        if ((rc = clb_ns_couple(ctx, id))) return rc;
        if (ts.stabilize_steps > 0 && (rc = clb_ade_stabilize(ctx, id, ts.stabilize_steps))) return rc;
    }
    return CLB_OK;
}

}  // namespace clb_driver

#endif /* CLB_DRIVER_HH */

/*
 * oracle.cpp -- CPU FP64 restatement of the CompLaB3D-GEMS coupled reaction
 * step.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * benchmark's CPU-baseline leg load this library as the checker.  Nothing in
 * the product links, loads or calls it.
 *
 * It is written the way the reference runs, not the way the GPU path runs:
 * Palabos ordering (collide every cell, data processors, then stream), the
 * populations of every cell stored, bounce-back cells stored and swapped in
 * place (full-way bounce-back), scalar "lattices" held as plain fields.  The
 * GPU path instead fuses stream+collide+kinetics in a pull kernel and keeps no
 * bounce-back cells; agreement between the two is the parity evidence.
 *
 * Citations are file:line in the CompLaB3D-GEMS tree (v1.4.0).  Palabos is not
 * part of that tree; the Palabos pieces restated here (BGK collision, the
 * regularized pressure / advection-diffusion boundaries, streaming at a
 * non-periodic edge, ValueTracer, getStoredAverageEnergy) are written from the
 * published LBM method and are pinned by closed forms and invariants in the
 * tests, not by reference runs.
 */
#include "clb.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#include <omp.h>

namespace {

const double THRD = 1e-12;         // complab3d_processors.hh:60
const double COMPLAB_THRD = 1e-14; // complab_functions.hh:74

// D3Q7 (complab3d_processors.hh:193-211): rest 1/4, faces 1/8, cs^2 = 1/4
const int Q7 = 7;
const int C7[7][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
const double W7[7] = {0.25,0.125,0.125,0.125,0.125,0.125,0.125};
const int OPP7[7] = {0,2,1,4,3,6,5};

// D3Q19: rest 1/3, faces 1/18, edges 1/36, cs^2 = 1/3
const int Q19 = 19;
const int C19[19][3] = {
    {0,0,0},
    {1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
    {1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0},
    {1,0,1},{-1,0,-1},{1,0,-1},{-1,0,1},
    {0,1,1},{0,-1,-1},{0,1,-1},{0,-1,1}};
const double W19[19] = {1.0/3,
    1.0/18,1.0/18,1.0/18,1.0/18,1.0/18,1.0/18,
    1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36,1.0/36};
inline int opp19(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

enum { ADE_BB = 0, ADE_PORE = 1, ADE_BFILM = 2 };
enum { NS_BGK = 0, NS_BIO = 1, NS_BB = 2, NS_NODYN = 3 };

/* ValueTracer of Palabos (complab.cpp:1097 constructs it with u=1, L=1000):
 * window of L/u/2 = 500 samples, converged when the population standard
 * deviation over the window divided by |mean| is below epsilon. */
struct Tracer {
    std::vector<double> v;
    size_t window = 500;
    void reset() { v.clear(); }
    void take(double x) { v.push_back(x); if (v.size() > window) v.erase(v.begin()); }
    bool converged(double eps) const {
        if (v.size() < window) return false;
        double m = 0; for (double x : v) m += x; m /= (double) v.size();
        double s = 0; for (double x : v) s += (x - m) * (x - m); s = std::sqrt(s / (double) v.size());
        double r = s / std::fabs(m);
        if (std::isnan(r)) return true;                 // "simulation diverged"
        return r < eps;
    }
};

struct Case {
    clb_case_params p;
    int nx, ny, nz; size_t N;
    std::vector<int8_t> geom;
    std::vector<int> mask, age, dist;
    std::vector<int8_t> adeCls, nsCls;
    std::vector<std::vector<double>> f;     // mobile species: 7*N pre-collision
    std::vector<std::vector<double>> imm;   // immobile species: N
    std::vector<std::vector<double>> rhoL, rhoR;  // imposed plane densities (ny*nz)
    std::vector<std::vector<double>> B;     // CA biomass per microbe
    std::vector<double> Btot;
    std::vector<int8_t> phase;
    std::vector<size_t> reopened;           // voxels reopened this step (ADE rebuild after the stream)
    std::vector<int8_t> adeClsCollide;      // ADE classes at this step's collide
    std::vector<double> g;                  // D3Q19 shifted populations, 19*N
    std::vector<double> u;                  // 3*N velocity read by the ADE lattices
    double nsOmega = 1.0, nsBioOmega = 1.0, deltaP = 0.0;
    Tracer tracer;
    double lastE = -1.0;
    int nsSaturate = 0;
    int clogged = 0;                        // percolationFlag (complab.cpp:3113, 3119, 3176)
    bool resolveFresh = true;               // the in-step re-solves' tracer ns_convg2 (complab.cpp:2166) is still empty
    long long oldTotMask = 0;
    long long iT = 0;
    int64_t counters[CLB_NCOUNTERS] = {0,0,0,0,0,0,0,0,0,0,0,0};
    std::vector<double> share;              // dissolution: per-face product share parked per species, S*N
    uint64_t caSweepCounter = 0;
    std::string err;
    // symbolic laws, one program per microbe; in use when every microbe has one
    std::vector<clb_sym_program> sym;
    uint32_t symMask = 0;
    // the abiotic symbolic law (<symbolic><abiotic_file>), in use when symAbOn
    clb_sym_program symAb{};
    bool symAbOn = false;
    // rate capture: the step's requested increments (accumulateRates), S*N
    std::vector<double> rate;
    bool rateOn = false;

    size_t id(int x, int y, int z) const { return (size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * (size_t) z); }
    bool inside(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz; }
    bool solidOrBB(int m) const { return m == CLB_SOLID || m == CLB_WALL; }
    bool isPore(int m) const { return m == CLB_PORE; }
};

double densityF(const Case &c, int s, size_t i) {
    const double *f = c.f[s].data();
    double r = 0; for (int q = 0; q < Q7; ++q) r += f[(size_t) q * c.N + i];
    return r;
}
double conc(const Case &c, int s, size_t i) { return c.p.mobile[s] ? densityF(c, s, i) : c.imm[s][i]; }

/* deposit dC on a species at a voxel: the D3Q7 weights 1/4, 1/8 (part1:151-153) */
void deposit(Case &c, int s, size_t i, double dC) {
    if (c.p.mobile[s]) { for (int q = 0; q < Q7; ++q) c.f[s][(size_t) q * c.N + i] += W7[q] * dC; }
    else c.imm[s][i] += dC;
}

/* ------------------------------------------------------------------------------------
 *  advection-diffusion lattice (Palabos AdvectionDiffusionBGKdynamics, complab.cpp:2307)
 * ------------------------------------------------------------------------------------ */
int adeClassFor(const Case &c, int s, int x, size_t i) {
    if (x == 0 && c.p.left_btype[s] == CLB_BC_CLOSED) return ADE_BB;            // complab_functions.hh:431
    if (x == c.nx - 1 && c.p.right_btype[s] == CLB_BC_CLOSED) return ADE_BB;
    return c.adeCls[i];
}

/* regularized first-order advection-diffusion Dirichlet plane: the population
 * entering from outside is fixed by the imposed density, then every population
 * is projected onto (rho, j): f_i = w_i (rho + c_i.j / cs^2). */
void adeDirichlet(double *fl, int inward, double rhoD) {
    int unknown = inward > 0 ? 1 : 2;
    double s = 0; for (int q = 0; q < Q7; ++q) if (q != unknown) s += fl[q];
    fl[unknown] = rhoD - s;
    double j[3] = {0,0,0};
    for (int q = 0; q < Q7; ++q) for (int a = 0; a < 3; ++a) j[a] += C7[q][a] * fl[q];
    for (int q = 0; q < Q7; ++q) {
        double cj = C7[q][0]*j[0] + C7[q][1]*j[1] + C7[q][2]*j[2];
        fl[q] = W7[q] * (rhoD + 4.0 * cj);
    }
}

void adeCollide(Case &c, int s) {
    const size_t N = c.N;
    double *f = c.f[s].data();
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        int cls = adeClassFor(c, s, x, i);
        double fl[7]; for (int q = 0; q < Q7; ++q) fl[q] = f[(size_t) q * N + i];
        if (cls == ADE_BB) {                                          // BounceBack::collide
            for (int q = 1; q < Q7; q += 2) std::swap(fl[q], fl[q + 1]);
        } else {
            if (x == 0 && c.p.left_btype[s] != CLB_BC_CLOSED)
                adeDirichlet(fl, +1, c.rhoL[s][(size_t) y + (size_t) c.ny * z]);
            else if (x == c.nx - 1 && c.p.right_btype[s] != CLB_BC_CLOSED)
                adeDirichlet(fl, -1, c.rhoR[s][(size_t) y + (size_t) c.ny * z]);
            double om = (cls == ADE_BFILM) ? c.p.omega_biofilm[s] : c.p.omega_pore[s];
            double rho = 0; for (int q = 0; q < Q7; ++q) rho += fl[q];
            const double *u = &c.u[0];
            double ux = u[i], uy = u[N + i], uz = u[2 * N + i];
            for (int q = 0; q < Q7; ++q) {
                double cu = C7[q][0]*ux + C7[q][1]*uy + C7[q][2]*uz;
                double feq = W7[q] * rho * (1.0 + 4.0 * cu);
                fl[q] = fl[q] * (1.0 - om) + om * feq;
            }
        }
        for (int q = 0; q < Q7; ++q) f[(size_t) q * N + i] = fl[q];
    }
}

/* streaming with Palabos' edge rule: a population that would arrive from
 * outside the block is the cell's own post-collision opposite one. */
void adeStream(Case &c, int s) {
    const size_t N = c.N;
    std::vector<double> nf(7 * N);
    double *f = c.f[s].data();
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        for (int q = 0; q < Q7; ++q) {
            int xs = x - C7[q][0], ys = y - C7[q][1], zs = z - C7[q][2];
            nf[(size_t) q * N + i] = c.inside(xs, ys, zs) ? f[(size_t) q * N + c.id(xs, ys, zs)]
                                                         : f[(size_t) OPP7[q] * N + i];
        }
    }
    c.f[s].swap(nf);
    /* FlatAdiabaticBoundaryFunctional3D (complab_functions.hh:384,389): after the
     * stream, the Neumann plane takes the density of the layer inside it.  An
     * inner cell that is bounce-back reports BounceBack's stored density, 0
     * (updateSoluteSolidDynamics3D installs BounceBack() with no density). */
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) {
        size_t k = (size_t) y + (size_t) c.ny * z;
        size_t il = c.id(1, y, z), ir = c.id(c.nx - 2, y, z);
        if (c.p.left_btype[s] == CLB_BC_NEUMANN) c.rhoL[s][k] = adeClassFor(c, s, 1, il) == ADE_BB ? 0.0 : densityF(c, s, il);
        if (c.p.right_btype[s] == CLB_BC_NEUMANN) c.rhoR[s][k] = adeClassFor(c, s, c.nx - 2, ir) == ADE_BB ? 0.0 : densityF(c, s, ir);
    }
}

/* ------------------------------------------------------------------------------------
 *  kinetics
 * ------------------------------------------------------------------------------------ */
/* the thermodynamic factor (complab3d_thermo.hh:215-255, Jin & Bethke 2003):
 * ln Q = sum nu_s ln max(conc_scale C_s, cmin); dG = dG0 + RT ln Q + dGloss;
 * f = dG + atp dGatp; F_T = max(0, 1 - exp(f / (chi RT))), 0 for f >= 0,
 * 1 where f / (chi RT) < -700 */
double thermoFactor(const clb_case_params &p, const double *C, double *dGout) {
    double lnQ = 0.0;
    for (int s = 0; s < p.n_species; ++s) {
        if (p.thm_nu[s] == 0.0) continue;
        double a = C[s] * p.thm_conc_scale;
        if (!(a > p.thm_cmin)) a = p.thm_cmin;
        lnQ += p.thm_nu[s] * std::log(a);
    }
    const double dG = p.thm_dg0 + p.thm_rt * lnQ + p.thm_dgloss;
    if (dGout) *dGout = dG;
    const double chi = p.thm_chi > 0.0 ? p.thm_chi : 1.0;
    const double x = (dG + p.thm_threshold) / (chi * p.thm_rt);
    if (x >= 0.0) return 0.0;
    if (x < -700.0) return 1.0;
    const double F = 1.0 - std::exp(x);
    return F < 0.0 ? 0.0 : (F > 1.0 ? 1.0 : F);
}

/* a .sym law compiled to postfix code (clb_sym_program), evaluated as
 * evalNode / evaluate do (complab3d_symbolic.hh:310-348, 766-804): variables
 * held in their ranges, then each rate in file order, NaN -> 0, +-1e300,
 * stored unscaled for later rates, leaving times unit_scale */
void symEvaluate(const clb_sym_program &P, const double *avail, double bmass, double *out) {
    double v[CLB_SYM_MAXVARS + CLB_SYM_MAXRATES];
    for (int i = 0; i < P.n_vars; ++i) {
        const int s = P.var_species[i];
        double x = s >= 0 ? avail[s] : bmass;
        if (P.hi[i] > P.lo[i]) { if (x < P.lo[i]) x = P.lo[i]; else if (x > P.hi[i]) x = P.hi[i]; }
        v[i] = x;
    }
    int pc = 0;
    for (int r = 0; r < P.n_rates; ++r) {
        std::vector<double> st;
        for (; pc < P.rate_end[r]; ++pc) {
            const int op = P.op[pc], arg = P.arg[pc];
            if (op == CLB_SYM_CONST) { st.push_back(P.konst[arg]); continue; }
            if (op == CLB_SYM_VAR) { st.push_back(v[arg]); continue; }
            double &top = st.back();
            switch (op) {
            case CLB_SYM_NEG: top = -top; continue;
            case CLB_SYM_EXP: top = std::exp(top); continue;
            case CLB_SYM_LOG: top = top > 0.0 ? std::log(top) : 0.0; continue;
            case CLB_SYM_LOG10: top = top > 0.0 ? std::log10(top) : 0.0; continue;
            case CLB_SYM_SQRT: top = top > 0.0 ? std::sqrt(top) : 0.0; continue;
            case CLB_SYM_ABS: top = std::fabs(top); continue;
            case CLB_SYM_TANH: top = std::tanh(top); continue;
            default: break;
            }
            const double y = st.back(); st.pop_back();
            const double x = st.back();
            double z = 0.0;
            switch (op) {
            case CLB_SYM_ADD: z = x + y; break;
            case CLB_SYM_SUB: z = x - y; break;
            case CLB_SYM_MUL: z = x * y; break;
            case CLB_SYM_DIV: z = std::fabs(y) < 1e-300 ? 0.0 : x / y; break;
            case CLB_SYM_POW: z = (x < 0.0 && y != std::floor(y)) ? 0.0 : std::pow(x, y); break;
            case CLB_SYM_MIN: z = x < y ? x : y; break;
            case CLB_SYM_MAX: z = x > y ? x : y; break;
            }
            st.back() = z;
        }
        double y = st.empty() ? 0.0 : st.back();
        if (y != y) y = 0.0;
        if (y > 1e300) y = 1e300; else if (y < -1e300) y = -1e300;
        v[P.n_vars + r] = y;
        out[r] = y * P.unit_scale;
    }
}

/* run_symbolic3D (complab3d_processors_symbolic.hh:104-182): the increments of
 * one voxel -- every organism present evaluates its law; substrate rates x
 * gate x dt are requests on one budget, a debit held to what the voxel has;
 * growth x gate x biomass x dt is the organism's increment */
void symbioticIncrements(const Case &c, const double *C, const double *B, double *dC, double *dB) {
    const clb_case_params &p = c.p;
    const double thrd = THRD;
    double avail[CLB_MAX_SPECIES], req[CLB_MAX_SPECIES], bm[CLB_MAX_MICROBES];
    for (int s = 0; s < p.n_species; ++s) { avail[s] = C[s] > thrd ? C[s] : 0.0; req[s] = 0.0; dC[s] = 0.0; }
    bool any = false;
    for (int m = 0; m < p.n_microbes; ++m) { bm[m] = B[m] > thrd ? B[m] : 0.0; dB[m] = 0.0; if (bm[m] > thrd) any = true; }
    if (!any) return;
    for (int m = 0; m < p.n_microbes; ++m) {
        if (!(bm[m] > thrd)) continue;
        const clb_sym_program &P = c.sym[(size_t) m];
        if (P.n_rates <= 0) continue;
        double out[CLB_SYM_MAXRATES];
        symEvaluate(P, avail, bm[m], out);
        const double ft = (p.thm_on && (p.thm_microbe < 0 || p.thm_microbe == m)) ? thermoFactor(p, avail, nullptr) : 1.0;
        for (int r = 0; r < P.n_rates; ++r) {
            const int rs = P.rate_species[r];
            if (rs >= 0) req[rs] += out[r] * ft * p.dt;
            else dB[m] += out[r] * ft * bm[m] * p.dt;
        }
    }
    for (int s = 0; s < p.n_species; ++s) {
        double draw = req[s];
        if (std::fabs(draw) <= thrd) continue;
        const double headroom = avail[s] + dC[s];
        if (draw + headroom < 0.0) draw = -headroom;
        dC[s] += draw;
    }
}

/* run_symbolic_abiotic3D (complab3d_processors_abiotic_learned.hh:60-135): the
 * abiotic .sym law at one open voxel, no biomass; each substrate's rate x dt
 * is a request held to what the voxel has (a budget of its own), added to the
 * voxel's abiotic increment when it is above the threshold */
void symAbioticAdd(const Case &c, const double *C, double *dAb) {
    const clb_case_params &p = c.p;
    const double thrd = THRD;
    double avail[CLB_MAX_SPECIES], req[CLB_MAX_SPECIES], out[CLB_SYM_MAXRATES];
    for (int s = 0; s < p.n_species; ++s) { avail[s] = C[s] > thrd ? C[s] : 0.0; req[s] = 0.0; }
    symEvaluate(c.symAb, avail, 0.0, out);
    for (int r = 0; r < c.symAb.n_rates; ++r) {
        const int rs = c.symAb.rate_species[r];                      // -1: the extent line itself
        if (rs >= 0) req[rs] += out[r] * p.dt;
    }
    for (int s = 0; s < p.n_species; ++s) {
        double draw = req[s];
        if (std::fabs(draw) <= thrd) continue;
        if (draw + avail[s] < 0.0) draw = -avail[s];
        if (std::fabs(draw) > thrd) dAb[s] += draw;
    }
}

/* some microbe has a symbolic law */
bool symActive(const Case &c) { return c.p.n_microbes > 0 && c.symMask != 0; }

/* dual-Monod restatement of defineRxnKinetics (config/kinetics/defineKinetics.biotic.hh:90-126):
 * growth = mu_max * S/(Ks+S) [* E/(Ke+E)] * B ; bioR = growth - kd*B ; subsR_s += nu_s*growth */
void bioticLaw(const clb_case_params &p, const double *C, const double *B, double *subsR, double *bioR, int mask,
               uint32_t skip = 0) {
    for (int s = 0; s < p.n_species; ++s) subsR[s] = 0;
    for (int m = 0; m < p.n_microbes; ++m) bioR[m] = 0;
    if (mask < 2) return;
    double S = p.donor >= 0 ? std::max(C[p.donor], 0.0) : 0.0;
    double E = p.acceptor >= 0 ? std::max(C[p.acceptor], 0.0) : 1.0;
    for (int m = 0; m < p.n_microbes; ++m) {
        if ((skip >> m) & 1u) continue;                                 // a microbe with a symbolic law
        double Bm = std::max(B[m], 0.0);
        if (Bm <= 0.0) continue;
        double mu = p.mu_max[m] * (S / (p.ks[m] + S));
        if (p.acceptor >= 0) mu *= E / (p.ke[m] + E);
        double growth = mu * Bm;
        bioR[m] += growth - p.kdecay[m] * Bm;
        for (int s = 0; s < p.n_species; ++s) subsR[s] += p.bio_stoich[m][s] * growth;
    }
    for (int s = 0; s < p.n_species; ++s) if (!std::isfinite(subsR[s])) subsR[s] = 0;
    for (int m = 0; m < p.n_microbes; ++m) if (!std::isfinite(bioR[m])) bioR[m] = 0;
    if (p.thm_on) {                                                    // part1:127-140
        const double F = thermoFactor(p, C, nullptr);
        if (F < 1.0) {
            for (int s = 0; s < p.n_species; ++s) subsR[s] *= F;
            for (int m = 0; m < p.n_microbes; ++m) bioR[m] *= F;
        }
    }
}

/* mass-action restatement of defineAbioticRxnKinetics
 * (ABC_sweep_package/kinetics/defineAbioticKinetics.hh:166-210,
 *  examples/13_precipitation/kinetics/defineAbioticKinetics.hh) */
void abioticLaw(const clb_case_params &p, const double *C, double *subsR) {
    for (int s = 0; s < p.n_species; ++s) subsR[s] = 0;
    for (int r = 0; r < p.n_rxn; ++r) {
        double a = p.rxn_r1[r] >= 0 ? std::max(C[p.rxn_r1[r]], 0.0) : 1.0;
        double b = p.rxn_r2[r] >= 0 ? std::max(C[p.rxn_r2[r]], 0.0) : 1.0;
        if (p.rxn_min_conc > 0 && ((p.rxn_r1[r] >= 0 && a <= p.rxn_min_conc) || (p.rxn_r2[r] >= 0 && b <= p.rxn_min_conc))) continue;
        double R = p.rxn_k[r] * a * b;
        if (p.rxn_cap_dt > 0 && p.rxn_cap_frac > 0) {
            double sc = std::min(p.rxn_r1[r] >= 0 ? a : 1e300, p.rxn_r2[r] >= 0 ? b : 1e300);
            double cap = sc * p.rxn_cap_frac / p.rxn_cap_dt;
            if (R > cap) R = cap;
        }
        for (int s = 0; s < p.n_species; ++s) subsR[s] += p.rxn_stoich[r][s] * R;
    }
    for (int s = 0; s < p.n_species; ++s) if (!std::isfinite(subsR[s])) subsR[s] = 0;
    if (p.abiotic_rate_scale != 1.0)                                   // part1:247-249
        for (int s = 0; s < p.n_species; ++s) subsR[s] *= p.abiotic_rate_scale;
}

/* the clamp dC = max(R dt, -C) (README.md:87-88; symbolic processor 161-176),
 * only for a debit and only against what the voxel holds.  The bind counter
 * counts debits beyond the availability by more than 1e-12 of it: a debit that
 * equals the availability but for rounding (the dissolution gather's own
 * positivity rule leaves exactly such debits) is capped but not counted. */
double clampInc(Case &c, double dC, double cur) {
    if (!c.p.clamp || dC >= 0) return dC;
    double lim = -std::max(cur, 0.0);
    if (dC < lim) {
        if (dC < lim - 1e-12 * std::fabs(lim)) {
#pragma omp atomic
            c.counters[2]++;
        }
        return lim;
    }
    return dC;
}

/* run_kinetics (part1:85-187) + update_rxnLattices (part1:406-497) */
/* reactionBox = x 1..nx-2 (complab.cpp:2337) */
void bioticStep(Case &c) {
    const clb_case_params &p = c.p;
    const int S = p.n_species, M = p.n_microbes;
    std::vector<double> dC((size_t) S * c.N, 0.0), dB((size_t) std::max(M,1) * c.N, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        int mask = c.mask[i];
        if (c.solidOrBB(mask)) continue;
        double C[CLB_MAX_SPECIES], Bv[CLB_MAX_MICROBES], sr[CLB_MAX_SPECIES], br[CLB_MAX_MICROBES];
        for (int s = 0; s < S; ++s) { C[s] = conc(c, s, i); if (C[s] < THRD) C[s] = 0; }
        for (int m = 0; m < M; ++m) { Bv[m] = c.B[m][i]; if (Bv[m] < THRD) Bv[m] = 0; }
        // the compiled law for the microbes without a symbolic law (run_kinetics),
        // then the symbolic organisms' increments (run_symbolic3D), into the
        // same increments (complab.cpp:2348-2393)
        if (p.bio_law != CLB_BIO_NONE) {
            bioticLaw(p, C, Bv, sr, br, mask, c.symMask);
            for (int s = 0; s < S; ++s) { double d = sr[s] * p.dt; if (d > THRD || d < -THRD) dC[(size_t) s * c.N + i] += d; }
            for (int m = 0; m < M; ++m) { double d = br[m] * p.dt; if (d > THRD || d < -THRD) dB[(size_t) m * c.N + i] += d; }
        }
        if (symActive(c)) {
            symbioticIncrements(c, C, Bv, sr, br);
            for (int s = 0; s < S; ++s) { double d = sr[s]; if (d > THRD || d < -THRD) dC[(size_t) s * c.N + i] += d; }
            for (int m = 0; m < M; ++m) { double d = br[m]; if (d > THRD || d < -THRD) dB[(size_t) m * c.N + i] += d; }
        }
    }
    if (c.rateOn) for (size_t k = 0; k < dC.size(); ++k) c.rate[k] += dC[k];     // complab.cpp:2450
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        if (c.solidOrBB(c.mask[i])) continue;
        for (int s = 0; s < S; ++s) {
            double d = dC[(size_t) s * c.N + i];
            if (d > THRD || d < -THRD) deposit(c, s, i, clampInc(c, d, conc(c, s, i)));
        }
        for (int m = 0; m < M; ++m) {
            double d = dB[(size_t) m * c.N + i];
            if (d > THRD || d < -THRD) {
                if (d < -c.B[m][i]) d = -c.B[m][i];                       // biomass never negative
                c.B[m][i] += d;
            }
        }
    }
}

/* run_abiotic_kinetics (part1:199-297) or surfaceAbioticKinetics3D
 * (precipitationVOP.hh:53-153), then update_abiotic_rxnLattices (part1:306-404) */
/* the dissolution rate law of declared phase q (defineDissolutionRate for
 * phase id q + 1; the reference's examples 14 and 15,
 * kinetics/defineAbioticKinetics.hh, have orders 1 and 1):
 * rate = k max(C[acid],0)^a max(mineral,0)^m, mineralR = -rate,
 * subsR[s] = stoich[s] rate; non-finite rates read 0.  x^1 is x and x^0 is 1
 * exactly, other orders std::pow. */
static double orderPow(double x, double e) { return e == 1.0 ? x : (e == 0.0 ? 1.0 : std::pow(x, e)); }
void dissolutionLaw(const clb_case_params &p, int q, const double *water, double mineral, double *subsR, double &mineralR) {
    const int S = p.n_species;
    const double H = p.dissol_acid[q] >= 0 ? orderPow(std::max(water[p.dissol_acid[q]], 0.0), p.dissol_acid_order[q]) : 1.0;
    const double rate = p.dissol_k[q] * H * orderPow(std::max(mineral, 0.0), p.dissol_mineral_order[q]);
    mineralR = -rate;
    for (int s = 0; s < S; ++s) subsR[s] = p.dissol_stoich[q][s] * rate;
    if (!std::isfinite(mineralR)) mineralR = 0.0;
    for (int s = 0; s < S; ++s) if (!std::isfinite(subsR[s])) subsR[s] = 0.0;
}

/* the declared phase of phase id pid, -1 for none (DissolutionConfig::byId) */
static int phaseIndex(const clb_case_params &p, int pid) { return (pid >= 1 && pid <= p.dissol_n_phases) ? pid - 1 : -1; }

static const int DLOC[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};

/* surfaceDissolutionKinetics3D (dissolutionVOP.hh:378-569): every voxel of the
 * declared phase with mineral left averages the water of its open face
 * neighbours in the reaction box, takes the rate law, removes the mineral
 * (at most half of it per step, products scaled alike) into its own increment,
 * and parks one share per open face of every product */
void dissolutionPark(Case &c, std::vector<double> &dC) {
    const clb_case_params &p = c.p;
    const int S = p.n_species;
    std::fill(c.share.begin(), c.share.end(), 0.0);
    double water[CLB_MAX_SPECIES], subsR[CLB_MAX_SPECIES];
    for (int x = 1; x < c.nx - 1; ++x) for (int y = 0; y < c.ny; ++y) for (int z = 0; z < c.nz; ++z) {
        size_t i = c.id(x, y, z);
        const int q = phaseIndex(p, c.phase[i]);
        if (q < 0) continue;
        const int sM = p.dissol_species[q];
        const double mineral = c.imm[sM][i];
        if (!(mineral > THRD)) continue;
        size_t openN[6]; int nOpen = 0;
        for (int d = 0; d < 6; ++d) {
            int X = x + DLOC[d][0], Y = y + DLOC[d][1], Z = z + DLOC[d][2];
            if (X < 1 || X > c.nx - 2 || Y < 0 || Y > c.ny - 1 || Z < 0 || Z > c.nz - 1) continue;
            size_t j = c.id(X, Y, Z);
            if (c.solidOrBB(c.mask[j])) continue;
            openN[nOpen++] = j;
        }
        if (nOpen == 0 && p.dissol_surface_only != 0) continue;
        for (int s = 0; s < S; ++s) water[s] = 0.0;
        if (nOpen > 0) {
            for (int k = 0; k < nOpen; ++k) for (int s = 0; s < S; ++s) water[s] += conc(c, s, openN[k]);
            for (int s = 0; s < S; ++s) water[s] /= (double) nOpen;
        }
        for (int s = 0; s < S; ++s) if (water[s] < 0.0) water[s] = 0.0;
        double mineralR = 0.0;
        dissolutionLaw(p, q, water, mineral, subsR, mineralR);
        if (!(mineralR < 0.0)) continue;
        double dMineral = mineralR * p.dt;
        double scale = 1.0;
        const double maxRemove = 0.5 * mineral;                     // MAX_DISSOL_FRACTION, dissolutionVOP.hh:485
        if (-dMineral > maxRemove) scale = maxRemove / (-dMineral);
        dMineral *= scale;
        if (!(-dMineral > THRD)) continue;
        dC[(size_t) sM * c.N + i] += dMineral;
        if (nOpen > 0) {
            for (int s = 0; s < S; ++s) {
                if (s == sM) continue;
                const double released = subsR[s] * p.dt * scale;
                if (std::fabs(released) <= THRD) continue;
                c.share[(size_t) s * c.N + i] = released / (double) nOpen;
                c.counters[8] += nOpen;
            }
        }
    }
}

/* dissolutionGather3D (dissolutionVOP.hh:607-720): every open voxel of the
 * reaction box collects one share from each face neighbour of the phase; a
 * consuming share may not take more than the voxel holds plus what is already
 * pending (the positivity rule, 669-683) */
void dissolutionGather(Case &c, std::vector<double> &dC) {
    const clb_case_params &p = c.p;
    const int S = p.n_species;
    const double thrdG = 1e-14;
    for (int x = 1; x < c.nx - 1; ++x) for (int y = 0; y < c.ny; ++y) for (int z = 0; z < c.nz; ++z) {
        size_t i = c.id(x, y, z);
        if (c.solidOrBB(c.mask[i])) continue;
        for (int d = 0; d < 6; ++d) {
            int X = x + DLOC[d][0], Y = y + DLOC[d][1], Z = z + DLOC[d][2];
            if (X < 1 || X > c.nx - 2 || Y < 0 || Y > c.ny - 1 || Z < 0 || Z > c.nz - 1) continue;
            size_t j = c.id(X, Y, Z);
            const int q = phaseIndex(p, c.phase[j]);
            if (q < 0) continue;
            const int sM = p.dissol_species[q];
            for (int s = 0; s < S; ++s) {
                if (s == sM) continue;
                const double share = c.share[(size_t) s * c.N + j];
                if (std::fabs(share) <= thrdG) continue;
                double give = share;
                if (give < 0.0) {
                    const double have = conc(c, s, i);
                    const double pend = dC[(size_t) s * c.N + i];
                    const double headroom = (have > 0.0 ? have : 0.0) + pend;
                    if (give + headroom < 0.0) { give = -headroom; c.counters[10]++; }
                    if (give > 0.0) give = 0.0;
                }
                c.counters[9]++;
                if (!(std::fabs(give) > thrdG)) continue;
                dC[(size_t) s * c.N + i] += give;
            }
        }
    }
}

/* run_abiotic_kinetics + (dissolution park and gather) + update_abiotic_rxnLattices
 * (part1:199-404; complab.cpp:2461-2545): one increment per species and voxel */
void abioticStep(Case &c) {
    const clb_case_params &p = c.p;
    const int S = p.n_species;
    std::vector<double> dC((size_t) S * c.N, 0.0);
    if (p.n_rxn > 0) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
            size_t i = c.id(x, y, z);
            int mask = c.mask[i];
            if (c.solidOrBB(mask)) continue;
            if (p.surface_only) {
                bool iface = false;
                for (int q = 1; q < Q7 && !iface; ++q) {
                    int X = x + C7[q][0], Y = y + C7[q][1], Z = z + C7[q][2];
                    if (!c.inside(X, Y, Z)) continue;
                    if (c.solidOrBB(c.mask[c.id(X, Y, Z)])) iface = true;
                }
                if (!iface) continue;
            }
            double C[CLB_MAX_SPECIES], sr[CLB_MAX_SPECIES];
            for (int s = 0; s < S; ++s) { C[s] = conc(c, s, i); if (C[s] < THRD) C[s] = 0; }
            abioticLaw(p, C, sr);
            for (int s = 0; s < S; ++s) { double d = sr[s] * p.dt; if (d > THRD || d < -THRD) dC[(size_t) s * c.N + i] += d; }
        }
    }
    if (c.symAbOn) {                                                    // complab.cpp:2489-2493, same increments
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
            size_t i = c.id(x, y, z);
            if (c.solidOrBB(c.mask[i])) continue;
            double C[CLB_MAX_SPECIES], d[CLB_MAX_SPECIES];
            for (int s = 0; s < S; ++s) { C[s] = conc(c, s, i); if (C[s] < THRD) C[s] = 0; d[s] = 0.0; }
            symAbioticAdd(c, C, d);
            for (int s = 0; s < S; ++s) dC[(size_t) s * c.N + i] += d[s];
        }
    }
    if (p.dissol_on) {
        dissolutionPark(c, dC);
        dissolutionGather(c, dC);
    }
    if (c.rateOn) for (size_t k = 0; k < dC.size(); ++k) c.rate[k] += dC[k];     // complab.cpp:2523
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        int mask = c.mask[i];
        for (int s = 0; s < S; ++s) {
            bool applies = !c.solidOrBB(mask) || !p.mobile[s];                // part1:327-330
            if (!applies) continue;
            double d = dC[(size_t) s * c.N + i];
            if (d > THRD || d < -THRD) deposit(c, s, i, clampInc(c, d, conc(c, s, i)));
        }
    }
}

/* ------------------------------------------------------------------------------------
 *  equilibrium speciation: EquilibriumChemistry<T> (part4_eqsolver.hh:186-1279)
 * ------------------------------------------------------------------------------------ */
const double MIN_CONC = 1e-30, MAX_CONC = 10.0, MIN_LOG_C = -30.0, MAX_LOG_C = 1.0;

struct Eq {
    int ns, nc;
    const clb_case_params *p;
    bool isEq(int i) const { for (int j = 0; j < nc; ++j) if (std::fabs(p->eq_stoich[i][j]) > 1e-10) return true; return false; }

    // calc_species, part4:450-504
    void species(const double *logC, const double *init, double *out) const {
        for (int i = 0; i < ns; ++i) {
            if (!isEq(i)) { out[i] = init[i] > MIN_CONC ? std::min(init[i], MAX_CONC) : MIN_CONC; continue; }
            double lc = p->eq_logk[i];
            for (int j = 0; j < nc; ++j) lc += p->eq_stoich[i][j] * logC[j];
            lc = std::max(std::min(lc, MAX_LOG_C), MIN_LOG_C);
            double v = std::pow(10.0, lc);
            if (!std::isfinite(v)) v = init[i] > MIN_CONC ? init[i] : MIN_CONC;
            out[i] = std::max(std::min(v, MAX_CONC), MIN_CONC);
        }
    }
    // calc_component_totals, part4:528-546
    void totals(const double *c, double *T) const {
        for (int j = 0; j < nc; ++j) {
            double t = 0;
            for (int i = 0; i < ns; ++i) if (isEq(i)) t += p->eq_stoich[i][j] * std::max(std::min(c[i], MAX_CONC), MIN_CONC);
            T[j] = std::max(t, MIN_CONC);
        }
    }
    // pcf_residual, part4:593-734
    void pcf(const double *logC, const double *T, const double *init, double *f) const {
        double sp[CLB_MAX_SPECIES];
        species(logC, init, sp);
        for (int j = 0; j < nc; ++j) {
            double pos = 0, neg = 0;
            for (int i = 0; i < ns; ++i) {
                if (!isEq(i)) continue;
                double mu = p->eq_stoich[i][j], ci = std::max(sp[i], MIN_CONC);
                if (mu > 1e-15) pos += mu * ci; else if (mu < -1e-15) neg += std::fabs(mu) * ci;
            }
            double SR, SP;
            if (T[j] >= 0) { SR = pos; SP = T[j] + neg; } else { SR = std::fabs(T[j]) + pos; SP = neg; }
            SR = std::max(SR, MIN_CONC); SP = std::max(SP, MIN_CONC);
            double mu0 = 1.0;
            for (int i = 0; i < ns; ++i) { if (!isEq(i)) continue; double mu = p->eq_stoich[i][j]; if (mu > 1e-10 && mu < mu0) mu0 = mu; }
            double v = (1.0 / mu0) * (std::log10(SP) - std::log10(SR));
            v = std::max(std::min(v, 10.0), -10.0);
            if (!std::isfinite(v)) v = 0.0;
            if (T[j] <= 1e-18) v = 0.0;
            f[j] = v;
        }
    }
    static double norm(const std::vector<double> &v) { double s = 0; for (double x : v) s += x * x; return std::sqrt(s); }
    static double dot(const std::vector<double> &a, const std::vector<double> &b) { double s = 0; for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i]; return s; }
    // qr_decomposition, part4:825-876 (modified Gram-Schmidt against the projected columns)
    static void qr(const std::vector<std::vector<double>> &cols, std::vector<std::vector<double>> &Q,
                   std::vector<std::vector<double>> &R, double &cond) {
        size_t m = cols.size();
        if (m == 0) { cond = 1.0; return; }
        size_t n = cols[0].size();
        Q.assign(m, std::vector<double>()); R.assign(m, std::vector<double>(m, 0.0));
        double rmax = 0, rmin = std::numeric_limits<double>::max();
        for (size_t j = 0; j < m; ++j) {
            Q[j] = cols[j];
            for (size_t i = 0; i < j; ++i) {
                R[i][j] = dot(Q[i], cols[j]);
                for (size_t k = 0; k < n; ++k) Q[j][k] -= R[i][j] * Q[i][k];
            }
            R[j][j] = norm(Q[j]);
            if (R[j][j] > 1e-15) for (size_t k = 0; k < n; ++k) Q[j][k] /= R[j][j];
            double a = std::fabs(R[j][j]);
            if (a > rmax) rmax = a;
            if (a > 1e-30 && a < rmin) rmin = a;
        }
        cond = (rmin > 1e-30) ? rmax / rmin : std::numeric_limits<double>::max();
    }
    // solve_equilibrium_anderson, part4:960-1180
    void anderson(const double *init, const double *T, double *out, int &iters, double &res, bool &conv) const {
        iters = 0; res = 0; conv = false;
        std::vector<double> om(nc);
        for (int j = 0; j < nc; ++j) {
            double cc = 1e-7;
            int si = p->eq_comp_species[j];
            if (si >= 0 && si < ns) cc = init[si];
            cc = std::max(std::min(cc, MAX_CONC), MIN_CONC);
            om[j] = std::log10(cc);
        }
        std::vector<double> f0(nc), onew(nc);
        pcf(om.data(), T, init, f0.data());
        for (int j = 0; j < nc; ++j) onew[j] = std::max(std::min(om[j] + f0[j], MAX_LOG_C), MIN_LOG_C);
        std::vector<std::vector<double>> oh, fh;
        oh.push_back(om); fh.push_back(f0);
        om = onew;
        const int depth = std::max(1, p->eq_anderson_depth);
        for (int it = 1; it < p->eq_max_iter; ++it) {
            iters = it;
            std::vector<double> fk(nc);
            pcf(om.data(), T, init, fk.data());
            double fn = norm(fk);
            res = fn;
            if (fn < p->eq_tol) { conv = true; species(om.data(), init, out); return; }
            oh.push_back(om); fh.push_back(fk);
            int mk = std::min(depth, (int) fh.size() - 1);
            if (mk < 1) { for (int j = 0; j < nc; ++j) onew[j] = om[j] + fk[j]; }
            else {
                std::vector<std::vector<double>> DF, DX;
                size_t hs = fh.size();
                for (int i = 0; i < mk; ++i) {
                    size_t ix = hs - mk - 1 + i;
                    std::vector<double> a(nc), b(nc);
                    for (int j = 0; j < nc; ++j) { a[j] = fh[ix + 1][j] - fh[ix][j]; b[j] = oh[ix + 1][j] - oh[ix][j]; }
                    DF.push_back(a); DX.push_back(b);
                }
                std::vector<std::vector<double>> Q, R; double cond;
                qr(DF, Q, R, cond);
                while (cond > 1e10 && DF.size() > 1) {
                    DF.erase(DF.begin()); DX.erase(DX.begin());
                    qr(DF, Q, R, cond);
                }
                size_t m = DF.size();
                std::vector<double> qtf(m), gam(m, 0.0);
                for (size_t i = 0; i < m; ++i) qtf[i] = dot(Q[i], fk);
                for (int i = (int) m - 1; i >= 0; --i) {                      // solve_upper_triangular, part4:899-919
                    gam[i] = qtf[i];
                    for (size_t j = i + 1; j < m; ++j) gam[i] -= R[i][j] * gam[j];
                    if (std::fabs(R[i][i]) > 1e-30) gam[i] /= R[i][i];
                }
                std::vector<double> dxg(nc, 0.0), dfg(nc, 0.0);
                for (size_t i = 0; i < m; ++i) for (int j = 0; j < nc; ++j) { dxg[j] += DX[i][j] * gam[i]; dfg[j] += DF[i][j] * gam[i]; }
                for (int j = 0; j < nc; ++j) onew[j] = om[j] - dxg[j] + 1.0 * (fk[j] - dfg[j]);
            }
            for (int j = 0; j < nc; ++j) {
                onew[j] = std::max(std::min(onew[j], MAX_LOG_C), MIN_LOG_C);
                if (!std::isfinite(onew[j])) onew[j] = om[j];
            }
            om = onew;
            while (oh.size() > (size_t) (depth + 1)) { oh.erase(oh.begin()); fh.erase(fh.begin()); }
        }
        species(om.data(), init, out);
    }
    /* Newton-Raphson on the log10 component activities (the solver of
     * part4_eqsolver_NR_reserve.hh:356-512, restated): F_j = sum_i S_ij c_i - T_j,
     * J_jk = ln10 sum_i S_ij S_ik c_i, step limited to 1 decade, stop when
     * max_j |F_j| / max(|T_j|, 1e-30) < tol. Starts from the same guess as AA-PCF. */
    void newton(const double *init, const double *T, double *out, int &iters, double &res, bool &conv) const {
        iters = 0; res = 0; conv = false;
        double om[CLB_MAX_COMP], sp[CLB_MAX_SPECIES];
        for (int j = 0; j < nc; ++j) {
            double cc = 1e-7;
            int si = p->eq_comp_species[j];
            if (si >= 0 && si < ns) cc = init[si];
            cc = std::max(std::min(cc, MAX_CONC), MIN_CONC);
            om[j] = std::log10(cc);
        }
        const double LN10 = 2.302585092994045684;
        for (int it = 0; it < p->eq_max_iter; ++it) {
            species(om, init, sp);
            double F[CLB_MAX_COMP], J[CLB_MAX_COMP][CLB_MAX_COMP];
            double rmax = 0;
            for (int j = 0; j < nc; ++j) {
                double t = 0;
                for (int i = 0; i < ns; ++i) if (isEq(i)) t += p->eq_stoich[i][j] * sp[i];
                F[j] = t - T[j];
                double r = std::fabs(F[j]) / std::max(std::fabs(T[j]), 1e-30);
                if (T[j] <= 1e-18) { r = 0; }
                rmax = std::max(rmax, r);
                for (int k = 0; k < nc; ++k) {
                    double a = 0;
                    for (int i = 0; i < ns; ++i) if (isEq(i)) a += p->eq_stoich[i][j] * p->eq_stoich[i][k] * sp[i];
                    J[j][k] = LN10 * a;
                }
            }
            iters = it; res = rmax;
            if (rmax < p->eq_tol) { conv = true; break; }
            // components whose total is negligible are frozen (the PCF zero-total guard, part4:731)
            for (int j = 0; j < nc; ++j) if (T[j] <= 1e-18) { for (int k = 0; k < nc; ++k) J[j][k] = (k == j); F[j] = 0; }
            // Gaussian elimination with partial pivoting (NR_reserve.hh:414)
            double A[CLB_MAX_COMP][CLB_MAX_COMP + 1];
            for (int j = 0; j < nc; ++j) { for (int k = 0; k < nc; ++k) A[j][k] = J[j][k]; A[j][nc] = -F[j]; }
            bool sing = false;
            for (int k = 0; k < nc; ++k) {
                int pv = k; for (int r = k + 1; r < nc; ++r) if (std::fabs(A[r][k]) > std::fabs(A[pv][k])) pv = r;
                if (std::fabs(A[pv][k]) < 1e-300) { sing = true; break; }
                if (pv != k) for (int cI = 0; cI <= nc; ++cI) std::swap(A[k][cI], A[pv][cI]);
                for (int r = k + 1; r < nc; ++r) {
                    double fct = A[r][k] / A[k][k];
                    for (int cI = k; cI <= nc; ++cI) A[r][cI] -= fct * A[k][cI];
                }
            }
            if (sing) break;
            double d[CLB_MAX_COMP];
            for (int k = nc - 1; k >= 0; --k) { double s = A[k][nc]; for (int cI = k + 1; cI < nc; ++cI) s -= A[k][cI] * d[cI]; d[k] = s / A[k][k]; }
            double dmax = 0; for (int k = 0; k < nc; ++k) dmax = std::max(dmax, std::fabs(d[k]));
            double sc = dmax > 1.0 ? 1.0 / dmax : 1.0;
            for (int k = 0; k < nc; ++k) om[k] = std::max(std::min(om[k] + sc * d[k], MAX_LOG_C), MIN_LOG_C);
            iters = it + 1;
        }
        species(om, init, out);
    }
    // calculate_species_concentrations, part4:1195-1222
    void solve(const double *c, double *out, int &iters, double &res, bool &conv) const {
        if (nc == 0 || ns == 0) { for (int i = 0; i < ns; ++i) out[i] = c[i]; conv = true; iters = 0; res = 0; return; }
        double T[CLB_MAX_COMP];
        totals(c, T);
        if (p->eq_solver == CLB_EQ_NEWTON) newton(c, T, out, iters, res, conv);
        else anderson(c, T, out, iters, res, conv);
        for (int i = 0; i < ns; ++i) {
            if (!std::isfinite(out[i])) out[i] = c[i] > MIN_CONC ? std::min(c[i], MAX_CONC) : MIN_CONC;
            out[i] = std::max(std::min(out[i], MAX_CONC), MIN_CONC);
        }
    }
};

/* run_equilibrium_biotic (part4:1359-1560), limiter at 1472-1485 */
void equilibriumStep(Case &c) {
    const clb_case_params &p = c.p;
    Eq eq{p.n_species, p.eq_nc, &p};
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        if (c.solidOrBB(c.mask[i])) continue;
        double C[CLB_MAX_SPECIES], E[CLB_MAX_SPECIES];
        for (int s = 0; s < p.n_species; ++s) C[s] = std::max(conc(c, s, i), MIN_CONC);
        int it; double res; bool conv;
        eq.solve(C, E, it, res, conv);
        if (!conv) {
#pragma omp atomic
            c.counters[3]++;
        }
        for (int s = 0; s < p.n_species; ++s) {
            double dC = E[s] - C[s];
            double mx = std::max(1e-9, 0.1 * std::fabs(C[s]));
            if (std::fabs(dC) > mx) dC = dC > 0 ? mx : -mx;
            if (C[s] + dC < 1e-20) dC = 1e-20 - C[s];
            if (std::fabs(dC) > THRD) deposit(c, s, i, dC);
        }
    }
}

/* ------------------------------------------------------------------------------------
 *  CA biofilm spill
 * ------------------------------------------------------------------------------------ */
/* updateLocalMaskNtotalLattices3D (part2:71-259); one pore code, so the
 * biofilm->pore relabel needs no neighbour */
void maskTotalUpdate(Case &c) {
    const clb_case_params &p = c.p;
    const double thrdB = p.thrd_bfilm_frac * p.bmax;
    for (size_t i = 0; i < c.N; ++i) {
        int mask = c.mask[i];
        if (c.solidOrBB(mask)) continue;
        double bm = 0; int nm = 0;
        for (int m = 0; m < p.n_microbes; ++m) { double v = c.B[m][i]; if (v > COMPLAB_THRD) { bm += v; nm += CLB_BIOFILM + m; } }
        if (std::fabs(bm - c.Btot[i]) > COMPLAB_THRD) c.Btot[i] = bm;
        bool pore = c.isPore(mask);
        if (!pore && bm < thrdB) c.mask[i] = CLB_PORE;
        else if (pore && bm >= thrdB && nm > 0) c.mask[i] = nm;
    }
}

/* the spill's random choices: splitmix64 of (seed, sweep, voxel, draw) */
uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
uint32_t caDraw(uint64_t seed, uint64_t sweep, uint64_t vox, uint32_t k) {
    uint64_t h = mix64(seed ^ mix64(sweep * 0x100000001b3ULL + 0x51ed27ULL));
    h = mix64(h ^ (vox * 0x9e3779b97f4a7c15ULL));
    h = mix64(h + k);
    return (uint32_t) (h >> 32);
}

/* pushExcessBiomass3D (part1:500-732) / halfPushExcessBiomass3D (735-970) +
 * pullExcessBiomass3D (973-1150), restated as a two-phase rule: every
 * over-full voxel decides its moves from the sweep's snapshot of mask, total
 * and age and writes them to a six-direction outbox; every voxel then gathers
 * what its neighbours sent.  The reference's serial in-place sweep with
 * std::rand has no reproducible outcome; this one does. */
void caSweep(Case &c) {
    const clb_case_params &p = c.p;
    const int M = p.n_microbes;
    const size_t N = c.N;
    std::vector<double> out((size_t) M * 6 * N, 0.0);
    const uint64_t sweep = c.caSweepCounter++;
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        int mask = c.mask[i];
        if (c.solidOrBB(mask)) continue;
        if (p.ca_method == CLB_CA_HALF && c.isPore(mask)) continue;
        double bMt = c.Btot[i];
        if (!(bMt > p.bmax)) continue;
        double bMd;
        if (p.ca_method == CLB_CA_HALF) bMd = bMt * 0.5;
        else { bMd = bMt - p.bmax; if (!(bMd > THRD)) continue; }
        // candidate neighbours in the order +x,-x,+y,-y,+z,-z (part1:537-548), bdryGap = 1
        int cand[6]; int nc = 0;
        for (int d = 0; d < 6; ++d) {
            int q = d + 1;
            int X = x + C7[q][0], Y = y + C7[q][1], Z = z + C7[q][2];
            if (q == 1 && x == c.nx - 2) continue;
            if (q == 2 && x == 1) continue;
            if (q == 3 && y == c.ny - 1) continue;
            if (q == 4 && y == 0) continue;
            if (q == 5 && z == c.nz - 1) continue;
            if (q == 6 && z == 0) continue;
            if (!c.inside(X, Y, Z)) continue;
            if (c.solidOrBB(c.mask[c.id(X, Y, Z)])) continue;
            cand[nc++] = d;
        }
        if (nc == 0) continue;                                   // [CA-ROBUST] leave it at the cap
        uint32_t k = 0;
        int perm[6]; for (int a = 0; a < nc; ++a) perm[a] = a;
        for (int a = nc - 1; a > 0; --a) { int b = (int) (caDraw(p.ca_seed, sweep, i, k++) % (uint32_t) (a + 1)); std::swap(perm[a], perm[b]); }
        double Bl[CLB_MAX_MICROBES]; for (int m = 0; m < M; ++m) Bl[m] = c.B[m][i];
        bool done = false;
        for (int a = 0; a < nc; ++a) {
            int d = cand[perm[a]], q = d + 1;
            size_t j = c.id(x + C7[q][0], y + C7[q][1], z + C7[q][2]);
            int nm = c.mask[j];
            double nb = c.isPore(nm) ? 0.0 : c.Btot[j];
            if (nb < p.bmax) {
                double cap = p.bmax - nb, part = bMd;
                if (bMd > cap) { part = cap; bMd -= cap; } else { bMd = 0; done = true; }
                for (int m = 0; m < M; ++m) {
                    double sh = Bl[m] / bMt * part;
                    if (sh > THRD) { Bl[m] -= sh; out[((size_t) m * 6 + d) * N + i] += sh; }
                }
                bMt -= part;
            }
            if (done) break;
        }
        if (!done) {
            int id0 = c.age[i];
            int young[6], ny = 0;
            for (int a = 0; a < nc; ++a) {
                int q = cand[a] + 1;
                int id1 = c.age[c.id(x + C7[q][0], y + C7[q][1], z + C7[q][2])];
                if (id0 > id1) young[ny++] = cand[a];
            }
            int d;
            if (ny > 1) d = young[caDraw(p.ca_seed, sweep, i, k++) % (uint32_t) ny];
            else if (ny == 1) d = young[0];
            else d = cand[nc > 1 ? caDraw(p.ca_seed, sweep, i, k++) % (uint32_t) nc : 0];
            for (int m = 0; m < M; ++m) {
                double sh = Bl[m] / bMt * bMd;
                Bl[m] -= sh; out[((size_t) m * 6 + d) * N + i] += sh;
            }
        }
        for (int m = 0; m < M; ++m) c.B[m][i] = Bl[m];
    }
    // gather, in the fixed order +x,-x,+y,-y,+z,-z of the sender's direction
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        for (int m = 0; m < M; ++m) {
            double add = 0;
            for (int d = 0; d < 6; ++d) {
                int q = d + 1;
                int X = x - C7[q][0], Y = y - C7[q][1], Z = z - C7[q][2];
                if (!c.inside(X, Y, Z)) continue;
                add += out[((size_t) m * 6 + d) * N + c.id(X, Y, Z)];
            }
            if (add != 0.0) c.B[m][i] += add;
        }
    }
}

/* updateAgeDistance3D (part2:840-965): one sweep in place, in the
 * processor's loop order (x outermost, z innermost, part2:862-867), so a voxel
 * reads the ages its -x, -y and -z face neighbours already took this sweep and
 * the old ages of its +x, +y and +z neighbours. */
void ageSweep(Case &c) {
    const double Bmax = c.p.bmax;
    for (int x = 0; x < c.nx; ++x) for (int y = 0; y < c.ny; ++y) for (int z = 0; z < c.nz; ++z) {
        size_t i = c.id(x, y, z);
        double B = c.Btot[i];
        if (!(B > COMPLAB_THRD)) continue;
        int del[6]; int nd = 0;
        for (int q = 1; q < Q7; ++q) {
            int X = x + C7[q][0], Y = y + C7[q][1], Z = z + C7[q][2];
            if (!c.inside(X, Y, Z)) continue;
            if (c.dist[c.id(X, Y, Z)] > 0) del[nd++] = q;
        }
        int a = c.age[i];
        int na = a; bool upd = true;
        if (a == 0) na = 1;
        else if (a == 1 && (B - Bmax) > -COMPLAB_THRD) {
            for (int t = 0; t < nd; ++t) {
                size_t j = c.id(x + C7[del[t]][0], y + C7[del[t]][1], z + C7[del[t]][2]);
                if (c.dist[j] > 0 && c.age[j] == 0) { upd = false; break; }
            }
            if (upd) na = 2;
        } else {
            int cnt = 0;
            for (int t = 0; t < nd; ++t) {
                size_t j = c.id(x + C7[del[t]][0], y + C7[del[t]][1], z + C7[del[t]][2]);
                if (c.age[j] >= a && (c.Btot[j] - Bmax) > -COMPLAB_THRD) ++cnt;
            }
            if (cnt == nd) na = a + 1; else upd = false;
        }
        if (upd) c.age[i] = na;
    }
}

long long sumAge(const Case &c) { long long s = 0; for (int v : c.age) s += v; return s; }
long long sumMask(const Case &c) { long long s = 0; for (int v : c.mask) s += v; return s; }
double maxBtot(const Case &c) { double m = -std::numeric_limits<double>::infinity(); for (double v : c.Btot) { if (std::isnan(v)) return v; m = std::max(m, v); } return m; }

/* the CA block of the main loop, complab.cpp:2793-3055 */
int caBlock(Case &c, int *sweepsOut) {
    const clb_case_params &p = c.p;
    int sweeps = 0;
    maskTotalUpdate(c);
    double gmax = maxBtot(c);
    if (!std::isfinite(gmax)) { c.err = "non-finite biomass"; return CLB_E_NONFINITE; }
    if (gmax - p.bmax > THRD) {
        double prev = gmax; int stall = 0; int wc = 0;
        while (gmax - p.bmax > THRD) {
            caSweep(c);
            maskTotalUpdate(c);
            gmax = maxBtot(c);
            ++sweeps; c.counters[6]++;
            if (!std::isfinite(gmax)) { c.err = "non-finite biomass"; return CLB_E_NONFINITE; }
            if (gmax > prev - THRD) { if (++stall >= 100) { c.err = "CA spill stalled"; if (sweepsOut) *sweepsOut = sweeps; return CLB_E_CA_STALL; } }
            else stall = 0;
            prev = gmax;
            if (wc % 50 == 0) {
                long long diff = 1; int wc1 = 0;
                while (diff != 0) {
                    long long o = sumAge(c);
                    ageSweep(c);
                    diff = sumAge(c) - o;
                    if (diff != 0 && ++wc1 > 1000) { c.err = "age relaxation did not settle"; return CLB_E_AGE_SWEEPS; }
                }
            }
            if (wc > 1000) { c.err = "CA spill did not settle"; return CLB_E_CA_SWEEPS; }
            ++wc;
        }
    }
    if (sweepsOut) *sweepsOut = sweeps;
    return CLB_OK;
}

/* ------------------------------------------------------------------------------------
 *  D3Q19 flow (IncBGKdynamics + BounceBack + NoDynamics + regularized pressure planes)
 * ------------------------------------------------------------------------------------ */
inline double feqInc(int q, double rb, const double *j, double jsq) {
    double cj = C19[q][0]*j[0] + C19[q][1]*j[1] + C19[q][2]*j[2];
    return W19[q] * (rb + 3.0 * cj + 4.5 * cj * cj - 1.5 * jsq);
}

/* regularized pressure plane (Latt et al. 2008) for incompressible BGK:
 * normal momentum from the imposed density, tangential momentum zero, the
 * unknown populations by bounce-back of their non-equilibrium part, then all
 * 19 rebuilt from (rho, j, Pi_neq). dir = +1 at x=0 (unknowns have c_x=+1). */
void nsPressurePlane(double *fl, int dir, double rhoBar, double *jout) {
    double r0 = 0, rout = 0;
    for (int q = 0; q < Q19; ++q) {
        if (C19[q][0] == 0) r0 += fl[q];
        else if (C19[q][0] == -dir) rout += fl[q];
    }
    double j[3] = {dir * (rhoBar - r0 - 2.0 * rout), 0.0, 0.0};
    double jsq = j[0] * j[0];
    double feq[19];
    for (int q = 0; q < Q19; ++q) feq[q] = feqInc(q, rhoBar, j, jsq);
    for (int q = 0; q < Q19; ++q) if (C19[q][0] == dir) fl[q] = feq[q] + (fl[opp19(q)] - feq[opp19(q)]);
    double Pi[3][3] = {{0,0,0},{0,0,0},{0,0,0}};
    for (int q = 0; q < Q19; ++q) {
        double ne = fl[q] - feq[q];
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) Pi[a][b] += C19[q][a] * C19[q][b] * ne;
    }
    for (int q = 0; q < Q19; ++q) {
        double s = 0;
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) s += (C19[q][a] * C19[q][b] - (a == b ? 1.0 / 3.0 : 0.0)) * Pi[a][b];
        fl[q] = feq[q] + W19[q] * 4.5 * s;
    }
    jout[0] = j[0]; jout[1] = 0; jout[2] = 0;
}

double nsRhoBarLeft() { return 0.0; }
double nsRhoBarRight(const Case &c) { return -c.deltaP * 3.0; }   // 1 - deltaP*invCs2 (complab_functions.hh:339)

/* one collideAndStream; returns the mean of |j|^2/2 over the collided fluid cells
 * (the statistic getStoredAverageEnergy reports) */
double nsCollideStream(Case &c) {
    const size_t N = c.N;
    // the energy of each cell, summed afterwards in cell order, so that threads
    // do not change the sum
    std::vector<double> ecell(N, 0.0);
    std::vector<char> ehas(N, 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        int cls = c.nsCls[i];
        double fl[19]; for (int q = 0; q < Q19; ++q) fl[q] = c.g[(size_t) q * N + i];
        /* a no-dynamics grain cell reflects like a bounce-back cell: Palabos would let
         * populations stream through it, which leaks flow into grain interiors
         * along the D3Q19 edge links wherever a grain shell is only face-closed */
        if (cls == NS_BB || cls == NS_NODYN) { for (int q = 1; q < Q19; q += 2) std::swap(fl[q], fl[q + 1]); }
        else {
            double om = cls == NS_BIO ? c.nsBioOmega : c.nsOmega;
            double rb = 0, j[3] = {0,0,0};
            if (x == 0) { rb = nsRhoBarLeft(); nsPressurePlane(fl, +1, rb, j); }
            else if (x == c.nx - 1) { rb = nsRhoBarRight(c); nsPressurePlane(fl, -1, rb, j); }
            else {
                for (int q = 0; q < Q19; ++q) { rb += fl[q]; for (int a = 0; a < 3; ++a) j[a] += C19[q][a] * fl[q]; }
            }
            double jsq = j[0]*j[0] + j[1]*j[1] + j[2]*j[2];
            ecell[i] = 0.5 * jsq; ehas[i] = 1;
            for (int q = 0; q < Q19; ++q) fl[q] += om * (feqInc(q, rb, j, jsq) - fl[q]);
        }
        for (int q = 0; q < Q19; ++q) c.g[(size_t) q * N + i] = fl[q];
    }
    double esum = 0; long long en = 0;
    for (size_t i = 0; i < N; ++i) if (ehas[i]) { esum += ecell[i]; ++en; }
    std::vector<double> ng(19 * N);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        for (int q = 0; q < Q19; ++q) {
            int xs = x - C19[q][0], ys = y - C19[q][1], zs = z - C19[q][2];
            ng[(size_t) q * N + i] = c.inside(xs, ys, zs) ? c.g[(size_t) q * N + c.id(xs, ys, zs)]
                                                         : c.g[(size_t) opp19(q) * N + i];
        }
    }
    c.g.swap(ng);
    return en ? esum / (double) en : 0.0;
}

/* velocity a cell reports (computeVelocity): j for BGK cells, the boundary
 * momentum on the pressure planes, zero for bounce-back / no-dynamics */
void nsVelocity(const Case &c, size_t i, int x, double *uo) {
    int cls = c.nsCls[i];
    uo[0] = uo[1] = uo[2] = 0;
    if (cls == NS_BB || cls == NS_NODYN) return;
    double fl[19]; for (int q = 0; q < Q19; ++q) fl[q] = c.g[(size_t) q * c.N + i];
    if (x == 0) { nsPressurePlane(fl, +1, nsRhoBarLeft(), uo); return; }
    if (x == c.nx - 1) { nsPressurePlane(fl, -1, nsRhoBarRight(c), uo); return; }
    for (int q = 0; q < Q19; ++q) for (int a = 0; a < 3; ++a) uo[a] += C19[q][a] * fl[q];
}

/* NSdomainSetup dynamics choice (complab_functions.hh:296-345) from the static
 * geometry: pore BGK, wall bounce-back, solid no-dynamics, biofilm codes BGK
 * with omega_bio or bounce-back when perm_ratio <= 1e-14 */
void nsClassesFromGeometry(Case &c) {
    for (size_t i = 0; i < c.N; ++i) {
        int g = c.geom[i];
        if (g == CLB_WALL) c.nsCls[i] = NS_BB;
        else if (g == CLB_SOLID) c.nsCls[i] = NS_NODYN;
        else if (g == CLB_PORE) c.nsCls[i] = NS_BGK;
        else c.nsCls[i] = c.p.perm_ratio > COMPLAB_THRD ? NS_BIO : NS_BB;
    }
}

/* initializeAtEquilibrium with PressureGradient (complab_functions.hh:82-94, 341):
 * rho = 1 - deltaP*3/(nx-1)*x, u = 0; bounce-back and no-dynamics cells hold 0 */
void nsInit(Case &c, double deltaP) {
    c.deltaP = deltaP;
    nsClassesFromGeometry(c);
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        double rb = -deltaP * 3.0 / (double) (c.nx - 1) * (double) x;
        int cls = c.nsCls[i];
        for (int q = 0; q < Q19; ++q) c.g[(size_t) q * c.N + i] = (cls == NS_BB || cls == NS_NODYN) ? 0.0 : W19[q] * rb;
    }
}

/* updateNsLatticesDynamics3D (part2:748-818): only open voxels change.  The
 * flow lattice is between two collideAndStream calls here, so its cells hold
 * streamed populations.  A cell that goes from bounce-back / no-dynamics to
 * BGK keeps what its open neighbours streamed into it; the populations that
 * arrived from closed neighbours are set to rest (0 in the shifted form),
 * which is what those closed-closed links hold from initialisation. */
void nsUpdateDynamics(Case &c, double permRatio) {
    const double bioOm = 1.0 / (permRatio * (1.0 / c.nsOmega - 0.5) + 0.5);
    c.nsBioOmega = bioOm;
    std::vector<int8_t> old = c.nsCls;
    for (size_t i = 0; i < c.N; ++i) {
        int mask = c.mask[i];
        if (c.solidOrBB(mask)) continue;
        bool curIsNsOmega = (c.nsCls[i] == NS_BGK);
        if (!c.isPore(mask) && curIsNsOmega) c.nsCls[i] = permRatio <= COMPLAB_THRD ? NS_BB : NS_BIO;
        else if (c.isPore(mask) && !curIsNsOmega) c.nsCls[i] = NS_BGK;
    }
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        bool wasClosed = old[i] == NS_BB || old[i] == NS_NODYN;
        bool isOpen = c.nsCls[i] == NS_BGK || c.nsCls[i] == NS_BIO;
        if (!(wasClosed && isOpen)) continue;
        c.g[i] = 0.0;
        for (int q = 1; q < Q19; ++q) {
            int xs = x - C19[q][0], ys = y - C19[q][1], zs = z - C19[q][2];
            bool fromOpen = false;
            if (c.inside(xs, ys, zs)) { int cl = old[c.id(xs, ys, zs)]; fromOpen = (cl == NS_BGK || cl == NS_BIO); }
            if (fromOpen) continue;
            c.g[(size_t) q * c.N + i] = 0.0;                       // both directions, as rebuildReopened
            if (c.inside(xs, ys, zs)) c.g[(size_t) opp19(q) * c.N + c.id(xs, ys, zs)] = 0.0;
        }
    }
}

void nsCouple(Case &c) {
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
        size_t i = c.id(x, y, z);
        double uo[3]; nsVelocity(c, i, x, uo);
        for (int a = 0; a < 3; ++a) c.u[(size_t) a * c.N + i] = uo[a];
    }
}

int nsSolve(Case &c, int maxIter, int rule, double eps, int reset, int *iters, int *converged) {
    if (reset) { c.tracer.reset(); c.lastE = -1; }
    int it = 0; bool conv = false;
    for (; it < maxIter; ) {
        double E = nsCollideStream(c);
        ++it;
        if (!std::isfinite(E)) { if (iters) *iters = it; if (converged) *converged = 0; c.err = "non-finite flow energy"; return CLB_E_NONFINITE; }
        if (rule == CLB_CONV_TRACER) { c.tracer.take(E); if (c.tracer.converged(eps)) { conv = true; break; } }
        else { if (c.lastE > 0 && std::fabs(E - c.lastE) <= eps * std::fabs(E)) { c.lastE = E; conv = true; break; } c.lastE = E; }
    }
    if (iters) *iters = it;
    if (converged) *converged = conv ? 1 : 0;
    return CLB_OK;
}

/* reductions of complab.cpp:1136-1157 */
void nsStats(Case &c, double out[4]) {
    double sumN = 0, maxUx = -1e300, sumOut = 0, sumPB = 0; long long nBox = 0, nOut = 0, nPB = 0;
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        double uo[3]; nsVelocity(c, i, x, uo);
        double n = std::sqrt(uo[0]*uo[0] + uo[1]*uo[1] + uo[2]*uo[2]);
        sumN += n; ++nBox;
        maxUx = std::max(maxUx, uo[0]);
        if (x == c.nx - 2) { sumOut += uo[0]; ++nOut; }
        int g = c.geom[i];
        if (g >= CLB_PORE) { sumPB += n; ++nPB; }
    }
    out[0] = sumN / (double) nBox; out[1] = maxUx; out[2] = sumOut / (double) nOut;
    out[3] = nPB ? sumPB / (double) nPB : 0.0;
}

/* ------------------------------------------------------------------------------------
 *  precipitation (precipitationVOP.hh:166-227) and the dynamics switches
 * ------------------------------------------------------------------------------------ */
void soluteSolidSwitch(Case &c) {          // updateSoluteSolidDynamics3D, part2:585-647
    for (size_t i = 0; i < c.N; ++i) {
        int mask = c.mask[i];
        bool want = c.solidOrBB(mask);
        bool isBB = c.adeCls[i] == ADE_BB;
        if (want && !isBB) c.adeCls[i] = ADE_BB;
        else if (!want && isBB) { c.adeCls[i] = c.isPore(mask) ? ADE_PORE : ADE_BFILM; c.reopened.push_back(i); }
    }
}
/* a voxel reopened during this step keeps, after the stream, the populations
 * its open neighbours sent; on every link to a neighbour closed at the
 * collide the rest value w_i c0 (which closed-closed links hold from set-up)
 * is set in both directions: the population arriving now, and the one the
 * closed neighbour returns at the next stream (the neighbour's stored
 * population that came from this voxel while it was closed).  A neighbour
 * open at the collide that closes in this same step (sealed by precipitation)
 * keeps its population toward this voxel, which it returns at the next
 * stream.  Parity unpinned: Palabos would return the closed neighbour's
 * stored population itself. */
void rebuildReopened(Case &c, int s) {
    for (size_t i : c.reopened) {
        int x = (int) (i % c.nx), y = (int) ((i / c.nx) % c.ny), z = (int) (i / ((size_t) c.nx * c.ny));
        c.f[s][i] = W7[0] * c.p.c0[s];
        for (int q = 1; q < Q7; ++q) {
            int xs = x - C7[q][0], ys = y - C7[q][1], zs = z - C7[q][2];
            bool fromOpen = c.inside(xs, ys, zs) && c.adeClsCollide[c.id(xs, ys, zs)] != ADE_BB
                            && !(xs == 0 && c.p.left_btype[s] == CLB_BC_CLOSED) && !(xs == c.nx - 1 && c.p.right_btype[s] == CLB_BC_CLOSED);
            if (fromOpen) continue;
            c.f[s][(size_t) q * c.N + i] = W7[q] * c.p.c0[s];
            if (c.inside(xs, ys, zs)) c.f[s][(size_t) OPP7[q] * c.N + c.id(xs, ys, zs)] = W7[q] * c.p.c0[s];
        }
    }
}
void soluteOmegaSwitch(Case &c) {          // updateSoluteDynamics3D, part2:485-553
    for (size_t i = 0; i < c.N; ++i) {
        int mask = c.mask[i];
        if (c.solidOrBB(mask)) continue;
        if (c.adeCls[i] == ADE_BB) continue;
        c.adeCls[i] = c.isPore(mask) ? ADE_PORE : ADE_BFILM;
    }
}
bool precipConvert(Case &c) {
    const clb_case_params &p = c.p;
    bool changed = false;
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        if (c.solidOrBB(c.mask[i])) continue;
        double P = conc(c, p.precip_species, i);
        if (P >= p.precip_max_rho) {
            c.mask[i] = CLB_SOLID;
            if (p.precip_phase_id > 0) c.phase[i] = (int8_t) p.precip_phase_id;
            changed = true;
        }
    }
    return changed;
}

/* dissolNodeConversion3D (dissolutionVOP.hh:744-812) over the reaction box: a
 * voxel of the phase whose mineral is below reopen_fraction x full_density
 * turns pore and loses its phase; returns the voxels reopened */
int dissolConvert(Case &c) {
    const clb_case_params &p = c.p;
    int n = 0;
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        size_t i = c.id(x, y, z);
        const int q = phaseIndex(p, c.phase[i]);
        if (q < 0 || !(p.dissol_full[q] > 0.0)) continue;
        if (c.imm[p.dissol_species[q]][i] < p.dissol_reopen_frac * p.dissol_full[q]) {
            c.mask[i] = CLB_PORE;
            c.phase[i] = 0;
            ++n;
        }
    }
    return n;
}

/* an in-step flow re-solve (complab.cpp:3087-3093, 3162-3168, 3221-3226): the
 * loop's own tracer ns_convg2, created empty before the main loop and never
 * reset, with the re-solve tolerance and cap */
int nsResolveSolve(Case &c, int *conv) {
    int it;
    const int rc = nsSolve(c, c.p.ns_max_iter, c.p.ns_conv_rule, c.p.ns_conv_eps, c.resolveFresh ? 1 : 0, &it, conv);
    c.resolveFresh = false;
    return rc;
}

/* re-solve after a mask change (maskPath, complab.cpp:3085-3130) or a
 * precipitation seal (3159-3178).  A re-solve that runs out of iterations
 * with no flow through the outlet plane x = nx-2 clogs the case; one whose
 * energy is not finite clogs it too on the mask-change path, where the
 * velocity is then not coupled (3119-3121) */
int nsResolve(Case &c, double permRatio, bool maskPath) {
    nsUpdateDynamics(c, permRatio);
    int conv = 0;
    int rc = nsResolveSolve(c, &conv);
    if (rc == CLB_E_NONFINITE && maskPath) { c.counters[4]++; c.clogged = 1; c.err.clear(); return CLB_OK; }
    if (rc) return rc;
    c.counters[4]++;
    if (!conv) {
        c.nsSaturate = 1;
        double st[4]; nsStats(c, st);
        if (st[2] > THRD) c.nsSaturate = 0;
        else c.clogged = 1;
    }
    if (!(maskPath && c.clogged)) nsCouple(c);
    return CLB_OK;
}

/* ------------------------------------------------------------------------------------
 *  one coupled step, complab.cpp:2178-3262
 * ------------------------------------------------------------------------------------ */
int coupledStep(Case &c) {
    const clb_case_params &p = c.p;
    if (c.rateOn) std::fill(c.rate.begin(), c.rate.end(), 0.0);
    c.adeClsCollide = c.adeCls;
    for (int s = 0; s < p.n_species; ++s) if (p.mobile[s]) adeCollide(c, s);
    if ((p.bio_law != CLB_BIO_NONE || symActive(c)) && p.n_microbes > 0) bioticStep(c);
    if (p.n_rxn > 0 || p.dissol_on || c.symAbOn) abioticStep(c);
    if (p.eq_on) equilibriumStep(c);
    int rc = CLB_OK;
    if (p.n_microbes > 0) {
        int sw = 0;
        rc = caBlock(c, &sw);
        if (rc) return rc;
        long long nm = sumMask(c);
        if (nm != c.oldTotMask) {
            c.oldTotMask = nm;
            c.counters[5]++;
            ageSweep(c);
            if (p.ade_update_interval > 0 && c.iT % p.ade_update_interval == 0) soluteOmegaSwitch(c);
            if (p.ns_update_interval > 0 && c.iT % p.ns_update_interval == 0 && p.flow_on && !c.nsSaturate) {
                rc = nsResolve(c, p.perm_ratio, true);
                if (rc) return rc;
            }
        }
    }
    if (p.precip_on && p.precip_interval > 0 && c.iT % p.precip_interval == 0) {
        long long pre = sumMask(c);
        precipConvert(c);
        long long post = sumMask(c);
        if (pre != post) {
            c.counters[5]++;
            if (p.flow_on && !c.nsSaturate) { rc = nsResolve(c, p.precip_perm_ratio, false); if (rc) return rc; }
            soluteSolidSwitch(c);
        }
    }
    if (p.dissol_on && p.dissol_interval > 0 && c.iT % p.dissol_interval == 0) {  // complab.cpp:3203-3247
        const int n = dissolConvert(c);
        if (n > 0) {
            c.counters[5]++;
            c.counters[11] += n;
            if (p.flow_on) {
                nsUpdateDynamics(c, p.perm_ratio);
                c.nsSaturate = 0;
                int conv;
                rc = nsResolveSolve(c, &conv);
                if (rc) return rc;
                c.counters[4]++;
                nsCouple(c);
                c.clogged = 0;                          // reopening can undo a clog (complab.cpp:3234)
            }
            soluteSolidSwitch(c);
        }
    }
    for (int s = 0; s < p.n_species; ++s) if (p.mobile[s]) { adeStream(c, s); rebuildReopened(c, s); }
    c.reopened.clear();
    c.iT++; c.counters[7]++;
    return c.clogged ? CLB_CLOGGED : CLB_OK;            // the loop ends after this step's stream (complab.cpp:3261)
}

/* only transport + kinetics: collide, biotic, abiotic, stream */
void transportStep(Case &c) {
    const clb_case_params &p = c.p;
    if (c.rateOn) std::fill(c.rate.begin(), c.rate.end(), 0.0);
    for (int s = 0; s < p.n_species; ++s) if (p.mobile[s]) adeCollide(c, s);
    if ((p.bio_law != CLB_BIO_NONE || symActive(c)) && p.n_microbes > 0) bioticStep(c);
    if (p.n_rxn > 0 || p.dissol_on || c.symAbOn) abioticStep(c);
    for (int s = 0; s < p.n_species; ++s) if (p.mobile[s]) adeStream(c, s);
}

/* ------------------------------------------------------------------------------------
 *  geometry (complab3d_geometry.hh)
 * ------------------------------------------------------------------------------------ */
struct Rng {                                   // complab3d_geometry.hh:77-86
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ? seed : 1) {}
    uint64_t next() { s ^= s >> 12; s ^= s << 25; s ^= s >> 27; return s * 2685821657736338717ULL; }
    double uniform() { return (double) (next() >> 11) / 9007199254740992.0; }
    int below(int n) { return n > 0 ? (int) (uniform() * (double) n) : 0; }
};

}  // namespace

/* ======================================================================================
 *  C entry points (orc_*), used by the tests through ctypes
 * ====================================================================================== */
extern "C" {

void *orc_case_create(const clb_case_params *p, const int8_t *geometry, const double *biomass0) {
    Case *c = new Case();
    c->p = *p;
    c->nx = p->nx; c->ny = p->ny; c->nz = p->nz;
    c->N = (size_t) c->nx * c->ny * c->nz;
    const size_t N = c->N;
    c->geom.assign(geometry, geometry + N);
    c->mask.resize(N); c->age.resize(N); c->dist.resize(N); c->adeCls.resize(N); c->nsCls.resize(N);
    c->phase.assign(N, 0);
    for (size_t i = 0; i < N; ++i) c->mask[i] = geometry[i];
    // age: createAgeDomain3D (part3:347-367)
    for (size_t i = 0; i < N; ++i) { int g = geometry[i]; c->age[i] = (g == CLB_SOLID || g == CLB_WALL) ? -1 : (g == CLB_PORE ? 0 : 1); }
    // dist > 0 exactly on open voxels of x<nx-1, y<ny-1, z<nz-1 (calculateDistanceFromSolid loops, complab_functions.hh:226-235)
    for (int z = 0; z < c->nz; ++z) for (int y = 0; y < c->ny; ++y) for (int x = 0; x < c->nx; ++x) {
        size_t i = c->id(x, y, z); int g = geometry[i];
        bool in = x < c->nx - 1 && y < c->ny - 1 && z < c->nz - 1;
        c->dist[i] = (in && g != CLB_SOLID && g != CLB_WALL) ? 1 : (g == CLB_SOLID && in ? -1 : 0);
    }
    // solute dynamics (complab_functions.hh:347-442, complab.cpp:1614-1656): solids and walls bounce back
    for (size_t i = 0; i < N; ++i) { int g = geometry[i]; c->adeCls[i] = (g == CLB_SOLID || g == CLB_WALL) ? ADE_BB : (g == CLB_PORE ? ADE_PORE : ADE_BFILM); }
    c->f.resize(p->n_species); c->imm.resize(p->n_species);
    c->rhoL.resize(p->n_species); c->rhoR.resize(p->n_species);
    for (int s = 0; s < p->n_species; ++s) {
        double c0 = p->c0[s];
        if (p->mobile[s]) {
            c->f[s].resize(7 * N);
            for (int q = 0; q < Q7; ++q) for (size_t i = 0; i < N; ++i) c->f[s][(size_t) q * N + i] = W7[q] * c0;
        } else {
            c->imm[s].assign(N, c0);
        }
        c->rhoL[s].assign((size_t) c->ny * c->nz, p->left_btype[s] == CLB_BC_DIRICHLET ? p->left_bc[s] : c0);
        c->rhoR[s].assign((size_t) c->ny * c->nz, p->right_btype[s] == CLB_BC_DIRICHLET ? p->right_bc[s] : c0);
    }
    // dissolution (complab.cpp:1667-1699): each voxel takes the id of the first
    // grain phase on its material (initPhaseLattice3D over the grain phases),
    // then every grain phase seeds its inventory on its material, in order
    if (p->dissol_on) {
        c->share.assign((size_t) p->n_species * N, 0.0);
        for (size_t i = 0; i < N; ++i)
            for (int q = 0; q < p->dissol_n_phases; ++q)
                if (!p->dissol_is_precipitate[q] && geometry[i] == p->dissol_material[q]) { c->phase[i] = (int8_t) (q + 1); break; }
        for (int q = 0; q < p->dissol_n_phases; ++q) {
            if (p->dissol_is_precipitate[q] || !(p->dissol_initial[q] > 0.0)) continue;
            for (size_t i = 0; i < N; ++i)
                if (geometry[i] == p->dissol_material[q]) c->imm[p->dissol_species[q]][i] = p->dissol_initial[q];
        }
    }
    c->B.resize(p->n_microbes);
    c->Btot.assign(N, 0.0);
    for (int m = 0; m < p->n_microbes; ++m) {
        c->B[m].assign(N, 0.0);
        for (size_t i = 0; i < N; ++i) {
            if (biomass0) c->B[m][i] = biomass0[(size_t) m * N + i];
            else if (geometry[i] == CLB_BIOFILM + m) c->B[m][i] = p->b0[m];   // initializeScalarLattice3D
        }
    }
    for (size_t i = 0; i < N; ++i) { double s = 0; for (int m = 0; m < p->n_microbes; ++m) s += c->B[m][i]; c->Btot[i] = s; }
    c->g.assign(19 * N, 0.0);
    c->u.assign(3 * N, 0.0);
    c->nsOmega = 1.0 / p->ns_tau;
    c->nsBioOmega = 1.0 / (p->perm_ratio * (1.0 / c->nsOmega - 0.5) + 0.5);
    nsClassesFromGeometry(*c);
    c->oldTotMask = sumMask(*c);
    return c;
}
void orc_case_destroy(void *h) { delete (Case *) h; }
const char *orc_last_error(void *h) { return ((Case *) h)->err.c_str(); }

int orc_set_omegas(void *h, const double *omp, const double *omb, double dt) {
    Case &c = *(Case *) h;
    for (int s = 0; s < c.p.n_species; ++s) { c.p.omega_pore[s] = omp[s]; c.p.omega_biofilm[s] = omb[s]; }
    c.p.dt = dt;
    return 0;
}

int orc_ns_init(void *h, double deltaP) { nsInit(*(Case *) h, deltaP); return 0; }
int orc_ns_solve(void *h, int maxIter, int rule, double eps, int reset, int *iters, int *conv) {
    return nsSolve(*(Case *) h, maxIter, rule, eps, reset, iters, conv);
}
int orc_ns_couple(void *h) { nsCouple(*(Case *) h); return 0; }
int orc_ns_stats(void *h, double *out) { nsStats(*(Case *) h, out); return 0; }
int orc_ns_update_dynamics(void *h, double permRatio) { nsUpdateDynamics(*(Case *) h, permRatio); return 0; }

/* stabilization pre-loop + reset (complab.cpp:2033-2041, part3:261-310) */
int orc_ade_stabilize(void *h, int steps) {
    Case &c = *(Case *) h;
    for (int t = 0; t < steps; ++t)
        for (int s = 0; s < c.p.n_species; ++s) if (c.p.mobile[s]) { adeCollide(c, s); adeStream(c, s); }
    for (int s = 0; s < c.p.n_species; ++s) {
        if (!c.p.mobile[s]) continue;
        double c0 = c.p.c0[s]; if (c0 < THRD && c0 > -THRD) c0 = 0;
        for (size_t i = 0; i < c.N; ++i) if (c.geom[i] >= CLB_PORE)
            for (int q = 0; q < Q7; ++q) c.f[s][(size_t) q * c.N + i] = W7[q] * c0;
    }
    return 0;
}

int orc_run_steps(void *h, int n, int *done) {
    Case &c = *(Case *) h;
    int k = 0, rc = 0;
    if (c.clogged) { if (done) *done = 0; return CLB_CLOGGED; }
    for (; k < n; ++k) { rc = coupledStep(c); if (rc) { if (rc == CLB_CLOGGED) ++k; break; } }
    if (done) *done = k;
    return rc;
}
int orc_ade_steps(void *h, int n) { Case &c = *(Case *) h; for (int k = 0; k < n; ++k) transportStep(c); return 0; }
int orc_equilibrium(void *h) { equilibriumStep(*(Case *) h); return 0; }
int orc_ca_spill(void *h, int *sweeps) { return caBlock(*(Case *) h, sweeps); }
/* one push + gather sweep of the CA rule, without the mask/total update */
int orc_ca_sweep(void *h) { caSweep(*(Case *) h); return 0; }
int orc_precip_convert(void *h) { return precipConvert(*(Case *) h) ? 1 : 0; }

int orc_read_field(void *h, int field, int index, double *dst) {
    Case &c = *(Case *) h;
    const size_t N = c.N;
    switch (field) {
    case CLB_FIELD_CONC:
        if (index < 0 || index >= c.p.n_species) return CLB_E_ARG;
        for (size_t i = 0; i < N; ++i) dst[i] = conc(c, index, i);
        return 0;
    case CLB_FIELD_BIOMASS: if (index < 0 || index >= c.p.n_microbes) return CLB_E_ARG; for (size_t i = 0; i < N; ++i) dst[i] = c.B[index][i]; return 0;
    case CLB_FIELD_BTOTAL: for (size_t i = 0; i < N; ++i) dst[i] = c.Btot[i]; return 0;
    case CLB_FIELD_MASK: for (size_t i = 0; i < N; ++i) dst[i] = c.mask[i]; return 0;
    case CLB_FIELD_VEL: if (index < 0 || index > 2) return CLB_E_ARG; for (size_t i = 0; i < N; ++i) dst[i] = c.u[(size_t) index * N + i]; return 0;
    case CLB_FIELD_AGE: for (size_t i = 0; i < N; ++i) dst[i] = c.age[i]; return 0;
    case CLB_FIELD_PHASE: for (size_t i = 0; i < N; ++i) dst[i] = c.phase[i]; return 0;
    case CLB_FIELD_DG:
    case CLB_FIELD_FT:                                                  /* computeEnergyField3D */
        if (!c.p.thm_on || index < 0 || index >= c.p.n_microbes) return CLB_E_ARG;
        for (size_t i = 0; i < N; ++i) {
            dst[i] = 0.0;
            const int x = (int) (i % (size_t) c.nx);
            if (x == 0 || x == c.nx - 1 || c.solidOrBB(c.mask[i])) continue;
            double C[CLB_MAX_SPECIES], dG = 0.0;
            for (int s = 0; s < c.p.n_species; ++s) { const double v = conc(c, s, i); C[s] = v > 0.0 ? v : 0.0; }
            const double F = thermoFactor(c.p, C, &dG);
            dst[i] = field == CLB_FIELD_FT ? F : dG;
        }
        return 0;
    case CLB_FIELD_RATE:                                                /* writeRateVTI(..., 1/dt) */
        if (!c.rateOn || index < 0 || index >= c.p.n_species) return CLB_E_ARG;
        for (size_t i = 0; i < N; ++i) dst[i] = c.rate[(size_t) index * N + i] * (1.0 / c.p.dt);
        return 0;
    case CLB_FIELD_NSRHO:
        /* computeDensity: the post-stream populations of a BGK cell, the
         * imposed density of a pressure-plane cell, 1 at closed cells */
        for (size_t i = 0; i < N; ++i) {
            const int x = (int) (i % (size_t) c.nx);
            double r = 0;
            if (c.nsCls[i] == NS_BB || c.nsCls[i] == NS_NODYN) r = 0;
            else if (x == 0) r = nsRhoBarLeft();
            else if (x == c.nx - 1) r = nsRhoBarRight(c);
            else for (int q = 0; q < Q19; ++q) r += c.g[(size_t) q * N + i];
            dst[i] = 1.0 + r;
        }
        return 0;
    }
    return CLB_E_ARG;
}
int orc_write_field(void *h, int field, int index, const double *src) {
    Case &c = *(Case *) h;
    const size_t N = c.N;
    if (field == CLB_FIELD_CONC) {
        if (index < 0 || index >= c.p.n_species) return CLB_E_ARG;
        if (c.p.mobile[index]) {
            /* open voxels take the field at rest; bounce-back voxels keep what they
             * hold (the background c0 they were seeded with) */
            for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
                size_t i = c.id(x, y, z);
                if (adeClassFor(c, index, x, i) == ADE_BB) continue;
                for (int q = 0; q < Q7; ++q) c.f[index][(size_t) q * N + i] = W7[q] * src[i];
            }
            // a Neumann plane imposes the density its inner layer holds now
            for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) {
                size_t k = (size_t) y + (size_t) c.ny * z;
                size_t il = c.id(1, y, z), ir = c.id(c.nx - 2, y, z);
                if (c.p.left_btype[index] == CLB_BC_NEUMANN) c.rhoL[index][k] = adeClassFor(c, index, 1, il) == ADE_BB ? 0.0 : src[il];
                if (c.p.right_btype[index] == CLB_BC_NEUMANN) c.rhoR[index][k] = adeClassFor(c, index, c.nx - 2, ir) == ADE_BB ? 0.0 : src[ir];
            }
        }
        else for (size_t i = 0; i < N; ++i) c.imm[index][i] = src[i];
        return 0;
    }
    if (field == CLB_FIELD_BIOMASS) {
        if (index < 0 || index >= c.p.n_microbes) return CLB_E_ARG;
        for (size_t i = 0; i < N; ++i) c.B[index][i] = src[i];
        for (size_t i = 0; i < N; ++i) { double s = 0; for (int m = 0; m < c.p.n_microbes; ++m) s += c.B[m][i]; c.Btot[i] = s; }
        return 0;
    }
    if (field == CLB_FIELD_VEL) {
        if (index < 0 || index > 2) return CLB_E_ARG;
        for (size_t i = 0; i < N; ++i) c.u[(size_t) index * N + i] = src[i];
        return 0;
    }
    if (field == CLB_FIELD_AGE) { for (size_t i = 0; i < N; ++i) c.age[i] = (int) std::lround(src[i]); return 0; }
    if (field == CLB_FIELD_MASK) { for (size_t i = 0; i < N; ++i) c.mask[i] = (int) std::lround(src[i]); return 0; }
    return CLB_E_ARG;
}
/* the age relaxation of complab.cpp:3025-3033 (to the fixed point) or one sweep */
int orc_age_relax(void *h, int toFixedPoint, int *sweeps) {
    Case &c = *(Case *) h;
    int k = 0;
    while (true) {
        long long o = sumAge(c);
        ageSweep(c);
        ++k;
        if (!toFixedPoint || sumAge(c) == o) break;
        if (k > 100000) return CLB_E_AGE_SWEEPS;
    }
    if (sweeps) *sweeps = k;
    return 0;
}
int orc_species_totals(void *h, double *out) {
    Case &c = *(Case *) h;
    for (int s = 0; s < c.p.n_species; ++s) {
        double t = 0;
        for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
            size_t i = c.id(x, y, z);
            if (!c.p.mobile[s] || !c.solidOrBB(c.mask[i])) t += conc(c, s, i);   // a mineral stays in the voxels it sealed
        }
        out[s] = t;
    }
    return 0;
}
/* diagnostics row (complab.cpp:2736-2785): per species then per microbe the
 * total over the open voxels of x = 1..nx-2, the minimum and maximum there, and
 * held -- what the lattice holds beyond the total: PopulationSum3D over the
 * whole box (part3:505-536) counts the x = 0 and x = nx-1 cells and the
 * populations a bounce-back cell holds.  Here held is the content of the
 * plane cells open to the species plus, in every closed cell, the populations
 * that arrived in the last stream from a cell open at its collide; the rest
 * values closed cells keep among themselves from set-up are left out (they
 * never meet an open cell).  Immobile species 0; microbes: biomass in the two
 * planes. */
int orc_field_stats(void *h, double *out) {
    Case &c = *(Case *) h;
    const int S = c.p.n_species;
    for (int r = 0; r < S + c.p.n_microbes; ++r) {
        double tot = 0, lo = 1e300, hi = -1e300, held = 0;
        for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
            const size_t i = c.id(x, y, z);
            const double v = r < S ? conc(c, r, i) : c.B[r - S][i];
            if (x > 0 && x < c.nx - 1) {
                const bool mob = r < S && c.p.mobile[r];
                if (mob && (c.solidOrBB(c.mask[i]) || adeClassFor(c, r, x, i) == ADE_BB)) continue;
                tot += v; lo = std::min(lo, v); hi = std::max(hi, v);
            }
        }
        if (r >= S) {
            for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) {
                held += c.B[r - S][c.id(0, y, z)];
                held += c.B[r - S][c.id(c.nx - 1, y, z)];
            }
        } else if (c.p.mobile[r]) {
            const double *f = c.f[r].data();
            for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
                const size_t i = c.id(x, y, z);
                if (adeClassFor(c, r, x, i) != ADE_BB) {
                    if (x == 0 || x == c.nx - 1) held += densityF(c, r, i);
                    continue;
                }
                for (int q = 1; q < Q7; ++q) {
                    const int xs = x - C7[q][0], ys = y - C7[q][1], zs = z - C7[q][2];
                    if (!c.inside(xs, ys, zs)) continue;
                    const size_t k = c.id(xs, ys, zs);
                    bool srcOpen;
                    if (xs == 0 && c.p.left_btype[r] == CLB_BC_CLOSED) srcOpen = false;
                    else if (xs == c.nx - 1 && c.p.right_btype[r] == CLB_BC_CLOSED) srcOpen = false;
                    else srcOpen = (c.adeClsCollide.empty() ? c.adeCls[k] : c.adeClsCollide[k]) != ADE_BB;
                    if (srcOpen) held += f[(size_t) q * c.N + i];
                }
            }
        }
        if (hi < lo) { lo = 0; hi = 0; }
        out[4 * r + 0] = tot; out[4 * r + 1] = lo; out[4 * r + 2] = hi; out[4 * r + 3] = held;
    }
    return 0;
}
int orc_counters(void *h, int64_t *out) {
    Case &c = *(Case *) h;
    for (int k = 0; k < CLB_NCOUNTERS; ++k) out[k] = c.counters[k];
    int64_t open = 0, bio = 0;                       // MaskedScalarCounts3D over the reaction box
    for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 1; x < c.nx - 1; ++x) {
        const int m = c.mask[c.id(x, y, z)];
        if (m >= CLB_PORE) ++open;
        if (m >= CLB_BIOFILM) ++bio;
    }
    out[0] = open; out[1] = bio;
    return 0;
}

/* one equilibrium solve on a composition (for the golden comparisons) */
int orc_eq_solve(const clb_case_params *p, const double *cin, double *cout, int *iters, double *res, int *conv) {
    Eq eq{p->n_species, p->eq_nc, p};
    double C[CLB_MAX_SPECIES];
    for (int s = 0; s < p->n_species; ++s) C[s] = cin[s];
    int it; double r; bool cv;
    eq.solve(C, cout, it, r, cv);
    if (iters) *iters = it;
    if (res) *res = r;
    if (conv) *conv = cv ? 1 : 0;
    return 0;
}
/* run_equilibrium_biotic on one voxel (limiter included) */
int orc_eq_voxel(const clb_case_params *p, const double *cin, double *cout) {
    Eq eq{p->n_species, p->eq_nc, p};
    double C[CLB_MAX_SPECIES], E[CLB_MAX_SPECIES];
    for (int s = 0; s < p->n_species; ++s) C[s] = std::max(cin[s], MIN_CONC);
    int it; double r; bool cv;
    eq.solve(C, E, it, r, cv);
    for (int s = 0; s < p->n_species; ++s) {
        double dC = E[s] - C[s];
        double mx = std::max(1e-9, 0.1 * std::fabs(C[s]));
        if (std::fabs(dC) > mx) dC = dC > 0 ? mx : -mx;
        if (C[s] + dC < 1e-20) dC = 1e-20 - C[s];
        cout[s] = cin[s] + (std::fabs(dC) > THRD ? dC : 0.0);
    }
    return 0;
}

/* the biotic rates of microbe m from a symbolic law (NULL: back to the compiled
 * law); m = CLB_SYM_ABIOTIC sets or clears the abiotic law */
int orc_set_symbolic(void *h, int m, const clb_sym_program *prog) {
    Case &c = *(Case *) h;
    if (m == CLB_SYM_ABIOTIC) {                                         // the abiotic law
        if (prog) { c.symAb = *prog; c.symAbOn = true; }
        else { memset(&c.symAb, 0, sizeof(clb_sym_program)); c.symAbOn = false; }
        return 0;
    }
    if (m < 0 || m >= c.p.n_microbes) return CLB_E_ARG;
    if (c.sym.size() < (size_t) c.p.n_microbes) c.sym.resize((size_t) c.p.n_microbes);
    if (prog) { c.sym[(size_t) m] = *prog; c.symMask |= 1u << m; }
    else { memset(&c.sym[(size_t) m], 0, sizeof(clb_sym_program)); c.symMask &= ~(1u << m); }
    return 0;
}

/* rate capture on or off (the increments of each following step, readable as
 * CLB_FIELD_RATE) */
int orc_rate_capture(void *h, int on) {
    Case &c = *(Case *) h;
    c.rateOn = on != 0;
    if (c.rateOn) c.rate.assign((size_t) std::max(c.p.n_species, 1) * c.N, 0.0);
    return 0;
}

/* one law's outputs at one composition, for the tests */
int orc_sym_evaluate(const clb_sym_program *prog, const double *avail, double bmass, double *out) {
    symEvaluate(*prog, avail, bmass, out);
    return 0;
}

/* the thermodynamic factor and dG of one composition, for the tests */
int orc_thermo_factor(const clb_case_params *p, const double *C, double *F, double *dG) {
    *F = thermoFactor(*p, C, dG);
    return 0;
}

/* a kinetics evaluation on one voxel (biotic then abiotic law), for the tests */
int orc_rates(const clb_case_params *p, const double *C, const double *B, int mask, double *subsR, double *bioR, double *abioR) {
    bioticLaw(*p, C, B, subsR, bioR, mask);
    abioticLaw(*p, C, abioR);
    return 0;
}

/* threads of the per-cell loops (n <= 0 leaves them); returns the count in use.
 * Results do not depend on it: every loop writes per cell, and the sums run
 * in cell order on one thread */
int orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); return omp_get_max_threads(); }
int orc_rng_draws(uint64_t seed, int n, uint64_t *out) { Rng r(seed); for (int k = 0; k < n; ++k) out[k] = r.next(); return 0; }

/* complab_geom::generate "spheres" (complab3d_geometry.hh:138-160) + addWalls (89-102) */
int orc_geom_spheres(int nx, int ny, int nz, double radius, int count, uint64_t seed, int walls_yz, int8_t *out) {
    const size_t N = (size_t) nx * ny * nz;
    for (size_t i = 0; i < N; ++i) out[i] = CLB_PORE;
    Rng rng(seed);
    for (int k = 0; k < count; ++k) {
        int cx = rng.below(nx), cy = rng.below(ny), cz = rng.below(nz);
        for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x) {
            double dx = x - cx, dy = y - cy, dz = (double) (z - cz);
            if (dx * dx + dy * dy + dz * dz <= radius * radius) out[(size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z)] = CLB_SOLID;
        }
    }
    if (walls_yz)
        for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x)
            if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1) out[(size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z)] = CLB_WALL;
    return 0;
}

/* the sphere-pack builder of the configurations: spheres from the Rng until the
 * open fraction is <= porosity, walls on y/z, grain shell -> wall, pockets not
 * connected to both x faces -> wall, percolation required; lattice with the
 * duplicated ghost x columns (readGeometry, complab_functions.hh:178-207) */
int orc_geom_pack(int nxf, int ny, int nz, double radius, double porosity, uint64_t seed, int kind, int8_t *lat, double *phi) {
    (void) kind;
    const size_t Nf = (size_t) nxf * ny * nz;
    std::vector<int8_t> g(Nf, CLB_PORE);
    auto at = [&](int x, int y, int z) -> int8_t & { return g[(size_t) x + (size_t) nxf * ((size_t) y + (size_t) ny * z)]; };
    for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nxf; ++x)
        if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1) at(x, y, z) = CLB_WALL;
    long long open = 0; for (auto v : g) open += v == CLB_PORE;
    Rng rng(seed);
    const int R = (int) std::ceil(radius);
    int guard = 0;
    while ((double) open / (double) Nf > porosity && guard < 10000000) {
        ++guard;
        int cx = rng.below(nxf), cy = rng.below(ny), cz = rng.below(nz);
        for (int z = std::max(0, cz - R); z <= std::min(nz - 1, cz + R); ++z)
            for (int y = std::max(0, cy - R); y <= std::min(ny - 1, cy + R); ++y)
                for (int x = std::max(0, cx - R); x <= std::min(nxf - 1, cx + R); ++x) {
                    double dx = x - cx, dy = y - cy, dz = z - cz;
                    if (dx * dx + dy * dy + dz * dz <= radius * radius && at(x, y, z) == CLB_PORE) { at(x, y, z) = CLB_SOLID; --open; }
                }
    }
    // keep only the cluster spanning x = 0 .. nxf-1 (6-connected flood fill)
    std::vector<int> lab(Nf, 0);
    std::vector<size_t> st;
    int cur = 0;
    std::vector<char> span(1, 0);
    for (size_t s0 = 0; s0 < Nf; ++s0) {
        if (g[s0] != CLB_PORE || lab[s0]) continue;
        ++cur; bool inl = false, outl = false;
        st.clear(); st.push_back(s0); lab[s0] = cur;
        while (!st.empty()) {
            size_t i = st.back(); st.pop_back();
            int x = (int) (i % nxf), y = (int) ((i / nxf) % ny), z = (int) (i / ((size_t) nxf * ny));
            if (x == 0) inl = true;
            if (x == nxf - 1) outl = true;
            for (int q = 1; q < Q7; ++q) {
                int X = x + C7[q][0], Y = y + C7[q][1], Z = z + C7[q][2];
                if (X < 0 || Y < 0 || Z < 0 || X >= nxf || Y >= ny || Z >= nz) continue;
                size_t j = (size_t) X + (size_t) nxf * ((size_t) Y + (size_t) ny * Z);
                if (g[j] == CLB_PORE && !lab[j]) { lab[j] = cur; st.push_back(j); }
            }
        }
        span.push_back(inl && outl);
    }
    bool any = false;
    for (size_t i = 0; i < Nf; ++i) if (g[i] == CLB_PORE) { if (!span[lab[i]]) g[i] = CLB_WALL; else any = true; }
    // grain shell: a solid voxel face-adjacent to pore becomes bounce-back
    std::vector<int8_t> h = g;
    for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nxf; ++x) {
        if (at(x, y, z) != CLB_SOLID) continue;
        for (int q = 1; q < Q7; ++q) {
            int X = x + C7[q][0], Y = y + C7[q][1], Z = z + C7[q][2];
            if (X < 0 || Y < 0 || Z < 0 || X >= nxf || Y >= ny || Z >= nz) continue;
            if (at(X, Y, Z) == CLB_PORE) { h[(size_t) x + (size_t) nxf * ((size_t) y + (size_t) ny * z)] = CLB_WALL; break; }
        }
    }
    g.swap(h);
    long long op = 0; for (auto v : g) op += v == CLB_PORE;
    if (phi) *phi = (double) op / (double) Nf;
    const int nx = nxf + 2;
    for (int z = 0; z < nz; ++z) for (int y = 0; y < ny; ++y) for (int x = 0; x < nx; ++x) {
        int xs = x == 0 ? 0 : (x == nx - 1 ? nxf - 1 : x - 1);
        lat[(size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z)] = at(xs, y, z);
    }
    return any ? CLB_OK : CLB_E_GEOMETRY;
}

}  // extern "C"

/* the convergence rule alone, for its test: the values fed one at a time to a
 * Tracer as the solve feeds the mean energies; *stop = the 1-based count at
 * which it first reports convergence, 0 if it never does */
extern "C" int orc_tracer_stop(const double *vals, int n, double eps, int *stop) {
    Tracer t;
    *stop = 0;
    for (int k = 0; k < n; ++k) {
        t.take(vals[k]);
        if (t.converged(eps)) { *stop = k + 1; break; }
    }
    return 0;
}

/* layout of clb_case_params as this compiler lays it out, for the ctypes mirror */
#include <cstddef>
extern "C" int orc_params_layout(int64_t *out) {
    out[0] = (int64_t) sizeof(clb_case_params);
    out[1] = (int64_t) offsetof(clb_case_params, dt);
    out[2] = (int64_t) offsetof(clb_case_params, bio_stoich);
    out[3] = (int64_t) offsetof(clb_case_params, ca_seed);
    out[4] = (int64_t) offsetof(clb_case_params, eq_stoich);
    out[5] = (int64_t) offsetof(clb_case_params, precip_on);
    out[6] = (int64_t) offsetof(clb_case_params, dissol_stoich);
    out[7] = (int64_t) offsetof(clb_case_params, thm_nu);
    return 0;
}

// clb_lattice.hip -- the lattice kernels of the coupled step on the compact
// open-voxel layout (gfx950).
//
// Only open voxels (geometry pore or biofilm) carry populations, one thread
// each, in three ranges: interior, x = 0 plane, x = nx-1 plane (clb_device.h).
// Populations are structure-of-arrays: f[slot][q][i] (ADE, full form, density
// = sum f) and g[q][i] (D3Q19, shifted form, density = 1 + sum g).
//
// Streaming is a pull: a launch reads the post-collision populations of the
// previous launch (`cur`) from its neighbours, through a neighbour table, and
// writes its own post-collision populations to `out`.  Bounce-back cells
// have no storage and no thread: a population that would come from one is the
// voxel's own opposite population from two launches back, still in `out`
// until this thread overwrites it.  That is the full-way bounce-back of a
// Palabos BounceBack cell (swap on collide, stream back out) without the cell.
// A launch pulls with the classes of the previous launch (`clsPull`) and
// collides with the current ones (`cls`), so a voxel that closes or reopens
// between two launches streams exactly as the reference's in-step change does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clb_device.h"
#include "clb_common.h"

namespace clb {

// ------------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------------
// The dispatcher deals workgroups round-robin over the 8 XCDs, each with its
// own L2.  Renumbering the blocks so that XCD x runs one contiguous run of the
// voxel order keeps the partly-used cache lines of the neighbour gathers (and
// the +-y/+-z neighbours themselves) in the L2 that fetched them.  A bijection
// of [0, gridDim.x) whatever the dispatch order.
__device__ __forceinline__ int xcdBlockOf(int b, int G) {
    const int q = G >> 3, r = G & 7, x = b & 7, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
__device__ __forceinline__ int xcdBlock() { return xcdBlockOf(blockIdx.x, gridDim.x); }

// compact neighbour n (>= 0) closed for a species: bounce-back class, or a
// closed x plane (complab_functions.hh:431-432)
__device__ __forceinline__ bool closedFor(const uint8_t *cls, int n, int Nint, int Nleft, int lbt, int rbt) {
    if (cls[n] == ADE_BB) return true;
    if (n >= Nint) return (n < Nint + Nleft) ? (lbt == BC_CLOSED_DEV) : (rbt == BC_CLOSED_DEV);
    return false;
}

template <typename R, typename T>
__device__ __forceinline__ void pull7(const AdeDev<R> &a, const R *__restrict__ cur, const R *__restrict__ old,
                                      int i, int lbt, int rbt, T fl[7]) {
    const long long No = a.Nopen;
    if (a.first) {
#pragma unroll
        for (int q = 0; q < 7; ++q) fl[q] = (T) cur[pidx<7>(q, i)];
        return;
    }
    fl[0] = (T) cur[pidx<7>(0, i)];
#pragma unroll
    for (int q = 1; q < 7; ++q) {
        const int n = a.nb[(q - 1) * No + i];
        if (n == NB_OUTSIDE) fl[q] = (T) cur[pidx<7>(opp7(q), i)];            // edge: own opposite
        else if (n == NB_CLOSED || closedFor(a.clsPull, n, a.Nint, a.Nleft, lbt, rbt))
            fl[q] = (T) old[pidx<7>(opp7(q), i)];                             // bounce-back
        else fl[q] = (T) cur[pidx<7>(q, n)];
    }
}

// first launch after (re)initialisation: the state is pre-collision, and a
// closed neighbour holds the species' rest population w c0; hand it to the
// slot the bounce-back rule reads at the next launch.  After the
// stabilization reset (planesOnly) that holds for the closed x planes alone:
// the reset rewrites every pore voxel, the planes' included, and leaves the
// walls with what they hold (stabilizeADElattice3D, part3:261-310).
template <typename R>
__device__ __forceinline__ void firstHandOver(const AdeDev<R> &a, R *cur, int i, int lbt, int rbt, double c0,
                                              bool planesOnly) {
    const long long No = a.Nopen;
#pragma unroll
    for (int q = 1; q < 7; ++q) {
        const int n = a.nb[(q - 1) * No + i];
        const bool closed = planesOnly ? (n >= a.Nint && closedFor(a.clsPull, n, a.Nint, a.Nleft, lbt, rbt))
                                       : (n == NB_CLOSED || (n >= 0 && closedFor(a.clsPull, n, a.Nint, a.Nleft, lbt, rbt)));
        if (closed) cur[pidx<7>(opp7(q), i)] = (R) (W7F * c0);
    }
}

// BGK with the first-order advection-diffusion equilibrium w_i rho (1 + c_i.u / cs^2),
// in the compute type T (double in FP64 mode, float in FP32 mode)
template <typename T>
__device__ __forceinline__ void bgk7(T f[7], T om, T ux, T uy, T uz) {
    T r = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) r += f[q];
    const T t0 = (T) 4 * r;
    const T c1 = (T) 1 - om;
    const T wr = (T) W7R, wf = (T) W7F;
    f[0] = f[0] * c1 + om * (wr * r);
    f[1] = f[1] * c1 + om * (wf * (r + t0 * ux));
    f[2] = f[2] * c1 + om * (wf * (r - t0 * ux));
    f[3] = f[3] * c1 + om * (wf * (r + t0 * uy));
    f[4] = f[4] * c1 + om * (wf * (r - t0 * uy));
    f[5] = f[5] * c1 + om * (wf * (r + t0 * uz));
    f[6] = f[6] * c1 + om * (wf * (r - t0 * uz));
}

// regularized first-order advection-diffusion Dirichlet completion
template <typename T>
__device__ __forceinline__ void adeDirichlet(T fl[7], int atLeft, T rhoD) {
    const int unk = atLeft ? 1 : 2;
    T s = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) if (q != unk) s += fl[q];
    fl[unk] = rhoD - s;
    const T jx = fl[1] - fl[2], jy = fl[3] - fl[4], jz = fl[5] - fl[6];
    const T wr = (T) W7R, wf = (T) W7F, c = (T) 4;
    fl[0] = wr * rhoD;
    fl[1] = wf * (rhoD + c * jx);
    fl[2] = wf * (rhoD - c * jx);
    fl[3] = wf * (rhoD + c * jy);
    fl[4] = wf * (rhoD - c * jy);
    fl[5] = wf * (rhoD + c * jz);
    fl[6] = wf * (rhoD - c * jz);
}

// ------------------------------------------------------------------------------------------
// kinetics (the rate laws of the reference's compiled-kinetics headers)
// ------------------------------------------------------------------------------------------
// the thermodynamic factor F_T of complab3d_thermo.hh:215-255 (Jin & Bethke):
// ln Q over the block's species with activities floored at cmin, dG, the
// driving force past the ATP threshold, 0 at or past it, 1 where exp underflows
__device__ double thermoGate(const KinDev &k, const double *C, double *dGout = nullptr) {
    double lnQ = 0.0;
    for (int s = 0; s < k.S; ++s) {
        if (k.thmNu[s] == 0.0) continue;
        double a = C[s] * k.thmScale;
        if (!(a > k.thmCmin)) a = k.thmCmin;
        lnQ += k.thmNu[s] * log(a);
    }
    const double dG = k.thmDG0 + k.thmRT * lnQ + k.thmLoss;
    if (dGout) *dGout = dG;
    const double x = (dG + k.thmThr) / ((k.thmChi > 0.0 ? k.thmChi : 1.0) * k.thmRT);
    if (x >= 0.0) return 0.0;
    if (x < -700.0) return 1.0;
    const double F = 1.0 - exp(x);
    return F < 0.0 ? 0.0 : (F > 1.0 ? 1.0 : F);
}

// dual Monod, config/kinetics/defineKinetics.biotic.hh:90-126; with the
// thermodynamic gate the whole rate vector is scaled by F_T (part1:127-140)
// (microbes in `skip` have a symbolic law and are left out)
__device__ void bioticLaw(const KinDev &k, const double *C, const double *B, double *sr, double *br, int mask,
                          unsigned int skip = 0u) {
    for (int s = 0; s < k.S; ++s) sr[s] = 0;
    for (int m = 0; m < k.M; ++m) br[m] = 0;
    if (mask < 2) return;
    const double S = k.donor >= 0 ? fmax(C[k.donor], 0.0) : 0.0;
    const double E = k.acceptor >= 0 ? fmax(C[k.acceptor], 0.0) : 1.0;
    for (int m = 0; m < k.M; ++m) {
        if ((skip >> m) & 1u) continue;
        const double Bm = fmax(B[m], 0.0);
        if (Bm <= 0.0) continue;
        double mu = k.mu_max[m] * (S / (k.ks[m] + S));
        if (k.acceptor >= 0) mu *= E / (k.ke[m] + E);
        const double growth = mu * Bm;
        br[m] += growth - k.kdecay[m] * Bm;
        for (int s = 0; s < k.S; ++s) sr[s] += k.bio_stoich[m][s] * growth;
    }
    for (int s = 0; s < k.S; ++s) if (!isfinite(sr[s])) sr[s] = 0;
    for (int m = 0; m < k.M; ++m) if (!isfinite(br[m])) br[m] = 0;
    if (k.thm) {
        const double F = thermoGate(k, C);
        if (F < 1.0) {
            for (int s = 0; s < k.S; ++s) sr[s] *= F;
            for (int m = 0; m < k.M; ++m) br[m] *= F;
        }
    }
}

// mass action with the ABC per-step cap, ABC_sweep_package/kinetics/defineAbioticKinetics.hh:166-210
__device__ void abioticLaw(const KinDev &k, const double *C, double *sr) {
    for (int s = 0; s < k.S; ++s) sr[s] = 0;
    for (int r = 0; r < k.nrxn; ++r) {
        const int r1 = k.r1[r], r2 = k.r2[r];
        const double a = r1 >= 0 ? fmax(C[r1], 0.0) : 1.0;
        const double b = r2 >= 0 ? fmax(C[r2], 0.0) : 1.0;
        if (k.min_conc > 0 && ((r1 >= 0 && a <= k.min_conc) || (r2 >= 0 && b <= k.min_conc))) continue;
        double R = k.rk[r] * a * b;
        if (k.cap_dt > 0 && k.cap_frac > 0) {
            const double sc = fmin(r1 >= 0 ? a : 1e300, r2 >= 0 ? b : 1e300);
            const double cap = sc * k.cap_frac / k.cap_dt;
            if (R > cap) R = cap;
        }
        for (int s = 0; s < k.S; ++s) sr[s] += k.rxn_stoich[r][s] * R;
    }
    for (int s = 0; s < k.S; ++s) if (!isfinite(sr[s])) sr[s] = 0;
    if (k.rate_scale != 1.0) for (int s = 0; s < k.S; ++s) sr[s] *= k.rate_scale;
}

// the clamp dC = max(R dt, -C), only for a debit and only against what is
// there; a bind is counted when the debit exceeds the availability by more
// than 1e-12 of it (one equal to it but for rounding is capped, not counted)
__device__ __forceinline__ double clampInc(int clamp, double d, double cur, unsigned int &binds) {
    if (!clamp || d >= 0) return d;
    const double lim = -fmax(cur, 0.0);
    if (d < lim) { if (d < lim - 1e-12 * fabs(lim)) ++binds; return lim; }
    return d;
}

// ------------------------------------------------------------------------------------------
// advection-diffusion
// ------------------------------------------------------------------------------------------
// one organism's symbolic law at one voxel (clb_sym_program; evalNode and
// evaluate of complab3d_symbolic.hh:310-348, 766-804): the variables, held in
// their ranges, then each rate's postfix code in file order
__device__ void symEvaluate(const clb_sym_program *P, const double *avail, double bmass, double *out) {
    double v[CLB_SYM_MAXVARS + CLB_SYM_MAXRATES];
    const int nv = P->n_vars, nr = P->n_rates;
    for (int i = 0; i < nv; ++i) {
        const int s = P->var_species[i];
        double x = s >= 0 ? avail[s] : bmass;
        if (P->hi[i] > P->lo[i]) {
            if (x < P->lo[i]) x = P->lo[i];
            else if (x > P->hi[i]) x = P->hi[i];
        }
        v[i] = x;
    }
    int pc = 0;
    for (int r = 0; r < nr; ++r) {
        double st[CLB_SYM_STACK];
        int sp = 0;
        for (; pc < P->rate_end[r]; ++pc) {
            const int op = P->op[pc], arg = P->arg[pc];
            if (op == CLB_SYM_CONST) { st[sp++] = P->konst[arg]; continue; }
            if (op == CLB_SYM_VAR) { st[sp++] = v[arg]; continue; }
            if (op == CLB_SYM_NEG) { st[sp - 1] = -st[sp - 1]; continue; }
            if (op <= CLB_SYM_POW || op >= CLB_SYM_MIN) {            // two operands
                const double y = st[--sp], x = st[sp - 1];
                double z;
                switch (op) {
                case CLB_SYM_ADD: z = x + y; break;
                case CLB_SYM_SUB: z = x - y; break;
                case CLB_SYM_MUL: z = x * y; break;
                case CLB_SYM_DIV: z = fabs(y) < 1e-300 ? 0.0 : x / y; break;
                case CLB_SYM_POW: z = (x < 0.0 && y != floor(y)) ? 0.0 : pow(x, y); break;
                case CLB_SYM_MIN: z = x < y ? x : y; break;
                default: z = x > y ? x : y; break;                    // CLB_SYM_MAX
                }
                st[sp - 1] = z;
                continue;
            }
            const double x = st[sp - 1];
            double z;
            switch (op) {
            case CLB_SYM_EXP: z = exp(x); break;
            case CLB_SYM_LOG: z = x > 0.0 ? log(x) : 0.0; break;
            case CLB_SYM_LOG10: z = x > 0.0 ? log10(x) : 0.0; break;
            case CLB_SYM_SQRT: z = x > 0.0 ? sqrt(x) : 0.0; break;
            case CLB_SYM_ABS: z = fabs(x); break;
            default: z = tanh(x); break;                              // CLB_SYM_TANH
            }
            st[sp - 1] = z;
        }
        double y = sp > 0 ? st[sp - 1] : 0.0;
        if (y != y) y = 0.0;
        if (y > 1e300) y = 1e300; else if (y < -1e300) y = -1e300;
        v[nv + r] = y;
        out[r] = y * P->unit_scale;
    }
}

// the symbolic organisms of one voxel (run_symbolic3D,
// complab3d_processors_symbolic.hh:104-182): every organism present evaluates
// its law on the voxel's chemistry; substrate rates x gate x dt are requests
// on one shared budget, a debit held to what the voxel has; the growth rate x
// gate x biomass x dt is the organism's increment.  dC and dB come out as
// increments of the step.
__device__ void symbioticIncrements(const KinDev &k, const double *C, const double *B, double *dC, double *dB) {
    const double thrd = 1e-12;
    double avail[CLB_DEV_MAXS], req[CLB_DEV_MAXS], bm[CLB_DEV_MAXM];
    for (int s = 0; s < k.S; ++s) { avail[s] = C[s] > thrd ? C[s] : 0.0; req[s] = 0.0; dC[s] = 0.0; }
    bool any = false;
    for (int m = 0; m < k.M; ++m) { bm[m] = B[m] > thrd ? B[m] : 0.0; dB[m] = 0.0; if (bm[m] > thrd) any = true; }
    if (!any) return;
    for (int m = 0; m < k.M; ++m) {
        if (!(bm[m] > thrd)) continue;
        const clb_sym_program *P = k.sym + m;
        if (P->n_rates <= 0) continue;
        double out[CLB_SYM_MAXRATES];
        symEvaluate(P, avail, bm[m], out);
        const double ft = (k.thm && (k.thmMicrobe < 0 || k.thmMicrobe == m)) ? thermoGate(k, avail) : 1.0;
        for (int r = 0; r < P->n_rates; ++r) {
            const int rs = P->rate_species[r];
            if (rs >= 0) req[rs] += out[r] * ft * k.dt;
            else dB[m] += out[r] * ft * bm[m] * k.dt;
        }
    }
    for (int s = 0; s < k.S; ++s) {
        double draw = req[s];
        if (fabs(draw) <= thrd) continue;
        const double headroom = avail[s] + dC[s];
        if (draw + headroom < 0.0) draw = -headroom;
        dC[s] += draw;
    }
}

// the abiotic symbolic law at one open voxel (run_symbolic_abiotic3D,
// complab3d_processors_abiotic_learned.hh:60-135): no biomass; each species'
// rate x dt is a request held to what the voxel has, added to the voxel's
// abiotic increment dAb when it is above the threshold
__device__ void symAbioticIncrements(const KinDev &k, const double *C, double *dAb) {
    const double thrd = 1e-12;
    const clb_sym_program *P = k.symAb;
    double avail[CLB_DEV_MAXS], req[CLB_DEV_MAXS], out[CLB_SYM_MAXRATES];
    for (int s = 0; s < k.S; ++s) { avail[s] = C[s] > thrd ? C[s] : 0.0; req[s] = 0.0; }
    symEvaluate(P, avail, 0.0, out);
    for (int r = 0; r < P->n_rates; ++r) {
        const int rs = P->rate_species[r];                           // -1: the extent line itself
        if (rs >= 0) req[rs] += out[r] * k.dt;
    }
    for (int s = 0; s < k.S; ++s) {
        double draw = req[s];
        if (fabs(draw) <= thrd) continue;
        if (draw + avail[s] < 0.0) draw = -avail[s];
        if (fabs(draw) > thrd) dAb[s] += draw;
    }
}

// a face neighbour of compact voxel i in the reaction box belongs to a
// declared phase: i collects that neighbour's parked products (the
// condition of the gather below and, seen from the phase voxel, the open
// neighbours whose water the park pass averages).  A voxel of a phase
// itself counts too: the park reads its mineral.
template <typename R>
__device__ __forceinline__ bool nextToPhase(const AdeDev<R> &a, const KinDev &k, int i) {
    const int own = k.phase[a.list[i]];
    bool near = own > 0 && own <= k.dPhases;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int n = a.nb[(opp7(d + 1) - 1) * a.Nopen + i];
        if (n < 0 || n >= a.Nint) continue;
        const int pid = k.phase[a.list[n]];
        if (pid > 0 && pid <= k.dPhases) near = true;
    }
    return near;
}

// biotic and abiotic kinetics of one open interior voxel on its post-collision
// populations f (every mobile species in registers): rates from the densities,
// the clamp, the D3Q7 deposit w_i dC, immobile species and biomass in place
// (run_kinetics + update_rxnLattices, part1:85-187, 406-497; the abiotic pair
// part1:199-404; surfaceAbioticKinetics3D's face gate)
template <typename R, int MAXS, bool DISSOL, bool SYM = false>
__device__ __forceinline__ void voxelKinetics(const AdeDev<R> &a, const KinDev &k, int i, R (&f)[MAXS][7],
                                              long long v, int mask, R (&bpre)[CLB_DEV_MAXM], bool havePre) {
    const long long No = a.Nopen;
    if (mask != 0 && mask != 1) {
        // species loops run to the static bound so that the slot selection in
        // concOf / depositTo resolves to fixed registers (a dynamic trip count
        // left f indexed at run time: 256 VGPRs instead of 164 at 4 species)
        unsigned int binds = 0;
        double C[CLB_DEV_MAXS], sr[CLB_DEV_MAXS];
        const int S = k.S;
        const long long N = a.N;
        auto concOf = [&](int s) -> double {
            const int sl = a.slot[s];
            if (a.mobile[s]) {
                double r = 0;
#pragma unroll
                for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == sl) {
#pragma unroll
                    for (int q = 0; q < 7; ++q) r += (double) f[m2][q];
                }
                return r;
            }
            return (double) a.imm[(long long) sl * N + v];
        };
        auto depositTo = [&](int s, double d) {
            const int sl = a.slot[s];
            if (a.mobile[s]) {
#pragma unroll
                for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == sl) {
#pragma unroll
                    for (int q = 0; q < 7; ++q) f[m2][q] = (R) ((double) f[m2][q] + w7(q) * d);
                }
            } else {
                R *p = a.imm + (long long) sl * N + v;
                *p = (R) ((double) *p + d);
            }
        };
        // the step's requested increments, kept for the rate capture
        // (accumulateRates, complab.cpp:2101-2105: the biotic and the abiotic
        // increments before the clamp; the symbolic variant only)
        double dReq[CLB_DEV_MAXS];
#pragma unroll
        for (int s = 0; s < CLB_DEV_MAXS; ++s) dReq[s] = 0.0;
        if (k.bio && k.M > 0) {
            double Bv[CLB_DEV_MAXM], br[CLB_DEV_MAXM];
            
#pragma unroll
            for (int s = 0; s < CLB_DEV_MAXS; ++s) { if (s < S) { C[s] = concOf(s); if (C[s] < 1e-12) C[s] = 0; } }
#pragma unroll
            for (int m = 0; m < CLB_DEV_MAXM; ++m) {
                if (m >= k.M) continue;
                Bv[m] = havePre ? (double) bpre[m] : (double) a.B[(long long) m * N + v];
                if (Bv[m] < 1e-12) Bv[m] = 0;
            }
            // the step's increments: rate x dt of the compiled law and, in the
            // symbolic variant, the symbolic organisms' increments after their
            // shared budget, each kept when above the threshold (run_kinetics
            // then run_symbolic3D into the same increments, complab.cpp:2348-2393)
            if (SYM) {
                double sc[CLB_DEV_MAXS], sb[CLB_DEV_MAXM];
                if (k.compiled) bioticLaw(k, C, Bv, sr, br, mask, k.symMask);
                else {
#pragma unroll
                    for (int s = 0; s < CLB_DEV_MAXS; ++s) sr[s] = 0.0;
#pragma unroll
                    for (int m = 0; m < CLB_DEV_MAXM; ++m) br[m] = 0.0;
                }
                symbioticIncrements(k, C, Bv, sc, sb);
                auto kept = [](double x) { return (x > 1e-12 || x < -1e-12) ? x : 0.0; };
#pragma unroll
                for (int s = 0; s < CLB_DEV_MAXS; ++s) if (s < S) sr[s] = kept(sr[s] * k.dt) + kept(sc[s]);
#pragma unroll
                for (int m = 0; m < CLB_DEV_MAXM; ++m) if (m < k.M) br[m] = kept(br[m] * k.dt) + kept(sb[m]);
            } else {
                bioticLaw(k, C, Bv, sr, br, mask);
#pragma unroll
                for (int s = 0; s < CLB_DEV_MAXS; ++s) if (s < S) sr[s] *= k.dt;
#pragma unroll
                for (int m = 0; m < CLB_DEV_MAXM; ++m) if (m < k.M) br[m] *= k.dt;
            }
#pragma unroll
            for (int s = 0; s < CLB_DEV_MAXS; ++s) {
                if (s >= S) continue;
                const double d = sr[s];
                if (SYM) dReq[s] = d;
                if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
            }
#pragma unroll
            for (int m = 0; m < CLB_DEV_MAXM; ++m) {
                if (m >= k.M) continue;
                double d = br[m];
                if (d > 1e-12 || d < -1e-12) {
                    R *p = a.B + (long long) m * N + v;
                    const double b = havePre ? (double) bpre[m] : (double) *p;
                    if (d < -b) d = -b;                               // biomass never negative
                    *p = (R) (b + d);
                    if (havePre) bpre[m] = (R) (b + d);
                }
            }
        }
        // the surface gate of the abiotic rates (precipitationVOP.hh:78-88) and,
        // in the transport launch of a dissolution step, whether the voxel is
        // next to the dissolving phase (nextToPhase): there the abiotic
        // increment waits for the parked products, elsewhere the gather adds
        // nothing and the increment is final here.  One pass over the faces.
        const bool wantNear = !DISSOL && k.nearPhase == 1;
        bool react = k.nrxn > 0, deferred = false;
        if (react && (k.surface_only || wantNear)) {
            bool surf = false, near = false;
            if (wantNear) { const int own = k.phase[v]; near = own > 0 && own <= k.dPhases; }
#pragma unroll
            for (int q = 1; q < 7; ++q) {
                const int n = a.nb[(opp7(q) - 1) * No + i];                // the voxel at +c_q
                if (n == NB_OUTSIDE) continue;
                if (n == NB_CLOSED) { surf = true; continue; }
                const long long vn = a.list[n];
                const int nm = a.mask[vn];
                if (nm == 0 || nm == 1) surf = true;
                if (wantNear && n < a.Nint) { const int pid = k.phase[vn]; if (pid > 0 && pid <= k.dPhases) near = true; }
            }
            if (k.surface_only) react = surf;
            deferred = react && near;
        }
        const bool symAb = SYM && k.symAb != nullptr;
        if ((k.nrxn > 0 || DISSOL || symAb) && !deferred) {
            // one increment per species: the abiotic rates, then (dissolution) the
            // product shares the face neighbours of the phase parked this step,
            // applied together (update_abiotic_rxnLattices, part1:306-404)
            double dAb[CLB_DEV_MAXS];
#pragma unroll
            for (int s = 0; s < CLB_DEV_MAXS; ++s) dAb[s] = 0.0;
            if (k.nrxn > 0) {
                if (react) {
#pragma unroll
                    for (int s = 0; s < CLB_DEV_MAXS; ++s) { if (s < S) { C[s] = concOf(s); if (C[s] < 1e-12) C[s] = 0; } }
                    abioticLaw(k, C, sr);
#pragma unroll
                    for (int s = 0; s < CLB_DEV_MAXS; ++s) {
                        if (s >= S) continue;
                        const double d = sr[s] * k.dt;
                        if (d > 1e-12 || d < -1e-12) dAb[s] += d;
                    }
                }
            }
            if (symAb) {                                                  // complab.cpp:2489-2493, same increments
#pragma unroll
                for (int s = 0; s < CLB_DEV_MAXS; ++s) { if (s < S) { C[s] = concOf(s); if (C[s] < 1e-12) C[s] = 0; } }
                symAbioticIncrements(k, C, dAb);
            }
            if (DISSOL) {                                                 // dissolutionGather3D, dissolutionVOP.hh:607-720
                unsigned int collected = 0, clamped = 0;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    const int n = a.nb[(opp7(d + 1) - 1) * No + i];        // the face neighbour at +dloc[d]
                    if (n < 0 || n >= a.Nint) continue;                    // outside the reaction box or closed geometry
                    const int pid = k.phase[a.list[n]];
                    if (pid <= 0 || pid > k.dPhases) continue;            // a declared phase parks
                    const int sM = k.dMineral[pid - 1];
#pragma unroll
                    for (int s = 0; s < CLB_DEV_MAXS; ++s) {
                        if (s >= S || s == sM) continue;                   // the mineral stays where it is
                        const double share = k.share[(long long) s * No + n];
                        if (fabs(share) <= 1e-14) continue;
                        double give = share;
                        if (give < 0.0) {
                            const double have = concOf(s);
                            const double headroom = (have > 0.0 ? have : 0.0) + dAb[s];
                            if (give + headroom < 0.0) { give = -headroom; ++clamped; }
                            if (give > 0.0) give = 0.0;
                        }
                        ++collected;
                        if (!(fabs(give) > 1e-14)) continue;
                        dAb[s] += give;
                    }
                }
                if (collected) atomicAdd(k.dcount + 1, (unsigned long long) collected);
                if (clamped) atomicAdd(k.dcount + 2, (unsigned long long) clamped);
            }
#pragma unroll
            for (int s = 0; s < CLB_DEV_MAXS; ++s) {
                if (s >= S) continue;
                const double d = dAb[s];
                if (SYM) dReq[s] += d;
                if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
            }
        }
        if (SYM && k.rate) {
#pragma unroll
            for (int s = 0; s < CLB_DEV_MAXS; ++s) if (s < S) k.rate[(long long) s * N + v] = dReq[s] * (1.0 / k.dt);
        }
        if (binds) atomicAdd(a.clampCount, (unsigned long long) binds);
    }
}

// the kinetics of the fused transport kernel: the increments of
// voxelKinetics<R, MAXS, false, false> bit for bit (the same operations in
// the same order), with each species' rate formed as it is deposited instead
// of in arrays over every species, so that the populations are about all the
// registers the kernel keeps live.  The thermodynamic gate reads the
// concentrations through `conc`.
template <typename F>
__device__ __forceinline__ double thermoGateOf(const KinDev &k, F conc) {
    double lnQ = 0.0;
#pragma unroll
    for (int s = 0; s < CLB_DEV_MAXS; ++s) {
        if (s >= k.S || k.thmNu[s] == 0.0) continue;
        double a = conc(s) * k.thmScale;
        if (!(a > k.thmCmin)) a = k.thmCmin;
        lnQ += k.thmNu[s] * log(a);
    }
    const double dG = k.thmDG0 + k.thmRT * lnQ + k.thmLoss;
    const double x = (dG + k.thmThr) / ((k.thmChi > 0.0 ? k.thmChi : 1.0) * k.thmRT);
    if (x >= 0.0) return 0.0;
    if (x < -700.0) return 1.0;
    const double Fv = 1.0 - exp(x);
    return Fv < 0.0 ? 0.0 : (Fv > 1.0 ? 1.0 : Fv);
}
template <typename R, int MAXS>
__device__ __forceinline__ void voxelKineticsFused(const AdeDev<R> &a, const KinDev &k, int i, R (&f)[MAXS][7],
                                                   long long v, int mask, R (&bpre)[CLB_DEV_MAXM]) {
    if (mask == 0 || mask == 1) return;
    const long long No = a.Nopen;
    unsigned int binds = 0;
    const int S = k.S;
    const long long N = a.N;
    auto concOf = [&](int s) -> double {
        const int sl = a.slot[s];
        if (a.mobile[s]) {
            double r = 0;
#pragma unroll
            for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == sl) {
#pragma unroll
                for (int q = 0; q < 7; ++q) r += (double) f[m2][q];
            }
            return r;
        }
        return (double) a.imm[(long long) sl * N + v];
    };
    auto depositTo = [&](int s, double d) {
        const int sl = a.slot[s];
        if (a.mobile[s]) {
#pragma unroll
            for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == sl) {
#pragma unroll
                for (int q = 0; q < 7; ++q) f[m2][q] = (R) ((double) f[m2][q] + w7(q) * d);
            }
        } else {
            R *p = a.imm + (long long) sl * N + v;
            *p = (R) ((double) *p + d);
        }
    };
    auto thr = [](double c) { return c < 1e-12 ? 0.0 : c; };
    if (k.bio && k.M > 0) {
        // bioticLaw: the growth and net biomass rate of each microbe
        double g[CLB_DEV_MAXM], br[CLB_DEV_MAXM];
        const double Sx = k.donor >= 0 ? fmax(thr(concOf(k.donor)), 0.0) : 0.0;
        const double Ex = k.acceptor >= 0 ? fmax(thr(concOf(k.acceptor)), 0.0) : 1.0;
#pragma unroll
        for (int m = 0; m < CLB_DEV_MAXM; ++m) {
            g[m] = 0.0; br[m] = 0.0;
            if (m >= k.M) continue;
            double Bv = (double) bpre[m];
            if (Bv < 1e-12) Bv = 0;
            const double Bm = fmax(Bv, 0.0);
            if (Bm <= 0.0) continue;
            double mu = k.mu_max[m] * (Sx / (k.ks[m] + Sx));
            if (k.acceptor >= 0) mu *= Ex / (k.ke[m] + Ex);
            const double growth = mu * Bm;
            br[m] += growth - k.kdecay[m] * Bm;
            g[m] = growth;
        }
        const double F = k.thm ? thermoGateOf(k, [&](int s) { return thr(concOf(s)); }) : 1.0;
#pragma unroll
        for (int s = 0; s < CLB_DEV_MAXS; ++s) {
            if (s >= S) continue;
            double sr = 0;
#pragma unroll
            for (int m = 0; m < CLB_DEV_MAXM; ++m) if (m < k.M) sr += k.bio_stoich[m][s] * g[m];
            if (!isfinite(sr)) sr = 0;
            if (k.thm && F < 1.0) sr *= F;
            const double d = sr * k.dt;
            if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
        }
#pragma unroll
        for (int m = 0; m < CLB_DEV_MAXM; ++m) {
            if (m >= k.M) continue;
            double b0 = br[m];
            if (!isfinite(b0)) b0 = 0;
            if (k.thm && F < 1.0) b0 *= F;
            double d = b0 * k.dt;
            if (d > 1e-12 || d < -1e-12) {
                R *p = a.B + (long long) m * N + v;
                const double b = (double) bpre[m];
                if (d < -b) d = -b;                                   // biomass never negative
                *p = (R) (b + d);
                bpre[m] = (R) (b + d);
            }
        }
    }
    // the surface gate of the abiotic rates and, in the transport launch of a
    // dissolution step, whether the increment waits for the parked products
    const bool wantNear = k.nearPhase == 1;
    bool react = k.nrxn > 0, deferred = false;
    if (react && (k.surface_only || wantNear)) {
        bool surf = false, near = false;
        if (wantNear) { const int own = k.phase[v]; near = own > 0 && own <= k.dPhases; }
#pragma unroll
        for (int q = 1; q < 7; ++q) {
            const int n = a.nb[(opp7(q) - 1) * No + i];                // the voxel at +c_q
            if (n == NB_OUTSIDE) continue;
            if (n == NB_CLOSED) { surf = true; continue; }
            const long long vn = a.list[n];
            const int nm = a.mask[vn];
            if (nm == 0 || nm == 1) surf = true;
            if (wantNear && n < a.Nint) { const int pid = k.phase[vn]; if (pid > 0 && pid <= k.dPhases) near = true; }
        }
        if (k.surface_only) react = surf;
        deferred = react && near;
    }
    if (k.nrxn > 0 && !deferred && react) {
        // abioticLaw: each reaction's rate (an excluded one contributes +-0)
        double Rr[CLB_DEV_MAXR];
#pragma unroll
        for (int r = 0; r < CLB_DEV_MAXR; ++r) {
            Rr[r] = 0.0;
            if (r >= k.nrxn) continue;
            const int r1 = k.r1[r], r2 = k.r2[r];
            const double av = r1 >= 0 ? fmax(thr(concOf(r1)), 0.0) : 1.0;
            const double bv = r2 >= 0 ? fmax(thr(concOf(r2)), 0.0) : 1.0;
            if (k.min_conc > 0 && ((r1 >= 0 && av <= k.min_conc) || (r2 >= 0 && bv <= k.min_conc))) continue;
            double Rv = k.rk[r] * av * bv;
            if (k.cap_dt > 0 && k.cap_frac > 0) {
                const double sc = fmin(r1 >= 0 ? av : 1e300, r2 >= 0 ? bv : 1e300);
                const double cap = sc * k.cap_frac / k.cap_dt;
                if (Rv > cap) Rv = cap;
            }
            Rr[r] = Rv;
        }
#pragma unroll
        for (int s = 0; s < CLB_DEV_MAXS; ++s) {
            if (s >= S) continue;
            double sr = 0;
#pragma unroll
            for (int r = 0; r < CLB_DEV_MAXR; ++r) if (r < k.nrxn) sr += k.rxn_stoich[r][s] * Rr[r];
            if (!isfinite(sr)) sr = 0;
            if (k.rate_scale != 1.0) sr *= k.rate_scale;
            const double d = sr * k.dt;
            if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
        }
    }
    if (binds) atomicAdd(a.clampCount, (unsigned long long) binds);
}

// the step's mask and total update of one voxel from its biomass b
// (updateLocalMaskNtotalLattices3D: the total, then pore <-> biofilm code at
// the threshold), and the total and mask it leaves for the reductions
template <typename R>
__device__ __forceinline__ void voxelMaskTotal(const AdeDev<R> &a, long long idx, int mask, const R (&b)[CLB_DEV_MAXM],
                                               double &bt, long long &ms) {
    if (mask != 0 && mask != 1) {
        double bm = 0; int nm = 0;
#pragma unroll
        for (int m = 0; m < CLB_DEV_MAXM; ++m) {
            if (m >= a.M) continue;
            const double v = (double) b[m];
            if (v > 1e-14) { bm += v; nm += 3 + m; }
        }
        double t = (double) a.Btot[idx];
        if (fabs(bm - t) > 1e-14) { a.Btot[idx] = (R) bm; t = (double) (R) bm; }
        const bool pore = (mask == 2);
        if (!pore && bm < a.thrdB) { mask = 2; a.mask[idx] = 2; }
        else if (pore && bm >= a.thrdB && nm > 0) { mask = nm; a.mask[idx] = (int8_t) nm; }
        bt = t;
    } else bt = (double) a.Btot[idx];
    ms = mask;
}
// a block's max total (NaN wins) and mask sum into partMax/partMask[slot]
__device__ __forceinline__ void blockMaskTotal(double bt, long long ms, double *partMax, long long *partMask, int slot) {
    __shared__ double sMax[4];
    __shared__ long long sMask[4];
    for (int off = 32; off > 0; off >>= 1) {
        bt = nanMax(bt, __shfl_down(bt, off, 64));
        ms += __shfl_down(ms, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sMax[wv] = bt; sMask[wv] = ms; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = -1e300; long long s = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { m = nanMax(m, sMax[w]); s += sMask[w]; }
        partMax[slot] = m; partMask[slot] = s;
    }
}

// an x-plane voxel for one mobile species: pull, the Dirichlet or Neumann
// plane rule, BGK.  The Neumann plane takes the inner layer's density after
// the stream (FlatAdiabaticBoundaryFunctional3D; a bounce-back or closed inner
// cell reports 0), which the interior kernel wrote before this one ran: the
// inner voxel's own bounce-back slots are overwritten by that kernel, so the
// planes cannot run in the same launch
template <typename R>
__device__ __forceinline__ void adePlaneVoxel(const AdeDev<R> &a, int i, int ms) {
    const int cls = a.cls[i];
    if (cls == ADE_BB) return;
    const long long No = a.Nopen;
    const bool left = i < a.Nint + a.Nleft;
    const int Np = a.Nopen - a.Nint;
    const R ux = a.u[i], uy = a.u[No + i], uz = a.u[2 * No + i];
    const int s = a.mobSpecies[ms];
    const int bt = left ? a.lbt[s] : a.rbt[s];
    if (bt == BC_CLOSED_DEV) return;
    R *cur = a.cur + (long long) ms * 7 * a.NoP;
    R *old = a.out + (long long) ms * 7 * a.NoP;
    R f[7];
    pull7<R, R>(a, cur, old, i, a.lbt[s], a.rbt[s], f);
    if (a.first) firstHandOver<R>(a, cur, i, a.lbt[s], a.rbt[s], a.c0[s], false);   // a plane voxel has no plane neighbour
    if (a.reset) {
        double c0 = a.c0[s];
        if (c0 < 1e-12 && c0 > -1e-12) c0 = 0;
        for (int q = 0; q < 7; ++q) f[q] = (R) (w7(q) * c0);
    }
    double rhoD = left ? a.lbc[s] : a.rbc[s];
    if (bt == BC_NEUMANN_DEV) {
        const int inner = a.nb[(left ? 1 : 0) * No + i];
        rhoD = 0.0;
        if (inner >= 0 && a.cls[inner] != ADE_BB) rhoD = a.rhoN[(long long) ms * Np + (i - a.Nint)];
    }
    adeDirichlet<R>(f, left ? 1 : 0, (R) rhoD);
    bgk7<R>(f, (R) (cls == ADE_BFILM ? a.omB[s] : a.omP[s]), ux, uy, uz);
    for (int q = 0; q < 7; ++q) old[pidx<7>(q, i)] = f[q];
}

// The transport pull's full-way bounce-back: what comes back to a voxel from
// a closed neighbour is its own opposite population from two launches back,
// which sits in its own slot of `out` until the thread overwrites it; reading
// it there fetches whole cache lines for the few lanes of a wave with that
// closed link.  Instead each population buffer has a store beside it: per
// mobile slot and tile of 64 interior voxels (one wave), 64 x 6 slots in which
// voxel i's closed links (pull kind 2), q ascending, take consecutive slots
// after those of the tile's earlier voxels, so a wave reads and writes its
// bounce-back values densely.  The store always mirrors the own slots of its
// buffer (still written every launch): the split kinetics kernels write it
// beside the own slots they change, and k_ade_bb_fill rebuilds it after every
// other change of the populations between launches and of the pull kinds.
// Links to a closed x plane (kind 3) keep the own-slot read.  A case uses the
// store (AdeDev::bbOn) only where it pays: few closed links per voxel and no
// per-step equilibrium (whose update of every voxel's own slots would have
// to be mirrored each step).

// interior: pull + BGK collide + biotic and abiotic kinetics with the clamp
// and the D3Q7 deposit, every mobile species of a voxel in one thread.  BB:
// the bounce-back store may be in use (then AdeDev::bbOn decides); without it
// the kernel carries none of the store's code
template <typename R, int MAXS, bool BB>
__device__ __forceinline__ void adeBulkVoxel(const AdeDev<R> &a, const KinDev &k, int i, double &bt, long long &mt) {
    const bool useBB = BB && a.bbOn;
    const bool in = i < a.Nint;
    const unsigned int kinds = in ? a.kinds[i] : 0u;
    const long long bb = useBB ? adeBBBase(i, kinds, in) : 0;   // every lane of the wave
    if (!in) return;
    const int cls = a.cls[i];
    if (cls == ADE_BB) {
        if (a.maskTotal) { const long long v = a.list[i]; bt = (double) a.Btot[v]; mt = a.mask[v]; }
        return;
    }
    const long long No = a.Nopen;
    // where each moving population comes from, shared by every species: a
    // neighbour's cur slot, the voxel's own opposite cur slot (grid edge), the
    // bounce-back store (closed neighbour), or its own opposite old slot (a
    // neighbour on an x plane, closed only for the species whose plane is)
    int nbq[6];
#pragma unroll
    for (int q = 1; q < 7; ++q) nbq[q - 1] = a.nb[(q - 1) * No + i];
    int kind[6];                       // 0 neighbour, 1 edge, 2 bounce-back, 3 plane neighbour (species rule)
#pragma unroll
    for (int q = 1; q < 7; ++q) kind[q - 1] = (kinds >> (2 * (q - 1))) & 3;
    const int Np = a.Nopen - a.Nint;
    R f[MAXS][7];
    const R ux = a.u[i], uy = a.u[No + i], uz = a.u[2 * No + i];
    const long long v = a.list[i];
    const int vmask = a.mask[v];
    // the biomass the kinetics reads, loaded with the populations
    R bpre[CLB_DEV_MAXM];
    const bool bioOn = k.on && k.bio;
#pragma unroll
    for (int m = 0; m < CLB_DEV_MAXM; ++m) bpre[m] = ((bioOn || a.maskTotal) && m < a.M) ? a.B[(long long) m * a.N + v] : (R) 0;
    // every species' populations first: no store comes between the loads, so
    // they are all in flight together
#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (ms >= a.Sm) continue;
        const R *__restrict__ cur = a.cur + (long long) ms * 7 * a.NoP;
        const R *__restrict__ old = a.out + (long long) ms * 7 * a.NoP;
        if (a.first) {
#pragma unroll
            for (int q = 0; q < 7; ++q) f[ms][q] = cur[pidx<7>(q, i)];
        } else {
            const bool lClosed = a.mlbt[ms] == BC_CLOSED_DEV, rClosed = a.mrbt[ms] == BC_CLOSED_DEV;
            const R *__restrict__ bbo = a.bbOut + (long long) ms * a.bbStride;
            f[ms][0] = cur[pidx<7>(0, i)];
            long long j = bb;
#pragma unroll
            for (int q = 1; q < 7; ++q) {
                const int n = nbq[q - 1];
                const int k0 = kind[q - 1];
                const int kd = k0 == 3 ? (((n < a.Nint + a.Nleft) ? lClosed : rClosed) ? 2 : 0) : k0;
                const bool st = useBB && k0 == 2;
                const R *base = st ? bbo : (kd == 2 ? old : cur);
                const long long off = st ? j : (kd == 0 ? pidx<7>(q, n) : pidx<7>(opp7(q), i));
                f[ms][q] = base[off];
                if (st) ++j;
            }
        }
    }
#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (ms >= a.Sm) continue;
        // the Neumann planes impose this voxel's density after the stream
        if (nbq[0] >= a.Nint && nbq[0] < a.Nint + a.Nleft && a.mlbt[ms] == BC_NEUMANN_DEV) {
            double r = 0;
#pragma unroll
            for (int q = 0; q < 7; ++q) r += (double) f[ms][q];
            a.rhoN[(long long) ms * Np + (nbq[0] - a.Nint)] = r;
        }
        if (nbq[1] >= a.Nint + a.Nleft && a.mrbt[ms] == BC_NEUMANN_DEV) {
            double r = 0;
#pragma unroll
            for (int q = 0; q < 7; ++q) r += (double) f[ms][q];
            a.rhoN[(long long) ms * Np + (nbq[1] - a.Nint)] = r;
        }
        if (a.first) firstHandOver<R>(a, a.cur + (long long) ms * 7 * a.NoP, i, a.mlbt[ms], a.mrbt[ms], a.mc0[ms], false);
        if (a.reset) {
            double c0 = a.mc0[ms];
            if (c0 < 1e-12 && c0 > -1e-12) c0 = 0;
            if (!a.first) firstHandOver<R>(a, a.cur + (long long) ms * 7 * a.NoP, i, a.mlbt[ms], a.mrbt[ms], c0, true);
#pragma unroll
            for (int q = 0; q < 7; ++q) f[ms][q] = (R) (w7(q) * c0);
        }
        bgk7<R>(f[ms], (R) (cls == ADE_BFILM ? a.momB[ms] : a.momP[ms]), ux, uy, uz);
    }

    if (k.on) voxelKineticsFused<R, MAXS>(a, k, i, f, v, vmask, bpre);

    // streaming stores: nothing reads these lines again in this launch (2-3 %
    // faster than plain stores on the 258x128x128 pack; the flow kernel, whose
    // working set is smaller, is slower with them)
#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (ms >= a.Sm) continue;
        R *o = a.out + (long long) ms * 7 * a.NoP;
#pragma unroll
        for (int q = 0; q < 7; ++q) __builtin_nontemporal_store(f[ms][q], &o[pidx<7>(q, i)]);
        // what goes back to this voxel from its closed neighbours two launches on
        if (useBB) {
            R *bo = a.bbOut + (long long) ms * a.bbStride;
            long long j = bb;
#pragma unroll
            for (int q = 1; q < 7; ++q)
                if (kind[q - 1] == 2) bo[j++] = f[ms][opp7(q)];
        }
    }
    if (a.maskTotal) voxelMaskTotal<R>(a, v, vmask, bpre, bt, mt);
}

template <typename R, int MAXS, bool BB>
__global__ void __launch_bounds__(256) k_ade_bulk(AdeDev<R> a, KinDev k) {
    if (halted(a.halt)) return;
    const int blk = xcdBlock();
    double bt = -1e300; long long mt = 0;
    adeBulkVoxel<R, MAXS, BB>(a, k, blk * blockDim.x + threadIdx.x, bt, mt);
    if (a.maskTotal) blockMaskTotal(bt, mt, a.partMax, a.partMask, blk);
}

// the planes after the interior, one thread per plane voxel and mobile species
// (blockIdx.y), so the short dependent load chains of the few plane voxels overlap
// one plane block (pb) for mobile slot ms; the planes' mask/total partials
// follow the interior's gInt block partials
template <typename R>
__device__ __forceinline__ void adePlanesBlock(const AdeDev<R> &a, int pb, int ms, int gInt) {
    const int i = a.Nint + pb * blockDim.x + threadIdx.x;
    if (i < a.Nopen && ms < a.Sm) adePlaneVoxel<R>(a, i, ms);
    if (a.maskTotal && ms == 0) {
        double bt = -1e300; long long mt = 0;
        if (i < a.Nopen) {
            const long long v = a.list[i];
            R b[CLB_DEV_MAXM];
#pragma unroll
            for (int m = 0; m < CLB_DEV_MAXM; ++m) b[m] = m < a.M ? a.B[(long long) m * a.N + v] : (R) 0;
            voxelMaskTotal<R>(a, v, a.mask[v], b, bt, mt);
        }
        blockMaskTotal(bt, mt, a.partMax, a.partMask, gInt + pb);
    }
}
template <typename R>
__global__ void __launch_bounds__(256) k_ade_planes(AdeDev<R> a) {
    if (halted(a.halt)) return;
    const int i = a.Nint + blockIdx.x * blockDim.x + threadIdx.x;
    const int ms = blockIdx.y;
    if (i < a.Nopen && ms < a.Sm) adePlaneVoxel<R>(a, i, ms);
    if (a.maskTotal && blockIdx.y == 0) {        // the planes' mask and total, after the interior's partials
        double bt = -1e300; long long mt = 0;
        if (i < a.Nopen) {
            const long long v = a.list[i];
            R b[CLB_DEV_MAXM];
#pragma unroll
            for (int m = 0; m < CLB_DEV_MAXM; ++m) b[m] = m < a.M ? a.B[(long long) m * a.N + v] : (R) 0;
            voxelMaskTotal<R>(a, v, a.mask[v], b, bt, mt);
        }
        blockMaskTotal(bt, mt, a.partMax, a.partMask, (int) ((a.Nint + 255) / 256) + (int) blockIdx.x);
    }
}


// the kinetics alone on the stored post-collision state (the split form of
// the fused kernel, for callers that apply the reference's processors one by
// one); bit-identical to the fused kernel's kinetics
template <typename R, int MAXS, bool DISSOL, bool SYM = false>
__device__ __forceinline__ void kineticsVoxel(const AdeDev<R> &a, const KinDev &k, int i) {
    if (a.cls[i] == ADE_BB) return;
    const long long v = a.list[i];
    const int mask = a.mask[v];
    if (mask == 0 || mask == 1) return;                  // no kinetics on solid / bounce-back voxels
    if (k.nearPhase == 2 && !nextToPhase(a, k, i)) return;
    const long long No = a.Nopen;
    R f[MAXS][7];
#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (ms >= a.Sm) continue;
        const R *c = a.cur + (long long) ms * 7 * a.NoP;
#pragma unroll
        for (int q = 0; q < 7; ++q) f[ms][q] = c[pidx<7>(q, i)];
    }
    R nopre[CLB_DEV_MAXM] = {};
    voxelKinetics<R, MAXS, DISSOL, SYM>(a, k, i, f, v, mask, nopre, false);
    const unsigned int kd = a.bbOn ? a.kinds[i] : 0u;
    const long long j0 = adeKind2Mask(kd) ? adeBBSlotOf(a.kinds, i) : 0;
#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (ms >= a.Sm) continue;
        R *c = a.cur + (long long) ms * 7 * a.NoP;
#pragma unroll
        for (int q = 0; q < 7; ++q) c[pidx<7>(q, i)] = f[ms][q];
        adeBBPut<R>(a, kd, j0, ms, f[ms]);                     // the store mirrors the own slots
    }
}
template <typename R, int MAXS, bool DISSOL, bool SYM>
__global__ void __launch_bounds__(256) k_kinetics(AdeDev<R> a, KinDev k) {
    if (halted(a.halt)) return;
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i >= a.Nint) return;
    kineticsVoxel<R, MAXS, DISSOL, SYM>(a, k, i);
}
// the same over a list of voxels whose length the device holds (the voxels
// next to the dissolving phase): a grid-stride loop of a fixed grid
template <typename R, int MAXS, bool DISSOL>
__global__ void __launch_bounds__(256) k_kinetics_listed(AdeDev<R> a, KinDev k, const int *sel, const unsigned int *count) {
    if (halted(a.halt)) return;
    const unsigned int n = *count;
    for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        kineticsVoxel<R, MAXS, DISSOL>(a, k, sel[t]);
}

// the compact voxels next to the dissolving phase (nextToPhase), appended in
// any order: one atomic per wave
template <typename R>
__global__ void __launch_bounds__(256) k_near_phase_list(AdeDev<R> a, KinDev k, int *sel, unsigned int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool near = i < a.Nint && nextToPhase(a, k, i);
    const unsigned long long b = __ballot(near);
    if (!b) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long) b) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned int) __popcll(b));
    base = __shfl(base, leader, 64);
    if (near) sel[base + __popcll(b & ((1ull << lane) - 1))] = i;
}

// ------------------------------------------------------------------------------------------
// mineral dissolution (dissolutionVOP.hh)
// ------------------------------------------------------------------------------------------
// density of species s at compact voxel n on the stored (post-collision) state
template <typename R>
__device__ __forceinline__ double storedConc(const AdeDev<R> &a, int s, int n) {
    const int sl = a.slot[s];
    if (a.mobile[s]) {
        const R *c = a.cur + (long long) sl * 7 * a.NoP;
        double r = 0;
#pragma unroll
        for (int q = 0; q < 7; ++q) r += (double) c[pidx<7>(q, n)];
        return r;
    }
    return (double) a.imm[(long long) sl * a.N + a.list[n]];
}

// x^e for the dissolution rate law: exactly x for order 1 and 1 for order 0
__device__ __forceinline__ double orderPow(double x, double e) {
    return e == 1.0 ? x : (e == 0.0 ? 1.0 : pow(x, e));
}

// surfaceDissolutionKinetics3D (dissolutionVOP.hh:378-569): a voxel of a
// declared phase with mineral left averages the water of its open face
// neighbours in the reaction box, takes its phase's rate law
// (defineDissolutionRate: rate = k max(C[acid],0)^a max(mineral,0)^m),
// removes the mineral (at most half of it per step; products scaled alike)
// and parks one share per open face of every product for the gather.  Shares
// of a phase voxel are rewritten every step, so a voxel that stops dissolving
// parks zeros.
template <typename R>
__global__ void __launch_bounds__(256) k_dissol_park(AdeDev<R> a, KinDev k) {
    if (halted(a.halt)) return;
    const int i = xcdBlock() * blockDim.x + threadIdx.x;
    if (i >= a.Nint) return;
    const long long v = a.list[i];
    const int pid = k.phase[v];
    if (pid <= 0 || pid > k.dPhases) return;
    const int q = pid - 1;
    const int sM = k.dMineral[q];
    const long long No = a.Nopen;
    const int S = k.S;
    for (int s = 0; s < S; ++s) k.share[(long long) s * No + i] = 0.0;
    R *pm = a.imm + (long long) a.slot[sM] * a.N + v;
    const double mineral = (double) *pm;
    if (!(mineral > 1e-12)) return;
    int openN[6];
    int nOpen = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int n = a.nb[(opp7(d + 1) - 1) * No + i];
        if (n < 0 || n >= a.Nint) continue;
        const int m = a.mask[a.list[n]];
        if (m == 0 || m == 1) continue;                       // still rock
        openN[nOpen++] = n;
    }
    if (nOpen == 0 && k.dSurfaceOnly) return;
    double water[CLB_DEV_MAXS], subsR[CLB_DEV_MAXS];
    for (int s = 0; s < S; ++s) water[s] = 0.0;
    if (nOpen > 0) {
        for (int kk = 0; kk < nOpen; ++kk)
            for (int s = 0; s < S; ++s) water[s] += storedConc<R>(a, s, openN[kk]);
        for (int s = 0; s < S; ++s) water[s] /= (double) nOpen;
    }
    for (int s = 0; s < S; ++s) if (water[s] < 0.0) water[s] = 0.0;
    const double H = k.dAcid[q] >= 0 ? orderPow(fmax(water[k.dAcid[q]], 0.0), k.dAo[q]) : 1.0;
    const double rate = k.dK[q] * H * orderPow(fmax(mineral, 0.0), k.dMo[q]);
    double mineralR = -rate;
    for (int s = 0; s < S; ++s) { subsR[s] = k.dStoich[q][s] * rate; if (!isfinite(subsR[s])) subsR[s] = 0.0; }
    if (!isfinite(mineralR)) mineralR = 0.0;
    if (!(mineralR < 0.0)) return;
    double dMineral = mineralR * k.dt;
    double scale = 1.0;
    const double maxRemove = 0.5 * mineral;                   // MAX_DISSOL_FRACTION, dissolutionVOP.hh:485
    if (-dMineral > maxRemove) scale = maxRemove / (-dMineral);
    dMineral *= scale;
    if (!(-dMineral > 1e-12)) return;
    *pm = (R) (mineral + dMineral);
    if (nOpen > 0) {
        unsigned long long parked = 0;
        for (int s = 0; s < S; ++s) {
            if (s == sM) continue;
            const double released = subsR[s] * k.dt * scale;
            if (fabs(released) <= 1e-12) continue;
            k.share[(long long) s * No + i] = released / (double) nOpen;
            parked += (unsigned long long) nOpen;
        }
        if (parked) atomicAdd(k.dcount, parked);
    }
}

// dissolNodeConversion3D (dissolutionVOP.hh:744-812): a voxel of a declared
// phase whose inventory is below reopen_fraction x its full_density turns
// pore and loses its phase
template <typename R>
__global__ void k_dissol_convert(AdeDev<R> a, int8_t *phase, DissolConv c, unsigned int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nint) return;
    const long long v = a.list[i];
    const int pid = phase[v];
    if (pid <= 0 || pid > c.n) return;
    if ((double) a.imm[(long long) c.slot[pid - 1] * a.N + v] < c.threshold[pid - 1]) {
        a.mask[v] = 2;
        phase[v] = 0;
        atomicAdd(count, 1u);
    }
}

// the x planes: pull + regularized Dirichlet (or Neumann) completion + BGK;
// density after the stream into a dense field (closed voxels keep 0)
// the energy snapshot of one interior voxel (computeEnergyField3D,
// complab3d_energyfield.hh:104-150): the gate's dG (which 0) or F_T (1) at the
// standing densities (negative ones read 0); open interior voxels only
template <typename R>
__global__ void __launch_bounds__(256) k_energy_field(AdeDev<R> a, KinDev k, double *dst, int which) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nint) return;
    const long long v = a.list[i];
    const int m = a.mask[v];
    if (m == 0 || m == 1) return;
    double C[CLB_DEV_MAXS];
    for (int s = 0; s < a.S; ++s) {
        double r = 0;
        if (!a.mobile[s]) r = (double) a.imm[(long long) a.slot[s] * a.N + v];
        else if (a.reset) {
            r = a.c0[s];
            if (r < 1e-12 && r > -1e-12) r = 0;
        } else if (!closedFor(a.cls, i, a.Nint, a.Nleft, a.lbt[s], a.rbt[s])) {
            const int ms = a.slot[s];
            double fl[7];
            pull7<R, double>(a, a.cur + (long long) ms * 7 * a.NoP, a.out + (long long) ms * 7 * a.NoP, i, a.lbt[s], a.rbt[s], fl);
            for (int q = 0; q < 7; ++q) r += fl[q];
        }
        C[s] = r > 0.0 ? r : 0.0;
    }
    double dG = 0.0;
    const double F = thermoGate(k, C, &dG);
    dst[v] = which ? F : dG;
}
template <typename R>
hipError_t launchEnergyField(const AdeDev<R> &a, const KinDev &k, double *dst, int which, hipStream_t st) {
    if (a.Nint > 0) hipLaunchKernelGGL(k_energy_field<R>, dim3((unsigned int) ((a.Nint + 255) / 256)), dim3(256), 0, st, a, k, dst, which);
    return hipGetLastError();
}

template <typename R>
__global__ void k_ade_density(AdeDev<R> a, int ms, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen) return;
    const int s = a.mobSpecies[ms];
    const long long v = a.list[i];
    if (closedFor(a.cls, i, a.Nint, a.Nleft, a.lbt[s], a.rbt[s])) { dst[v] = 0.0; return; }
    double r = 0;
    if (a.reset) {
        r = a.c0[s];
        if (r < 1e-12 && r > -1e-12) r = 0;
    } else {
        double fl[7];
        pull7<R, double>(a, a.cur + (long long) ms * 7 * a.NoP, a.out + (long long) ms * 7 * a.NoP, i, a.lbt[s], a.rbt[s], fl);
        for (int q = 0; q < 7; ++q) r += fl[q];
    }
    dst[v] = r;
}

// the source kind of each moving population of an interior voxel under the
// pull classes, two bits per direction (0 neighbour, 1 grid edge, 2 bounce-
// back, 3 x-plane neighbour whose rule depends on the species' boundary type);
// recomputed whenever the pull classes change, so the collide kernel reads one
// word instead of gathering its neighbours' classes
__global__ void __launch_bounds__(256) k_ade_kinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint, uint16_t *kinds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nint) return;
    unsigned int code = 0;
    for (int q = 1; q < 7; ++q) {
        const int n = nb[(long long) (q - 1) * Nopen + i];
        const unsigned int k = n == NB_OUTSIDE ? 1u : (n == NB_CLOSED ? 2u : (clsPull[n] == ADE_BB ? 2u : (n >= Nint ? 3u : 0u)));
        code |= k << (2 * (q - 1));
    }
    kinds[i] = (uint16_t) code;
}

// per block, for one mobile species: the mass the interior total does not
// report -- the content of the open x-plane voxels and the populations in
// flight into closed voxels (written at the last collide by a voxel open at
// it, toward a voxel closed now: the bounce-back rule returns them two
// launches later) -- and the smallest and largest concentration over the
// voxels the total counts.  parts[3 * b + 0] held, + 1 minus the minimum,
// + 2 the maximum.  Before the first launch a closed neighbour holds the rest
// population w c0 toward the voxel instead.
template <typename R>
__global__ void __launch_bounds__(256) k_ade_held(AdeDev<R> a, int ms, double *parts) {
    __shared__ double sH[4], sLo[4], sHi[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = a.mobSpecies[ms];
    const long long No = a.Nopen;
    const R *cur = a.cur + (long long) ms * 7 * a.NoP;
    double h = 0, lo = -1e300, hi = -1e300;
    if (i < a.Nopen) {
        const int lbt = a.lbt[s], rbt = a.rbt[s];
        if (!closedFor(a.cls, i, a.Nint, a.Nleft, lbt, rbt)) {
            double v = 0;
            if (a.reset) { v = a.c0[s]; if (v < 1e-12 && v > -1e-12) v = 0; }
            else {
                double fl[7];
                pull7<R, double>(a, cur, a.out + (long long) ms * 7 * a.NoP, i, lbt, rbt, fl);
                for (int q = 0; q < 7; ++q) v += fl[q];
            }
            if (i >= a.Nint) h += v;
            else {
                const int m = a.mask[a.list[i]];
                if (m != 0 && m != 1) { lo = -v; hi = v; }
            }
        }
        if (!closedFor(a.clsPull, i, a.Nint, a.Nleft, lbt, rbt)) {
            // a pending stabilization reset has rewritten the closed x planes
            // (pore voxels) with the rest populations, not the walls
            double c0r = a.c0[s];
            if (c0r < 1e-12 && c0r > -1e-12) c0r = 0;
            for (int q = 1; q < 7; ++q) {
                const int d = a.nb[(opp7(q) - 1) * No + i];          // the voxel at x + c_q
                if (d == NB_CLOSED || (d >= 0 && closedFor(a.cls, d, a.Nint, a.Nleft, lbt, rbt))) {
                    if (a.first) h += W7F * a.c0[s];
                    else if (a.reset && d >= a.Nint) h += W7F * c0r;
                    else h += (double) cur[pidx<7>(q, i)];
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        h += __shfl_down(h, off, 64);
        lo = fmax(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sH[wv] = h; sLo[wv] = lo; sHi[wv] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double H = 0, L = -1e300, U = -1e300;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { H += sH[w]; L = fmax(L, sLo[w]); U = fmax(U, sHi[w]); }
        parts[3 * blockIdx.x] = H; parts[3 * blockIdx.x + 1] = L; parts[3 * blockIdx.x + 2] = U;
    }
}

// populations at rest for a concentration field (closed voxels: the background c0)
template <typename R>
__global__ void k_ade_set(AdeDev<R> a, int ms, const double *src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen) return;
    const int s = a.mobSpecies[ms];
    R *cur = a.cur + (long long) ms * 7 * a.NoP;
    const double c = closedFor(a.cls, i, a.Nint, a.Nleft, a.lbt[s], a.rbt[s]) ? a.c0[s] : src[a.list[i]];
    for (int q = 0; q < 7; ++q) cur[pidx<7>(q, i)] = (R) (w7(q) * c);
}

// a voxel reopened this step: rest population, and the rest value on every
// link to a neighbour that was closed at the last collide, in both directions:
// what arrives from it at the next launch (read from `out`) and what the
// closed neighbour returns at the launch after (read from `cur`, the slot the
// voxel did not write while it was closed); the grid edge alike.  On a link to
// a neighbour that was open at the last collide, the `cur` slot takes that
// neighbour's population toward the voxel from the launch before: should the
// neighbour close in this same step, it returns to the voxel what the voxel
// reflected to it while closed, which is that population

template <typename R>
__global__ void k_ade_reopen(AdeDev<R> a, const uint8_t *reopened) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen || !reopened[i]) return;
    const long long No = a.Nopen;
    for (int ms = 0; ms < a.Sm; ++ms) {
        const int s = a.mobSpecies[ms];
        R *cur = a.cur + (long long) ms * 7 * a.NoP;
        R *old = a.out + (long long) ms * 7 * a.NoP;
        cur[pidx<7>(0, i)] = (R) (W7R * a.c0[s]);
        for (int q = 1; q < 7; ++q) {
            const int n = a.nb[(q - 1) * No + i];
            const bool closedSrc = n < 0 || closedFor(a.clsPull, n, a.Nint, a.Nleft, a.lbt[s], a.rbt[s]);
            if (closedSrc) {
                old[pidx<7>(opp7(q), i)] = (R) (W7F * a.c0[s]);
                cur[pidx<7>(opp7(q), i)] = (R) (W7F * a.c0[s]);
            } else {
                cur[pidx<7>(opp7(q), i)] = old[pidx<7>(q, n)];
            }
        }
    }
}

// total of a mobile species over the open interior (x = 1..nx-2, live mask open)
template <typename R>
__global__ void __launch_bounds__(256) k_ade_total(AdeDev<R> a, int ms, double *parts) {
    __shared__ double sS[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (i < a.Nint) {
        const int s = a.mobSpecies[ms];
        const int m = a.mask[a.list[i]];
        if (m != 0 && m != 1 && !closedFor(a.cls, i, a.Nint, a.Nleft, a.lbt[s], a.rbt[s])) {
            if (a.reset) { v = a.c0[s]; if (v < 1e-12 && v > -1e-12) v = 0; }
            else {
                double fl[7];
                pull7<R, double>(a, a.cur + (long long) ms * 7 * a.NoP, a.out + (long long) ms * 7 * a.NoP, i, a.lbt[s], a.rbt[s], fl);
                for (int q = 0; q < 7; ++q) v += fl[q];
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sS[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) { double s2 = 0; for (int w = 0; w < (int) (blockDim.x >> 6); ++w) s2 += sS[w]; parts[blockIdx.x] = s2; }
}

// ------------------------------------------------------------------------------------------
// D3Q19 incompressible BGK (IncBGKdynamics) with regularized pressure planes
// ------------------------------------------------------------------------------------------
template <typename R, typename T>
__device__ __forceinline__ void pull19(const NsDev<R> &a, int i, T fl[19]) {
    const long long No = a.Nopen;
    if (a.first) {
#pragma unroll
        for (int q = 0; q < 19; ++q) fl[q] = (T) a.cur[pidx<19>(q, i)];
        return;
    }
    fl[0] = (T) a.cur[pidx<19>(0, i)];
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const int n = a.nb[(q - 1) * No + i];
        if (n == NB_OUTSIDE) fl[q] = (T) a.cur[pidx<19>(opp19(q), i)];
        else if (n == NB_CLOSED || a.clsPull[n] >= NS_BB) fl[q] = (T) a.out[pidx<19>(opp19(q), i)];
        else fl[q] = (T) a.cur[pidx<19>(q, n)];
    }
}
// The dense bounce-back store of the interior flow pull.  A population that
// comes back from a closed neighbour is the voxel's own opposite population
// from two launches back; the pull used to read it from the voxel's own slot
// of `out`, which fetches whole cache lines for the few lanes of a wave with
// that closed link.  Instead every tile of 64 interior voxels (one wave) owns
// 64 x 18 slots of a store paired with each population buffer: voxel i's
// closed links (pull kind 2), q ascending, take consecutive slots after those
// of the tile's earlier voxels, so a wave reads and writes its bounce-back
// values densely.  The slot layout depends on the pull kinds alone; the store
// is rebuilt from the own slots of the buffer (still written every launch)
// whenever the kinds change or voxels reopen (k_ns_bb_fill).
__device__ __forceinline__ unsigned long long nsKind2Mask(unsigned long long kd) {
    return (kd >> 1) & ~kd & 0x555555555ull;                       // 2-bit fields equal to 2
}
// first slot of voxel i (every lane of the wave takes part: a prefix sum)
__device__ __forceinline__ long long nsBBBase(int i, unsigned long long kd, bool valid) {
    const int c = valid ? __popcll(nsKind2Mask(kd)) : 0;
    int x = c;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return (long long) (i >> 6) * (64 * 18) + (x - c);
}

// the interior's pull from the precomputed source kinds and the voxel's
// neighbour record (loaded beforehand, a chunk ahead: nsBulkBlocks)
template <typename R, typename T>
__device__ __forceinline__ void pull19k(const NsDev<R> &a, int i, T fl[19], unsigned long long kd, long long bb,
                                        const int2 w0, const int2 w1, const int2 w2) {
    const long long No = a.Nopen;
    if (a.first) {
#pragma unroll
        for (int q = 0; q < 19; ++q) fl[q] = (T) a.cur[pidx<19>(q, i)];
        return;
    }
    const unsigned int flags = (unsigned int) w2.y;
    int nbq[18];
#pragma unroll
    for (int q = 1; q < 19; ++q) nbq[q - 1] = nsRecSource(q, i, w0.x, w0.y, w1.x, w1.y, w2.x, flags);
    if (flags & NS_REC_FAR_MASK) {
#pragma unroll
        for (int q = 1; q < 19; ++q)
            if ((flags >> (q - 1)) & 1u) nbq[q - 1] = a.nb[(q - 1) * No + i];
    }
    fl[0] = (T) a.cur[pidx<19>(0, i)];
    long long j = bb;
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const unsigned int k = (unsigned int) (kd >> (2 * (q - 1))) & 3u;
        const R *base = k == 2 ? a.bbOut : a.cur;
        const long long off = k == 0 ? pidx<19>(q, nbq[q - 1]) : (k == 2 ? j : pidx<19>(opp19(q), i));
        fl[q] = (T) base[off];
        if (k == 2) ++j;
    }
}

template <typename R>
__device__ __forceinline__ void nsFirstHandOver(const NsDev<R> &a, int i) {
    const long long No = a.Nopen;
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const int n = a.nb[(q - 1) * No + i];
        if (n == NB_CLOSED || (n >= 0 && a.clsPull[n] >= NS_BB)) a.cur[pidx<19>(opp19(q), i)] = (R) 0;
    }
}
// the same for an interior voxel, whose bounce-back values of the launch
// after next come from the store paired with cur
template <typename R>
__device__ __forceinline__ void nsFirstHandOverBB(const NsDev<R> &a, int i, unsigned long long kd, long long bb) {
    long long j = bb;
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        if (((kd >> (2 * (q - 1))) & 3ull) != 2ull) continue;
        a.cur[pidx<19>(opp19(q), i)] = (R) 0;
        a.bbCur[j++] = (R) 0;
    }
}

template <typename T>
__device__ __forceinline__ T feqInc(int q, T rb, T jx, T jy, T jz, T jsq) {
    const T cj = (T) D19C[q][0] * jx + (T) D19C[q][1] * jy + (T) D19C[q][2] * jz;
    return (T) w19(q) * (rb + (T) 3 * cj + (T) 4.5 * cj * cj - (T) 1.5 * jsq);
}

// dir = +1 at x = 0 (unknowns c_x = +1), -1 at x = nx-1; tangential momentum 0
__device__ void nsPressurePlane(double fl[19], int dir, double rb, double &jx) {
    double r0 = 0, rout = 0;
    for (int q = 0; q < 19; ++q) {
        if (D19C[q][0] == 0) r0 += fl[q];
        else if (D19C[q][0] == -dir) rout += fl[q];
    }
    jx = dir * (rb - r0 - 2.0 * rout);
    const double jsq = jx * jx;
    double feq[19];
    for (int q = 0; q < 19; ++q) feq[q] = feqInc<double>(q, rb, jx, 0.0, 0.0, jsq);
    for (int q = 0; q < 19; ++q) if (D19C[q][0] == dir) fl[q] = feq[q] + (fl[opp19(q)] - feq[opp19(q)]);
    double P[3][3] = {{0,0,0},{0,0,0},{0,0,0}};
    for (int q = 0; q < 19; ++q) {
        const double ne = fl[q] - feq[q];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) P[i][j] += D19C[q][i] * D19C[q][j] * ne;
    }
    for (int q = 0; q < 19; ++q) {
        double s = 0;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += (D19C[q][i] * D19C[q][j] - (i == j ? 1.0 / 3.0 : 0.0)) * P[i][j];
        fl[q] = feq[q] + w19(q) * 4.5 * s;
    }
}

__device__ __forceinline__ void blockEnergy(double e, double cnt, double *partE, double *partN, int slot) {
    __shared__ double sE[4], sN[4];
    for (int off = 32; off > 0; off >>= 1) { e += __shfl_down(e, off, 64); cnt += __shfl_down(cnt, off, 64); }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sE[wv] = e; sN[wv] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double E = 0, C = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { E += sE[w]; C += sN[w]; }
        partE[slot] = E; partN[slot] = C;
    }
}

// interior: pull + BGK of the 256 voxels of chunk c (one per thread), whose
// class, pull kinds and neighbour record were loaded a chunk ahead; the
// energy statistic |j|^2/2 accumulated into e, cnt
struct NsAhead {
    int cls;
    unsigned long long kd;
    int2 w0, w1, w2;
};
template <typename R>
__device__ __forceinline__ void nsLoadAhead(const NsDev<R> &a, int c, NsAhead &h) {
    const int i = c * 256 + (int) threadIdx.x;
    const bool in = i < a.Nint;
    h.cls = in ? (int) a.cls[i] : NS_BB;
    h.kd = in ? a.kinds[i] : 0ull;
    if (in && !a.first) {
        h.w0 = a.nbr[i]; h.w1 = a.nbr[(long long) a.Nint + i]; h.w2 = a.nbr[2 * (long long) a.Nint + i];
    } else {
        h.w0 = h.w1 = h.w2 = make_int2(0, 0);
    }
}
template <typename R>
__device__ __forceinline__ void nsBulkChunk(const NsDev<R> &a, int c, const NsAhead &h, double &e, double &cnt) {
    const int i = c * 256 + (int) threadIdx.x;
    const bool in = i < a.Nint;
    const long long bb = nsBBBase(i, h.kd, in);           // every lane of the wave
    if (!in || h.cls >= NS_BB) return;
    // the collision in FP64 whatever the storage: an FP32 lattice rounds only
    // where it stores (the kernel is bound by its memory traffic, not its
    // arithmetic)
    double fl[19];
    pull19k<R, double>(a, i, fl, h.kd, bb, h.w0, h.w1, h.w2);
    if (a.first) nsFirstHandOverBB<R>(a, i, h.kd, bb);
    double rb = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        rb += fl[q];
        jx += (double) D19C[q][0] * fl[q]; jy += (double) D19C[q][1] * fl[q]; jz += (double) D19C[q][2] * fl[q];
    }
    const double jsq = jx * jx + jy * jy + jz * jz;
    e += 0.5 * jsq; cnt += 1;
    const double om = h.cls == NS_BIO ? a.omBio : a.om;
#pragma unroll
    for (int q = 0; q < 19; ++q) fl[q] = fl[q] + om * (feqInc<double>(q, rb, jx, jy, jz, jsq) - fl[q]);
    // streaming stores: the next launch reads these lines, but from beyond
    // what the caches hold of one iteration's traffic (130.6 against 139.0 us
    // per iteration on the 258x128x128 pack, 79.3 against 82.9 in FP32)
#pragma unroll
    for (int q = 0; q < 19; ++q) __builtin_nontemporal_store((R) fl[q], &a.out[pidx<19>(q, i)]);
    // what goes back to this voxel from its closed neighbours two launches
    // on: the new population opposite each closed link
    long long j = bb;
#pragma unroll
    for (int q = 1; q < 19; ++q)
        if (((h.kd >> (2 * (q - 1))) & 3ull) == 2ull) a.bbOut[j++] = (R) fl[opp19(q)];
}
// The interior in G blocks of 256 threads, block lb taking the contiguous
// chunks [lb C / G, (lb + 1) C / G) of the C chunks of 256 interior voxels.
// A thread loads the class, kinds and neighbour record of its voxel in the
// next chunk before it pulls in the current one, so those loads are in
// flight beside the current gathers: one dependent round of memory per chunk
// instead of two.
// Without AHEAD the loads of a chunk come right before its pull (fewer live
// registers: more waves per SIMD); with G = C that is one chunk per block.
template <typename R, bool AHEAD>
__device__ __forceinline__ void nsBulkBlocks(const NsDev<R> &a, double *partE, double *partN, int lb, int G) {
    const int C = (a.Nint + 255) / 256;
    const int c0 = (int) ((long long) lb * C / G), c1 = (int) ((long long) (lb + 1) * C / G);
    double e = 0, cnt = 0;
    if (AHEAD) {
        NsAhead h;
        if (c0 < c1) nsLoadAhead<R>(a, c0, h);
        for (int c = c0; c < c1; ++c) {
            NsAhead nx = h;
            if (c + 1 < c1) nsLoadAhead<R>(a, c + 1, nx);
            nsBulkChunk<R>(a, c, h, e, cnt);
            h = nx;
        }
    } else {
        for (int c = c0; c < c1; ++c) {
            NsAhead h;
            nsLoadAhead<R>(a, c, h);
            nsBulkChunk<R>(a, c, h, e, cnt);
        }
    }
    blockEnergy(e, cnt, partE, partN, lb);
}

// the pressure planes
template <typename R>
__device__ __forceinline__ void nsPlaneBlock(const NsDev<R> &a, double *partE, double *partN, int b, int slot0) {
    const int i = a.Nint + b * blockDim.x + threadIdx.x;
    double e = 0, cnt = 0;
    if (i < a.Nopen) {
        const int cls = a.cls[i];
        if (cls < NS_BB) {
            double fl[19];
            pull19<R, double>(a, i, fl);
            if (a.first) nsFirstHandOver<R>(a, i);
            const bool left = i < a.Nint + a.Nleft;
            const double rb = left ? 0.0 : a.rhoBarRight;
            double jx;
            nsPressurePlane(fl, left ? +1 : -1, rb, jx);
            const double jsq = jx * jx;
            e = 0.5 * jsq; cnt = 1;
            const double om = cls == NS_BIO ? a.omBio : a.om;
            const long long No = a.Nopen;
            for (int q = 0; q < 19; ++q) a.out[pidx<19>(q, i)] = (R) (fl[q] + om * (feqInc<double>(q, rb, jx, 0.0, 0.0, jsq) - fl[q]));
        }
    }
    blockEnergy(e, cnt, partE, partN, slot0 + b);
}

// one flow iteration in one launch: the first gP blocks take the pressure
// planes (few blocks with a long serial chain each, so they start first and
// run beside the interior), the rest the interior in XCD order; block partials
// at [interior blocks][plane blocks]
template <typename R, bool AHEAD>
__global__ void __launch_bounds__(256) k_ns_step(NsDev<R> a, double *partE, double *partN, int gP) {
    if (*a.done) return;
    const int gB = (int) gridDim.x - gP;
    if ((int) blockIdx.x < gP) nsPlaneBlock<R>(a, partE, partN, blockIdx.x, gB);
    else nsBulkBlocks<R, AHEAD>(a, partE, partN, xcdBlockOf((int) blockIdx.x - gP, gB), gB);
}

// the convergence rule, evaluated on the device after every iteration.  The
// partials are summed by 1024 threads, thread t taking t, t + 1024, ... (all
// of its loads issued together), then in a fixed tree; the window statistics
// run in the ring's order (oldest first).  The ring's head and fill and each
// thread's sample of the window's W - 1 older values are loaded beside the
// partials (thread 0 then overwrites only the head slot, which no thread
// reads), so the statistics wait on no further memory round trip.
__device__ __forceinline__ void nsTracerBlock(NsTracer *t, const double *partE, const double *partN, int nparts, int rule,
                                              double eps, int *done) {
    __shared__ double sE[16], sN[16];
    __shared__ double win[16];
    __shared__ double sVal;
    if (*done) return;
    const int tid = threadIdx.x;
    const int W = NS_TRACER_WINDOW;
    static_assert(NS_TRACER_WINDOW <= 512, "one window sample per thread of the first eight waves");
    int h0 = 0, c0 = 0;
    double xr = 0.0;
    if (rule == 0) {
        h0 = t->head; c0 = t->count;
        if (tid < W - 1) xr = t->ring[(h0 + 1 + tid) % W];
    }
    double E = 0, C = 0;
    for (int i0 = tid; i0 < nparts; i0 += 1024 * 8) {
        double e[8], c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = i0 + 1024 * k;
            e[k] = i < nparts ? partE[i] : 0.0;
            c[k] = i < nparts ? partN[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) { E += e[k]; C += c[k]; }
    }
    for (int off = 32; off > 0; off >>= 1) { E += __shfl_down(E, off, 64); C += __shfl_down(C, off, 64); }
    if ((tid & 63) == 0) { sE[tid >> 6] = E; sN[tid >> 6] = C; }
    __syncthreads();
    if (tid == 0) {
        double Es = 0, Cs = 0;
        for (int w = 0; w < 16; ++w) { Es += sE[w]; Cs += sN[w]; }
        const double val = Cs > 0 ? Es / Cs : 0.0;
        t->iters += 1;
        t->last = val;
        sVal = val;
        if (!isfinite(val)) { t->nonfinite = 1; *done = 1; }
        else if (rule == 0) {                 // ValueTracer(1, 1000, eps): window 500
            t->ring[h0] = val;
            t->head = (h0 + 1) % W;
            if (c0 < W) t->count = c0 + 1;
        } else {
            if (t->prevE > 0 && fabs(val - t->prevE) <= eps * fabs(val)) { t->converged = 1; *done = 1; }
            t->prevE = val;
        }
    }
    __syncthreads();
    const double val = sVal;
    if (rule != 0 || !isfinite(val) || c0 + 1 < W) return;
    // the window's mean and population standard deviation, both sums as a
    // tree over the first 512 threads (wave sums, then the eight wave
    // partials in order; threads 512 and up add zeros): rounding-level apart
    // from the sequential sums
    const double x = tid < W - 1 ? xr : (tid == W - 1 ? val : 0.0);
    const int lane = tid & 63, wv = tid >> 6;
    double sx = x;
    for (int off = 32; off > 0; off >>= 1) sx += __shfl_down(sx, off, 64);
    if (lane == 0 && wv < 8) win[wv] = sx;
    __syncthreads();
    if (tid == 0) {
        double m = 0;
        for (int w = 0; w < 8; ++w) m += win[w];
        sVal = m / (double) W;
    }
    __syncthreads();
    const double m = sVal;
    const double d = tid < W ? x - m : 0.0;
    double sd = d * d;
    for (int off = 32; off > 0; off >>= 1) sd += __shfl_down(sd, off, 64);
    __syncthreads();
    if (lane == 0 && wv < 8) win[8 + wv] = sd;
    __syncthreads();
    if (tid == 0) {
        double v = 0;
        for (int w = 0; w < 8; ++w) v += win[8 + w];
        const double r = sqrt(v / (double) W) / fabs(m);
        if (isnan(r) || r < eps) { t->converged = 1; *done = 1; }
    }
}
__global__ void __launch_bounds__(1024) k_ns_tracer(NsTracer *t, const double *partE, const double *partN, int nparts, int rule, double eps, int *done) {
    nsTracerBlock(t, partE, partN, nparts, rule, eps, done);
}

// velocity of the current flow state (computeVelocity) into the compact field
// read by the ADE lattices; optionally the box reductions of the Peclet rescale
// the flow density computeDensity reports, into a dense field: 1 + the sum of
// the post-stream populations at an open interior voxel, the imposed density
// on an open pressure plane, 1 at closed voxels
__global__ void k_fill_u32(unsigned int *dst, unsigned int n, unsigned int v) {
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}
__global__ void k_fill_d(double *dst, long long n, double v) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}
hipError_t launchFillD(double *dst, size_t n, double v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_d, dim3((unsigned int) ((n + 255) / 256)), dim3(256), 0, st, dst, (long long) n, v);
    return hipGetLastError();
}
template <typename R>
__global__ void __launch_bounds__(256) k_ns_density(NsDev<R> a, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen || a.cls[i] >= NS_BB) return;
    double r;
    if (i >= a.Nint) r = i < a.Nint + a.Nleft ? 0.0 : a.rhoBarRight;
    else {
        double fl[19];
        pull19<R, double>(a, i, fl);
        r = 0;
        for (int q = 0; q < 19; ++q) r += fl[q];
    }
    dst[a.list[i]] = 1.0 + r;
}

template <typename R, typename RA>
__global__ void __launch_bounds__(256) k_ns_velocity(NsDev<R> a, RA *u, double *parts) {
    __shared__ double red[4][4], r4[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};      // sum|u| box, max ux box, sum ux outlet, sum|u| open box, count open box
    bool inBox = false;
    if (i < a.Nopen) {
        double ux = 0, uy = 0, uz = 0;
        if (a.cls[i] < NS_BB) {
            double fl[19];
            pull19<R, double>(a, i, fl);
            if (i >= a.Nint) {
                const bool left = i < a.Nint + a.Nleft;
                nsPressurePlane(fl, left ? +1 : -1, left ? 0.0 : a.rhoBarRight, ux);
            } else {
                for (int q = 0; q < 19; ++q) { ux += D19C[q][0] * fl[q]; uy += D19C[q][1] * fl[q]; uz += D19C[q][2] * fl[q]; }
            }
        }
        if (u) { u[i] = (RA) ux; u[a.Nopen + i] = (RA) uy; u[2 * (long long) a.Nopen + i] = (RA) uz; }
        inBox = i < a.Nint;
        const double n = sqrt(ux * ux + uy * uy + uz * uz);
        if (inBox) {
            v[0] = n; v[1] = ux;
            if (a.geom[a.list[i]] >= 2) { v[3] = n; v[4] = 1; }
            const int x = (int) (a.list[i] % a.nx);
            if (x == a.nx - 2) v[2] = ux;
        }
    }
    if (!parts) return;
    if (!inBox) v[1] = -1e300;
    for (int off = 32; off > 0; off >>= 1) {
        v[0] += __shfl_down(v[0], off, 64);
        v[1] = fmax(v[1], __shfl_down(v[1], off, 64));
        v[2] += __shfl_down(v[2], off, 64);
        v[3] += __shfl_down(v[3], off, 64);
        v[4] += __shfl_down(v[4], off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { red[0][wv] = v[0]; red[1][wv] = v[1]; red[2][wv] = v[2]; red[3][wv] = v[3]; r4[wv] = v[4]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = -1e300, s2 = 0, s3 = 0, s4 = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { s0 += red[0][w]; s1 = fmax(s1, red[1][w]); s2 += red[2][w]; s3 += red[3][w]; s4 += r4[w]; }
        parts[blockIdx.x * 5 + 0] = s0; parts[blockIdx.x * 5 + 1] = s1; parts[blockIdx.x * 5 + 2] = s2;
        parts[blockIdx.x * 5 + 3] = s3; parts[blockIdx.x * 5 + 4] = s4;
    }
}

// initializeAtEquilibrium(PressureGradient): rho = 1 - 3 dP x/(nx-1), u = 0
template <typename R>
__global__ void k_ns_init(NsDev<R> a, double deltaP) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen) return;
    const int x = (int) (a.list[i] % a.nx);
    const double rb = -deltaP * 3.0 / (double) (a.nx - 1) * (double) x;
    const bool closed = a.cls[i] >= NS_BB;
    for (int q = 0; q < 19; ++q) {
        const R v = closed ? (R) 0 : (R) (w19(q) * rb);
        a.cur[pidx<19>(q, i)] = v;
        a.out[pidx<19>(q, i)] = v;
    }
}

// updateNsLatticesDynamics3D (part2:748-818) from the live mask
__global__ void k_ns_update_dyn(const int8_t *mask, const int *list, uint8_t *cls, int Nopen, double permRatio, unsigned int *changed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nopen) return;
    const int m = mask[list[i]];
    if (m == 0 || m == 1) return;
    const int c = cls[i];
    const bool isNsOmega = (c == NS_BGK);
    int nc = c;
    if (m != 2 && isNsOmega) nc = permRatio <= 1e-14 ? NS_BB : NS_BIO;
    else if (m == 2 && !isNsOmega) nc = NS_BGK;
    if (nc != c) { cls[i] = (uint8_t) nc; atomicAdd(changed, 1u); }
}
// a cell reopened in the flow lattice: rest (0) on every link to a neighbour
// that was closed at the last collide, in both directions (as for the
// advection-diffusion lattices)
template <typename R>
__global__ void k_ns_reopen(NsDev<R> a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nopen) return;
    if (!(a.clsPull[i] >= NS_BB && a.cls[i] < NS_BB)) return;
    const long long No = a.Nopen;
    a.cur[pidx<19>(0, i)] = (R) 0;
    for (int q = 1; q < 19; ++q) {
        const int n = a.nb[(q - 1) * No + i];
        if (n < 0 || a.clsPull[n] >= NS_BB) { a.out[pidx<19>(opp19(q), i)] = (R) 0; a.cur[pidx<19>(opp19(q), i)] = (R) 0; }
        else a.cur[pidx<19>(opp19(q), i)] = a.out[pidx<19>(q, n)];
    }
}

// compact <-> dense staging
template <typename R>
__global__ void k_scatter(const R *src, const int *list, int n, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[list[i]] = (double) src[i];
}
template <typename R>
__global__ void k_gather(const double *src, const int *list, int n, R *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (R) src[list[i]];
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
static inline unsigned int nblk(long long n, int b) { return (unsigned int) ((n + b - 1) / b); }
template <typename R>
hipError_t launchNsDensity(const NsDev<R> &a, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_d, dim3(nblk(a.N, 256)), dim3(256), 0, st, dst, a.N, 1.0);
    if (a.Nopen > 0) hipLaunchKernelGGL(k_ns_density<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, dst);
    return hipGetLastError();
}
template hipError_t launchNsDensity<double>(const NsDev<double> &, double *, hipStream_t);
template hipError_t launchNsDensity<float>(const NsDev<float> &, double *, hipStream_t);

// interior blocks of a flow launch: one per chunk of 256 voxels, or in the
// prefetching mode (NsDev::shape > 0) that many blocks looping over chunks
static unsigned int nsInteriorBlocks(int Nint, int shape) {
    const unsigned int g = nblk(Nint, 256);
    return shape > 0 && (unsigned int) shape < g ? (unsigned int) shape : g;
}
unsigned int nsStepBlocks(int Nopen, int Nint, int shape) { return nsInteriorBlocks(Nint, shape) + nblk(Nopen - Nint, 256); }

template <typename R, bool BB>
static void launchAdeBulk(const AdeDev<R> &a, const KinDev &k, unsigned int g, hipStream_t st) {
    if (a.Sm <= 1) hipLaunchKernelGGL((k_ade_bulk<R, 1, BB>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 2) hipLaunchKernelGGL((k_ade_bulk<R, 2, BB>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 3) hipLaunchKernelGGL((k_ade_bulk<R, 3, BB>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 4) hipLaunchKernelGGL((k_ade_bulk<R, 4, BB>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 6) hipLaunchKernelGGL((k_ade_bulk<R, 6, BB>), dim3(g), dim3(256), 0, st, a, k);
    else hipLaunchKernelGGL((k_ade_bulk<R, 8, BB>), dim3(g), dim3(256), 0, st, a, k);
}
template <typename R>
hipError_t launchAdeStep(const AdeDev<R> &a, const KinDev &k, hipStream_t st, hipEvent_t afterInterior) {
    const int Np = a.Nopen - a.Nint;
    if (a.Nint > 0) {
        const unsigned int g = nblk(a.Nint, 256);
        if (a.bbOn) launchAdeBulk<R, true>(a, k, g, st);
        else launchAdeBulk<R, false>(a, k, g, st);
    }
    if (afterInterior) {
        const hipError_t e = hipEventRecord(afterInterior, st);
        if (e != hipSuccess) return e;
    }
    if (Np > 0 && a.Sm > 0) hipLaunchKernelGGL(k_ade_planes<R>, dim3(nblk(Np, 256), a.Sm), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <typename R, bool DISSOL>
static void kineticsListedVariants(const AdeDev<R> &a, const KinDev &k, const int *sel, const unsigned int *count,
                                   hipStream_t st) {
    const unsigned int g = nblk(a.Nint, 256) < 1024u ? nblk(a.Nint, 256) : 1024u;
    if (a.Sm <= 1) hipLaunchKernelGGL((k_kinetics_listed<R, 1, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
    else if (a.Sm <= 2) hipLaunchKernelGGL((k_kinetics_listed<R, 2, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
    else if (a.Sm <= 3) hipLaunchKernelGGL((k_kinetics_listed<R, 3, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
    else if (a.Sm <= 4) hipLaunchKernelGGL((k_kinetics_listed<R, 4, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
    else if (a.Sm <= 6) hipLaunchKernelGGL((k_kinetics_listed<R, 6, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
    else hipLaunchKernelGGL((k_kinetics_listed<R, 8, DISSOL>), dim3(g), dim3(256), 0, st, a, k, sel, count);
}
template <typename R>
hipError_t launchKineticsListed(const AdeDev<R> &a, const KinDev &k, const int *sel, const unsigned int *count,
                                hipStream_t st) {
    if (a.Nint <= 0 || !k.on) return hipSuccess;
    if (k.dissol) kineticsListedVariants<R, true>(a, k, sel, count, st);
    else kineticsListedVariants<R, false>(a, k, sel, count, st);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNearPhaseList(const AdeDev<R> &a, const KinDev &k, int *sel, unsigned int *count, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, count, 1u, 0u);
    if (a.Nint > 0) hipLaunchKernelGGL(k_near_phase_list<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, k, sel, count);
    return hipGetLastError();
}
template <typename R, bool DISSOL, bool SYM>
static void kineticsVariants(const AdeDev<R> &a, const KinDev &k, hipStream_t st) {
    const unsigned int g = nblk(a.Nint, 256);
    if (a.Sm <= 1) hipLaunchKernelGGL((k_kinetics<R, 1, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 2) hipLaunchKernelGGL((k_kinetics<R, 2, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 3) hipLaunchKernelGGL((k_kinetics<R, 3, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 4) hipLaunchKernelGGL((k_kinetics<R, 4, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 6) hipLaunchKernelGGL((k_kinetics<R, 6, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
    else hipLaunchKernelGGL((k_kinetics<R, 8, DISSOL, SYM>), dim3(g), dim3(256), 0, st, a, k);
}
template <typename R>
hipError_t launchKinetics(const AdeDev<R> &a, const KinDev &k, hipStream_t st) {
    if (a.Nint <= 0 || !k.on) return hipSuccess;
    if (k.dissol) kineticsVariants<R, true, false>(a, k, st);           // symbolic laws are refused with dissolution
    else if ((k.symOn && k.bio) || k.symAb || k.rate) kineticsVariants<R, false, true>(a, k, st);   // symbolic laws, rate capture
    else kineticsVariants<R, false, false>(a, k, st);
    return hipGetLastError();
}
template <typename R>
hipError_t launchDissolPark(const AdeDev<R> &a, const KinDev &k, hipStream_t st) {
    if (a.Nint <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dissol_park<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, k);
    return hipGetLastError();
}
template <typename R>
hipError_t launchDissolConvert(const AdeDev<R> &a, int8_t *phase, const DissolConv &c, unsigned int *count, hipStream_t st) {
    if (a.Nint <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dissol_convert<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, phase, c, count);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeDensity(const AdeDev<R> &a, int ms, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_density<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, ms, dst);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeSet(const AdeDev<R> &a, int ms, const double *src, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_set<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, ms, src);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeReopen(const AdeDev<R> &a, const uint8_t *re, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_reopen<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, re);
    return hipGetLastError();
}
// the flow's counterpart: 0 neighbour, 1 grid edge, 2 bounce-back (a closed
// neighbour or one whose pull class is bounce-back), 2 bits per D3Q19 direction
__global__ void __launch_bounds__(256) k_ns_kinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint,
                                                  unsigned long long *kinds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nint) return;
    unsigned long long code = 0;
    for (int q = 1; q < 19; ++q) {
        const int n = nb[(long long) (q - 1) * Nopen + i];
        const unsigned long long k = n == NB_OUTSIDE ? 1ull : ((n == NB_CLOSED || clsPull[n] >= NS_BB) ? 2ull : 0ull);
        code |= k << (2 * (q - 1));
    }
    kinds[i] = code;
}
// the dense bounce-back store `bb` from the own slots of the population
// buffer `g` it pairs with (nsBBBase), for the current pull kinds
template <typename R>
__global__ void __launch_bounds__(256) k_ns_bb_fill(NsDev<R> a, const R *g, R *bb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.Nint;
    const unsigned long long kd = in ? a.kinds[i] : 0ull;
    long long j = nsBBBase(i, kd, in);
    if (!in) return;
    for (int q = 1; q < 19; ++q)
        if (((kd >> (2 * (q - 1))) & 3ull) == 2ull) bb[j++] = g[pidx<19>(opp19(q), i)];
}
template <typename R>
hipError_t launchNsBBFill(const NsDev<R> &a, const R *g, R *bb, hipStream_t st) {
    if (a.Nint > 0) hipLaunchKernelGGL(k_ns_bb_fill<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, g, bb);
    return hipGetLastError();
}
hipError_t launchNsKinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint, unsigned long long *kinds, hipStream_t st) {
    if (Nint > 0) hipLaunchKernelGGL(k_ns_kinds, dim3(nblk(Nint, 256)), dim3(256), 0, st, nb, clsPull, Nopen, Nint, kinds);
    return hipGetLastError();
}
// the bounce-back store `bb` of every mobile slot from the own slots of the
// population buffer `g` it pairs with (adeBBBase), for the current pull kinds
template <typename R>
__global__ void __launch_bounds__(256) k_ade_bb_fill(AdeDev<R> a, const R *g, R *bb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.Nint;
    const unsigned int kd = in ? a.kinds[i] : 0u;
    const long long j0 = adeBBBase(i, kd, in);
    if (!in) return;
    for (int ms = 0; ms < a.Sm; ++ms) {
        const R *gs = g + (long long) ms * 7 * a.NoP;
        R *bs = bb + (long long) ms * a.bbStride;
        long long j = j0;
#pragma unroll
        for (int q = 1; q < 7; ++q)
            if (((kd >> (2 * (q - 1))) & 3u) == 2u) bs[j++] = gs[pidx<7>(opp7(q), i)];
    }
}
template <typename R>
hipError_t launchAdeBBFill(const AdeDev<R> &a, const R *g, R *bb, hipStream_t st) {
    if (a.bbOn && a.Nint > 0 && a.Sm > 0) hipLaunchKernelGGL(k_ade_bb_fill<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, g, bb);
    return hipGetLastError();
}
hipError_t launchAdeKinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint, uint16_t *kinds, hipStream_t st) {
    if (Nint > 0) hipLaunchKernelGGL(k_ade_kinds, dim3(nblk(Nint, 256)), dim3(256), 0, st, nb, clsPull, Nopen, Nint, kinds);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeHeld(const AdeDev<R> &a, int ms, double *parts, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_held<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, ms, parts);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeTotal(const AdeDev<R> &a, int ms, double *parts, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_total<R>, dim3(nblk(a.Nint, 256)), dim3(256), 0, st, a, ms, parts);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsStep(const NsDev<R> &a, double *partE, double *partN, NsTracer *t, int rule, double eps, hipStream_t st,
                        hipEvent_t evStart, hipEvent_t evEnd) {
    const unsigned int gB = nsInteriorBlocks(a.Nint, a.shape), gP = nblk(a.Nopen - a.Nint, 256);
    if (evStart) { const hipError_t e = hipEventRecord(evStart, st); if (e != hipSuccess) return e; }
    if (gB + gP) {
        if (a.shape > 0) hipLaunchKernelGGL((k_ns_step<R, true>), dim3(gB + gP), dim3(256), 0, st, a, partE, partN, (int) gP);
        else hipLaunchKernelGGL((k_ns_step<R, false>), dim3(gB + gP), dim3(256), 0, st, a, partE, partN, (int) gP);
    }
    if (evEnd) { const hipError_t e = hipEventRecord(evEnd, st); if (e != hipSuccess) return e; }
    hipLaunchKernelGGL(k_ns_tracer, dim3(1), dim3(1024), 0, st, t, partE, partN, (int) (gB + gP), rule, eps, a.done);
    return hipGetLastError();
}
template <typename R, typename RA>
hipError_t launchNsVelocity(const NsDev<R> &a, RA *u, double *parts, hipStream_t st) {
    hipLaunchKernelGGL((k_ns_velocity<R, RA>), dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, u, parts);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsInit(const NsDev<R> &a, double dP, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_init<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a, dP);
    return hipGetLastError();
}
hipError_t launchNsUpdateDyn(const int8_t *mask, const int *list, uint8_t *cls, int Nopen, double pr, unsigned int *changed, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_update_dyn, dim3(nblk(Nopen, 256)), dim3(256), 0, st, mask, list, cls, Nopen, pr, changed);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsReopen(const NsDev<R> &a, const uint8_t *, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_reopen<R>, dim3(nblk(a.Nopen, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}
template <typename R>
hipError_t launchScatterCompact(const R *src, const int *list, int n, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_scatter<R>, dim3(nblk(n, 256)), dim3(256), 0, st, src, list, n, dst);
    return hipGetLastError();
}
template <typename R>
hipError_t launchGatherCompact(const double *src, const int *list, int n, R *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_gather<R>, dim3(nblk(n, 256)), dim3(256), 0, st, src, list, n, dst);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// the batched form: every member of a batch in one launch
// ------------------------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_ns_step_group(const NsMember<R> *__restrict__ E, const int2 *__restrict__ map) {
    const int2 e = map[xcdBlockOf((int) blockIdx.x, (int) gridDim.x)];
    const NsMember<R> &g = E[e.x];
    if (!g.active || *g.a.done) return;
    if (e.y < g.gP) nsPlaneBlock<R>(g.a, g.partE, g.partN, e.y, g.gB);
    else nsBulkBlocks<R, false>(g.a, g.partE, g.partN, e.y - g.gP, g.gB);
}
template <typename R>
__global__ void __launch_bounds__(1024) k_ns_tracer_group(const NsMember<R> *__restrict__ E) {
    const NsMember<R> &g = E[blockIdx.x];
    if (!g.active) return;
    nsTracerBlock(g.tracer, g.partE, g.partN, g.gB + g.gP, g.rule, g.eps, g.a.done);
}
template <typename R>
hipError_t launchNsStepGroup(const NsMember<R> *E, int nMembers, const int2 *map, int nBlocks, hipStream_t st) {
    if (nBlocks > 0) hipLaunchKernelGGL(k_ns_step_group<R>, dim3(nBlocks), dim3(256), 0, st, E, map);
    if (nMembers > 0) hipLaunchKernelGGL(k_ns_tracer_group<R>, dim3(nMembers), dim3(1024), 0, st, E);
    return hipGetLastError();
}
template <typename R>
__global__ void k_ns_group_status(const NsMember<R> *__restrict__ E, int n, long long *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const NsTracer *t = E[j].tracer;
    out[4 * j + 0] = t->iters;
    out[4 * j + 1] = t->converged;
    out[4 * j + 2] = t->nonfinite;
    out[4 * j + 3] = *E[j].a.done;
}
template <typename R>
hipError_t launchNsGroupStatus(const NsMember<R> *E, int nMembers, long long *out, hipStream_t st) {
    if (nMembers > 0) hipLaunchKernelGGL(k_ns_group_status<R>, dim3((nMembers + 63) / 64), dim3(64), 0, st, E, nMembers, out);
    return hipGetLastError();
}
template <typename R, int MAXS>
__global__ void __launch_bounds__(256) k_ade_bulk_group(const AdeMember<R> *__restrict__ E, const int2 *__restrict__ map) {
    const int2 e = map[xcdBlockOf((int) blockIdx.x, (int) gridDim.x)];
    const AdeMember<R> &g = E[e.x];
    if (halted(g.a.halt)) return;
    double bt = -1e300; long long mt = 0;
    adeBulkVoxel<R, MAXS, true>(g.a, g.k, e.y * (int) blockDim.x + (int) threadIdx.x, bt, mt);
    if (g.a.maskTotal) blockMaskTotal(bt, mt, g.a.partMax, g.a.partMask, e.y);
}
template <typename R>
__global__ void __launch_bounds__(256) k_ade_planes_group(const AdeMember<R> *__restrict__ E, const int2 *__restrict__ map) {
    const int2 e = map[blockIdx.x];
    const AdeMember<R> &g = E[e.x];
    if (halted(g.a.halt)) return;
    adePlanesBlock<R>(g.a, e.y & 0xffff, e.y >> 16, g.gInt);
}
template <typename R>
hipError_t launchAdeStepGroup(const AdeMember<R> *E, const int2 *mapInt, int nInt, const int2 *mapPln, int nPln, int maxSm,
                              int tile, hipStream_t st) {
    if (nInt > 0) {
        const dim3 g(nInt), b(tile);
        if (maxSm <= 1) hipLaunchKernelGGL((k_ade_bulk_group<R, 1>), g, b, 0, st, E, mapInt);
        else if (maxSm <= 2) hipLaunchKernelGGL((k_ade_bulk_group<R, 2>), g, b, 0, st, E, mapInt);
        else if (maxSm <= 3) hipLaunchKernelGGL((k_ade_bulk_group<R, 3>), g, b, 0, st, E, mapInt);
        else if (maxSm <= 4) hipLaunchKernelGGL((k_ade_bulk_group<R, 4>), g, b, 0, st, E, mapInt);
        else if (maxSm <= 6) hipLaunchKernelGGL((k_ade_bulk_group<R, 6>), g, b, 0, st, E, mapInt);
        else hipLaunchKernelGGL((k_ade_bulk_group<R, 8>), g, b, 0, st, E, mapInt);
    }
    if (nPln > 0) hipLaunchKernelGGL(k_ade_planes_group<R>, dim3(nPln), dim3(tile), 0, st, E, mapPln);
    return hipGetLastError();
}

#define INST(R) \
    template hipError_t launchAdeStep<R>(const AdeDev<R> &, const KinDev &, hipStream_t, hipEvent_t); \
    template hipError_t launchKinetics<R>(const AdeDev<R> &, const KinDev &, hipStream_t); \
    template hipError_t launchEnergyField<R>(const AdeDev<R> &, const KinDev &, double *, int, hipStream_t); \
    template hipError_t launchKineticsListed<R>(const AdeDev<R> &, const KinDev &, const int *, const unsigned int *, hipStream_t); \
    template hipError_t launchNearPhaseList<R>(const AdeDev<R> &, const KinDev &, int *, unsigned int *, hipStream_t); \
    template hipError_t launchDissolPark<R>(const AdeDev<R> &, const KinDev &, hipStream_t); \
    template hipError_t launchDissolConvert<R>(const AdeDev<R> &, int8_t *, const DissolConv &, unsigned int *, hipStream_t); \
    template hipError_t launchAdeDensity<R>(const AdeDev<R> &, int, double *, hipStream_t); \
    template hipError_t launchAdeSet<R>(const AdeDev<R> &, int, const double *, hipStream_t); \
    template hipError_t launchAdeReopen<R>(const AdeDev<R> &, const uint8_t *, hipStream_t); \
    template hipError_t launchAdeTotal<R>(const AdeDev<R> &, int, double *, hipStream_t); \
    template hipError_t launchAdeHeld<R>(const AdeDev<R> &, int, double *, hipStream_t); \
    template hipError_t launchNsStep<R>(const NsDev<R> &, double *, double *, NsTracer *, int, double, hipStream_t, hipEvent_t, hipEvent_t); \
    template hipError_t launchNsVelocity<R, R>(const NsDev<R> &, R *, double *, hipStream_t); \
    template hipError_t launchNsInit<R>(const NsDev<R> &, double, hipStream_t); \
    template hipError_t launchNsReopen<R>(const NsDev<R> &, const uint8_t *, hipStream_t); \
    template hipError_t launchNsBBFill<R>(const NsDev<R> &, const R *, R *, hipStream_t); \
    template hipError_t launchAdeBBFill<R>(const AdeDev<R> &, const R *, R *, hipStream_t); \
    template hipError_t launchNsStepGroup<R>(const NsMember<R> *, int, const int2 *, int, hipStream_t); \
    template hipError_t launchNsGroupStatus<R>(const NsMember<R> *, int, long long *, hipStream_t); \
    template hipError_t launchAdeStepGroup<R>(const AdeMember<R> *, const int2 *, int, const int2 *, int, int, int, hipStream_t); \
    template hipError_t launchScatterCompact<R>(const R *, const int *, int, double *, hipStream_t); \
    template hipError_t launchGatherCompact<R>(const double *, const int *, int, R *, hipStream_t);
INST(double)
INST(float)
// FP32 transport lattices fed by an FP64 flow solve
template hipError_t launchNsVelocity<double, float>(const NsDev<double> &, float *, double *, hipStream_t);

}  // namespace clb

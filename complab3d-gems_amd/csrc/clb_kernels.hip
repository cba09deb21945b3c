// clb_kernels.hip -- HIP kernels of the coupled reaction step (gfx950).
//
// Storage (one case):
//   ADE populations  f[slot][q][N], two buffers, slot = mobile species, full form
//                    (density = sum f), q = 0..6 (rest, +x,-x,+y,-y,+z,-z)
//   D3Q19 flow       g[q][N], two buffers, shifted form (density = 1 + sum g)
//   scalars          immobile species, biomass, velocity: [k][N]
//   voxel bytes      live mask, ADE class, flow class (+ the class seen by the
//                    previous launch's pull), static geometry
// N = nx*ny*nz, index x + nx*(y + ny*z), x fastest; one thread per voxel.
//
// Streaming is a pull inside the fused kernel: a launch reads the
// post-collision populations of the previous launch (`cur`) from the
// neighbours and writes its own post-collision populations to `out`.
// Bounce-back cells are never stored or visited: a population that would come
// from a bounce-back neighbour is the voxel's own opposite population from two
// launches back, which is still in `out` until this thread overwrites it.  That
// is exactly the full-way bounce-back of a Palabos BounceBack cell (swap on
// collide, stream back out), without the cell.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clb_device.h"
#include "clb_eq_device.h"

namespace clb {

__device__ __constant__ int D7C[7][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
__device__ __constant__ int D19C[19][3] = {
    {0,0,0},
    {1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
    {1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0},
    {1,0,1},{-1,0,-1},{1,0,-1},{-1,0,1},
    {0,1,1},{0,-1,-1},{0,1,-1},{0,-1,1}};

#define W7R 0.25
#define W7F 0.125
__device__ __forceinline__ double w7(int q) { return q == 0 ? W7R : W7F; }
__device__ __forceinline__ int opp7(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__device__ __forceinline__ int opp19(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__device__ __forceinline__ double w19(int q) { return q == 0 ? (1.0 / 3.0) : (q < 7 ? (1.0 / 18.0) : (1.0 / 36.0)); }

__device__ __forceinline__ void coords(long long idx, int nx, int ny, int &x, int &y, int &z) {
    long long t = idx / nx;
    x = (int) (idx - t * nx);
    z = (int) (t / ny);
    y = (int) (t - (long long) z * ny);
}

// is species slot closed (bounce-back) at voxel (x, class byte)?
__device__ __forceinline__ bool adeClosed(int cls, int x, int nx, int lbt, int rbt) {
    return cls == ADE_BB || (x == 0 && lbt == BC_CLOSED_DEV) || (x == nx - 1 && rbt == BC_CLOSED_DEV);
}

// ------------------------------------------------------------------------------------------
// pull of one D3Q7 population set at voxel idx (coordinates x,y,z)
// ------------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void pull7(const R *__restrict__ cur, const R *__restrict__ old,
                                      const uint8_t *__restrict__ clsPull, long long N, int nx, int ny, int nz,
                                      long long idx, int x, int y, int z, int lbt, int rbt, int first,
                                      double fl[7]) {
    if (first) {
#pragma unroll
        for (int q = 0; q < 7; ++q) fl[q] = (double) cur[(long long) q * N + idx];
        return;
    }
    fl[0] = (double) cur[idx];
#pragma unroll
    for (int q = 1; q < 7; ++q) {
        const int xs = x - D7C[q][0], ys = y - D7C[q][1], zs = z - D7C[q][2];
        if (xs < 0 || ys < 0 || zs < 0 || xs >= nx || ys >= ny || zs >= nz) {
            fl[q] = (double) cur[(long long) opp7(q) * N + idx];          // edge: own opposite
        } else {
            const long long n = idx - (D7C[q][0] + (long long) nx * (D7C[q][1] + (long long) ny * D7C[q][2]));
            if (adeClosed(clsPull[n], xs, nx, lbt, rbt)) fl[q] = (double) old[(long long) opp7(q) * N + idx];
            else fl[q] = (double) cur[(long long) q * N + n];
        }
    }
}

// regularized first-order advection-diffusion Dirichlet completion
__device__ __forceinline__ void adeDirichlet(double fl[7], int atLeft, double rhoD) {
    const int unk = atLeft ? 1 : 2;
    double s = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) if (q != unk) s += fl[q];
    fl[unk] = rhoD - s;
    const double jx = fl[1] - fl[2], jy = fl[3] - fl[4], jz = fl[5] - fl[6];
    fl[0] = W7R * rhoD;
    fl[1] = W7F * (rhoD + 4.0 * jx);
    fl[2] = W7F * (rhoD - 4.0 * jx);
    fl[3] = W7F * (rhoD + 4.0 * jy);
    fl[4] = W7F * (rhoD - 4.0 * jy);
    fl[5] = W7F * (rhoD + 4.0 * jz);
    fl[6] = W7F * (rhoD - 4.0 * jz);
}

// ------------------------------------------------------------------------------------------
// kinetics (the rate laws of the reference's compiled-kinetics headers)
// ------------------------------------------------------------------------------------------
__device__ void bioticLaw(const KinDev &k, const double *C, const double *B, double *sr, double *br, int mask) {
    for (int s = 0; s < k.S; ++s) sr[s] = 0;
    for (int m = 0; m < k.M; ++m) br[m] = 0;
    if (mask < 2) return;
    const double S = k.donor >= 0 ? fmax(C[k.donor], 0.0) : 0.0;
    const double E = k.acceptor >= 0 ? fmax(C[k.acceptor], 0.0) : 1.0;
    for (int m = 0; m < k.M; ++m) {
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
}

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

__device__ __forceinline__ double clampInc(int clamp, double d, double cur, unsigned int &binds) {
    if (!clamp || d >= 0) return d;
    const double lim = -fmax(cur, 0.0);
    if (d < lim) { ++binds; return lim; }
    return d;
}

// ------------------------------------------------------------------------------------------
// Neumann planes: density of the inner layer after this launch's stream
// ------------------------------------------------------------------------------------------
template <typename R>
__global__ void k_ade_neumann(AdeDev<R> a) {
    const long long t = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long) a.ny * a.nz;
    if (t >= 2 * plane) return;
    const int side = (int) (t / plane);               // 0 left (x=1), 1 right (x=nx-2)
    const long long k = t - side * plane;
    const int y = (int) (k % a.ny), z = (int) (k / a.ny);
    const int x = side == 0 ? 1 : a.nx - 2;
    const long long idx = x + (long long) a.nx * (y + (long long) a.ny * z);
    for (int ms = 0; ms < a.Sm; ++ms) {
        const int s = a.mobSpecies[ms];
        const int bt = side == 0 ? a.lbt[s] : a.rbt[s];
        if (bt != BC_NEUMANN_DEV) continue;
        double r = 0;
        // a bounce-back inner cell reports its stored density, 0
        if (!adeClosed(a.cls[idx], x, a.nx, a.lbt[s], a.rbt[s])) {
            double fl[7];
            const R *cur = a.cur + (long long) ms * 7 * a.N;
            const R *old = a.out + (long long) ms * 7 * a.N;
            pull7<R>(cur, old, a.clsPull, a.N, a.nx, a.ny, a.nz, idx, x, y, z, a.lbt[s], a.rbt[s], a.first, fl);
            for (int q = 0; q < 7; ++q) r += fl[q];
        }
        a.rhoN[(long long) (side * a.Smax + ms) * plane + k] = r;
    }
}

// ------------------------------------------------------------------------------------------
// fused ADE step: stream (pull) + boundary planes + BGK collide + biotic and
// abiotic kinetics with the clamp and the D3Q7 deposit
// ------------------------------------------------------------------------------------------
template <typename R, int MAXS>
__global__ void __launch_bounds__(256) k_ade_step(AdeDev<R> a, KinDev k) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    const int cls = a.cls[idx];
    if (cls == ADE_BB) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    const long long N = a.N;
    const long long plane = (long long) a.ny * a.nz;
    const long long pk = y + (long long) a.ny * z;

    double f[MAXS][7];
    double rho[MAXS];
    bool act[MAXS];
    double ux = 0, uy = 0, uz = 0;
    if (a.u) { ux = (double) a.u[idx]; uy = (double) a.u[N + idx]; uz = (double) a.u[2 * N + idx]; }

#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        act[ms] = false;
        rho[ms] = 0;
        if (ms >= a.Sm) continue;
        const int s = a.mobSpecies[ms];
        const int lbt = a.lbt[s], rbt = a.rbt[s];
        if (adeClosed(cls, x, a.nx, lbt, rbt)) continue;
        act[ms] = true;
        R *cur = a.cur + (long long) ms * 7 * N;
        R *old = a.out + (long long) ms * 7 * N;
        pull7<R>(cur, old, a.clsPull, N, a.nx, a.ny, a.nz, idx, x, y, z, lbt, rbt, a.first, f[ms]);
        if (a.first) {
            // hand the stored population of each closed neighbour to this voxel's
            // buffer slot that the bounce-back rule reads at the next launch
#pragma unroll
            for (int q = 1; q < 7; ++q) {
                const int xs = x - D7C[q][0], ys = y - D7C[q][1], zs = z - D7C[q][2];
                if (xs < 0 || ys < 0 || zs < 0 || xs >= a.nx || ys >= a.ny || zs >= a.nz) continue;
                const long long n = idx - (D7C[q][0] + (long long) a.nx * (D7C[q][1] + (long long) a.ny * D7C[q][2]));
                if (adeClosed(a.clsPull[n], xs, a.nx, lbt, rbt))
                    cur[(long long) opp7(q) * N + idx] = cur[(long long) opp7(q) * N + n];
            }
        }
        if (a.reset && a.geom[idx] >= 2) {
            double c0 = a.c0[s];
            if (c0 < 1e-12 && c0 > -1e-12) c0 = 0;
#pragma unroll
            for (int q = 0; q < 7; ++q) f[ms][q] = w7(q) * c0;
        }
        if (x == 0 || x == a.nx - 1) {
            const int left = (x == 0);
            const int bt = left ? lbt : rbt;
            double rhoD = left ? a.lbc[s] : a.rbc[s];
            if (bt == BC_NEUMANN_DEV) rhoD = a.rhoN[(long long) ((left ? 0 : 1) * a.Smax + ms) * plane + pk];
            adeDirichlet(f[ms], left, rhoD);
        }
        const double om = (cls == ADE_BFILM) ? a.omB[s] : a.omP[s];
        double r = 0;
#pragma unroll
        for (int q = 0; q < 7; ++q) r += f[ms][q];
        const double t0 = 4.0 * r;
        const double fe0 = W7R * r;
        f[ms][0] = f[ms][0] * (1.0 - om) + om * fe0;
        const double fe1 = W7F * (r + t0 * ux), fe2 = W7F * (r - t0 * ux);
        const double fe3 = W7F * (r + t0 * uy), fe4 = W7F * (r - t0 * uy);
        const double fe5 = W7F * (r + t0 * uz), fe6 = W7F * (r - t0 * uz);
        f[ms][1] = f[ms][1] * (1.0 - om) + om * fe1;
        f[ms][2] = f[ms][2] * (1.0 - om) + om * fe2;
        f[ms][3] = f[ms][3] * (1.0 - om) + om * fe3;
        f[ms][4] = f[ms][4] * (1.0 - om) + om * fe4;
        f[ms][5] = f[ms][5] * (1.0 - om) + om * fe5;
        f[ms][6] = f[ms][6] * (1.0 - om) + om * fe6;
        rho[ms] = r;
    }

    // -------- kinetics over the reaction box x = 1..nx-2 on open voxels --------
    if (k.on && x > 0 && x < a.nx - 1) {
        const int mask = a.mask[idx];
        if (mask != 0 && mask != 1) {
            unsigned int binds = 0;
            double C[8], sr[8];
            const int S = k.S;
            auto concOf = [&](int s) -> double {
                const int ms = a.slot[s];
                if (a.mobile[s]) {
                    double r = 0;
#pragma unroll
                    for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == ms) {
#pragma unroll
                        for (int q = 0; q < 7; ++q) r += f[m2][q];
                    }
                    return r;
                }
                return (double) a.imm[(long long) ms * N + idx];
            };
            auto depositTo = [&](int s, double d) {
                const int ms = a.slot[s];
                if (a.mobile[s]) {
#pragma unroll
                    for (int m2 = 0; m2 < MAXS; ++m2) if (m2 == ms) {
#pragma unroll
                        for (int q = 0; q < 7; ++q) f[m2][q] += w7(q) * d;
                    }
                } else {
                    R *p = a.imm + (long long) ms * N + idx;
                    *p = (R) ((double) *p + d);
                }
            };
            if (k.bio && k.M > 0) {
                double Bv[CLB_DEV_MAXM], br[CLB_DEV_MAXM];
                for (int s = 0; s < S; ++s) { C[s] = concOf(s); if (C[s] < 1e-12) C[s] = 0; }
                for (int m = 0; m < k.M; ++m) { Bv[m] = (double) a.B[(long long) m * N + idx]; if (Bv[m] < 1e-12) Bv[m] = 0; }
                bioticLaw(k, C, Bv, sr, br, mask);
                for (int s = 0; s < S; ++s) {
                    const double d = sr[s] * k.dt;
                    if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
                }
                for (int m = 0; m < k.M; ++m) {
                    double d = br[m] * k.dt;
                    if (d > 1e-12 || d < -1e-12) {
                        R *p = a.B + (long long) m * N + idx;
                        const double b = (double) *p;
                        if (d < -b) d = -b;
                        *p = (R) (b + d);
                    }
                }
            }
            if (k.nrxn > 0) {
                bool react = true;
                if (k.surface_only) {
                    react = false;
#pragma unroll
                    for (int q = 1; q < 7; ++q) {
                        const int X = x + D7C[q][0], Y = y + D7C[q][1], Z = z + D7C[q][2];
                        if (X < 0 || Y < 0 || Z < 0 || X >= a.nx || Y >= a.ny || Z >= a.nz) continue;
                        const int nm = a.mask[idx + D7C[q][0] + (long long) a.nx * (D7C[q][1] + (long long) a.ny * D7C[q][2])];
                        if (nm == 0 || nm == 1) react = true;
                    }
                }
                if (react) {
                    for (int s = 0; s < S; ++s) { C[s] = concOf(s); if (C[s] < 1e-12) C[s] = 0; }
                    abioticLaw(k, C, sr);
                    for (int s = 0; s < S; ++s) {
                        const double d = sr[s] * k.dt;
                        if (d > 1e-12 || d < -1e-12) depositTo(s, clampInc(k.clamp, d, concOf(s), binds));
                    }
                }
            }
            if (binds) atomicAdd(a.clampCount, (unsigned long long) binds);
        }
    }

#pragma unroll
    for (int ms = 0; ms < MAXS; ++ms) {
        if (!act[ms]) continue;
        R *o = a.out + (long long) ms * 7 * N;
#pragma unroll
        for (int q = 0; q < 7; ++q) o[(long long) q * N + idx] = (R) f[ms][q];
    }
}

// density after the stream, for reads and reductions (closed voxels: 0)
template <typename R>
__global__ void k_ade_density(AdeDev<R> a, int ms, double *dst) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    const int s = a.mobSpecies[ms];
    if (adeClosed(a.cls[idx], x, a.nx, a.lbt[s], a.rbt[s])) { dst[idx] = 0.0; return; }
    double fl[7];
    pull7<R>(a.cur + (long long) ms * 7 * a.N, a.out + (long long) ms * 7 * a.N, a.clsPull, a.N, a.nx, a.ny, a.nz,
             idx, x, y, z, a.lbt[s], a.rbt[s], a.first, fl);
    double r = 0;
    for (int q = 0; q < 7; ++q) r += fl[q];
    dst[idx] = r;
}

// writes populations at rest for a concentration field into `cur` (every voxel)
template <typename R>
__global__ void k_ade_set(AdeDev<R> a, int ms, const double *src) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    R *cur = a.cur + (long long) ms * 7 * a.N;
    const double c = src[idx];
    for (int q = 0; q < 7; ++q) cur[(long long) q * a.N + idx] = (R) (w7(q) * c);
}

// a voxel reopened this step: links from closed neighbours carry the rest value
template <typename R>
__global__ void k_ade_reopen(AdeDev<R> a, const uint8_t *reopened) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N || !reopened[idx]) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    for (int ms = 0; ms < a.Sm; ++ms) {
        const int s = a.mobSpecies[ms];
        R *old = a.out + (long long) ms * 7 * a.N;
        a.cur[(long long) ms * 7 * a.N + idx] = (R) (W7R * a.c0[s]);
        for (int q = 1; q < 7; ++q) {
            const int xs = x - D7C[q][0], ys = y - D7C[q][1], zs = z - D7C[q][2];
            bool closedSrc = true;
            if (!(xs < 0 || ys < 0 || zs < 0 || xs >= a.nx || ys >= a.ny || zs >= a.nz)) {
                const long long n = idx - (D7C[q][0] + (long long) a.nx * (D7C[q][1] + (long long) a.ny * D7C[q][2]));
                closedSrc = adeClosed(a.clsPull[n], xs, a.nx, a.lbt[s], a.rbt[s]);
            }
            if (closedSrc) old[(long long) opp7(q) * a.N + idx] = (R) (W7F * a.c0[s]);
        }
    }
}

// ------------------------------------------------------------------------------------------
// D3Q19 incompressible BGK with regularized pressure planes
// ------------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void pull19(const NsDev<R> &a, long long idx, int x, int y, int z, double fl[19]) {
    const long long N = a.N;
    if (a.first) {
#pragma unroll
        for (int q = 0; q < 19; ++q) fl[q] = (double) a.cur[(long long) q * N + idx];
        return;
    }
    fl[0] = (double) a.cur[idx];
#pragma unroll
    for (int q = 1; q < 19; ++q) {
        const int xs = x - D19C[q][0], ys = y - D19C[q][1], zs = z - D19C[q][2];
        if (xs < 0 || ys < 0 || zs < 0 || xs >= a.nx || ys >= a.ny || zs >= a.nz) {
            fl[q] = (double) a.cur[(long long) opp19(q) * N + idx];
        } else {
            const long long n = idx - (D19C[q][0] + (long long) a.nx * (D19C[q][1] + (long long) a.ny * D19C[q][2]));
            if (a.clsPull[n] >= NS_BB) fl[q] = (double) a.out[(long long) opp19(q) * N + idx];
            else fl[q] = (double) a.cur[(long long) q * N + n];
        }
    }
}

__device__ __forceinline__ double feqInc(int q, double rb, double jx, double jy, double jz, double jsq) {
    const double cj = D19C[q][0] * jx + D19C[q][1] * jy + D19C[q][2] * jz;
    return w19(q) * (rb + 3.0 * cj + 4.5 * cj * cj - 1.5 * jsq);
}

// dir = +1 at x = 0 (unknowns c_x = +1), -1 at x = nx-1
__device__ void nsPressurePlane(double fl[19], int dir, double rb, double &jx) {
    double r0 = 0, rout = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        if (D19C[q][0] == 0) r0 += fl[q];
        else if (D19C[q][0] == -dir) rout += fl[q];
    }
    jx = dir * (rb - r0 - 2.0 * rout);
    const double jsq = jx * jx;
    double feq[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) feq[q] = feqInc(q, rb, jx, 0.0, 0.0, jsq);
#pragma unroll
    for (int q = 0; q < 19; ++q) if (D19C[q][0] == dir) fl[q] = feq[q] + (fl[opp19(q)] - feq[opp19(q)]);
    double P[3][3] = {{0,0,0},{0,0,0},{0,0,0}};
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const double ne = fl[q] - feq[q];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) P[i][j] += D19C[q][i] * D19C[q][j] * ne;
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        double s = 0;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += (D19C[q][i] * D19C[q][j] - (i == j ? 1.0 / 3.0 : 0.0)) * P[i][j];
        fl[q] = feq[q] + w19(q) * 4.5 * s;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_ns_step(NsDev<R> a, double *partE, double *partN) {
    __shared__ double sE[256 / 64], sN[256 / 64];
    if (*a.done) return;
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0, cnt = 0;
    if (idx < a.N) {
        const int cls = a.cls[idx];
        if (cls < NS_BB) {
            int x, y, z;
            coords(idx, a.nx, a.ny, x, y, z);
            double fl[19];
            pull19<R>(a, idx, x, y, z, fl);
            if (a.first) {
#pragma unroll
                for (int q = 1; q < 19; ++q) {
                    const int xs = x - D19C[q][0], ys = y - D19C[q][1], zs = z - D19C[q][2];
                    if (xs < 0 || ys < 0 || zs < 0 || xs >= a.nx || ys >= a.ny || zs >= a.nz) continue;
                    const long long n = idx - (D19C[q][0] + (long long) a.nx * (D19C[q][1] + (long long) a.ny * D19C[q][2]));
                    if (a.clsPull[n] >= NS_BB) a.cur[(long long) opp19(q) * a.N + idx] = a.cur[(long long) opp19(q) * a.N + n];
                }
            }
            double rb = 0, jx = 0, jy = 0, jz = 0;
            if (x == 0) { rb = 0.0; nsPressurePlane(fl, +1, rb, jx); }
            else if (x == a.nx - 1) { rb = a.rhoBarRight; nsPressurePlane(fl, -1, rb, jx); }
            else {
#pragma unroll
                for (int q = 0; q < 19; ++q) {
                    rb += fl[q];
                    jx += D19C[q][0] * fl[q]; jy += D19C[q][1] * fl[q]; jz += D19C[q][2] * fl[q];
                }
            }
            const double jsq = jx * jx + jy * jy + jz * jz;
            e = 0.5 * jsq; cnt = 1;
            const double om = cls == NS_BIO ? a.omBio : a.om;
#pragma unroll
            for (int q = 0; q < 19; ++q) {
                const double v = fl[q] + om * (feqInc(q, rb, jx, jy, jz, jsq) - fl[q]);
                a.out[(long long) q * a.N + idx] = (R) v;
            }
        }
    }
    // fixed-order block reduction of the energy statistic
    for (int off = 32; off > 0; off >>= 1) { e += __shfl_down(e, off, 64); cnt += __shfl_down(cnt, off, 64); }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sE[wv] = e; sN[wv] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double E = 0, C = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { E += sE[w]; C += sN[w]; }
        partE[blockIdx.x] = E; partN[blockIdx.x] = C;
    }
}

// the convergence rule, evaluated on the device after every iteration
__global__ void k_ns_tracer(NsTracer *t, const double *partE, const double *partN, int nparts, int rule, double eps, int *done) {
    __shared__ double sE[64], sN[64];
    if (*done) return;
    const int tid = threadIdx.x;   // 64 threads
    double E = 0, C = 0;
    for (int i = tid; i < nparts; i += 64) { E += partE[i]; C += partN[i]; }
    sE[tid] = E; sN[tid] = C;
    __syncthreads();
    if (tid == 0) {
        double Es = 0, Cs = 0;
        for (int i = 0; i < 64; ++i) { Es += sE[i]; Cs += sN[i]; }
        const double val = Cs > 0 ? Es / Cs : 0.0;
        t->iters += 1;
        t->last = val;
        if (!isfinite(val)) { t->nonfinite = 1; *done = 1; return; }
        if (rule == 0) {
            const int W = NS_TRACER_WINDOW;
            t->ring[t->head] = val;
            t->head = (t->head + 1) % W;
            if (t->count < W) t->count += 1;
            if (t->count >= W) {
                double m = 0;
                for (int i = 0; i < W; ++i) m += t->ring[(t->head + i) % W];
                m /= (double) W;
                double s = 0;
                for (int i = 0; i < W; ++i) { const double d = t->ring[(t->head + i) % W] - m; s += d * d; }
                s = sqrt(s / (double) W);
                const double r = s / fabs(m);
                if (isnan(r) || r < eps) { t->converged = 1; *done = 1; }
            }
        } else {
            if (t->prevE > 0 && fabs(val - t->prevE) <= eps * fabs(val)) { t->converged = 1; *done = 1; }
            t->prevE = val;
        }
    }
}

// velocity of the current flow state (computeVelocity): j on open voxels,
// zero on closed ones; optionally the box reductions of the Peclet rescale
template <typename R>
__global__ void __launch_bounds__(256) k_ns_velocity(NsDev<R> a, R *u, const int8_t *geom, double *parts) {
    __shared__ double red[5][4];
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};      // sum|u| box, max ux box, sum ux outlet, sum|u| open, count open
    bool inBox = false, outlet = false;
    double ux = 0, uy = 0, uz = 0;
    if (idx < a.N) {
        int x, y, z;
        coords(idx, a.nx, a.ny, x, y, z);
        if (a.cls[idx] < NS_BB) {
            double fl[19];
            pull19<R>(a, idx, x, y, z, fl);
            if (x == 0) nsPressurePlane(fl, +1, 0.0, ux);
            else if (x == a.nx - 1) nsPressurePlane(fl, -1, a.rhoBarRight, ux);
            else {
                for (int q = 0; q < 19; ++q) { ux += D19C[q][0] * fl[q]; uy += D19C[q][1] * fl[q]; uz += D19C[q][2] * fl[q]; }
            }
        }
        if (u) { u[idx] = (R) ux; u[a.N + idx] = (R) uy; u[2 * a.N + idx] = (R) uz; }
        inBox = x > 0 && x < a.nx - 1;
        outlet = x == a.nx - 2;
        const double n = sqrt(ux * ux + uy * uy + uz * uz);
        if (inBox) { v[0] = n; v[1] = ux; if (geom[idx] >= 2) { v[3] = n; v[4] = 1; } }
        if (outlet) v[2] = ux;
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
    __shared__ double r4[4];
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
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    const double rb = -deltaP * 3.0 / (double) (a.nx - 1) * (double) x;
    const bool closed = a.cls[idx] >= NS_BB;
    for (int q = 0; q < 19; ++q) {
        const R v = closed ? (R) 0 : (R) (w19(q) * rb);
        a.cur[(long long) q * a.N + idx] = v;
        a.out[(long long) q * a.N + idx] = v;
    }
}

// NS dynamics from the live mask (updateNsLatticesDynamics3D); reopened cells
// get rest populations on links from closed neighbours
__global__ void k_ns_update_dyn(const int8_t *mask, uint8_t *cls, long long N, double permRatio, unsigned int *changed) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int m = mask[idx];
    if (m == 0 || m == 1) return;
    const int c = cls[idx];
    const bool isNsOmega = (c == NS_BGK);
    int nc = c;
    if (m != 2 && isNsOmega) nc = permRatio <= 1e-14 ? NS_BB : NS_BIO;
    else if (m == 2 && !isNsOmega) nc = NS_BGK;
    if (nc != c) { cls[idx] = (uint8_t) nc; atomicAdd(changed, 1u); }
}
template <typename R>
__global__ void k_ns_reopen(NsDev<R> a, const uint8_t *oldCls) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    if (!(oldCls[idx] >= NS_BB && a.cls[idx] < NS_BB)) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    // links from neighbours that were closed carry rest (0, shifted form); the
    // next pull reads them from this voxel's own slot of `out`
    a.cur[idx] = (R) 0;
    for (int q = 1; q < 19; ++q) {
        const int xs = x - D19C[q][0], ys = y - D19C[q][1], zs = z - D19C[q][2];
        bool closedSrc = true;
        if (!(xs < 0 || ys < 0 || zs < 0 || xs >= a.nx || ys >= a.ny || zs >= a.nz)) {
            const long long n = idx - (D19C[q][0] + (long long) a.nx * (D19C[q][1] + (long long) a.ny * D19C[q][2]));
            closedSrc = oldCls[n] >= NS_BB;
        }
        if (closedSrc) a.out[(long long) opp19(q) * a.N + idx] = (R) 0;
    }
}

// ------------------------------------------------------------------------------------------
// equilibrium speciation (FP64, one thread per voxel)
// ------------------------------------------------------------------------------------------
template <typename R, int MAXS>
__global__ void __launch_bounds__(128) k_equilibrium(AdeDev<R> a, EqDev e, unsigned int *nonconv) {
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.N) return;
    int x, y, z;
    coords(idx, a.nx, a.ny, x, y, z);
    if (x <= 0 || x >= a.nx - 1) return;
    const int mask = a.mask[idx];
    if (mask == 0 || mask == 1) return;
    const long long N = a.N;
    const int S = e.ns;
    double C[EQ_MAXS], E[EQ_MAXS];
    for (int s = 0; s < S; ++s) {
        double r;
        if (a.mobile[s]) {
            const R *o = a.cur + (long long) a.slot[s] * 7 * N + idx;
            r = 0; for (int q = 0; q < 7; ++q) r += (double) o[(long long) q * N];
        } else r = (double) a.imm[(long long) a.slot[s] * N + idx];
        C[s] = fmax(r, EQ_MIN_CONC);
    }
    const bool conv = eqSolve(e, C, E);
    if (!conv) atomicAdd(nonconv, 1u);
    for (int s = 0; s < S; ++s) {
        double dC = E[s] - C[s];
        const double mx = fmax(1e-9, 0.1 * fabs(C[s]));
        if (fabs(dC) > mx) dC = dC > 0 ? mx : -mx;
        if (C[s] + dC < 1e-20) dC = 1e-20 - C[s];
        if (fabs(dC) > 1e-12) {
            if (a.mobile[s]) {
                R *o = a.cur + (long long) a.slot[s] * 7 * N + idx;
                for (int q = 0; q < 7; ++q) o[(long long) q * N] = (R) ((double) o[(long long) q * N] + w7(q) * dC);
            } else {
                R *p = a.imm + (long long) a.slot[s] * N + idx;
                *p = (R) ((double) *p + dC);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// CA biofilm spill
// ------------------------------------------------------------------------------------------
// updateLocalMaskNtotalLattices3D + the global max of the total and the mask sum
template <typename R>
__global__ void __launch_bounds__(256) k_ca_mask_total(CaDev<R> c, double *partMax, long long *partMask) {
    __shared__ double sMax[4];
    __shared__ long long sMask[4];
    const long long idx = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    double bt = -1e300; long long ms = 0;
    if (idx < c.N) {
        int mask = c.mask[idx];
        if (mask != 0 && mask != 1) {
            double bm = 0; int nm = 0;
            for (int m = 0; m < c.M; ++m) { const double v = (double) c.B[(long long) m * c.N + idx]; if (v > 1e-14) { bm += v; nm += 3 + m; } }
            double t = (double) c.Btot[idx];
            if (fabs(bm - t) > 1e-14) { c.Btot[idx] = (R) bm; t = bm; }
            const bool pore = (mask == 2);
            if (!pore && bm < c.thrdB) { mask = 2; c.mask[idx] = 2; }
            else if (pore && bm >= c.thrdB && nm > 0) { mask = nm; c.mask[idx] = (int8_t) nm; }
        }
        bt = (double) c.Btot[idx];
        if (isnan(bt)) bt = __builtin_nan("");
        ms = mask;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(bt, off, 64);
        bt = (isnan(bt) || isnan(o)) ? __builtin_nan("") : fmax(bt, o);
        ms += __shfl_down(ms, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sMax[wv] = bt; sMask[wv] = ms; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = -1e300; long long s = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { m = (isnan(m) || isnan(sMax[w])) ? __builtin_nan("") : fmax(m, sMax[w]); s += sMask[w]; }
        partMax[blockIdx.x] = m; partMask[blockIdx.x] = s;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t caDraw(uint64_t seed, uint64_t sweep, uint64_t vox, uint32_t k) {
    uint64_t h = mix64(seed ^ mix64(sweep * 0x100000001b3ULL + 0x51ed27ULL));
    h = mix64(h ^ (vox * 0x9e3779b97f4a7c15ULL));
    h = mix64(h + k);
    return (uint32_t) (h >> 32);
}

template <typename R>
__global__ void __launch_bounds__(256) k_ca_push(CaDev<R> c, uint64_t sweep) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    c.pushed[i] = 0;
    const int mask = c.mask[i];
    if (mask == 0 || mask == 1) return;
    if (c.half && mask == 2) return;
    double bMt = (double) c.Btot[i];
    if (!(bMt > c.bmax)) return;
    double bMd;
    if (c.half) bMd = bMt * 0.5;
    else { bMd = bMt - c.bmax; if (!(bMd > 1e-12)) return; }
    int x, y, z;
    coords(i, c.nx, c.ny, x, y, z);
    int cand[6]; int nc = 0;
    for (int d = 0; d < 6; ++d) {
        const int q = d + 1;
        const int X = x + D7C[q][0], Y = y + D7C[q][1], Z = z + D7C[q][2];
        if (q == 1 && x == c.nx - 2) continue;
        if (q == 2 && x == 1) continue;
        if (q == 3 && y == c.ny - 1) continue;
        if (q == 4 && y == 0) continue;
        if (q == 5 && z == c.nz - 1) continue;
        if (q == 6 && z == 0) continue;
        if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
        const int nm = c.mask[X + (long long) c.nx * (Y + (long long) c.ny * Z)];
        if (nm == 0 || nm == 1) continue;
        cand[nc++] = d;
    }
    if (nc == 0) return;
    uint32_t kk = 0;
    int perm[6];
    for (int a = 0; a < nc; ++a) perm[a] = a;
    for (int a = nc - 1; a > 0; --a) {
        const int b = (int) (caDraw(c.seed, sweep, (uint64_t) i, kk++) % (uint32_t) (a + 1));
        const int t = perm[a]; perm[a] = perm[b]; perm[b] = t;
    }
    double Bl[CLB_DEV_MAXM], ob[CLB_DEV_MAXM][6];
    for (int m = 0; m < c.M; ++m) { Bl[m] = (double) c.B[(long long) m * c.N + i]; for (int d = 0; d < 6; ++d) ob[m][d] = 0; }
    bool done = false;
    for (int a = 0; a < nc; ++a) {
        const int d = cand[perm[a]], q = d + 1;
        const long long j = i + D7C[q][0] + (long long) c.nx * (D7C[q][1] + (long long) c.ny * D7C[q][2]);
        const int nm = c.mask[j];
        const double nb = (nm == 2) ? 0.0 : (double) c.Btot[j];
        if (nb < c.bmax) {
            const double cap = c.bmax - nb;
            double part = bMd;
            if (bMd > cap) { part = cap; bMd -= cap; } else { bMd = 0; done = true; }
            for (int m = 0; m < c.M; ++m) {
                const double sh = Bl[m] / bMt * part;
                if (sh > 1e-12) { Bl[m] -= sh; ob[m][d] += sh; }
            }
            bMt -= part;
        }
        if (done) break;
    }
    if (!done) {
        const int id0 = c.age[i];
        int young[6]; int ny = 0;
        for (int a = 0; a < nc; ++a) {
            const int q = cand[a] + 1;
            const int id1 = c.age[i + D7C[q][0] + (long long) c.nx * (D7C[q][1] + (long long) c.ny * D7C[q][2])];
            if (id0 > id1) young[ny++] = cand[a];
        }
        int d;
        if (ny > 1) d = young[caDraw(c.seed, sweep, (uint64_t) i, kk++) % (uint32_t) ny];
        else if (ny == 1) d = young[0];
        else d = cand[nc > 1 ? caDraw(c.seed, sweep, (uint64_t) i, kk++) % (uint32_t) nc : 0];
        for (int m = 0; m < c.M; ++m) {
            const double sh = Bl[m] / bMt * bMd;
            Bl[m] -= sh; ob[m][d] += sh;
        }
    }
    for (int m = 0; m < c.M; ++m) {
        c.B[(long long) m * c.N + i] = (R) Bl[m];
        for (int d = 0; d < 6; ++d) c.outbox[((long long) m * 6 + d) * c.N + i] = (R) ob[m][d];
    }
    c.pushed[i] = 1;
}

template <typename R>
__global__ void __launch_bounds__(256) k_ca_gather(CaDev<R> c) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    int x, y, z;
    coords(i, c.nx, c.ny, x, y, z);
    double add[CLB_DEV_MAXM];
    for (int m = 0; m < c.M; ++m) add[m] = 0;
    bool any = false;
    for (int d = 0; d < 6; ++d) {
        const int q = d + 1;
        const int X = x - D7C[q][0], Y = y - D7C[q][1], Z = z - D7C[q][2];
        if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
        const long long j = X + (long long) c.nx * (Y + (long long) c.ny * Z);
        if (!c.pushed[j]) continue;
        any = true;
        for (int m = 0; m < c.M; ++m) add[m] += (double) c.outbox[((long long) m * 6 + d) * c.N + j];
    }
    if (!any) return;
    for (int m = 0; m < c.M; ++m)
        if (add[m] != 0.0) c.B[(long long) m * c.N + i] = (R) ((double) c.B[(long long) m * c.N + i] + add[m]);
}

// updateAgeDistance3D, one Jacobi sweep (ageIn -> ageOut) + the sum of the increments
template <typename R>
__global__ void __launch_bounds__(256) k_ca_age(CaDev<R> c, const int *ageIn, int *ageOut, long long *partSum) {
    __shared__ long long sS[4];
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    long long sum = 0;
    if (i < c.N) {
        const int a = ageIn[i];
        int na = a;
        const double B = (double) c.Btot[i];
        if (B > 1e-14) {
            int x, y, z;
            coords(i, c.nx, c.ny, x, y, z);
            long long del[6]; int nd = 0;
            for (int q = 1; q < 7; ++q) {
                const int X = x + D7C[q][0], Y = y + D7C[q][1], Z = z + D7C[q][2];
                if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
                const long long j = X + (long long) c.nx * (Y + (long long) c.ny * Z);
                if (c.dist[j] > 0) del[nd++] = j;
            }
            if (a == 0) na = 1;
            else if (a == 1 && (B - c.bmax) > -1e-14) {
                bool upd = true;
                for (int t = 0; t < nd; ++t) if (c.dist[del[t]] > 0 && ageIn[del[t]] == 0) { upd = false; break; }
                if (upd) na = 2;
            } else {
                int cnt = 0;
                for (int t = 0; t < nd; ++t) if (ageIn[del[t]] >= a && ((double) c.Btot[del[t]] - c.bmax) > -1e-14) ++cnt;
                if (cnt == nd) na = a + 1;
            }
        }
        ageOut[i] = na;
        sum = na - a;
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sS[wv] = sum;
    __syncthreads();
    if (threadIdx.x == 0) { long long s = 0; for (int w = 0; w < (int) (blockDim.x >> 6); ++w) s += sS[w]; partSum[blockIdx.x] = s; }
}

// ------------------------------------------------------------------------------------------
// mask-driven switches and precipitation
// ------------------------------------------------------------------------------------------
// updateSoluteDynamics3D: pore/biofilm relaxation of open voxels
__global__ void k_ade_omega_switch(const int8_t *mask, uint8_t *cls, long long N) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int m = mask[i];
    if (m == 0 || m == 1) return;
    if (cls[i] == ADE_BB) return;
    cls[i] = m == 2 ? ADE_PORE : ADE_BFILM;
}
// updateSoluteSolidDynamics3D: bounce-back <-> BGK from the live mask
__global__ void k_ade_solid_switch(const int8_t *mask, uint8_t *cls, uint8_t *reopened, long long N, unsigned int *nre) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int m = mask[i];
    const bool want = (m == 0 || m == 1);
    const bool isBB = cls[i] == ADE_BB;
    reopened[i] = 0;
    if (want && !isBB) cls[i] = ADE_BB;
    else if (!want && isBB) { cls[i] = m == 2 ? ADE_PORE : ADE_BFILM; reopened[i] = 1; atomicAdd(nre, 1u); }
}
// precipNodeConversion3D over x = 1..nx-2
template <typename R>
__global__ void k_precip(AdeDev<R> a, int slot, double maxRho, int phaseId, int8_t *phase, unsigned int *changed) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    int x, y, z;
    coords(i, a.nx, a.ny, x, y, z);
    if (x <= 0 || x >= a.nx - 1) return;
    const int m = a.mask[i];
    if (m == 0 || m == 1) return;
    const double P = (double) a.imm[(long long) slot * a.N + i];
    if (P >= maxRho) {
        a.mask[i] = 0;
        if (phaseId > 0) phase[i] = (int8_t) phaseId;
        atomicAdd(changed, 1u);
    }
}

// species totals over open voxels of x = 1..nx-2 (per-block partials)
template <typename R>
__global__ void __launch_bounds__(256) k_species_total(AdeDev<R> a, int s, double *parts) {
    __shared__ double sS[4];
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (i < a.N) {
        int x, y, z;
        coords(i, a.nx, a.ny, x, y, z);
        const int m = a.mask[i];
        if (x > 0 && x < a.nx - 1 && m != 0 && m != 1) {
            if (a.mobile[s]) {
                const int ms = a.slot[s];
                if (!adeClosed(a.cls[i], x, a.nx, a.lbt[s], a.rbt[s])) {
                    double fl[7];
                    pull7<R>(a.cur + (long long) ms * 7 * a.N, a.out + (long long) ms * 7 * a.N, a.clsPull, a.N,
                             a.nx, a.ny, a.nz, i, x, y, z, a.lbt[s], a.rbt[s], a.first, fl);
                    for (int q = 0; q < 7; ++q) v += fl[q];
                }
            } else v = (double) a.imm[(long long) a.slot[s] * a.N + i];
        }
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sS[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) { double s2 = 0; for (int w = 0; w < (int) (blockDim.x >> 6); ++w) s2 += sS[w]; parts[blockIdx.x] = s2; }
}

// fixed-order final reduction of per-block partials (one block of 64 threads)
__global__ void k_final_sum(const double *parts, long long n, int stride, int comp, int isMax, double *out) {
    __shared__ double s[64];
    const int t = threadIdx.x;
    double v = isMax ? -1e300 : 0.0;
    for (long long i = t; i < n; i += 64) {
        const double p = parts[i * stride + comp];
        v = isMax ? ((isnan(v) || isnan(p)) ? __builtin_nan("") : fmax(v, p)) : v + p;
    }
    s[t] = v;
    __syncthreads();
    if (t == 0) {
        double r = isMax ? -1e300 : 0.0;
        for (int i = 0; i < 64; ++i) r = isMax ? ((isnan(r) || isnan(s[i])) ? __builtin_nan("") : fmax(r, s[i])) : r + s[i];
        *out = r;
    }
}
__global__ void k_final_sum_ll(const long long *parts, long long n, long long *out) {
    __shared__ long long s[64];
    const int t = threadIdx.x;
    long long v = 0;
    for (long long i = t; i < n; i += 64) v += parts[i];
    s[t] = v;
    __syncthreads();
    if (t == 0) { long long r = 0; for (int i = 0; i < 64; ++i) r += s[i]; *out = r; }
}

// scalar field <-> double staging
template <typename R>
__global__ void k_to_double(const R *src, double *dst, long long n) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double) src[i];
}
template <typename R>
__global__ void k_from_double(const double *src, R *dst, long long n) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (R) src[i];
}
template <typename T>
__global__ void k_int_to_double(const T *src, double *dst, long long n) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double) src[i];
}

// ------------------------------------------------------------------------------------------
// explicit instantiations through host launchers
// ------------------------------------------------------------------------------------------
static inline unsigned int nblk(long long n, int b) { return (unsigned int) ((n + b - 1) / b); }

template <typename R>
hipError_t launchAdeStep(const AdeDev<R> &a, const KinDev &k, hipStream_t st) {
    if (a.anyNeumann)
        hipLaunchKernelGGL(k_ade_neumann<R>, dim3(nblk(2LL * a.ny * a.nz, 256)), dim3(256), 0, st, a);
    const unsigned int g = nblk(a.N, 256);
    if (a.Sm <= 2) hipLaunchKernelGGL((k_ade_step<R, 2>), dim3(g), dim3(256), 0, st, a, k);
    else if (a.Sm <= 4) hipLaunchKernelGGL((k_ade_step<R, 4>), dim3(g), dim3(256), 0, st, a, k);
    else hipLaunchKernelGGL((k_ade_step<R, 8>), dim3(g), dim3(256), 0, st, a, k);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeDensity(const AdeDev<R> &a, int ms, double *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_density<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, ms, dst);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeSet(const AdeDev<R> &a, int ms, const double *src, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_set<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, ms, src);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAdeReopen(const AdeDev<R> &a, const uint8_t *re, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_reopen<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, re);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsStep(const NsDev<R> &a, double *partE, double *partN, NsTracer *t, int rule, double eps, hipStream_t st) {
    const unsigned int g = nblk(a.N, 256);
    hipLaunchKernelGGL(k_ns_step<R>, dim3(g), dim3(256), 0, st, a, partE, partN);
    hipLaunchKernelGGL(k_ns_tracer, dim3(1), dim3(64), 0, st, t, partE, partN, (int) g, rule, eps, a.done);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsVelocity(const NsDev<R> &a, R *u, const int8_t *geom, double *parts, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_velocity<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, u, geom, parts);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsInit(const NsDev<R> &a, double dP, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_init<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, dP);
    return hipGetLastError();
}
hipError_t launchNsUpdateDyn(const int8_t *mask, uint8_t *cls, long long N, double pr, unsigned int *changed, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_update_dyn, dim3(nblk(N, 256)), dim3(256), 0, st, mask, cls, N, pr, changed);
    return hipGetLastError();
}
template <typename R>
hipError_t launchNsReopen(const NsDev<R> &a, const uint8_t *oldCls, hipStream_t st) {
    hipLaunchKernelGGL(k_ns_reopen<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, oldCls);
    return hipGetLastError();
}
template <typename R>
hipError_t launchEquilibrium(const AdeDev<R> &a, const EqDev &e, unsigned int *nonconv, hipStream_t st) {
    hipLaunchKernelGGL((k_equilibrium<R, 8>), dim3(nblk(a.N, 128)), dim3(128), 0, st, a, e, nonconv);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaMaskTotal(const CaDev<R> &c, double *pm, long long *ps, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_mask_total<R>, dim3(nblk(c.N, 256)), dim3(256), 0, st, c, pm, ps);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaPushGather(const CaDev<R> &c, uint64_t sweep, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_push<R>, dim3(nblk(c.N, 256)), dim3(256), 0, st, c, sweep);
    hipLaunchKernelGGL(k_ca_gather<R>, dim3(nblk(c.N, 256)), dim3(256), 0, st, c);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaAge(const CaDev<R> &c, const int *in, int *out, long long *ps, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_age<R>, dim3(nblk(c.N, 256)), dim3(256), 0, st, c, in, out, ps);
    return hipGetLastError();
}
hipError_t launchOmegaSwitch(const int8_t *mask, uint8_t *cls, long long N, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_omega_switch, dim3(nblk(N, 256)), dim3(256), 0, st, mask, cls, N);
    return hipGetLastError();
}
hipError_t launchSolidSwitch(const int8_t *mask, uint8_t *cls, uint8_t *re, long long N, unsigned int *nre, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_solid_switch, dim3(nblk(N, 256)), dim3(256), 0, st, mask, cls, re, N, nre);
    return hipGetLastError();
}
template <typename R>
hipError_t launchPrecip(const AdeDev<R> &a, int slot, double maxRho, int phaseId, int8_t *phase, unsigned int *changed, hipStream_t st) {
    hipLaunchKernelGGL(k_precip<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, slot, maxRho, phaseId, phase, changed);
    return hipGetLastError();
}
template <typename R>
hipError_t launchSpeciesTotal(const AdeDev<R> &a, int s, double *parts, hipStream_t st) {
    hipLaunchKernelGGL(k_species_total<R>, dim3(nblk(a.N, 256)), dim3(256), 0, st, a, s, parts);
    return hipGetLastError();
}
hipError_t launchFinalSum(const double *parts, long long n, int stride, int comp, int isMax, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(64), 0, st, parts, n, stride, comp, isMax, out);
    return hipGetLastError();
}
hipError_t launchFinalSumLL(const long long *parts, long long n, long long *out, hipStream_t st) {
    hipLaunchKernelGGL(k_final_sum_ll, dim3(1), dim3(64), 0, st, parts, n, out);
    return hipGetLastError();
}
template <typename R>
hipError_t launchToDouble(const R *src, double *dst, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_double<R>, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}
template <typename R>
hipError_t launchFromDouble(const double *src, R *dst, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_from_double<R>, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}
hipError_t launchI8ToDouble(const int8_t *src, double *dst, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_int_to_double<int8_t>, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}
hipError_t launchI32ToDouble(const int *src, double *dst, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_int_to_double<int>, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

#define INST(R) \
    template hipError_t launchAdeStep<R>(const AdeDev<R> &, const KinDev &, hipStream_t); \
    template hipError_t launchAdeDensity<R>(const AdeDev<R> &, int, double *, hipStream_t); \
    template hipError_t launchAdeSet<R>(const AdeDev<R> &, int, const double *, hipStream_t); \
    template hipError_t launchAdeReopen<R>(const AdeDev<R> &, const uint8_t *, hipStream_t); \
    template hipError_t launchNsStep<R>(const NsDev<R> &, double *, double *, NsTracer *, int, double, hipStream_t); \
    template hipError_t launchNsVelocity<R>(const NsDev<R> &, R *, const int8_t *, double *, hipStream_t); \
    template hipError_t launchNsInit<R>(const NsDev<R> &, double, hipStream_t); \
    template hipError_t launchNsReopen<R>(const NsDev<R> &, const uint8_t *, hipStream_t); \
    template hipError_t launchEquilibrium<R>(const AdeDev<R> &, const EqDev &, unsigned int *, hipStream_t); \
    template hipError_t launchCaMaskTotal<R>(const CaDev<R> &, double *, long long *, hipStream_t); \
    template hipError_t launchCaPushGather<R>(const CaDev<R> &, uint64_t, hipStream_t); \
    template hipError_t launchCaAge<R>(const CaDev<R> &, const int *, int *, long long *, hipStream_t); \
    template hipError_t launchPrecip<R>(const AdeDev<R> &, int, double, int, int8_t *, unsigned int *, hipStream_t); \
    template hipError_t launchSpeciesTotal<R>(const AdeDev<R> &, int, double *, hipStream_t); \
    template hipError_t launchToDouble<R>(const R *, double *, long long, hipStream_t); \
    template hipError_t launchFromDouble<R>(const double *, R *, long long, hipStream_t);
INST(double)
INST(float)

}  // namespace clb

// clb_aux.hip -- the dense-grid kernels of the coupled step (gfx950): the CA
// biofilm spill, the mask-driven dynamics switches, precipitation, the
// reductions and the host<->device staging conversions.  Scalar fields
// (mask, biomass, immobile species, age) are dense, index x + nx*(y + ny*z).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "clb_device.h"
#include "clb_common.h"

namespace clb {

// ------------------------------------------------------------------------------------------
// CA biofilm spill
// ------------------------------------------------------------------------------------------
// updateLocalMaskNtotalLattices3D (part2:71-259) over the open voxels, with
// per-block partials of the global max of the total and of the mask sum.
// Voxels outside the compact list are closed geometry: their mask and total
// never move, so their max and sum are constants of the case.  In a queued
// batch (`check`) the kernel does nothing once an earlier step halted it.
// the update of one open voxel: its total from the microbes' biomass, its
// mask from the total; returns the total and the mask for the reductions
template <typename R>
__device__ __forceinline__ void caMaskTotalVoxel(const CaDev<R> &c, long long idx, double &bt, long long &ms) {
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
    ms = mask;
}
// the block's max total and mask sum into the partials
template <typename R>
__device__ __forceinline__ void caMaskTotalPartials(const CaDev<R> &c, double bt, long long ms) {
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
        c.partMax[blockIdx.x] = m; c.partMask[blockIdx.x] = s;
    }
}
template <typename R>
__global__ void __launch_bounds__(256) k_ca_mask_total(CaDev<R> c, int check) {
    if (check && c.ctl->halt) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double bt = -1e300; long long ms = 0;
    if (i < c.Nopen) caMaskTotalVoxel<R>(c, c.list[i], bt, ms);
    caMaskTotalPartials<R>(c, bt, ms);
}

// the global max and mask sum from the partials, published in the control
// block; with `check` also the step's decision: halt the queued batch when
// the CA must run (max total above Bmax), the mask moved, or the total is not
// finite (check 2: the step always needs the host, a precipitation step)
template <typename R>
__device__ __forceinline__ void caMaskFinalBlock(const CaDev<R> &c, int nparts, int check);
template <typename R>
__global__ void __launch_bounds__(1024) k_ca_mask_final(CaDev<R> c, int nparts, int check) {
    caMaskFinalBlock<R>(c, nparts, check);
}
// the batched form: member blockIdx.x's decision, check 2 on the steps its
// host must finish (precipitation / dissolution checks), 1 otherwise
template <typename R>
__global__ void __launch_bounds__(1024) k_ca_mask_final_group(const AdeMember<R> *__restrict__ E) {
    const AdeMember<R> &g = E[blockIdx.x];
    if (!g.ca) return;
    const long long it = g.iT0 + g.c.ctl->done;
    const bool due = (g.precipEvery > 0 && it % g.precipEvery == 0) || (g.dissolEvery > 0 && it % g.dissolEvery == 0);
    caMaskFinalBlock<R>(g.c, g.gInt + g.gPln, due ? 2 : 1);
}
template <typename R>
hipError_t launchCaMaskFinalGroup(const AdeMember<R> *E, int nMembers, hipStream_t st) {
    if (nMembers > 0) hipLaunchKernelGGL(k_ca_mask_final_group<R>, dim3(nMembers), dim3(1024), 0, st, E);
    return hipGetLastError();
}
template <typename R>
__global__ void k_ctl_reset_group(const AdeMember<R> *__restrict__ E, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !E[j].ca) return;
    StepCtl *c = E[j].c.ctl;
    c->halt = 0; c->done = 0; c->oldMask = E[j].oldMask;
}
template <typename R>
__global__ void k_ctl_gather_group(const AdeMember<R> *__restrict__ E, int n, StepCtl *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !E[j].ca) return;
    out[j] = *E[j].c.ctl;
}
template <typename R>
hipError_t launchCtlResetGroup(const AdeMember<R> *E, int nMembers, hipStream_t st) {
    if (nMembers > 0) hipLaunchKernelGGL(k_ctl_reset_group<R>, dim3((nMembers + 63) / 64), dim3(64), 0, st, E, nMembers);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCtlGatherGroup(const AdeMember<R> *E, int nMembers, StepCtl *out, hipStream_t st) {
    if (nMembers > 0) hipLaunchKernelGGL(k_ctl_gather_group<R>, dim3((nMembers + 63) / 64), dim3(64), 0, st, E, nMembers, out);
    return hipGetLastError();
}
template <typename R>
__device__ __forceinline__ void caMaskFinalBlock(const CaDev<R> &c, int nparts, int check) {
    __shared__ double sMax[16];
    __shared__ long long sMask[16];
    StepCtl *ctl = c.ctl;
    if (check && ctl->halt) return;
    // max and integer sum do not depend on the order: every load of a thread
    // is issued before any is used
    double mm = -1e300;
    long long ss = 0;
    for (int b = threadIdx.x; b < nparts; b += 1024 * 8) {
        double m[8];
        long long s[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = b + 1024 * k;
            m[k] = j < nparts ? c.partMax[j] : -1e300;
            s[k] = j < nparts ? c.partMask[j] : 0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) { mm = nanMax(mm, m[k]); ss += s[k]; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mm = nanMax(mm, __shfl_down(mm, off, 64));
        ss += __shfl_down(ss, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sMax[wv] = mm; sMask[wv] = ss; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double gm = c.gmaxClosed; long long gs = c.msumClosed;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { gm = nanMax(gm, sMax[w]); gs += sMask[w]; }
        ctl->gmax = gm;
        ctl->msum = gs;
        if (check) {
            const bool stop = check == 2 || !isfinite(gm) || gm - c.bmax > 1e-12 || gs != ctl->oldMask;
            if (stop) ctl->halt = 1;
            else ctl->done += 1;
        }
    }
}

// a new batch: no halt, no finished step, the accepted mask sum
__global__ void k_step_ctl_reset(StepCtl *ctl, long long oldMask) {
    ctl->halt = 0; ctl->done = 0; ctl->oldMask = oldMask;
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

// The spill sweep (pushExcessBiomass3D, part1:500-732, as a push and a
// gather over a snapshot), over the compact open voxels (a closed voxel
// neither pushes nor receives).  A voxel over Bmax pushes its excess into an
// outbox slot per face and takes the sweep's stamp (sweep + 1), so nothing
// is cleared between sweeps; then every voxel collects, in face order, the
// slots of the face neighbours holding the current stamp.
template <typename R>
__device__ __forceinline__ void caPushVoxel(const CaDev<R> &c, uint64_t sweep, long long i, uint32_t stamp) {
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
    c.pushed[i] = stamp;
}
template <typename R>
__global__ void __launch_bounds__(256) k_ca_push(CaDev<R> c, uint64_t sweep, uint32_t stamp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.Nopen) caPushVoxel<R>(c, sweep, c.list[i], stamp);
}

// the gather of voxel i from its pushing face neighbours, face order
template <typename R>
__device__ __forceinline__ void caGatherVoxel(const CaDev<R> &c, long long i, uint32_t stamp) {
    int x, y, z;
    coords(i, c.nx, c.ny, x, y, z);
    double add[CLB_DEV_MAXM];
    for (int m = 0; m < c.M; ++m) add[m] = 0;
    for (int d = 0; d < 6; ++d) {
        const int q = d + 1;
        const int X = x - D7C[q][0], Y = y - D7C[q][1], Z = z - D7C[q][2];
        if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
        const long long j = X + (long long) c.nx * (Y + (long long) c.ny * Z);
        if (c.pushed[j] != stamp) continue;
        for (int m = 0; m < c.M; ++m) add[m] += (double) c.outbox[((long long) m * 6 + d) * c.N + j];
    }
    for (int m = 0; m < c.M; ++m)
        if (add[m] != 0.0) c.B[(long long) m * c.N + i] = (R) ((double) c.B[(long long) m * c.N + i] + add[m]);
}
template <typename R>
__global__ void __launch_bounds__(256) k_ca_gather(CaDev<R> c, uint32_t stamp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.Nopen) caGatherVoxel<R>(c, c.list[i], stamp);
}
// the gather and then the mask/total update of the same voxel
// (updateLocalMaskNtotalLattices3D after the sweep, complab.cpp:2800-2810),
// with the block partials of the max total and the mask sum
template <typename R>
__global__ void __launch_bounds__(256) k_ca_gather_total(CaDev<R> c, uint32_t stamp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double bt = -1e300; long long ms = 0;
    if (i < c.Nopen) {
        const long long v = c.list[i];
        caGatherVoxel<R>(c, v, stamp);
        caMaskTotalVoxel<R>(c, v, bt, ms);
    }
    caMaskTotalPartials<R>(c, bt, ms);
}

// updateAgeDistance3D's rule at one voxel (every voxel that can hold biomass
// is in the compact list), from the ages it reads
template <typename R>
__device__ __forceinline__ int ageRule(const CaDev<R> &c, const int *ageIn, long long i, int a) {
    const double B = (double) c.Btot[i];
    if (!(B > 1e-14)) return a;
    int x, y, z;
    coords(i, c.nx, c.ny, x, y, z);
    long long del[6]; int nd = 0;
    for (int q = 1; q < 7; ++q) {
        const int X = x + D7C[q][0], Y = y + D7C[q][1], Z = z + D7C[q][2];
        if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
        const long long j = X + (long long) c.nx * (Y + (long long) c.ny * Z);
        if (c.dist[j] > 0) del[nd++] = j;
    }
    if (a == 0) return 1;
    if (a == 1 && (B - c.bmax) > -1e-14) {
        for (int t = 0; t < nd; ++t) if (ageIn[del[t]] == 0) return a;
        return 2;
    }
    int cnt = 0;
    for (int t = 0; t < nd; ++t) if (ageIn[del[t]] >= a && ((double) c.Btot[del[t]] - c.bmax) > -1e-14) ++cnt;
    return cnt == nd ? a + 1 : a;
}

template <typename R>
__device__ __forceinline__ void agePartial(long long sum, long long *partSum) {
    __shared__ long long sS[4];
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sS[wv] = sum;
    __syncthreads();
    if (threadIdx.x == 0) { long long s = 0; for (int w = 0; w < (int) (blockDim.x >> 6); ++w) s += sS[w]; partSum[blockIdx.x] = s; }
}

// one Jacobi sweep (ageIn -> ageOut) over the compact voxels + the sum of the
// increments; the other voxels hold the same age in both buffers
template <typename R>
__global__ void __launch_bounds__(256) k_ca_age(CaDev<R> c, const int *ageIn, int *ageOut, long long *partSum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long sum = 0;
    if (i < c.Nopen) {
        const long long v = c.list[i];
        const int a = ageIn[v];
        const int na = ageRule(c, ageIn, v, a);
        ageOut[v] = na;
        sum = na - a;
    }
    agePartial<R>(sum, partSum);
}

// the voxels whose age can move while the biomass stands still: those holding
// biomass, compacted in list order without atomics (per-block counts, one
// block's exclusive scan of them, an ordered scatter)
template <typename R>
__device__ __forceinline__ bool holdsBiomass(const CaDev<R> &c, int i, long long &v) {
    if (i >= c.Nopen) return false;
    v = c.list[i];
    return (double) c.Btot[v] > 1e-14;
}
// the kept voxels before this wave in the block, and the block's total
__device__ __forceinline__ unsigned int blockRank(bool keep, unsigned int &blockTotal) {
    __shared__ unsigned int sW[4];
    const unsigned long long b = __ballot(keep);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sW[wv] = (unsigned int) __popcll(b);
    __syncthreads();
    unsigned int before = 0, total = 0;
    for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { if (w < wv) before += sW[w]; total += sW[w]; }
    blockTotal = total;
    return before + (unsigned int) __popcll(b & ((1ull << lane) - 1ull));
}
template <typename R>
__global__ void __launch_bounds__(256) k_age_count(CaDev<R> c, unsigned int *blockCount) {
    long long v = 0;
    const bool keep = holdsBiomass(c, blockIdx.x * blockDim.x + threadIdx.x, v);
    unsigned int total = 0;
    blockRank(keep, total);
    if (threadIdx.x == 0) blockCount[blockIdx.x] = total;
}
// exclusive scan of the block counts in place; the sum into *total
__global__ void __launch_bounds__(1024) k_age_scan(unsigned int *blockCount, int nb, unsigned int *total) {
    __shared__ unsigned int sS[1024];
    const int t = threadIdx.x, per = (nb + 1023) / 1024, b0 = t * per;
    unsigned int s = 0;
    for (int k = 0; k < per; ++k) if (b0 + k < nb) s += blockCount[b0 + k];
    sS[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {             // inclusive Hillis-Steele scan
        const unsigned int add = t >= off ? sS[t - off] : 0u;
        __syncthreads();
        sS[t] += add;
        __syncthreads();
    }
    unsigned int run = t > 0 ? sS[t - 1] : 0u;
    for (int k = 0; k < per; ++k) if (b0 + k < nb) { const unsigned int x = blockCount[b0 + k]; blockCount[b0 + k] = run; run += x; }
    if (t == 1023) *total = sS[1023];
}
template <typename R>
__global__ void __launch_bounds__(256) k_age_scatter(CaDev<R> c, const unsigned int *blockOffset, int *blist) {
    long long v = 0;
    const bool keep = holdsBiomass(c, blockIdx.x * blockDim.x + threadIdx.x, v);
    unsigned int total = 0;
    const unsigned int r = blockRank(keep, total);
    if (keep) blist[blockOffset[blockIdx.x] + r] = (int) v;
}

// one sweep in place over those voxels, for the relaxation to the fixed
// point: a voxel may read ages this sweep already raised.  The rule is
// monotone -- an age only rises, and raising a neighbour never blocks a rise
// -- so any order of these updates climbs to the same least fixed point above
// the start as the Jacobi sweeps, and a sweep that raises nothing stands on
// it.  Sweep k counts the voxels it raised in raised[k]; a sweep queued after
// one that raised nothing does nothing.
// The voxel count is the compaction's, read on the device (*nDev), so the
// host queues the sweeps right behind the compaction; the fixed grid strides
// over the voxels.
template <typename R>
__global__ void __launch_bounds__(256) k_ca_age_inplace(CaDev<R> c, int *age, const int *blist, const unsigned int *nDev,
                                                        unsigned int *raised, int k) {
    if (k > 0 && raised[k - 1] == 0) return;
    const int n = (int) *nDev;
    unsigned int up = 0;
    for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const int i = i0 + (int) threadIdx.x;
        bool u = false;
        if (i < n) {
            const long long v = blist[i];
            const int a = age[v];
            const int na = ageRule(c, age, v, a);
            if (na != a) { age[v] = na; u = true; }
        }
        up += (unsigned int) __popcll(__ballot(u));
    }
    if ((threadIdx.x & 63) == 0 && up) atomicAdd(raised + k, up);
}

// The single in-place sweep of updateAgeDistance3D that follows a mask change
// (complab.cpp:3079).  The processor walks x outermost and z innermost
// (part2:862-867), so a voxel reads the ages its -x, -y and -z face neighbours
// took earlier in the same sweep and the old ages of its +x, +y and +z
// neighbours; the rule reads face neighbours only.  The sweep's result is
// therefore the unique solution of new[v] = rule(old[v], new[-faces], old[+faces])
// over the voxels holding biomass: triangular in the order x + y + z.  It is
// solved by relaxation: every launch re-evaluates each such voxel from the
// current `ageNew` of its -face neighbours (in place; each value a launch
// reads is an earlier launch's or this one's), and a launch that changes
// nothing stands on the solution -- the planes x + y + z = k are exact after
// k + 1 launches at the latest, so it terminates.  `ageNew` starts as a copy
// of `ageOld` (every other voxel keeps its age).
template <typename R>
__device__ __forceinline__ int ageRuleOrdered(const CaDev<R> &c, const int *ageOld, const int *ageNew, long long i) {
    const int a = ageOld[i];
    const double B = (double) c.Btot[i];
    if (!(B > 1e-14)) return a;
    int x, y, z;
    coords(i, c.nx, c.ny, x, y, z);
    long long del[6]; bool earlier[6]; int nd = 0;
    for (int q = 1; q < 7; ++q) {
        const int X = x + D7C[q][0], Y = y + D7C[q][1], Z = z + D7C[q][2];
        if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
        const long long j = X + (long long) c.nx * (Y + (long long) c.ny * Z);
        if (c.dist[j] > 0) { del[nd] = j; earlier[nd] = (D7C[q][0] + D7C[q][1] + D7C[q][2]) < 0; ++nd; }
    }
    if (a == 0) return 1;
    if (a == 1 && (B - c.bmax) > -1e-14) {
        for (int t = 0; t < nd; ++t) if ((earlier[t] ? ageNew[del[t]] : ageOld[del[t]]) == 0) return a;
        return 2;
    }
    int cnt = 0;
    for (int t = 0; t < nd; ++t) {
        const int an = earlier[t] ? ageNew[del[t]] : ageOld[del[t]];
        if (an >= a && ((double) c.Btot[del[t]] - c.bmax) > -1e-14) ++cnt;
    }
    return cnt == nd ? a + 1 : a;
}
// one relaxation launch over the voxels holding biomass; launch k counts the
// voxels it changed in changed[k], and a launch queued after one that changed
// nothing does nothing
template <typename R>
__global__ void __launch_bounds__(256) k_ca_age_ordered(CaDev<R> c, const int *ageOld, int *ageNew, const int *blist,
                                                        const unsigned int *nDev, unsigned int *changed, int k) {
    if (k > 0 && changed[k - 1] == 0) return;
    const int n = (int) *nDev;
    unsigned int chg = 0;
    for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const int i = i0 + (int) threadIdx.x;
        bool ch = false;
        if (i < n) {
            const long long v = blist[i];
            const int na = ageRuleOrdered(c, ageOld, ageNew, v);
            if (na != ageNew[v]) { ageNew[v] = na; ch = true; }
        }
        chg += (unsigned int) __popcll(__ballot(ch));
    }
    if ((threadIdx.x & 63) == 0 && chg) atomicAdd(changed + k, chg);
}

// ------------------------------------------------------------------------------------------
// mask-driven switches (compact ADE classes from the dense live mask) and precipitation
// ------------------------------------------------------------------------------------------
// updateSoluteDynamics3D (part2:485-553): pore/biofilm relaxation of open voxels
__global__ void k_ade_omega_switch(const int8_t *mask, const int *list, uint8_t *cls, int Nopen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nopen) return;
    const int m = mask[list[i]];
    if (m == 0 || m == 1) return;
    if (cls[i] == ADE_BB) return;
    cls[i] = m == 2 ? ADE_PORE : ADE_BFILM;
}
// updateSoluteSolidDynamics3D (part2:585-647): bounce-back <-> BGK from the live mask
__global__ void k_ade_solid_switch(const int8_t *mask, const int *list, uint8_t *cls, uint8_t *reopened, int Nopen,
                                   unsigned int *nre) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nopen) return;
    const int m = mask[list[i]];
    const bool want = (m == 0 || m == 1);
    const bool isBB = cls[i] == ADE_BB;
    reopened[i] = 0;
    if (want && !isBB) cls[i] = ADE_BB;
    else if (!want && isBB) { cls[i] = m == 2 ? ADE_PORE : ADE_BFILM; reopened[i] = 1; atomicAdd(nre, 1u); }
}
// precipNodeConversion3D (precipitationVOP.hh:166-227) over x = 1..nx-2
template <typename R>
__global__ void k_precip(int nx, int ny, long long N, int8_t *mask, const R *P, double maxRho, int phaseId,
                         int8_t *phase, unsigned int *changed) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x, y, z;
    coords(i, nx, ny, x, y, z);
    if (x <= 0 || x >= nx - 1) return;
    const int m = mask[i];
    if (m == 0 || m == 1) return;
    if ((double) P[i] >= maxRho) {
        mask[i] = 0;
        if (phaseId > 0) phase[i] = (int8_t) phaseId;
        atomicAdd(changed, 1u);
    }
}
// total of an immobile species over the open voxels of x = 1..nx-2 (per-block partials)
// per block, for a dense scalar field (an immobile species, a biomass): the
// sum over every voxel of the reaction box x = 1..nx-2 (a mineral keeps its
// inventory in the voxels it sealed), minus the minimum and the maximum there,
// and the sum over the x = 0 and x = nx-1 planes; parts[4 * b + 0..3]
template <typename R>
__global__ void __launch_bounds__(256) k_scalar_stats(int nx, int ny, long long N, const R *f, double *parts) {
    __shared__ double sv[4][4];
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    double t = 0, lo = -1e300, hi = -1e300, pl = 0;
    if (i < N) {
        int x, y, z;
        coords(i, nx, ny, x, y, z);
        const double v = (double) f[i];
        if (x == 0 || x == nx - 1) pl = v;
        else { t = v; lo = -v; hi = v; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_down(t, off, 64);
        pl += __shfl_down(pl, off, 64);
        lo = fmax(lo, __shfl_down(lo, off, 64));
        hi = fmax(hi, __shfl_down(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sv[wv][0] = t; sv[wv][1] = lo; sv[wv][2] = hi; sv[wv][3] = pl; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double T = 0, L = -1e300, U = -1e300, P = 0;
        for (int w = 0; w < (int) (blockDim.x >> 6); ++w) { T += sv[w][0]; L = fmax(L, sv[w][1]); U = fmax(U, sv[w][2]); P += sv[w][3]; }
        parts[4 * blockIdx.x] = T; parts[4 * blockIdx.x + 1] = L; parts[4 * blockIdx.x + 2] = U; parts[4 * blockIdx.x + 3] = P;
    }
}

// fixed-order final reduction of per-block partials (one block of 256 threads,
// four independent accumulators per thread so the loads overlap)
__global__ void __launch_bounds__(256) k_final_sum(const double *parts, long long n, int stride, int comp, int isMax, double *out) {
    __shared__ double s[256];
    const int t = threadIdx.x;
    const double id = isMax ? -1e300 : 0.0;
    double v[4] = {id, id, id, id};
    for (long long i = t; i < n; i += 1024) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long j = i + 256 * k;
            if (j < n) {
                const double p = parts[j * stride + comp];
                v[k] = isMax ? nanMax(v[k], p) : v[k] + p;
            }
        }
    }
    s[t] = isMax ? nanMax(nanMax(v[0], v[1]), nanMax(v[2], v[3])) : (v[0] + v[1]) + (v[2] + v[3]);
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] = isMax ? nanMax(s[t], s[t + w]) : s[t] + s[t + w];
        __syncthreads();
    }
    if (t == 0) *out = s[0];
}
__global__ void __launch_bounds__(1024) k_final_sum_ll(const long long *parts, long long n, long long *out) {
    __shared__ long long s[1024];
    const int t = threadIdx.x;
    long long acc = 0;
    for (long long i = t; i < n; i += 1024 * 8) {
        long long v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = i + 1024 * k < n ? parts[i + 1024 * k] : 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    s[t] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) *out = s[0];
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

// plain copy, 16 bytes per lane, grid-stride
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long n) {
    for (long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long) gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
hipError_t launchCopy16(const void *src, void *dst, long long n16, hipStream_t st) {
    const unsigned int g = (unsigned int) std::min<long long>((n16 + 255) / 256, 256LL * 64);
    hipLaunchKernelGGL(k_copy16, dim3(g), dim3(256), 0, st, (const uint4 *) src, (uint4 *) dst, n16);
    return hipGetLastError();
}
static inline unsigned int nblk(long long n, int b) { return (unsigned int) ((n + b - 1) / b); }

template <typename R>
hipError_t launchCaMaskTotal(const CaDev<R> &c, int check, hipStream_t st) {
    const unsigned int g = nblk(c.Nopen, 256);
    hipLaunchKernelGGL(k_ca_mask_total<R>, dim3(g), dim3(256), 0, st, c, check);
    hipLaunchKernelGGL(k_ca_mask_final<R>, dim3(1), dim3(1024), 0, st, c, (int) g, check);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaMaskFinal(const CaDev<R> &c, int nparts, int check, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_mask_final<R>, dim3(1), dim3(1024), 0, st, c, nparts, check);
    return hipGetLastError();
}
hipError_t launchStepCtlReset(StepCtl *ctl, long long oldMask, hipStream_t st) {
    hipLaunchKernelGGL(k_step_ctl_reset, dim3(1), dim3(1), 0, st, ctl, oldMask);
    return hipGetLastError();
}
// one spill sweep (push, gather); with `total` the gather also makes the
// mask/total update and the final reduction publishes the max total and the
// mask sum (what updateLocalMaskNtotalLattices3D after the sweep gives)
template <typename R>
hipError_t launchCaPushGather(const CaDev<R> &c, uint64_t sweep, hipStream_t st, bool total) {
    if (c.Nopen <= 0) return hipSuccess;
    const uint32_t stamp = (uint32_t) (sweep + 1);
    const unsigned int g = nblk(c.Nopen, 256);
    hipLaunchKernelGGL(k_ca_push<R>, dim3(g), dim3(256), 0, st, c, sweep, stamp);
    if (total) {
        hipLaunchKernelGGL(k_ca_gather_total<R>, dim3(g), dim3(256), 0, st, c, stamp);
        hipLaunchKernelGGL(k_ca_mask_final<R>, dim3(1), dim3(1024), 0, st, c, (int) g, 0);
    } else {
        hipLaunchKernelGGL(k_ca_gather<R>, dim3(g), dim3(256), 0, st, c, stamp);
    }
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaAge(const CaDev<R> &c, const int *in, int *out, long long *ps, hipStream_t st) {
    if (c.Nopen > 0) hipLaunchKernelGGL(k_ca_age<R>, dim3(nblk(c.Nopen, 256)), dim3(256), 0, st, c, in, out, ps);
    return hipGetLastError();
}
template <typename R>
hipError_t launchAgeCompact(const CaDev<R> &c, int *blist, unsigned int *blockScratch, unsigned int *count, hipStream_t st) {
    const unsigned int nb = nblk(c.Nopen, 256);
    if (nb == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_age_count<R>, dim3(nb), dim3(256), 0, st, c, blockScratch);
    hipLaunchKernelGGL(k_age_scan, dim3(1), dim3(1024), 0, st, blockScratch, (int) nb, count);
    hipLaunchKernelGGL(k_age_scatter<R>, dim3(nb), dim3(256), 0, st, c, blockScratch, blist);
    return hipGetLastError();
}
// the age sweeps' grid: enough workgroups for every voxel of the case to
// have a thread, at most 1024 (the compacted voxels hold biomass, a fraction)
static inline unsigned int ageGrid(int Nopen) { return std::max(1u, std::min(nblk(Nopen, 256), 1024u)); }
template <typename R>
hipError_t launchCaAgeInplace(const CaDev<R> &c, int *age, const int *blist, const unsigned int *nDev, unsigned int *raised,
                              int k, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_age_inplace<R>, dim3(ageGrid(c.Nopen)), dim3(256), 0, st, c, age, blist, nDev, raised, k);
    return hipGetLastError();
}
template <typename R>
hipError_t launchCaAgeOrdered(const CaDev<R> &c, const int *ageOld, int *ageNew, const int *blist, const unsigned int *nDev,
                              unsigned int *changed, int k, hipStream_t st) {
    hipLaunchKernelGGL(k_ca_age_ordered<R>, dim3(ageGrid(c.Nopen)), dim3(256), 0, st, c, ageOld, ageNew, blist, nDev, changed, k);
    return hipGetLastError();
}
hipError_t launchOmegaSwitch(const int8_t *mask, const int *list, uint8_t *cls, int Nopen, hipStream_t st) {
    hipLaunchKernelGGL(k_ade_omega_switch, dim3(nblk(Nopen, 256)), dim3(256), 0, st, mask, list, cls, Nopen);
    return hipGetLastError();
}
hipError_t launchSolidSwitch(const int8_t *mask, const int *list, uint8_t *cls, uint8_t *re, int Nopen, unsigned int *nre,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_ade_solid_switch, dim3(nblk(Nopen, 256)), dim3(256), 0, st, mask, list, cls, re, Nopen, nre);
    return hipGetLastError();
}
template <typename R>
hipError_t launchPrecip(int nx, int ny, long long N, int8_t *mask, const R *P, double maxRho, int phaseId, int8_t *phase,
                        unsigned int *changed, hipStream_t st) {
    hipLaunchKernelGGL(k_precip<R>, dim3(nblk(N, 256)), dim3(256), 0, st, nx, ny, N, mask, P, maxRho, phaseId, phase, changed);
    return hipGetLastError();
}
template <typename R>
hipError_t launchScalarStats(int nx, int ny, long long N, const R *f, double *parts, hipStream_t st) {
    hipLaunchKernelGGL(k_scalar_stats<R>, dim3(nblk(N, 256)), dim3(256), 0, st, nx, ny, N, f, parts);
    return hipGetLastError();
}
hipError_t launchFinalSum(const double *parts, long long n, int stride, int comp, int isMax, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, st, parts, n, stride, comp, isMax, out);
    return hipGetLastError();
}
hipError_t launchFinalSumLL(const long long *parts, long long n, long long *out, hipStream_t st) {
    hipLaunchKernelGGL(k_final_sum_ll, dim3(1), dim3(1024), 0, st, parts, n, out);
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
// nearest integer of each value (an age field written by the caller)
__global__ void k_double_to_i32(const double *src, int *dst, long long n) {
    const long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (int) llround(src[i]);
}
hipError_t launchDoubleToI32(const double *src, int *dst, long long n, hipStream_t st) {
    hipLaunchKernelGGL(k_double_to_i32, dim3(nblk(n, 256)), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

#define INST(R) \
    template hipError_t launchCaMaskTotal<R>(const CaDev<R> &, int, hipStream_t); \
    template hipError_t launchCaMaskFinal<R>(const CaDev<R> &, int, int, hipStream_t); \
    template hipError_t launchCaMaskFinalGroup<R>(const AdeMember<R> *, int, hipStream_t); \
    template hipError_t launchCtlResetGroup<R>(const AdeMember<R> *, int, hipStream_t); \
    template hipError_t launchCtlGatherGroup<R>(const AdeMember<R> *, int, StepCtl *, hipStream_t); \
    template hipError_t launchCaPushGather<R>(const CaDev<R> &, uint64_t, hipStream_t, bool); \
    template hipError_t launchCaAge<R>(const CaDev<R> &, const int *, int *, long long *, hipStream_t); \
    template hipError_t launchAgeCompact<R>(const CaDev<R> &, int *, unsigned int *, unsigned int *, hipStream_t); \
    template hipError_t launchCaAgeInplace<R>(const CaDev<R> &, int *, const int *, const unsigned int *, unsigned int *, int, hipStream_t); \
    template hipError_t launchCaAgeOrdered<R>(const CaDev<R> &, const int *, int *, const int *, const unsigned int *, unsigned int *, int, hipStream_t); \
    template hipError_t launchPrecip<R>(int, int, long long, int8_t *, const R *, double, int, int8_t *, unsigned int *, hipStream_t); \
    template hipError_t launchScalarStats<R>(int, int, long long, const R *, double *, hipStream_t); \
    template hipError_t launchToDouble<R>(const R *, double *, long long, hipStream_t); \
    template hipError_t launchFromDouble<R>(const double *, R *, long long, hipStream_t);
INST(double)
INST(float)

}  // namespace clb

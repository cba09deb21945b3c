// clb_equilibrium.hip -- the per-voxel equilibrium speciation step (gfx950),
// FP64 whatever the lattice precision, one thread per open interior voxel of
// the reaction box (complab.cpp:2566 applies run_equilibrium_biotic to the
// reaction box; part4_eqsolver.hh:1360-1500 is the per-cell update).
//
// The voxel's species totals (density of every mobile lattice, the immobile
// field) go through the solver; the change, limited to a tenth of the
// concentration (at least 1e-9) and kept above 1e-20, is added back to the
// populations by lattice weight.  Kernels are specialised on the species and
// component bounds and on the solver (clb_eq_device.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clb_device.h"
#include "clb_common.h"
#include "clb_eq_device.h"

namespace clb {

// NS >= ns and NC >= nc bound the per-voxel arrays at compile time (the
// solver's arrays then stay in registers); the launcher picks the smallest
// bounds that hold the case (4x2, 6x3, 8x4, 8x6, 8x8), and the kernel of its
// solver.  Up to 6 components the register file holds the solver (8x6
// Newton: no scratch, AA-PCF: 76 bytes a lane; 8x8: 1160 and 624).  The iteration is
// a chain of dependent FP64 operations, so waves in flight hide its latency:
// Newton with nc <= 3 is held to 168 VGPRs for 3 waves per SIMD (a 28-byte
// spill, 17 % faster than 2 waves at 220 VGPRs on the 512x256x256 carbonate
// case); the larger solvers keep every register they need.
template <int NS, int NC, int SOLVER>
constexpr int eqWavesPerSimd() { return (SOLVER == EQ_SOLVER_NEWTON && NC <= 3) ? 3 : 1; }

template <typename R, int NS, int NC, int SOLVER>
__global__ void __launch_bounds__(128, (eqWavesPerSimd<NS, NC, SOLVER>())) k_equilibrium(AdeDev<R> a, EqDev e, unsigned int *nonconv) {
    if (halted(a.halt)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Nint) return;
    const long long v = a.list[i];
    const int mask = a.mask[v];
    if (mask == 0 || mask == 1) return;
    const long long No = a.Nopen, N = a.N;
    const int S = e.ns;
    double C[NS], E[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s >= S) break;
        double r;
        if (a.mobile[s]) {
            const R *o = a.cur + (long long) a.slot[s] * 7 * a.NoP;
            r = 0;
#pragma unroll
            for (int q = 0; q < 7; ++q) r += (double) o[pidx<7>(q, i)];
        } else r = (double) a.imm[(long long) a.slot[s] * N + v];
        C[s] = fmax(r, EQ_MIN_CONC);
    }
    const bool conv = eqSolve<NS, NC, SOLVER>(e, C, E);
    if (!conv) atomicAdd(nonconv, 1u);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s >= S) break;
        double dC = E[s] - C[s];
        const double mx = fmax(1e-9, 0.1 * fabs(C[s]));                 // part4:1472-1485
        if (fabs(dC) > mx) dC = dC > 0 ? mx : -mx;
        if (C[s] + dC < 1e-20) dC = 1e-20 - C[s];
        if (fabs(dC) > 1e-12) {
            if (a.mobile[s]) {
                R *o = a.cur + (long long) a.slot[s] * 7 * a.NoP;
#pragma unroll
                for (int q = 0; q < 7; ++q) o[pidx<7>(q, i)] = (R) ((double) o[pidx<7>(q, i)] + w7(q) * dC);
            } else {
                R *p = a.imm + (long long) a.slot[s] * N + v;
                *p = (R) ((double) *p + dC);
            }
        }
    }
}

static inline unsigned int nblkEq(long long n, int b) { return (unsigned int) ((n + b - 1) / b); }

template <typename R, int SOLVER>
static void launchEqSolver(const AdeDev<R> &a, const EqDev &e, unsigned int *nonconv, hipStream_t st) {
    const dim3 g(nblkEq(a.Nint, 128)), b(128);
    if (e.ns <= 4 && e.nc <= 2) hipLaunchKernelGGL((k_equilibrium<R, 4, 2, SOLVER>), g, b, 0, st, a, e, nonconv);
    else if (e.ns <= 6 && e.nc <= 3) hipLaunchKernelGGL((k_equilibrium<R, 6, 3, SOLVER>), g, b, 0, st, a, e, nonconv);
    else if (e.nc <= 4) hipLaunchKernelGGL((k_equilibrium<R, EQ_MAXS, 4, SOLVER>), g, b, 0, st, a, e, nonconv);
    else if (e.nc <= 6) hipLaunchKernelGGL((k_equilibrium<R, EQ_MAXS, 6, SOLVER>), g, b, 0, st, a, e, nonconv);
    else hipLaunchKernelGGL((k_equilibrium<R, EQ_MAXS, EQ_MAXC, SOLVER>), g, b, 0, st, a, e, nonconv);
}

template <typename R>
hipError_t launchEquilibrium(const AdeDev<R> &a, const EqDev &e, unsigned int *nonconv, hipStream_t st) {
    if (e.ns > EQ_MAXS || e.nc > EQ_MAXC) return hipErrorInvalidValue;
    if (e.solver == EQ_SOLVER_NEWTON) launchEqSolver<R, EQ_SOLVER_NEWTON>(a, e, nonconv, st);
    else launchEqSolver<R, EQ_SOLVER_AAPCF>(a, e, nonconv, st);
    return hipGetLastError();
}

template hipError_t launchEquilibrium<double>(const AdeDev<double> &, const EqDev &, unsigned int *, hipStream_t);
template hipError_t launchEquilibrium<float>(const AdeDev<float> &, const EqDev &, unsigned int *, hipStream_t);

}  // namespace clb

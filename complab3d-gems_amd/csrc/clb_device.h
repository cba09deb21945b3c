// clb_device.h -- argument blocks shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clb.h"

#define CLB_DEV_MAXS 8
#define CLB_DEV_MAXM 2
#define CLB_DEV_MAXR 4
#define CLB_DEV_MAXP 4
#define NS_TRACER_WINDOW 500

namespace clb {

enum { ADE_BB = 0, ADE_PORE = 1, ADE_BFILM = 2 };
enum { NS_BGK = 0, NS_BIO = 1, NS_BB = 2, NS_NODYN = 3 };
enum { BC_DIRICHLET_DEV = 0, BC_NEUMANN_DEV = 1, BC_CLOSED_DEV = 2 };

// The open voxels (geometry pore or biofilm) of a case in compact order:
// [0, Nint) the interior x = 1..nx-2, then [Nint, Nint+Nleft) the x = 0 plane,
// then the x = nx-1 plane; each in x-fastest grid order.  list[i] is the grid
// index of compact voxel i.  nb[q-1][i] (q = 1..Q-1) is the compact index of
// the voxel a population of direction q is pulled from (grid index - c_q), or
// NB_CLOSED when that voxel is closed geometry, NB_OUTSIDE beyond the grid.
#define NB_CLOSED -1
#define NB_OUTSIDE -2

// Populations are stored in tiles of 64 voxels (one wavefront): population q
// of compact voxel i of a Q-population lattice lives at
// ((i / 64) * Q + q) * 64 + i % 64.  The Q loads or stores of one wave then
// fall in one contiguous stretch of Q x 64 values of a tile instead of Q
// arrays Nopen apart; on gfx950 the 19-array pattern of one thread per voxel
// streams at 4.7 TB/s and the tiled one at 5.6 TB/s.  A lattice holds
// popPad(Nopen) voxels' worth (the tail of the last tile is never touched).
__host__ __device__ __forceinline__ long long popPad(long long n) { return (n + 63) & ~63LL; }
template <int Q>
__host__ __device__ __forceinline__ long long pidx(int q, long long i) {
    return (((i >> 6) * Q + q) << 6) | (i & 63);
}

// step-batch control block: a batch of coupled steps is queued without a host
// round trip; the mask/total reduction of each step decides on the device
// whether the host must act (CA spill, mask change, non-finite biomass), and
// then every later kernel of the batch returns at once
struct StepCtl {
    int halt;                      // set by the step that needs the host
    int done;                      // steps of the batch finished without host action
    double gmax;                   // max total biomass of the last reduction
    long long msum;                // mask sum of the last reduction
    long long oldMask;             // the mask sum the host accepted last
};

// the advection-diffusion lattices of one case
template <typename R>
struct AdeDev {
    int nx, ny, nz;
    long long N;
    int Nopen, Nint, Nleft;
    long long NoP;                 // popPad(Nopen): the voxels one species' lattice holds
    const int *list;               // [Nopen]
    const int *nb;                 // [6][Nopen]
    int S, Sm, Smax, M;
    int mobile[CLB_DEV_MAXS];      // per species
    int slot[CLB_DEV_MAXS];        // per species: mobile slot or immobile slot
    int mobSpecies[CLB_DEV_MAXS];  // per mobile slot: species index
    R *cur, *out;                  // [Smax][7 x NoP] tiled (pidx<7>)
    // the bounce-back store paired with each population buffer: per mobile
    // slot (stride bbStride) and tile of 64 interior voxels, 64 x 6 slots
    // holding the voxels' own populations opposite their closed links (pull
    // kind 2), in lane order; the pull reads and writes bbOut
    R *bbCur, *bbOut;              // [Smax][bbStride]
    long long bbStride;            // popPad(Nint) x 6
    int bbOn;                      // the store is in use (else the pull reads the own slots of out)
    R *imm;                        // [Si][N] dense
    R *B;                          // [M][N] dense
    const R *u;                    // [3][Nopen]
    int8_t *mask;                  // dense live mask
    const uint8_t *cls, *clsPull;  // [Nopen]
    const uint16_t *kinds;         // [Nint] source kind of each moving population, 2 bits per direction
    double *rhoN;                  // [Smax][Nopen - Nint] Neumann plane densities
    int anyNeumann;
    int first, reset;
    double c0[CLB_DEV_MAXS], omP[CLB_DEV_MAXS], omB[CLB_DEV_MAXS];
    int lbt[CLB_DEV_MAXS], rbt[CLB_DEV_MAXS];
    double lbc[CLB_DEV_MAXS], rbc[CLB_DEV_MAXS];
    // the same per mobile slot, so that an unrolled slot loop reads them from
    // fixed argument offsets
    int mlbt[CLB_DEV_MAXS], mrbt[CLB_DEV_MAXS];
    double mc0[CLB_DEV_MAXS], momP[CLB_DEV_MAXS], momB[CLB_DEV_MAXS];
    unsigned long long *clampCount;
    const int *halt;               // StepCtl::halt of a queued batch, or null
    // the per-voxel mask and total update of the step (k_ca_mask_total's) done
    // by the transport kernels, with block partials for the final reduction
    int maskTotal;
    R *Btot;
    double thrdB;
    double *partMax;
    long long *partMask;
};

// the transport bounce-back store (AdeDev::bbCur/bbOut): a voxel's closed
// links are its pull kinds equal to 2 (2 bits per direction q = 1..6)
__host__ __device__ inline unsigned int adeKind2Mask(unsigned int kd) { return (kd >> 1) & ~kd & 0x555u; }

// rate-law parameters (biotic dual Monod, abiotic mass action)
// the reopen check of the declared phases (dissolNodeConversion3D): phase id
// q + 1 reopens below threshold[q] of the immobile slot slot[q]
struct DissolConv {
    int n;
    int slot[CLB_DEV_MAXP];
    double threshold[CLB_DEV_MAXP];
};

struct KinDev {
    int on, bio, S, M;
    double dt;
    int clamp;
    int donor, acceptor;
    double mu_max[CLB_DEV_MAXM], ks[CLB_DEV_MAXM], ke[CLB_DEV_MAXM], kdecay[CLB_DEV_MAXM];
    double bio_stoich[CLB_DEV_MAXM][CLB_DEV_MAXS];
    int nrxn;
    int r1[CLB_DEV_MAXR], r2[CLB_DEV_MAXR];
    double rk[CLB_DEV_MAXR];
    double rxn_stoich[CLB_DEV_MAXR][CLB_DEV_MAXS];
    double cap_frac, cap_dt, min_conc, rate_scale;
    int surface_only;
    // dissolution of the declared solid phases (dissolutionVOP.hh): phase
    // id p + 1 keeps its inventory in species dMineral[p] and dissolves by
    // rate = dK max(C[dAcid],0)^dAo max(mineral,0)^dMo, products dStoich
    int dissol;                    // the gather runs with the abiotic increment
    int dPhases, dSurfaceOnly;
    int dMineral[CLB_DEV_MAXP], dAcid[CLB_DEV_MAXP];
    double dK[CLB_DEV_MAXP], dAo[CLB_DEV_MAXP], dMo[CLB_DEV_MAXP];
    double dStoich[CLB_DEV_MAXP][CLB_DEV_MAXS];
    const int8_t *phase;           // dense phase ids
    double *share;                 // [S][Nopen] per-face product share parked this step
    unsigned long long *dcount;    // [0] shares parked, [1] collected, [2] collected and clamped
    // the voxels next to the phase (a face neighbour in the reaction box of
    // a declared phase) take the gather: 0 every voxel as usual; 1 (the transport
    // launch) their abiotic increment is left for the pass after the park;
    // 2 (that pass) only they are visited
    int nearPhase;
    // symbolic laws of the biotic rates, one program per microbe (device memory)
    const clb_sym_program *sym;
    int symOn;                     // some microbe has one
    unsigned int symMask;          // which microbes: the compiled law skips them
    int compiled;                  // the compiled biotic law is on (bio_law)
    // the abiotic symbolic law (device memory), null when there is none
    const clb_sym_program *symAb;
    // rate capture (split kinetics kernel only): the step's requested
    // increments of every species, before the clamp, / dt, dense [S][N]
    double *rate;
    // thermodynamic gate of the biotic rates (clb_case_params thm_*)
    int thm, thmMicrobe;
    double thmRT, thmDG0, thmLoss, thmThr, thmChi, thmCmin, thmScale;
    double thmNu[CLB_DEV_MAXS];
};

template <typename R>
struct NsDev {
    int nx, ny, nz;
    long long N;
    int Nopen, Nint, Nleft;
    long long NoP;                 // popPad(Nopen)
    const int *list;               // [Nopen]
    const int8_t *geom;            // dense geometry codes (the pore-space statistics)
    const int *nb;                 // [18][Nopen]
    R *cur, *out;                  // [19 x NoP] tiled (pidx<19>)
    // the dense bounce-back stores paired with cur and out (nsBBSlot): the
    // interior pull reads the populations that come back from closed
    // neighbours from bbOut and writes the new ones there
    R *bbCur, *bbOut;              // [popPad(Nint) x 18]
    const uint8_t *cls, *clsPull;  // [Nopen]
    const unsigned long long *kinds;  // [Nint] source kind of each moving population, 2 bits per direction
    const int2 *nbr;               // [3][Nint] compact neighbour record (NsNbRecord words), the interior's pull
    double om, omBio, rhoBarRight;
    int first;
    int *done;
    int shape;                     // interior launch shape: 0 a block per chunk of 256 voxels,
                                   // > 0 that many blocks looping over chunks, loads a chunk ahead
};

// components of D3Q19 direction q (the order of D19C in clb_common.h), for
// compile-time use on both sides
__host__ __device__ constexpr int d19x(int q) {
    return (q == 1 || q == 7 || q == 9 || q == 11 || q == 13) ? 1
         : ((q == 2 || q == 8 || q == 10 || q == 12 || q == 14) ? -1 : 0);
}
__host__ __device__ constexpr int d19y(int q) {
    return (q == 3 || q == 7 || q == 10 || q == 15 || q == 17) ? 1
         : ((q == 4 || q == 8 || q == 9 || q == 16 || q == 18) ? -1 : 0);
}
__host__ __device__ constexpr int d19z(int q) {
    return (q == 5 || q == 11 || q == 14 || q == 15 || q == 18) ? 1
         : ((q == 6 || q == 12 || q == 13 || q == 16 || q == 17) ? -1 : 0);
}

// The interior flow pull's neighbour indices in 24 bytes instead of 18 x 4.
// The interior compact order is the dense order of the interior voxels, so
// within one x row of the lattice the listed voxels are consecutive: with
// P(y', z') the compact position of (x, y', z') (the listed interior voxels
// before it in dense order), the source (x + sx, y', z') of a direction is
// P - 1 for sx = -1, P for sx = 0 and P + [(x, y', z') listed] for sx = +1,
// and i + sx in the voxel's own row.  The record holds the positions of the
// eight other rows a D3Q19 pull reaches: P(y, z-1) and P(y, z+1) whole, the
// rows at y -/+ 1 as int16 offsets from i (z) or from those two (z -/+ 1),
// four "listed" bits for the rows whose x -/+ 1 voxels are sources, and one
// bit per direction whose index comes from the full table instead (a source
// on an x plane, or an offset beyond int16).  Every index the record gives is
// checked at creation against the full table with the decoder below.
// Device layout: three int2 words per interior voxel, word k of voxel i at
// [k * Nint + i]: (rowM, rowP), (d01, d23), (d45, flags).
struct NsNbRecord {
    int32_t rowM, rowP;            // P(y, z-1), P(y, z+1)
    int32_t d01, d23, d45;         // int16 pairs: P(y-1, z) - i, P(y+1, z) - i; P(y-1, z-1) - rowM, P(y+1, z-1) - rowM;
                                   // P(y-1, z+1) - rowP, P(y+1, z+1) - rowP
    uint32_t flags;                // bits 0-17: direction q = 1..18 from the full table; bits 18-21: (x, y-1, z),
                                   // (x, y+1, z), (x, y, z-1), (x, y, z+1) listed
};
static_assert(sizeof(NsNbRecord) == 24, "three 8-byte words");
#define NS_REC_FAR_MASK 0x3ffffu

__host__ __device__ inline int nsRecLo(int w) { return (int) (short) (w & 0xffff); }
__host__ __device__ inline int nsRecHi(int w) { return w >> 16; }
// the compact index of the source of direction q (1..18) of interior voxel i,
// the pull's neighbour at (x, y, z) - c_q
__host__ __device__ inline int nsRecSource(int q, int i, int rowM, int rowP, int d01, int d23, int d45, unsigned int flags) {
    const int sx = -d19x(q), sy = -d19y(q), sz = -d19z(q);
    if (sy == 0 && sz == 0) return i + sx;
    int P, bit;
    if (sz == 0) { P = i + (sy < 0 ? nsRecLo(d01) : nsRecHi(d01)); bit = sy < 0 ? 18 : 19; }
    else if (sy == 0) { P = sz < 0 ? rowM : rowP; bit = sz < 0 ? 20 : 21; }
    else {
        P = sz < 0 ? rowM + (sy < 0 ? nsRecLo(d23) : nsRecHi(d23)) : rowP + (sy < 0 ? nsRecLo(d45) : nsRecHi(d45));
        bit = 0;                                           // D3Q19 has no source at x -/+ 1 in these rows
    }
    return P + (sx < 0 ? -1 : (sx > 0 ? (int) ((flags >> bit) & 1u) : 0));
}

struct NsTracer {
    double ring[NS_TRACER_WINDOW];
    int head, count;
    long long iters;
    int converged, nonfinite;
    double last, prevE;
};

template <typename R>
struct CaDev {
    int nx, ny, nz;
    long long N;
    int M;
    int Nopen;
    const int *list;               // the compact open voxels (the only ones whose mask or total moves)
    double gmaxClosed;             // max total over the other voxels (fixed at creation)
    long long msumClosed;          // mask sum over the other voxels (fixed at creation)
    StepCtl *ctl;
    double *partMax;               // [blocks]
    long long *partMask;           // [blocks]
    R *B;                          // [M][N]
    R *Btot;
    int8_t *mask;
    const int *age;
    const int8_t *dist;
    R *outbox;                     // [M][6][N]
    // the sweep that last pushed from each voxel (its stamp, sweep + 1): a
    // voxel pushed in the current sweep holds the current stamp
    uint32_t *pushed;
    double bmax, thrdB;
    int half;
    uint64_t seed;
};

#define EQ_MAXS 8
#define EQ_MAXC 8
#define EQ_MAXD 4
#define EQ_MIN_CONC 1e-30
#define EQ_MAX_CONC 10.0
#define EQ_SOLVER_AAPCF 0
#define EQ_SOLVER_NEWTON 1

struct EqDev {
    int ns, nc;
    int solver;                    // EQ_SOLVER_AAPCF or EQ_SOLVER_NEWTON
    double S[EQ_MAXS][EQ_MAXC];
    double logk[EQ_MAXS];
    int isEq[EQ_MAXS];
    int compSpecies[EQ_MAXC];
    double mu0[EQ_MAXC];           // smallest positive coefficient per component (PCF scale)
    double tol;
    int maxIter, depth;
};

// ---- the batched form: the members of a batch in single launches --------
// A launch covers every member: a table of member descriptors in device
// memory and a map of the launch's workgroups to (member, workgroup of the
// member's own launch), so each member computes exactly what its own launch
// computes (the same block partials in the same order).
template <typename R>
struct NsMember {
    NsDev<R> a;
    double *partE, *partN;         // the member's block partials (interior, then planes)
    NsTracer *tracer;
    int gP, gB;                    // plane blocks, interior blocks (one per 256 voxels)
    int rule;
    double eps;
    int active;                    // 0: the member sits this solve out
};
template <typename R>
struct AdeMember {
    AdeDev<R> a;
    KinDev k;
    CaDev<R> c;
    int gInt, gPln;                // interior and plane workgroups at the batch's tile size
    int ca;                        // 1: a CA microbe, the step's mask/total decision on the device
    long long iT0;                 // the member's steps at the start of the batch
    long long oldMask;             // the mask sum its host accepted last
    int precipEvery, dissolEvery;  // a step on these intervals needs the host: it halts the member
};

// host launchers: compact lattices (clb_lattice.hip)
// the interior kernel, then (when given) an event, then the planes
template <typename R> hipError_t launchAdeStep(const AdeDev<R> &, const KinDev &, hipStream_t, hipEvent_t afterInterior);
template <typename R> hipError_t launchKinetics(const AdeDev<R> &, const KinDev &, hipStream_t);
// the kinetics over a device-held list of voxels, and the list of the voxels next to the dissolving phase
template <typename R> hipError_t launchKineticsListed(const AdeDev<R> &, const KinDev &, const int *sel,
                                                     const unsigned int *count, hipStream_t);
template <typename R> hipError_t launchNearPhaseList(const AdeDev<R> &, const KinDev &, int *sel, unsigned int *count,
                                                    hipStream_t);
template <typename R> hipError_t launchDissolPark(const AdeDev<R> &, const KinDev &, hipStream_t);
template <typename R> hipError_t launchDissolConvert(const AdeDev<R> &, int8_t *phase, const DissolConv &,
                                                     unsigned int *count, hipStream_t);
template <typename R> hipError_t launchAdeDensity(const AdeDev<R> &, int, double *, hipStream_t);
template <typename R> hipError_t launchAdeSet(const AdeDev<R> &, int, const double *, hipStream_t);
template <typename R> hipError_t launchAdeReopen(const AdeDev<R> &, const uint8_t *, hipStream_t);
template <typename R> hipError_t launchAdeTotal(const AdeDev<R> &, int, double *, hipStream_t);
template <typename R> hipError_t launchAdeHeld(const AdeDev<R> &, int, double *, hipStream_t);
hipError_t launchAdeKinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint, uint16_t *kinds, hipStream_t);
hipError_t launchNsKinds(const int *nb, const uint8_t *clsPull, int Nopen, int Nint, unsigned long long *kinds, hipStream_t);
// evStart / evEnd, when given, are recorded around the flow kernel alone (not its tracer)
template <typename R> hipError_t launchNsStep(const NsDev<R> &, double *, double *, NsTracer *, int, double, hipStream_t,
                                              hipEvent_t evStart = nullptr, hipEvent_t evEnd = nullptr);
template <typename R, typename RA> hipError_t launchNsVelocity(const NsDev<R> &, RA *, double *, hipStream_t);
template <typename R> hipError_t launchNsInit(const NsDev<R> &, double, hipStream_t);
hipError_t launchNsUpdateDyn(const int8_t *, const int *, uint8_t *, int, double, unsigned int *, hipStream_t);
template <typename R> hipError_t launchNsReopen(const NsDev<R> &, const uint8_t *, hipStream_t);
// the dense bounce-back store `bb` rebuilt from the populations `g` it pairs
// with, for the current pull kinds
template <typename R> hipError_t launchNsBBFill(const NsDev<R> &, const R *g, R *bb, hipStream_t);
template <typename R> hipError_t launchAdeBBFill(const AdeDev<R> &, const R *g, R *bb, hipStream_t);
template <typename R> hipError_t launchEquilibrium(const AdeDev<R> &, const EqDev &, unsigned int *, hipStream_t);
template <typename R> hipError_t launchScatterCompact(const R *, const int *, int, double *, hipStream_t);
template <typename R> hipError_t launchGatherCompact(const double *, const int *, int, R *, hipStream_t);
unsigned int nsStepBlocks(int Nopen, int Nint, int shape);
// batched form: one flow iteration (and each member's tracer) of every active
// member; one transport launch (interior, then planes) of every member at
// tile size `tile`, with the largest member's species bound
template <typename R> hipError_t launchNsStepGroup(const NsMember<R> *E, int nMembers, const int2 *map, int nBlocks,
                                                   hipStream_t);
template <typename R> hipError_t launchAdeStepGroup(const AdeMember<R> *E, const int2 *mapInt, int nInt,
                                                    const int2 *mapPln, int nPln, int maxSm, int tile, hipStream_t);
template <typename R> hipError_t launchCaMaskFinalGroup(const AdeMember<R> *E, int nMembers, hipStream_t);
// the members' convergence state after a run of flow iterations: per member
// {tracer iterations, converged, non-finite, done}
template <typename R> hipError_t launchNsGroupStatus(const NsMember<R> *E, int nMembers, long long *out, hipStream_t);
// the CA members' step controls: a new batch of steps (no halt, nothing
// finished, the accepted mask sum), and a copy of each member's control
// (slot j for member j) for one read-back
template <typename R> hipError_t launchCtlResetGroup(const AdeMember<R> *E, int nMembers, hipStream_t);
template <typename R> hipError_t launchCtlGatherGroup(const AdeMember<R> *E, int nMembers, StepCtl *out, hipStream_t);
// host launchers: dense grids (clb_aux.hip)
template <typename R> hipError_t launchCaMaskTotal(const CaDev<R> &, int check, hipStream_t);
template <typename R> hipError_t launchCaMaskFinal(const CaDev<R> &, int nparts, int check, hipStream_t);
hipError_t launchStepCtlReset(StepCtl *, long long oldMask, hipStream_t);
template <typename R> hipError_t launchCaPushGather(const CaDev<R> &, uint64_t, hipStream_t, bool total);
template <typename R> hipError_t launchAgeCompact(const CaDev<R> &, int *blist, unsigned int *blockScratch, unsigned int *count,
                                                  hipStream_t);
template <typename R> hipError_t launchCaAgeInplace(const CaDev<R> &, int *age, const int *blist, const unsigned int *nDev,
                                                    unsigned int *raised, int k, hipStream_t);
template <typename R> hipError_t launchCaAge(const CaDev<R> &, const int *, int *, long long *, hipStream_t);
template <typename R> hipError_t launchCaAgeOrdered(const CaDev<R> &, const int *ageOld, int *ageNew, const int *blist,
                                                    const unsigned int *nDev, unsigned int *changed, int k, hipStream_t);
hipError_t launchOmegaSwitch(const int8_t *, const int *, uint8_t *, int, hipStream_t);
hipError_t launchSolidSwitch(const int8_t *, const int *, uint8_t *, uint8_t *, int, unsigned int *, hipStream_t);
template <typename R> hipError_t launchPrecip(int, int, long long, int8_t *, const R *, double, int, int8_t *, unsigned int *, hipStream_t);
template <typename R> hipError_t launchScalarStats(int, int, long long, const R *, double *, hipStream_t);
hipError_t launchFinalSum(const double *, long long, int, int, int, double *, hipStream_t);
hipError_t launchFinalSumLL(const long long *, long long, long long *, hipStream_t);
template <typename R> hipError_t launchToDouble(const R *, double *, long long, hipStream_t);
template <typename R> hipError_t launchNsDensity(const NsDev<R> &, double *dst, hipStream_t);
hipError_t launchFillD(double *dst, size_t n, double v, hipStream_t);
template <typename R> hipError_t launchEnergyField(const AdeDev<R> &, const KinDev &, double *dst, int which, hipStream_t);
template <typename R> hipError_t launchFromDouble(const double *, R *, long long, hipStream_t);
hipError_t launchI8ToDouble(const int8_t *, double *, long long, hipStream_t);
hipError_t launchI32ToDouble(const int *, double *, long long, hipStream_t);
hipError_t launchDoubleToI32(const double *, int *, long long, hipStream_t);
hipError_t launchCopy16(const void *, void *, long long, hipStream_t);

}  // namespace clb

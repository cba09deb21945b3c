// clb_host.cpp -- host driver of the coupled reaction step and the C-ABI (include/clb.h).
//
// A context owns one HIP stream on one device; a case owns every device field
// of one simulation.  The driver reproduces the order of the CompLaB3D main
// loop (complab.cpp:2178-3262) with device kernels; host code only makes the
// loop decisions that the reference makes on reduced scalars (the CA trigger,
// the mask-change trigger, the flow re-solve and its convergence).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <algorithm>
#include <vector>

#include "clb.h"
#include "clb_device.h"

using namespace clb;

namespace {

struct HipFail {
    hipError_t e;
    std::string what;
};

#define HIPT(call)                                                     \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) throw HipFail{e_, std::string(#call)};   \
    } while (0)

struct Status {
    int code;
    std::string msg;
};

// ------------------------------------------------------------------ checks
// the parameter block's own consistency (everything but the geometry)
// the grain phase seeded on a voxel of geometry code gv: the id (k + 1) of the
// first declared grain phase on that material, 0 for none (initPhaseLattice3D,
// dissolutionVOP.hh:300-326, over the grain phases only, complab.cpp:1672-1683)
int grainPhaseOf(const clb_case_params &p, int gv) {
    if (!p.dissol_on) return 0;
    for (int k = 0; k < p.dissol_n_phases && k < CLB_MAX_PHASES; ++k)
        if (!p.dissol_is_precipitate[k] && p.dissol_material[k] == gv) return k + 1;
    return 0;
}

Status checkParamBlock(const clb_case_params *p) {
    if (!p) return {CLB_E_ARG, "null parameters"};
    if (p->nx < 3 || p->ny < 1 || p->nz < 1) return {CLB_E_ARG, "lattice must be at least 3 x 1 x 1"};
    // voxel indices are 32-bit on the device and in the compact tables
    if ((long long) p->nx * p->ny * p->nz > (long long) INT_MAX) return {CLB_E_ARG, "lattice holds more than 2^31 - 1 voxels"};
    if (p->precision != CLB_FP64 && p->precision != CLB_FP32 && p->precision != CLB_FP32_ADE)
        return {CLB_E_ARG, "unknown precision"};
    if (p->n_species < 0 || p->n_species > CLB_MAX_SPECIES) return {CLB_E_ARG, "n_species out of range"};
    if (p->n_microbes < 0 || p->n_microbes > CLB_MAX_MICROBES) return {CLB_E_ARG, "n_microbes out of range"};
    if (p->n_rxn < 0 || p->n_rxn > CLB_MAX_RXN) return {CLB_E_ARG, "n_rxn out of range"};
    for (int r = 0; r < p->n_rxn; ++r)
        if (p->rxn_r1[r] < -1 || p->rxn_r1[r] >= p->n_species || p->rxn_r2[r] < -1 || p->rxn_r2[r] >= p->n_species)
            return {CLB_E_ARG, "reaction reactant index out of range"};
    if (p->eq_on && (p->eq_nc < 0 || p->eq_nc > CLB_MAX_COMP)) return {CLB_E_ARG, "eq_nc out of range"};
    if (p->eq_on)
        for (int j = 0; j < p->eq_nc; ++j)
            if (p->eq_comp_species[j] < -1 || p->eq_comp_species[j] >= p->n_species)
                return {CLB_E_ARG, "eq_comp_species index out of range"};
    if (p->thm_on) {                                               // complab3d_thermo.hh:490-508
        if (!(p->thm_chi > 0.0)) return {CLB_E_ARG, "thermodynamic gate: chi must be > 0"};
        if (!(p->thm_cmin > 0.0)) return {CLB_E_ARG, "thermodynamic gate: cmin must be > 0, or ln Q can be -inf"};
        if (!(p->thm_rt > 0.0)) return {CLB_E_ARG, "thermodynamic gate: R T must be > 0"};
        if (!(p->thm_conc_scale > 0.0)) return {CLB_E_ARG, "thermodynamic gate: conc_scale must be > 0"};
        if (p->thm_dgloss < 0.0) return {CLB_E_ARG, "thermodynamic gate: dGloss must be >= 0"};
        if (p->thm_microbe < -1 || p->thm_microbe >= p->n_microbes) return {CLB_E_ARG, "thm_microbe out of range"};
    }
    if (p->eq_on && (p->eq_anderson_depth < 1 || p->eq_anderson_depth > 4)) return {CLB_E_ARG, "eq_anderson_depth must be 1..4"};
    for (int s = 0; s < p->n_species; ++s) {
        if (p->left_btype[s] < 0 || p->left_btype[s] > 2 || p->right_btype[s] < 0 || p->right_btype[s] > 2)
            return {CLB_E_ARG, "boundary type must be 0, 1 or 2"};
    }
    if (p->precip_on && (p->precip_species < 0 || p->precip_species >= p->n_species || p->mobile[p->precip_species]))
        return {CLB_E_ARG, "precip_species must be an immobile species"};
    if (p->precip_on && (p->precip_phase_id < 0 || p->precip_phase_id > 127)) return {CLB_E_ARG, "precip_phase_id must be 0..127"};
    if (p->dissol_on) {
        if (p->dissol_n_phases < 1 || p->dissol_n_phases > CLB_MAX_PHASES)
            return {CLB_E_ARG, "dissol_n_phases must be 1..CLB_MAX_PHASES"};       // complab.cpp:533-537
        if (!(p->dissol_reopen_frac > 0.0) || p->dissol_reopen_frac >= 1.0)
            return {CLB_E_ARG, "dissol_reopen_frac must lie in (0, 1)"};             // complab.cpp:539-543
        if (p->dissol_interval < 1) return {CLB_E_ARG, "dissol_interval must be >= 1"};
        for (int k = 0; k < p->dissol_n_phases; ++k) {
            const int sM = p->dissol_species[k];
            if (sM < 0 || sM >= p->n_species || p->mobile[sM])                       // complab.cpp:556-560
                return {CLB_E_ARG, "a phase's dissol_species must be an immobile species"};
            if (!(p->dissol_full[k] > 0.0)) return {CLB_E_ARG, "a phase's dissol_full must be positive"};   // 561-566
            if (p->dissol_acid[k] < -1 || p->dissol_acid[k] >= p->n_species) return {CLB_E_ARG, "dissol_acid out of range"};
            if (!std::isfinite(p->dissol_k[k]) || !(p->dissol_acid_order[k] >= 0.0) || !(p->dissol_mineral_order[k] >= 0.0) ||
                !std::isfinite(p->dissol_acid_order[k]) || !std::isfinite(p->dissol_mineral_order[k]))
                return {CLB_E_ARG, "a phase's rate constant must be finite and its orders finite and >= 0"};
            if (!p->dissol_is_precipitate[k] && (p->dissol_material[k] < 0 || p->dissol_material[k] >= CLB_PORE))
                return {CLB_E_ARG, "a grain phase needs a closed material number (0 or 1)"};
            if (!p->dissol_is_precipitate[k] && !(p->dissol_initial[k] >= 0.0))
                return {CLB_E_ARG, "a grain phase's dissol_initial must be >= 0"};
            for (int s = 0; s < p->n_species; ++s)
                if (!p->mobile[s] && s != sM && p->dissol_stoich[k][s] != 0.0)
                    return {CLB_E_ARG, "dissolution products must be mobile species"};
        }
        if (p->precip_on && p->precip_phase_id > 0) {
            const int k = p->precip_phase_id - 1;
            if (k >= p->dissol_n_phases || !p->dissol_is_precipitate[k] || p->dissol_species[k] != p->precip_species)
                return {CLB_E_ARG, "precip_phase_id must name a precipitate phase whose substrate is precip_species"};
        }
    }
    if (p->bio_law != CLB_BIO_NONE) {
        if (p->donor >= p->n_species || p->acceptor >= p->n_species) return {CLB_E_ARG, "donor/acceptor index out of range"};
    }
    return {CLB_OK, ""};
}

// a symbolic law's program against the case it binds to: counts, bindings,
// op codes, slots and the stack depth every op leaves, before it can run
Status checkSymProgram(const clb_case_params &p, const clb_sym_program &P, bool abiotic) {
    if (P.n_ops < 0 || P.n_ops > CLB_SYM_MAXOPS || P.n_vars < 0 || P.n_vars > CLB_SYM_MAXVARS ||
        P.n_rates < 1 || P.n_rates > CLB_SYM_MAXRATES || P.n_const < 0 || P.n_const > CLB_SYM_MAXCONST)
        return {CLB_E_ARG, "symbolic law: a count is out of range"};
    if (!(P.unit_scale > 0.0) || !std::isfinite(P.unit_scale)) return {CLB_E_ARG, "symbolic law: unit_scale"};
    for (int i = 0; i < P.n_vars; ++i) {
        if (P.var_species[i] < -1 || P.var_species[i] >= p.n_species) return {CLB_E_ARG, "symbolic law: variable binding"};
        if (abiotic && P.var_species[i] < 0) return {CLB_E_ARG, "the abiotic law binds no biomass"};
    }
    int pc = 0;
    for (int r = 0; r < P.n_rates; ++r) {
        if (P.rate_species[r] < -1 || P.rate_species[r] >= p.n_species) return {CLB_E_ARG, "symbolic law: rate binding"};
        if (P.rate_end[r] < pc || P.rate_end[r] > P.n_ops) return {CLB_E_ARG, "symbolic law: rate code bounds"};
        int sp = 0;
        for (; pc < P.rate_end[r]; ++pc) {            // the stack each op leaves, checked before any run
            const int op = P.op[pc], arg = P.arg[pc];
            if (op == CLB_SYM_CONST) { if (arg < 0 || arg >= P.n_const) return {CLB_E_ARG, "symbolic law: constant index"}; ++sp; }
            else if (op == CLB_SYM_VAR) { if (arg < 0 || arg >= P.n_vars + r) return {CLB_E_ARG, "symbolic law: value slot"}; ++sp; }
            else if (op == CLB_SYM_NEG || (op >= CLB_SYM_EXP && op <= CLB_SYM_TANH)) { if (sp < 1) return {CLB_E_ARG, "symbolic law: stack"}; }
            else if ((op >= CLB_SYM_ADD && op <= CLB_SYM_POW) || op == CLB_SYM_MIN || op == CLB_SYM_MAX) {
                if (sp < 2) return {CLB_E_ARG, "symbolic law: stack"};
                --sp;
            } else return {CLB_E_ARG, "symbolic law: unknown op"};
            if (sp > CLB_SYM_STACK) return {CLB_E_ARG, "symbolic law: stack deeper than the evaluator's"};
        }
        if (sp != 1) return {CLB_E_ARG, "symbolic law: a rate's code must leave one value"};
    }
    return {CLB_OK, ""};
}

// ------------------------------------------------------------------ checkpoint layout
// A checkpoint block: ChkHeader, the parameter block, the compact voxel list
// (checked, not restored: a checkpoint loads into a case of the same
// geometry), then every persistent device field in chkSectionSizes() order.
// Scratch buffers of a step (reduction partials, CA outbox, dissolution
// shares, staging) are not part of it: every step rewrites them before reading.
struct ChkHeader {
    char magic[8];
    int32_t abi, rBytes, rnBytes, nx, ny, nz, Nopen, Nint, Nleft, Sm, Si, M, flowOn, dissolOn;
    uint8_t fCurIsA, gCurIsA, adeFirst, adeReset, adeClsDirty, nsFirst, nsClsDirty, pad0;
    double nsOmega, nsBioOmega, deltaP, gmaxClosed;
    int64_t iT, oldTotMask, msumClosed, launchNo;
    int32_t nsSaturate, batchLen, clogged, resolveFresh;
    uint64_t caSweep;
    int64_t cnt[CLB_NCOUNTERS];
    uint32_t symMask, symOn;
};
static const char kChkMagic[8] = {'C', 'L', 'B', 'C', 'H', 'K', '0', '2'};
enum { CHK_SYM_SECTION = 20 };                     // index of the symbolic programs in the section list
// the byte sizes of the sections, from the shape the header states
std::vector<size_t> chkSectionSizes(const ChkHeader &h) {
    std::vector<size_t> v;
    const size_t r = (size_t) h.rBytes, rn = (size_t) h.rnBytes;
    const size_t n = (size_t) h.nx * h.ny * h.nz, no = (size_t) h.Nopen, ni = (size_t) std::max(h.Nint, 1);
    const size_t sm = (size_t) std::max(h.Sm, 1), si = (size_t) std::max(h.Si, 1), m = (size_t) std::max(h.M, 1);
    const size_t np = (size_t) std::max(h.Nopen - h.Nint, 1);
    v.push_back(n); v.push_back(n); v.push_back(n); v.push_back(n);          // mask geom dist phase
    v.push_back(no); v.push_back(no); v.push_back(no); v.push_back(no);      // adeCls adeClsPull nsCls nsClsPull
    v.push_back(no);                                                          // reopened
    v.push_back(ni * sizeof(uint16_t));                                       // adeKinds
    const size_t nop = (size_t) popPad(h.Nopen);                              // tiled population lattices
    v.push_back(sm * 7 * nop * r); v.push_back(sm * 7 * nop * r);             // fA fB
    v.push_back(si * n * r); v.push_back(m * n * r); v.push_back(n * r);      // imm B Btot
    v.push_back(n * sizeof(int)); v.push_back(n * sizeof(int));               // age age2 (pointer order)
    v.push_back(h.M > 0 ? n * sizeof(uint32_t) : 0);                          // pushed (sweep stamps)
    v.push_back(3 * no * r);                                                  // u
    v.push_back(sm * np * sizeof(double));                                    // rhoN
    v.push_back(((size_t) std::max(h.M, 1) + 1) * sizeof(clb_sym_program));   // symDev  (CHK_SYM_SECTION)
    v.push_back(sizeof(unsigned long long));                                  // clampCount
    v.push_back(3 * sizeof(unsigned long long));                              // dcount
    v.push_back(sizeof(StepCtl));                                             // ctl
    if (h.flowOn) {
        v.push_back(19 * nop * rn); v.push_back(19 * nop * rn);               // gA gB
        v.push_back(ni * sizeof(unsigned long long));                         // nsKinds
        v.push_back(sizeof(NsTracer));                                        // tracer
        v.push_back(sizeof(int));                                             // nsDone
    }
    return v;
}
// a checkpoint block checked from its bytes alone: magic, the parameter
// block, the shape the header states against it, the total size, and every
// symbolic program it would upload
Status checkCheckpointBlock(const uint8_t *src, size_t n, ChkHeader *hOut, clb_case_params *pOut) {
    if (!src || n < sizeof(ChkHeader) + sizeof(clb_case_params)) return {CLB_E_ARG, "not a checkpoint (too short)"};
    ChkHeader h;
    memcpy(&h, src, sizeof(h));
    if (memcmp(h.magic, kChkMagic, 8) != 0) return {CLB_E_ARG, "not a checkpoint of this library"};
    clb_case_params q;
    memcpy(&q, src + sizeof(h), sizeof(q));
    Status s = checkParamBlock(&q);
    if (s.code) return {s.code, "checkpoint parameter block: " + s.msg};
    if (h.nx != q.nx || h.ny != q.ny || h.nz != q.nz || h.M != q.n_microbes || h.flowOn != q.flow_on ||
        h.dissolOn != q.dissol_on || h.Nopen < 1 || h.Nint < 0 || h.Nleft < 0 || h.Nint + h.Nleft > h.Nopen ||
        (long long) h.Nopen > (long long) h.nx * h.ny * h.nz || h.Sm < 0 || h.Si < 0 || h.Sm + h.Si != q.n_species ||
        (h.rBytes != 4 && h.rBytes != 8) || (h.rnBytes != 4 && h.rnBytes != 8))
        return {CLB_E_ARG, "checkpoint header does not match its parameter block"};
    const std::vector<size_t> sz = chkSectionSizes(h);
    size_t total = sizeof(ChkHeader) + sizeof(clb_case_params) + (size_t) h.Nopen * sizeof(int), symOff = 0;
    for (size_t k = 0; k < sz.size(); ++k) { if (k == CHK_SYM_SECTION) symOff = total; total += sz[k]; }
    if (n != total) return {CLB_E_ARG, "checkpoint size does not match the case"};
    const int slots = std::max(q.n_microbes, 1) + 1;
    for (int k = 0; k < slots; ++k) {
        const bool abiotic = k == slots - 1;
        const bool used = abiotic ? (h.symOn & 2u) != 0 : (k < q.n_microbes && ((h.symMask >> k) & 1u));
        if (!used) continue;
        clb_sym_program P;
        memcpy(&P, src + symOff + (size_t) k * sizeof(clb_sym_program), sizeof(P));
        Status t = checkSymProgram(q, P, abiotic);
        if (t.code) return {t.code, "checkpoint " + t.msg};
    }
    if (h.symMask >> std::max(q.n_microbes, 0)) return {CLB_E_ARG, "checkpoint symbolic mask names a microbe the case lacks"};
    if (hOut) *hOut = h;
    if (pOut) *pOut = q;
    return {CLB_OK, ""};
}

struct CaseBase {
    clb_case_params p;
    int inBatches = 0;                          // batches holding the case (it cannot be destroyed then)
    int nx = 0, ny = 0, nz = 0;
    long long N = 0;
    std::vector<void *> allocs;
    virtual ~CaseBase() {}
    virtual void nsInit(hipStream_t, double dP) = 0;
    virtual Status nsSolve(hipStream_t, int maxIter, int rule, double eps, int reset, int *iters, int *conv) = 0;
    virtual void nsCouple(hipStream_t) = 0;
    virtual void nsStats(hipStream_t, double out[4]) = 0;
    virtual void nsUpdateDynamics(hipStream_t, double perm) = 0;
    virtual void adeStabilize(hipStream_t, int steps) = 0;
    virtual Status runSteps(hipStream_t, int n, int *done) = 0;
    virtual void adeSteps(hipStream_t, int n) = 0;
    virtual void equilibrium(hipStream_t) = 0;
    virtual Status caBlock(hipStream_t, int *sweeps) = 0;
    virtual Status readField(hipStream_t, int field, int index, double *dst) = 0;
    virtual Status writeField(hipStream_t, int field, int index, const double *src) = 0;
    virtual void speciesTotals(hipStream_t, double *out) = 0;
    virtual Status speciesTotalsScratch(hipStream_t, double *buf, size_t cap, double *out) = 0;
    virtual void fieldStats(hipStream_t, double *out) = 0;
    virtual void counters(hipStream_t, int64_t out[CLB_NCOUNTERS]) = 0;
    virtual void kernelTime(hipStream_t, int reset, double *ms, int64_t *n) = 0;
    virtual void flowKernelTime(int reset, double *ms, int64_t *n) = 0;
    virtual void setTimeStride(int stride) = 0;
    virtual Status setTuning(int knob, int value) = 0;
    virtual void setRelaxation(const double *omp, const double *omb, double dt) = 0;
    // the processors one at a time (host facade of the reference's processor API)
    virtual void adeCollideStream(hipStream_t, int n) = 0;
    virtual void kinetics(hipStream_t, int which) = 0;
    virtual void maskTotalNow(hipStream_t, double *gmax, long long *msum) = 0;
    virtual void caSweep(hipStream_t) = 0;
    virtual Status ageUpdate(hipStream_t, bool toFixedPoint) = 0;
    virtual void soluteDynamics(hipStream_t) = 0;
    virtual void soluteSolid(hipStream_t, int *reopened) = 0;
    virtual void precipConvert(hipStream_t, int *changed) = 0;
    virtual Status dissolConvert(hipStream_t, int *reopened) = 0;
    virtual Status setSymbolic(int microbe, const clb_sym_program *prog) = 0;
    virtual Status rateCapture(hipStream_t, bool on) = 0;
    // checkpoint / restart of the device-resident state
    virtual size_t checkpointBytes() = 0;
    virtual Status checkpointSave(hipStream_t, uint8_t *dst, size_t n) = 0;
    virtual Status checkpointLoad(hipStream_t, const uint8_t *src, size_t n) = 0;
};

template <typename T>
T *dalloc(std::vector<void *> &allocs, size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw HipFail{e, "hipMalloc"};
    allocs.push_back(p);
    return (T *) p;
}

inline unsigned int nblk(long long n, int b) { return (unsigned int) ((n + b - 1) / b); }

// R: storage of the transport lattices, immobile species and biomass;
// RN: storage of the flow lattice
template <typename R, typename RN = R>
struct Case : CaseBase {
    // fields
    int8_t *mask = nullptr, *geom = nullptr, *dist = nullptr, *phase = nullptr;
    // compact open-voxel layout (clb_device.h): list, neighbour tables, classes
    int Nopen = 0, Nint = 0, Nleft = 0;
    int *list = nullptr, *nb7 = nullptr, *nb19 = nullptr;
    int2 *nsNbr = nullptr;                      // [3][Nint] the interior flow pull's neighbour record words
    uint8_t *adeCls = nullptr, *adeClsPull = nullptr, *nsCls = nullptr, *nsClsPull = nullptr, *reopened = nullptr;
    uint16_t *adeKinds = nullptr;
    unsigned long long *nsKinds = nullptr;
    R *fA = nullptr, *fB = nullptr, *imm = nullptr, *B = nullptr, *Btot = nullptr, *outbox = nullptr;
    R *fbbA = nullptr, *fbbB = nullptr;         // transport bounce-back stores paired with fA / fB
    bool adeBBOn = false;                       // the stores are in use (AdeDev::bbOn)
    uint32_t *pushed = nullptr;                 // the stamp of the sweep that last pushed from a voxel
    int *age = nullptr, *age2 = nullptr, *ageList = nullptr;
    unsigned int *ageCnt = nullptr;             // [0] biomass voxels, [1..8] voxels raised per queued sweep
    RN *gA = nullptr, *gB = nullptr;
    RN *bbA = nullptr, *bbB = nullptr;          // dense bounce-back stores paired with gA / gB
    R *u = nullptr;
    double *rhoN = nullptr, *parts = nullptr, *scal = nullptr, *stage = nullptr;
    long long *partsLL = nullptr, *scalLL = nullptr;
    unsigned long long *clampCount = nullptr;
    unsigned int *ucnt = nullptr;       // [0] changed, [1] reopened, [2] eq nonconverged
    int *nsDone = nullptr;
    NsTracer *tracer = nullptr;
    // state
    int Sm = 0, Si = 0;
    int mobSpecies[CLB_DEV_MAXS], slot[CLB_DEV_MAXS];
    bool fCurIsA = true, gCurIsA = true;
    bool adeFirst = true, adeReset = false, adeClsDirty = false;
    bool nsFirst = true, nsClsDirty = false;
    double nsOmega = 1, nsBioOmega = 1, deltaP = 0;
    long long iT = 0;
    int nsSaturate = 0;
    int nsShape = 0;                            // flow interior launch shape (NsDev::shape; clb_set_tuning)
    int clogged = 0;                            // percolationFlag: the run stopped for good
    bool resolveFresh = true;                   // the in-step re-solves' tracer is still empty
    long long oldTotMask = 0;
    uint64_t caSweep_ = 0;
    int64_t cnt[CLB_NCOUNTERS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
    std::vector<long long> evLaunch;            // the transport launch each event pair brackets
    size_t evUsed = 0;
    long long launchNo = 0;                     // transport launches queued so far
    int timeStride = 1;                         // time every timeStride-th launch
    double kms = 0;
    int64_t klaunch = 0;
    // the same for the flow kernel (k_ns_step, its tracer left out): every
    // timeStride-th iteration of nsSolve, working iterations only
    std::vector<std::pair<hipEvent_t, hipEvent_t>> nsEvPool;
    std::vector<long long> nsEvIter;            // the solve's iteration each pair brackets
    size_t nsEvUsed = 0;
    long long nsLaunchNo = 0;
    double nsKms = 0;
    int64_t nsKlaunch = 0;
    unsigned int nparts = 0;
    // queued step batches (CA cases): the control block on the device, its
    // pinned host mirror, the mask/total partials and the closed-voxel constants
    StepCtl *ctl = nullptr, *ctlHost = nullptr;
    // small read-backs of the host's decisions land in pinned memory, and the
    // host waits on them by polling an event rather than sleeping in a
    // stream synchronisation (each round trip of the CA loop and the age
    // sweeps pays that wake-up)
    unsigned int *pinU = nullptr;               // kPinU counters
    hipEvent_t evWait = nullptr;
    static constexpr int kPinU = 32;
    double *caPartMax = nullptr;
    long long *caPartMask = nullptr;
    // dissolution: per-face product shares [S][Nopen], device counters (parked,
    // collected, clamped)
    double *share = nullptr;
    unsigned long long *dcount = nullptr;
    // the voxels next to the dissolving phase (the gather's voxels), rebuilt
    // on the device after the phases change
    int *nearList = nullptr;
    unsigned int *nearCnt = nullptr;
    bool nearDirty = true;
    // symbolic laws: one program per microbe on the device; the biotic rates
    // come from them when every microbe has one
    clb_sym_program *symDev = nullptr;
    uint32_t symMask = 0;
    bool symOn = false;
    // the abiotic symbolic law, held after the microbes' programs
    bool symAb = false;
    // rate capture: the split kinetics kernel writes each step's requested
    // increments / dt here ([S][N], allocated at the first capture)
    double *rateBuf = nullptr;
    bool rateOn = false;
    size_t symSlots() const { return (size_t) std::max(p.n_microbes, 1) + 1; }
    clb_sym_program *symAbDev() const { return symDev + symSlots() - 1; }
    double gmaxClosed = -1e300;
    long long msumClosed = 0;
    int batchLen = 8;
    static constexpr int kBatchMax = 64;
    static constexpr size_t kEvPool = 256;

    ~Case() override {
        for (auto &e : evpool) { (void) hipEventDestroy(e.first); (void) hipEventDestroy(e.second); }
        for (auto &e : nsEvPool) { (void) hipEventDestroy(e.first); (void) hipEventDestroy(e.second); }
        for (void *q : allocs) (void) hipFree(q);
        if (ctlHost) (void) hipHostFree(ctlHost);
        if (pinU) (void) hipHostFree(pinU);
        if (evWait) (void) hipEventDestroy(evWait);
    }
    // waits for everything queued on st so far, polling
    void waitHost(hipStream_t st) {
        HIPT(hipEventRecord(evWait, st));
        hipError_t e;
        while ((e = hipEventQuery(evWait)) == hipErrorNotReady) {}
        if (e != hipSuccess) throw HipFail{e, "hipEventQuery"};
    }
    // n counters of the device at src, read back into pinU[slot..] and waited on
    const unsigned int *readU(hipStream_t st, const unsigned int *src, int n, int slot = 0) {
        HIPT(hipMemcpyAsync(pinU + slot, src, (size_t) n * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        waitHost(st);
        return pinU + slot;
    }

    R *fCur() const { return fCurIsA ? fA : fB; }
    R *fOut() const { return fCurIsA ? fB : fA; }
    long long adeBBStride() const { return std::max(popPad(Nint), 64LL) * 6; }
    RN *gCur() const { return gCurIsA ? gA : gB; }
    RN *gOut() const { return gCurIsA ? gB : gA; }

    AdeDev<R> ade(const int *halt = nullptr) const {
        AdeDev<R> a;
        memset(&a, 0, sizeof(a));
        a.nx = nx; a.ny = ny; a.nz = nz; a.N = N;
        a.Nopen = Nopen; a.Nint = Nint; a.Nleft = Nleft; a.NoP = popPad(Nopen); a.list = list; a.nb = nb7;
        a.S = p.n_species; a.Sm = Sm; a.Smax = Sm > 0 ? Sm : 1; a.M = p.n_microbes;
        for (int s = 0; s < CLB_DEV_MAXS; ++s) {
            a.mobile[s] = s < p.n_species ? p.mobile[s] : 0;
            a.slot[s] = slot[s];
            a.mobSpecies[s] = mobSpecies[s];
            a.c0[s] = p.c0[s]; a.omP[s] = p.omega_pore[s]; a.omB[s] = p.omega_biofilm[s];
            a.lbt[s] = p.left_btype[s]; a.rbt[s] = p.right_btype[s];
            a.lbc[s] = p.left_bc[s]; a.rbc[s] = p.right_bc[s];
        }
        for (int ms = 0; ms < Sm; ++ms) {
            const int s = mobSpecies[ms];
            a.mlbt[ms] = p.left_btype[s]; a.mrbt[ms] = p.right_btype[s];
            a.mc0[ms] = p.c0[s]; a.momP[ms] = p.omega_pore[s]; a.momB[ms] = p.omega_biofilm[s];
        }
        a.cur = fCur(); a.out = fOut(); a.imm = imm; a.B = B;
        a.bbCur = fCurIsA ? fbbA : fbbB; a.bbOut = fCurIsA ? fbbB : fbbA; a.bbStride = adeBBStride();
        a.bbOn = adeBBOn ? 1 : 0;
        a.u = u; a.mask = mask; a.cls = adeCls; a.clsPull = adeClsPull; a.kinds = adeKinds;
        a.rhoN = rhoN;
        a.first = adeFirst ? 1 : 0; a.reset = adeReset ? 1 : 0;
        a.clampCount = clampCount;
        a.halt = halt;
        a.maskTotal = 0;
        a.Btot = Btot; a.thrdB = p.thrd_bfilm_frac * p.bmax;
        a.partMax = caPartMax; a.partMask = caPartMask;
        return a;
    }
    KinDev kin(bool on) const {
        KinDev k;
        memset(&k, 0, sizeof(k));
        k.on = on ? 1 : 0;
        k.bio = (p.bio_law != CLB_BIO_NONE && p.n_microbes > 0) ? 1 : 0;
        k.S = p.n_species; k.M = p.n_microbes;
        k.dt = p.dt; k.clamp = p.clamp;
        k.donor = p.donor; k.acceptor = p.acceptor;
        for (int m = 0; m < CLB_DEV_MAXM; ++m) {
            k.mu_max[m] = p.mu_max[m]; k.ks[m] = p.ks[m]; k.ke[m] = p.ke[m]; k.kdecay[m] = p.kdecay[m];
            for (int s = 0; s < CLB_DEV_MAXS; ++s) k.bio_stoich[m][s] = p.bio_stoich[m][s];
        }
        k.nrxn = p.n_rxn;
        for (int r = 0; r < CLB_DEV_MAXR; ++r) {
            k.r1[r] = p.rxn_r1[r]; k.r2[r] = p.rxn_r2[r]; k.rk[r] = p.rxn_k[r];
            for (int s = 0; s < CLB_DEV_MAXS; ++s) k.rxn_stoich[r][s] = p.rxn_stoich[r][s];
        }
        k.cap_frac = p.rxn_cap_frac; k.cap_dt = p.rxn_cap_dt; k.min_conc = p.rxn_min_conc;
        k.rate_scale = p.abiotic_rate_scale; k.surface_only = p.surface_only;
        k.dissol = p.dissol_on ? 1 : 0;
        k.dPhases = p.dissol_on ? p.dissol_n_phases : 0;
        k.dSurfaceOnly = p.dissol_surface_only;
        for (int q = 0; q < CLB_DEV_MAXP; ++q) {
            k.dMineral[q] = p.dissol_species[q]; k.dAcid[q] = p.dissol_acid[q];
            k.dK[q] = p.dissol_k[q]; k.dAo[q] = p.dissol_acid_order[q]; k.dMo[q] = p.dissol_mineral_order[q];
            for (int s = 0; s < CLB_DEV_MAXS; ++s) k.dStoich[q][s] = p.dissol_stoich[q][s];
        }
        k.phase = phase; k.share = share; k.dcount = dcount;
        k.sym = symDev; k.symOn = symOn ? 1 : 0; k.symMask = symMask;
        k.compiled = k.bio;
        k.symAb = symAb ? symAbDev() : nullptr;
        k.rate = rateOn ? rateBuf : nullptr;
        if (symOn && p.n_microbes > 0) k.bio = 1;
        k.thm = p.thm_on ? 1 : 0; k.thmMicrobe = p.thm_microbe;
        k.thmRT = p.thm_rt; k.thmDG0 = p.thm_dg0; k.thmLoss = p.thm_dgloss; k.thmThr = p.thm_threshold;
        k.thmChi = p.thm_chi; k.thmCmin = p.thm_cmin; k.thmScale = p.thm_conc_scale;
        for (int s = 0; s < CLB_DEV_MAXS; ++s) k.thmNu[s] = p.thm_nu[s];
        if (!k.bio && k.nrxn == 0 && !k.dissol && !k.symAb) k.on = 0;
        return k;
    }
    NsDev<RN> ns() const {
        NsDev<RN> a;
        memset(&a, 0, sizeof(a));
        a.nx = nx; a.ny = ny; a.nz = nz; a.N = N;
        a.Nopen = Nopen; a.Nint = Nint; a.Nleft = Nleft; a.NoP = popPad(Nopen); a.list = list; a.nb = nb19; a.geom = geom;
        a.cur = gCur(); a.out = gOut();
        a.bbCur = gCurIsA ? bbA : bbB; a.bbOut = gCurIsA ? bbB : bbA;
        a.cls = nsCls; a.clsPull = nsClsPull; a.kinds = nsKinds; a.nbr = nsNbr;
        a.om = nsOmega; a.omBio = nsBioOmega; a.rhoBarRight = -deltaP * 3.0;
        a.first = nsFirst ? 1 : 0;
        a.done = nsDone;
        a.shape = nsShape;
        return a;
    }
    CaDev<R> ca() const {
        CaDev<R> c;
        memset(&c, 0, sizeof(c));
        c.nx = nx; c.ny = ny; c.nz = nz; c.N = N; c.M = p.n_microbes;
        c.Nopen = Nopen; c.list = list;
        c.gmaxClosed = gmaxClosed; c.msumClosed = msumClosed;
        c.ctl = ctl; c.partMax = caPartMax; c.partMask = caPartMask;
        c.B = B; c.Btot = Btot; c.mask = mask; c.age = age; c.dist = dist;
        c.outbox = outbox; c.pushed = pushed;
        c.bmax = p.bmax; c.thrdB = p.thrd_bfilm_frac * p.bmax;
        c.half = p.ca_method == CLB_CA_HALF ? 1 : 0;
        c.seed = p.ca_seed;
        return c;
    }
    EqDev eqdev() const {
        EqDev e;
        memset(&e, 0, sizeof(e));
        e.ns = p.n_species; e.nc = p.eq_nc; e.solver = p.eq_solver;
        for (int i = 0; i < p.n_species; ++i) {
            e.logk[i] = p.eq_logk[i];
            e.isEq[i] = 0;
            for (int j = 0; j < p.eq_nc; ++j) { e.S[i][j] = p.eq_stoich[i][j]; if (std::fabs(p.eq_stoich[i][j]) > 1e-10) e.isEq[i] = 1; }
        }
        for (int j = 0; j < p.eq_nc; ++j) {
            e.compSpecies[j] = p.eq_comp_species[j];
            double mu0 = 1.0;                                   // pcf_residual, part4:702-712
            for (int i = 0; i < p.n_species; ++i) {
                if (!e.isEq[i]) continue;
                double mu = p.eq_stoich[i][j];
                if (mu > 1e-10 && mu < mu0) mu0 = mu;
            }
            e.mu0[j] = mu0;
        }
        e.tol = p.eq_tol; e.maxIter = p.eq_max_iter; e.depth = p.eq_anderson_depth;
        return e;
    }

    // flow dynamics from the static geometry (NSdomainSetup, complab_functions.hh:296-345)
    int nsClassOf(int gv) const {
        if (gv == CLB_PORE) return NS_BGK;
        if (gv > CLB_PORE) return p.perm_ratio > 1e-14 ? NS_BIO : NS_BB;
        return gv == CLB_WALL ? NS_BB : NS_NODYN;
    }
    // the interior voxels' neighbour records (NsNbRecord, one word array after
    // another) from the full D3Q19 table; every index a record gives is
    // decoded here as the pull decodes it and checked against the table
    std::vector<int32_t> nsRecords(const std::vector<int> &lst, const std::vector<int> &t19,
                                   const int (*C)[3]) const {
        const size_t M = (size_t) std::max(Nint, 1);
        std::vector<int32_t> words(6 * M, 0);
        for (int q = 1; q < 19; ++q)
            if (C[q][0] != d19x(q) || C[q][1] != d19y(q) || C[q][2] != d19z(q))
                throw HipFail{hipErrorInvalidValue, "D3Q19 direction table mismatch"};
        // compact position of (x, y, z): the listed interior voxels before it
        // (the interior list is in dense order: one pass counts them all)
        std::vector<int> before((size_t) N + 1);
        {
            int k = 0;
            for (long long v = 0; v <= N; ++v) {
                while (k < Nint && lst[(size_t) k] < v) ++k;
                before[(size_t) v] = k;
            }
        }
        auto pos = [&](int x, int y, int z) -> long long {
            if (y < 0 || y >= ny || z < 0 || z >= nz) return 0;
            return before[(size_t) (x + (long long) nx * (y + (long long) ny * z))];
        };
        auto listedAt = [&](int y, int z, int x) -> unsigned int {
            if (y < 0 || y >= ny || z < 0 || z >= nz) return 0u;
            const size_t v = (size_t) (x + (long long) nx * (y + (long long) ny * z));
            return before[v + 1] > before[v] ? 1u : 0u;
        };
        auto fit16 = [](long long d) -> int { return (d < -32768 || d > 32767) ? 0 : (int) d; };
        auto pack = [](int lo, int hi) -> int32_t { return (int32_t) ((uint32_t) (uint16_t) lo | ((uint32_t) (uint16_t) hi << 16)); };
        for (int i = 0; i < Nint; ++i) {
            const long long v = lst[i];
            const int x = (int) (v % nx), y = (int) ((v / nx) % ny), z = (int) (v / ((long long) nx * ny));
            NsNbRecord r;
            memset(&r, 0, sizeof(r));
            const long long pM = pos(x, y, z - 1), pP = pos(x, y, z + 1);
            r.rowM = (int32_t) pM; r.rowP = (int32_t) pP;
            r.d01 = pack(fit16(pos(x, y - 1, z) - i), fit16(pos(x, y + 1, z) - i));
            r.d23 = pack(fit16(pos(x, y - 1, z - 1) - pM), fit16(pos(x, y + 1, z - 1) - pM));
            r.d45 = pack(fit16(pos(x, y - 1, z + 1) - pP), fit16(pos(x, y + 1, z + 1) - pP));
            r.flags = (listedAt(y - 1, z, x) << 18) | (listedAt(y + 1, z, x) << 19) |
                      (listedAt(y, z - 1, x) << 20) | (listedAt(y, z + 1, x) << 21);
            for (int q = 1; q < 19; ++q) {
                const int n = t19[(size_t) (q - 1) * Nopen + i];
                if (n < 0) continue;                                // closed or outside: never decoded
                if (n >= Nint || nsRecSource(q, i, r.rowM, r.rowP, r.d01, r.d23, r.d45, r.flags) != n)
                    r.flags |= 1u << (q - 1);
            }
            words[2 * (size_t) i] = r.rowM; words[2 * (size_t) i + 1] = r.rowP;
            words[2 * (M + i)] = r.d01; words[2 * (M + i) + 1] = r.d23;
            words[2 * (2 * M + i)] = r.d45; words[2 * (2 * M + i) + 1] = (int32_t) r.flags;
        }
        return words;
    }

    void create(const clb_case_params &pp, const int8_t *geometry, const double *biomass0, hipStream_t st) {
        p = pp;
        nx = p.nx; ny = p.ny; nz = p.nz; N = (long long) nx * ny * nz;
        Sm = 0; Si = 0;
        for (int s = 0; s < CLB_DEV_MAXS; ++s) { mobSpecies[s] = 0; slot[s] = 0; }
        for (int s = 0; s < p.n_species; ++s) {
            if (p.mobile[s]) { slot[s] = Sm; mobSpecies[Sm] = s; ++Sm; }
            else { slot[s] = Si; ++Si; }
        }
        // the compact open-voxel order: interior, x = 0 plane, x = nx-1 plane
        std::vector<int8_t> g(geometry, geometry + N), dh(N), ph(N, 0);
        std::vector<int> lst, d2c(N, -1);
        // voxels that can carry solute: open geometry, and the grain of a
        // dissolvable phase (closed until it dissolves and reopens)
        auto listed = [&](int gv) { return gv >= CLB_PORE || grainPhaseOf(p, gv) > 0; };
        for (int pass = 0; pass < 3; ++pass)
            for (long long v = 0; v < N; ++v) {
                const int x = (int) (v % nx);
                const bool want = pass == 0 ? (x > 0 && x < nx - 1) : (pass == 1 ? x == 0 : x == nx - 1);
                if (want && listed(g[v])) { d2c[v] = (int) lst.size(); lst.push_back((int) v); }
                if (pass == 0 && v == N - 1) Nint = (int) lst.size();
                if (pass == 1 && v == N - 1) Nleft = (int) lst.size() - Nint;
            }
        Nopen = (int) lst.size();
        if (Nopen == 0) throw HipFail{hipErrorInvalidValue, "the geometry has no open voxel"};
        auto nbTable = [&](int Q, const int (*C)[3]) {
            std::vector<int> t((size_t) (Q - 1) * Nopen);
            for (int i = 0; i < Nopen; ++i) {
                const long long v = lst[i];
                const int x = (int) (v % nx), y = (int) ((v / nx) % ny), z = (int) (v / ((long long) nx * ny));
                for (int q = 1; q < Q; ++q) {
                    const int xs = x - C[q][0], ys = y - C[q][1], zs = z - C[q][2];
                    int n = -2;
                    if (xs >= 0 && ys >= 0 && zs >= 0 && xs < nx && ys < ny && zs < nz) {
                        n = d2c[xs + (long long) nx * (ys + (long long) ny * zs)];
                        if (n < 0) n = -1;
                    }
                    t[(size_t) (q - 1) * Nopen + i] = n;
                }
            }
            return t;
        };
        static const int C7h[7][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        static const int C19h[19][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
            {1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0},{1,0,1},{-1,0,-1},{1,0,-1},{-1,0,1},{0,1,1},{0,-1,-1},{0,1,-1},{0,-1,1}};
        mask = dalloc<int8_t>(allocs, N); geom = dalloc<int8_t>(allocs, N);
        dist = dalloc<int8_t>(allocs, N); phase = dalloc<int8_t>(allocs, N);
        list = dalloc<int>(allocs, Nopen);
        nb7 = dalloc<int>(allocs, (size_t) 6 * Nopen);
        adeCls = dalloc<uint8_t>(allocs, Nopen); adeClsPull = dalloc<uint8_t>(allocs, Nopen);
        // whole tiles of 64 (adeBBSlotOf reads a tile's line); the tail stays 0
        adeKinds = dalloc<uint16_t>(allocs, (size_t) std::max(popPad(Nint), 64LL));
        HIPT(hipMemsetAsync(adeKinds, 0, (size_t) std::max(popPad(Nint), 64LL) * sizeof(uint16_t), st));
        nsCls = dalloc<uint8_t>(allocs, Nopen); nsClsPull = dalloc<uint8_t>(allocs, Nopen);
        reopened = dalloc<uint8_t>(allocs, Nopen);
        fA = dalloc<R>(allocs, (size_t) std::max(Sm, 1) * 7 * popPad(Nopen));
        fB = dalloc<R>(allocs, (size_t) std::max(Sm, 1) * 7 * popPad(Nopen));
        imm = dalloc<R>(allocs, (size_t) std::max(Si, 1) * N);
        const int M = p.n_microbes;
        B = dalloc<R>(allocs, (size_t) std::max(M, 1) * N);
        Btot = dalloc<R>(allocs, N);
        if (M > 0) {
            outbox = dalloc<R>(allocs, (size_t) M * 6 * N);
            pushed = dalloc<uint32_t>(allocs, N);
            HIPT(hipMemsetAsync(pushed, 0, N * sizeof(uint32_t), st));      // no sweep has pushed
        }
        age = dalloc<int>(allocs, N); age2 = dalloc<int>(allocs, N); ageList = dalloc<int>(allocs, std::max(Nopen, 1));
        ageCnt = dalloc<unsigned int>(allocs, 32);
        if (p.flow_on) {
            gA = dalloc<RN>(allocs, 19 * (size_t) popPad(Nopen)); gB = dalloc<RN>(allocs, 19 * (size_t) popPad(Nopen));
            bbA = dalloc<RN>(allocs, 18 * (size_t) std::max(popPad(Nint), 64LL)); bbB = dalloc<RN>(allocs, 18 * (size_t) std::max(popPad(Nint), 64LL));
            nb19 = dalloc<int>(allocs, (size_t) 18 * Nopen);
            nsNbr = dalloc<int2>(allocs, 3 * (size_t) std::max(Nint, 1));
            nsKinds = dalloc<unsigned long long>(allocs, std::max(Nint, 1));
        }
        u = dalloc<R>(allocs, 3 * (size_t) Nopen);
        rhoN = dalloc<double>(allocs, (size_t) std::max(Sm, 1) * std::max(Nopen - Nint, 1));
        nparts = nblk(N, 128);
        parts = dalloc<double>(allocs, (size_t) nparts * 5);
        partsLL = dalloc<long long>(allocs, nparts);
        scal = dalloc<double>(allocs, 16);
        scalLL = dalloc<long long>(allocs, 4);
        stage = dalloc<double>(allocs, N);
        clampCount = dalloc<unsigned long long>(allocs, 1);
        ucnt = dalloc<unsigned int>(allocs, 4);
        nsDone = dalloc<int>(allocs, 1);
        tracer = dalloc<NsTracer>(allocs, 1);
        ctl = dalloc<StepCtl>(allocs, 1);
        symDev = dalloc<clb_sym_program>(allocs, symSlots());
        HIPT(hipMemsetAsync(symDev, 0, symSlots() * sizeof(clb_sym_program), st));
        if (p.dissol_on) {
            share = dalloc<double>(allocs, (size_t) std::max(p.n_species, 1) * Nopen);
            nearList = dalloc<int>(allocs, std::max(Nint, 1));
            nearCnt = dalloc<unsigned int>(allocs, 1);
        }
        dcount = dalloc<unsigned long long>(allocs, 3);
        // one mask/total partial per transport workgroup, at the smallest
        // workgroup the batched form uses (64)
        caPartMax = dalloc<double>(allocs, nblk(Nint, 64) + nblk(Nopen - Nint, 64) + 1);
        caPartMask = dalloc<long long>(allocs, nblk(Nint, 64) + nblk(Nopen - Nint, 64) + 1);
        {
            void *h = nullptr;
            HIPT(hipHostMalloc(&h, sizeof(StepCtl), hipHostMallocDefault));
            ctlHost = (StepCtl *) h;
            memset(ctlHost, 0, sizeof(StepCtl));
            HIPT(hipHostMalloc(&h, kPinU * sizeof(unsigned int), hipHostMallocDefault));
            pinU = (unsigned int *) h;
            HIPT(hipEventCreateWithFlags(&evWait, hipEventDisableTiming));
        }

        // voxel bytes and scalar fields
        std::vector<uint8_t> ac(Nopen), nc(Nopen);
        std::vector<int> ag(N);
        for (long long i = 0; i < N; ++i) {
            const int gv = g[i];
            const int x = (int) (i % nx), y = (int) ((i / nx) % ny), z = (int) (i / ((long long) nx * ny));
            const bool closed = gv == CLB_SOLID || gv == CLB_WALL;
            ag[i] = closed ? -1 : (gv == CLB_PORE ? 0 : 1);                      // createAgeDomain3D
            const bool in = x < nx - 1 && y < ny - 1 && z < nz - 1;             // calculateDistanceFromSolid
            dh[i] = (in && !closed) ? 1 : ((gv == CLB_SOLID && in) ? -1 : 0);
        }
        for (int i = 0; i < Nopen; ++i) {
            const int gv = g[lst[i]];
            ac[i] = gv == CLB_PORE ? ADE_PORE : (gv > CLB_PORE ? ADE_BFILM : ADE_BB);
            nc[i] = nsClassOf(gv);
        }
        for (long long i = 0; i < N; ++i) ph[i] = (int8_t) grainPhaseOf(p, g[i]);      // initPhaseLattice3D
        HIPT(hipMemcpyAsync(mask, g.data(), N, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(geom, g.data(), N, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(dist, dh.data(), N, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(phase, ph.data(), N, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpy(list, lst.data(), (size_t) Nopen * sizeof(int), hipMemcpyHostToDevice));
        {
            std::vector<int> t7 = nbTable(7, C7h);
            HIPT(hipMemcpy(nb7, t7.data(), t7.size() * sizeof(int), hipMemcpyHostToDevice));
            // the transport bounce-back store pays where closed links are few
            // (measured: 0.73 per interior voxel 5 % faster, 1.6 per voxel 3 %
            // slower than the own-slot reads) and no equilibrium changes every
            // voxel's own slots each step
            long long closedLinks = 0;
            for (int i = 0; i < Nint; ++i)
                for (int q = 1; q < 7; ++q) {
                    const int n = t7[(size_t) (q - 1) * Nopen + i];
                    if (n == NB_CLOSED || (n >= 0 && ac[(size_t) n] == ADE_BB)) ++closedLinks;
                }
            adeBBOn = Sm > 0 && Nint > 0 && !p.eq_on && closedLinks <= (long long) Nint;
            if (adeBBOn) {
                fbbA = dalloc<R>(allocs, (size_t) Sm * adeBBStride());
                fbbB = dalloc<R>(allocs, (size_t) Sm * adeBBStride());
            }
            if (p.flow_on) {
                std::vector<int> t19 = nbTable(19, C19h);
                HIPT(hipMemcpy(nb19, t19.data(), t19.size() * sizeof(int), hipMemcpyHostToDevice));
                std::vector<int32_t> rec = nsRecords(lst, t19, C19h);
                HIPT(hipMemcpy(nsNbr, rec.data(), rec.size() * sizeof(int32_t), hipMemcpyHostToDevice));
            }
        }
        HIPT(hipMemcpyAsync(adeCls, ac.data(), Nopen, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(adeClsPull, ac.data(), Nopen, hipMemcpyHostToDevice, st));
        HIPT(launchAdeKinds(nb7, adeClsPull, Nopen, Nint, adeKinds, st));
        HIPT(hipMemcpyAsync(nsCls, nc.data(), Nopen, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(nsClsPull, nc.data(), Nopen, hipMemcpyHostToDevice, st));
        if (p.flow_on) HIPT(launchNsKinds(nb19, nsClsPull, Nopen, Nint, nsKinds, st));
        HIPT(hipMemcpyAsync(age, ag.data(), N * sizeof(int), hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(age2, ag.data(), N * sizeof(int), hipMemcpyHostToDevice, st));
        HIPT(hipMemsetAsync(reopened, 0, Nopen, st));
        // populations at rest with c0 (initializeAtEquilibrium + stabilizeADElattice3D seeding)
        {
            const long long nop = popPad(Nopen);
            std::vector<R> fl((size_t) 7 * nop);
            for (int ms = 0; ms < Sm; ++ms) {
                int s = mobSpecies[ms];
                for (int q = 0; q < 7; ++q) {
                    R v = (R) ((q == 0 ? 0.25 : 0.125) * p.c0[s]);
                    for (long long i = 0; i < nop; ++i) fl[(size_t) pidx<7>(q, i)] = v;
                }
                HIPT(hipMemcpy(fA + (size_t) ms * 7 * nop, fl.data(), fl.size() * sizeof(R), hipMemcpyHostToDevice));
                HIPT(hipMemcpy(fB + (size_t) ms * 7 * nop, fl.data(), fl.size() * sizeof(R), hipMemcpyHostToDevice));
            }
            std::vector<R> sc(N);
            for (int s = 0; s < p.n_species; ++s) {
                if (p.mobile[s]) continue;
                std::fill(sc.begin(), sc.end(), (R) p.c0[s]);
                // each grain phase's inventory on its material, in phase order (complab.cpp:1687-1698)
                for (int q = 0; p.dissol_on && q < p.dissol_n_phases; ++q) {
                    if (p.dissol_is_precipitate[q] || !(p.dissol_initial[q] > 0.0) || p.dissol_species[q] != s) continue;
                    for (long long i = 0; i < N; ++i) if (g[i] == p.dissol_material[q]) sc[i] = (R) p.dissol_initial[q];
                }
                HIPT(hipMemcpy(imm + (size_t) slot[s] * N, sc.data(), N * sizeof(R), hipMemcpyHostToDevice));
            }
            std::vector<double> tot(N, 0.0);
            for (int m = 0; m < M; ++m) {
                for (long long i = 0; i < N; ++i) {
                    double v = biomass0 ? biomass0[(size_t) m * N + i] : (g[i] == CLB_BIOFILM + m ? p.b0[m] : 0.0);
                    sc[i] = (R) v;
                    tot[i] += (double) sc[i];
                }
                HIPT(hipMemcpy(B + (size_t) m * N, sc.data(), N * sizeof(R), hipMemcpyHostToDevice));
            }
            for (long long i = 0; i < N; ++i) sc[i] = (R) tot[i];
            HIPT(hipMemcpy(Btot, sc.data(), N * sizeof(R), hipMemcpyHostToDevice));
            long long ms = 0;
            for (long long i = 0; i < N; ++i) ms += g[i];
            oldTotMask = ms;
            // voxels outside the compact list keep their mask and total for good
            gmaxClosed = -1e300; msumClosed = 0;
            for (long long i = 0; i < N; ++i) {
                if (d2c[i] >= 0) continue;
                msumClosed += g[i];
                const double t = (double) sc[i];
                gmaxClosed = (std::isnan(gmaxClosed) || std::isnan(t)) ? t + gmaxClosed : std::max(gmaxClosed, t);
            }
        }
        HIPT(hipMemsetAsync(u, 0, 3 * (size_t) Nopen * sizeof(R), st));
        HIPT(hipMemsetAsync(clampCount, 0, sizeof(unsigned long long), st));
        HIPT(hipMemsetAsync(ucnt, 0, 4 * sizeof(unsigned int), st));
        HIPT(hipMemsetAsync(nsDone, 0, sizeof(int), st));
        HIPT(hipMemsetAsync(tracer, 0, sizeof(NsTracer), st));
        HIPT(hipMemsetAsync(ctl, 0, sizeof(StepCtl), st));
        HIPT(hipMemsetAsync(dcount, 0, 3 * sizeof(unsigned long long), st));
        if (share) HIPT(hipMemsetAsync(share, 0, (size_t) std::max(p.n_species, 1) * Nopen * sizeof(double), st));
        nsOmega = 1.0 / p.ns_tau;
        nsBioOmega = 1.0 / (p.perm_ratio * (1.0 / nsOmega - 0.5) + 0.5);
        HIPT(hipStreamSynchronize(st));
    }

    void setRelaxation(const double *omp, const double *omb, double dt) override {
        for (int s = 0; s < p.n_species; ++s) { p.omega_pore[s] = omp[s]; p.omega_biofilm[s] = omb[s]; }
        p.dt = dt;
    }

    // ------------------------------------------------------------ symbolic laws
    Status setSymbolic(int microbe, const clb_sym_program *prog) override {
        const bool abiotic = microbe == CLB_SYM_ABIOTIC;
        if (!abiotic && (microbe < 0 || microbe >= p.n_microbes)) return {CLB_E_ARG, "microbe index out of range"};
        if (prog) {
            if (p.dissol_on) return {CLB_E_ARG, "symbolic rate laws are not combined with dissolution here"};
            Status chk = checkSymProgram(p, *prog, abiotic);
            if (chk.code) return chk;
            clb_sym_program *dst = abiotic ? symAbDev() : symDev + microbe;
            HIPT(hipMemcpy(dst, prog, sizeof(clb_sym_program), hipMemcpyHostToDevice));
            if (abiotic) symAb = true; else symMask |= 1u << microbe;
        } else if (abiotic) {
            HIPT(hipMemset(symAbDev(), 0, sizeof(clb_sym_program)));
            symAb = false;
        } else {
            HIPT(hipMemset(symDev + microbe, 0, sizeof(clb_sym_program)));
            symMask &= ~(1u << microbe);
        }
        symOn = p.n_microbes > 0 && symMask != 0;
        return {CLB_OK, ""};
    }

    Status rateCapture(hipStream_t st, bool on) override {
        if (on && p.dissol_on) return {CLB_E_ARG, "rate capture is not combined with dissolution here"};
        if (on && !rateBuf) {
            rateBuf = dalloc<double>(allocs, (size_t) std::max(p.n_species, 1) * N);
            HIPT(launchFillD(rateBuf, (size_t) std::max(p.n_species, 1) * N, 0.0, st));
        }
        rateOn = on;
        return {CLB_OK, ""};
    }

    // ------------------------------------------------------------ checkpoint
    // the device pointers of the sections, sized by chkSectionSizes
    struct ChkSection { void *ptr; size_t bytes; };
    std::vector<ChkSection> chkSections() const {
        void *ptrs[] = {mask, geom, dist, phase, adeCls, adeClsPull, nsCls, nsClsPull, reopened, adeKinds,
                        fA, fB, imm, B, Btot, age, age2, pushed, u, rhoN, symDev, clampCount, dcount, ctl,
                        gA, gB, nsKinds, tracer, nsDone};
        const std::vector<size_t> sz = chkSectionSizes(chkHeader());
        std::vector<ChkSection> v;
        for (size_t k = 0; k < sz.size(); ++k) if (sz[k]) v.push_back({ptrs[k], sz[k]});
        return v;
    }
    ChkHeader chkHeader() const {
        ChkHeader h;
        memset(&h, 0, sizeof(h));
        memcpy(h.magic, kChkMagic, 8);
        h.abi = CLB_ABI_VERSION; h.rBytes = (int32_t) sizeof(R); h.rnBytes = (int32_t) sizeof(RN);
        h.nx = nx; h.ny = ny; h.nz = nz; h.Nopen = Nopen; h.Nint = Nint; h.Nleft = Nleft;
        h.Sm = Sm; h.Si = Si; h.M = p.n_microbes; h.flowOn = p.flow_on; h.dissolOn = p.dissol_on;
        h.fCurIsA = fCurIsA; h.gCurIsA = gCurIsA; h.adeFirst = adeFirst; h.adeReset = adeReset;
        h.adeClsDirty = adeClsDirty; h.nsFirst = nsFirst; h.nsClsDirty = nsClsDirty;
        h.nsOmega = nsOmega; h.nsBioOmega = nsBioOmega; h.deltaP = deltaP; h.gmaxClosed = gmaxClosed;
        h.iT = iT; h.oldTotMask = oldTotMask; h.msumClosed = msumClosed; h.launchNo = launchNo;
        h.nsSaturate = nsSaturate; h.batchLen = batchLen; h.caSweep = caSweep_;
        h.clogged = clogged; h.resolveFresh = resolveFresh ? 1 : 0;
        for (int k = 0; k < CLB_NCOUNTERS; ++k) h.cnt[k] = cnt[k];
        h.symMask = symMask; h.symOn = (symOn ? 1u : 0u) | (symAb ? 2u : 0u);
        return h;
    }
    size_t checkpointBytes() override {
        size_t b = sizeof(ChkHeader) + sizeof(clb_case_params) + (size_t) Nopen * sizeof(int);
        for (const auto &sec : chkSections()) b += sec.bytes;
        return b;
    }
    Status checkpointSave(hipStream_t st, uint8_t *dst, size_t n) override {
        if (n < checkpointBytes()) return {CLB_E_ARG, "checkpoint buffer too small"};
        flushTimes(st);
        const ChkHeader h = chkHeader();
        uint8_t *w = dst;
        memcpy(w, &h, sizeof(h)); w += sizeof(h);
        memcpy(w, &p, sizeof(p)); w += sizeof(p);
        HIPT(hipMemcpyAsync(w, list, (size_t) Nopen * sizeof(int), hipMemcpyDeviceToHost, st));
        w += (size_t) Nopen * sizeof(int);
        for (const auto &sec : chkSections()) {
            HIPT(hipMemcpyAsync(w, sec.ptr, sec.bytes, hipMemcpyDeviceToHost, st));
            w += sec.bytes;
        }
        HIPT(hipStreamSynchronize(st));
        return {CLB_OK, ""};
    }
    Status checkpointLoad(hipStream_t st, const uint8_t *src, size_t n) override {
        ChkHeader h;
        clb_case_params q;
        Status chk = checkCheckpointBlock(src, n, &h, &q);      // nothing reaches the device unchecked
        if (chk.code) return chk;
        const ChkHeader mine = chkHeader();
        if (h.abi != mine.abi || h.rBytes != mine.rBytes || h.rnBytes != mine.rnBytes)
            return {CLB_E_ARG, "checkpoint of another library version or precision"};
        if (h.nx != nx || h.ny != ny || h.nz != nz || h.Nopen != Nopen || h.Nint != Nint || h.Nleft != Nleft ||
            h.Sm != Sm || h.Si != Si || h.M != mine.M || h.flowOn != mine.flowOn || h.dissolOn != mine.dissolOn)
            return {CLB_E_ARG, "checkpoint of a case of another shape (lattice, species, microbes or processes)"};
        for (int s2 = 0; s2 < p.n_species; ++s2)
            if ((q.mobile[s2] != 0) != (p.mobile[s2] != 0)) return {CLB_E_ARG, "checkpoint of a case with other mobile species"};
        if (n != checkpointBytes()) return {CLB_E_ARG, "checkpoint size does not match the case"};
        const uint8_t *r = src + sizeof(h) + sizeof(q);
        std::vector<int> lst(Nopen);
        HIPT(hipMemcpy(lst.data(), list, (size_t) Nopen * sizeof(int), hipMemcpyDeviceToHost));
        if (memcmp(lst.data(), r, (size_t) Nopen * sizeof(int)) != 0)
            return {CLB_E_ARG, "checkpoint of another geometry"};
        r += (size_t) Nopen * sizeof(int);
        for (const auto &sec : chkSections()) {
            HIPT(hipMemcpyAsync(sec.ptr, r, sec.bytes, hipMemcpyHostToDevice, st));
            r += sec.bytes;
        }
        HIPT(hipStreamSynchronize(st));
        p = q;
        fCurIsA = h.fCurIsA; gCurIsA = h.gCurIsA; adeFirst = h.adeFirst; adeReset = h.adeReset;
        adeClsDirty = h.adeClsDirty; nsFirst = h.nsFirst; nsClsDirty = h.nsClsDirty;
        nsOmega = h.nsOmega; nsBioOmega = h.nsBioOmega; deltaP = h.deltaP; gmaxClosed = h.gmaxClosed;
        iT = h.iT; oldTotMask = h.oldTotMask; msumClosed = h.msumClosed; launchNo = h.launchNo;
        nsSaturate = h.nsSaturate; batchLen = h.batchLen; caSweep_ = h.caSweep;
        clogged = h.clogged; resolveFresh = h.resolveFresh != 0;
        for (int k = 0; k < CLB_NCOUNTERS; ++k) cnt[k] = h.cnt[k];
        kms = 0; klaunch = 0;
        nearDirty = true;
        symMask = h.symMask; symOn = (h.symOn & 1u) != 0; symAb = (h.symOn & 2u) != 0;
        nsBBRebuild(st);                                 // derived from the loaded populations
        adeBBRebuild(st);
        HIPT(hipStreamSynchronize(st));
        return {CLB_OK, ""};
    }

    // the transport's bounce-back stores rebuilt from the own slots of their
    // population buffers (k_ade_bb_fill): after the pull kinds change, voxels
    // reopen, a field is written or a restart.  The kinetics and equilibrium
    // kernels that change the current populations between launches update
    // the store beside the own slots themselves (adeBBPut)
    void adeBBRebuild(hipStream_t st) {
        if (!adeBBOn) return;
        AdeDev<R> a = ade();
        HIPT(launchAdeBBFill<R>(a, fA, fbbA, st));
        HIPT(launchAdeBBFill<R>(a, fB, fbbB, st));
    }

    // ------------------------------------------------------------------ flow
    // both dense bounce-back stores rebuilt from their population buffers:
    // after the pull kinds change, voxels reopen, or the populations are set
    void nsBBRebuild(hipStream_t st) {
        if (!p.flow_on) return;
        NsDev<RN> a = ns();
        HIPT(launchNsBBFill<RN>(a, gA, bbA, st));
        HIPT(launchNsBBFill<RN>(a, gB, bbB, st));
    }
    void nsInit(hipStream_t st, double dP) override {
        if (!p.flow_on) return;
        deltaP = dP;
        // NSdomainSetup: dynamics from the static geometry
        std::vector<int8_t> g(N);
        std::vector<int> lst(Nopen);
        HIPT(hipMemcpy(g.data(), geom, N, hipMemcpyDeviceToHost));
        HIPT(hipMemcpy(lst.data(), list, (size_t) Nopen * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<uint8_t> nc(Nopen);
        for (int i = 0; i < Nopen; ++i) nc[i] = nsClassOf(g[lst[i]]);
        HIPT(hipMemcpyAsync(nsCls, nc.data(), Nopen, hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(nsClsPull, nc.data(), Nopen, hipMemcpyHostToDevice, st));
        HIPT(launchNsKinds(nb19, nsClsPull, Nopen, Nint, nsKinds, st));
        nsBioOmega = 1.0 / (p.perm_ratio * (1.0 / nsOmega - 0.5) + 0.5);
        gCurIsA = true;
        HIPT(launchNsInit<RN>(ns(), dP, st));
        nsBBRebuild(st);
        nsFirst = true;
        nsClsDirty = false;
        HIPT(hipStreamSynchronize(st));
    }

    Status nsSolve(hipStream_t st, int maxIter, int rule, double eps, int reset, int *iters, int *conv) override {
        if (!p.flow_on) { if (iters) *iters = 0; if (conv) *conv = 1; return {CLB_OK, ""}; }
        if (reset) HIPT(hipMemsetAsync(tracer, 0, sizeof(NsTracer), st));
        HIPT(hipMemsetAsync(nsDone, 0, sizeof(int), st));
        // the tracer keeps its window; only the per-call fields restart
        NsTracer th;
        HIPT(hipMemcpyAsync(&th, tracer, sizeof(NsTracer), hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
        long long base = th.iters;
        th.converged = 0; th.nonfinite = 0;
        HIPT(hipMemcpyAsync(tracer, &th, sizeof(NsTracer), hipMemcpyHostToDevice, st));
        const unsigned int g = nsStepBlocks(Nopen, Nint, nsShape);
        double *partE = parts, *partN = parts + g;
        int done = 0;
        long long ran = 0;
        bool parity = gCurIsA;
        bool launched = false;
        while (ran < maxIter) {
            const int batch = (int) std::min<long long>(100, maxIter - ran);
            bool curA = parity;
            for (int k = 0; k < batch; ++k) {
                NsDev<RN> a = ns();
                a.cur = curA ? gA : gB; a.out = curA ? gB : gA;
                a.bbCur = curA ? bbA : bbB; a.bbOut = curA ? bbB : bbA;
                const bool timed = nsLaunchNo++ % timeStride == 0;
                if (timed && nsEvUsed == nsEvPool.size()) {
                    hipEvent_t e0, e1;
                    HIPT(hipEventCreate(&e0)); HIPT(hipEventCreate(&e1));
                    nsEvPool.push_back({e0, e1}); nsEvIter.push_back(0);
                }
                if (timed) nsEvIter[nsEvUsed] = ran + k;
                HIPT(launchNsStep<RN>(a, partE, partN, tracer, rule, eps, st, timed ? nsEvPool[nsEvUsed].first : nullptr,
                                      timed ? nsEvPool[nsEvUsed].second : nullptr));
                if (timed) ++nsEvUsed;
                nsFirst = false;
                if (!launched && nsClsDirty) {
                    HIPT(hipMemcpyAsync(nsClsPull, nsCls, Nopen, hipMemcpyDeviceToDevice, st));
                    HIPT(launchNsKinds(nb19, nsClsPull, Nopen, Nint, nsKinds, st));
                    nsBBRebuild(st);
                    nsClsDirty = false;
                }
                launched = true;
                curA = !curA;
            }
            HIPT(hipMemcpyAsync(&th, tracer, sizeof(NsTracer), hipMemcpyDeviceToHost, st));
            HIPT(hipMemcpyAsync(&done, nsDone, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPT(hipStreamSynchronize(st));
            const long long eff = th.iters - base;   // launches that did work
            for (size_t q = 0; q < nsEvUsed; ++q) {
                if (nsEvIter[q] >= eff) continue;    // a launch after the stop returns at once
                float ms = 0;
                HIPT(hipEventElapsedTime(&ms, nsEvPool[q].first, nsEvPool[q].second));
                nsKms += ms; ++nsKlaunch;
            }
            nsEvUsed = 0;
            parity = ((eff & 1) == 0) ? gCurIsA : !gCurIsA;
            if (!done && eff < ran + batch) return {CLB_E_HIP, "flow iterations were lost"};
            ran = eff;
            if (done) break;
        }
        gCurIsA = parity;
        if (iters) *iters = (int) ran;
        if (conv) *conv = th.converged ? 1 : 0;
        if (th.nonfinite) return {CLB_E_NONFINITE, "non-finite flow energy"};
        return {CLB_OK, ""};
    }

    void nsCouple(hipStream_t st) override {
        if (!p.flow_on) return;
        HIPT((launchNsVelocity<RN, R>(ns(), u, nullptr, st)));
    }
    void nsStats(hipStream_t st, double out[4]) override {
        for (int k = 0; k < 4; ++k) out[k] = 0;
        if (!p.flow_on) return;
        const unsigned int g = nblk(Nopen, 256);
        HIPT((launchNsVelocity<RN, R>(ns(), nullptr, parts, st)));
        for (int c = 0; c < 5; ++c) HIPT(launchFinalSum(parts, g, 5, c, c == 1, scal + c, st));
        double h[5];
        HIPT(hipMemcpyAsync(h, scal, 5 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
        out[0] = h[0] / ((double) (nx - 2) * ny * nz);
        // closed voxels of the box report u = 0 (BounceBack/NoDynamics computeVelocity)
        out[1] = (Nint < (long long) (nx - 2) * ny * nz) ? std::max(h[1], 0.0) : h[1];
        out[2] = h[2] / ((double) ny * nz);
        out[3] = h[4] > 0 ? h[3] / h[4] : 0.0;
    }
    void nsUpdateDynamics(hipStream_t st, double perm) override {
        if (!p.flow_on) return;
        nsBioOmega = 1.0 / (perm * (1.0 / nsOmega - 0.5) + 0.5);
        HIPT(hipMemsetAsync(ucnt, 0, sizeof(unsigned int), st));
        HIPT(launchNsUpdateDyn(mask, list, nsCls, Nopen, perm, ucnt, st));
        const unsigned int ch = readU(st, ucnt, 1)[0];
        if (ch) {
            HIPT(launchNsReopen<RN>(ns(), nullptr, st));
            nsBBRebuild(st);
            nsClsDirty = true;
        }
    }

    // ------------------------------------------------------------------ transport
    // device time of the interior transport kernel (the fused stream, collide
    // and kinetics of every interior voxel), from a pair of events around it on
    // the case's stream; read back when the stream has drained
    void flushTimes(hipStream_t st) {
        if (!evUsed) return;
        HIPT(hipStreamSynchronize(st));
        for (size_t k = 0; k < evUsed; ++k) {
            float ms = 0;
            HIPT(hipEventElapsedTime(&ms, evpool[k].first, evpool[k].second));
            kms += ms; ++klaunch;
        }
        evUsed = 0;
    }
    // one transport launch with its kinetics.  With dissolution the abiotic
    // increment of a voxel next to the dissolving phase waits for the products
    // the phase voxels park: the fused launch does transport, the biotic
    // kinetics and every other voxel's abiotic kinetics, then the park pass
    // (which reads only the waiting voxels' water), then the abiotic kinetics
    // with the gather at the waiting voxels (the reference's order,
    // complab.cpp:2461-2545).
    void transportLaunch(hipStream_t st, bool kinOn, const int *halt = nullptr, bool maskTotal = false) {
        if (kinOn && (symOn || symAb || rateOn)) {
            // the symbolic laws run in the split kinetics kernel, biotic then
            // abiotic, on the transported state (the biomass changes there, so
            // the step's mask/total update follows it, not the transport)
            fusedLaunch(st, false, halt, false);
            if (rateOn) HIPT(launchFillD(rateBuf, (size_t) std::max(p.n_species, 1) * N, 0.0, st));
            HIPT(launchKinetics<R>(ade(halt), kin(true), st));
            return;
        }
        if (!kinOn || !p.dissol_on) { fusedLaunch(st, kinOn, halt, maskTotal); return; }
        if (nearDirty) {
            HIPT(launchNearPhaseList<R>(ade(), kin(true), nearList, nearCnt, st));
            nearDirty = false;
        }
        fusedLaunch(st, true, halt, maskTotal, 1);
        KinDev k = kin(true);
        k.bio = 0;
        k.nearPhase = 2;
        HIPT(launchDissolPark<R>(ade(halt), k, st));
        HIPT(launchKineticsListed<R>(ade(halt), k, nearList, nearCnt, st));
    }
    void fusedLaunch(hipStream_t st, bool kinOn, const int *halt = nullptr, bool maskTotal = false,
                     int nearPhase = 0) {
        const bool timed = launchNo % timeStride == 0;
        if (timed && evUsed == evpool.size()) {
            if (evpool.size() >= kEvPool) flushTimes(st);
            else {
                hipEvent_t e0, e1;
                HIPT(hipEventCreate(&e0)); HIPT(hipEventCreate(&e1));
                evpool.push_back({e0, e1}); evLaunch.push_back(0);
            }
        }
        AdeDev<R> a = ade(halt);
        a.maskTotal = maskTotal ? 1 : 0;
        KinDev k = kin(kinOn);
        k.dissol = 0;
        k.nearPhase = nearPhase;
        if (!k.bio && k.nrxn == 0) k.on = 0;
        if (timed) HIPT(hipEventRecord(evpool[evUsed].first, st));
        HIPT(launchAdeStep<R>(a, k, st, timed ? evpool[evUsed].second : nullptr));
        if (timed) { evLaunch[evUsed] = launchNo; ++evUsed; }
        ++launchNo;
        fCurIsA = !fCurIsA;
        const bool wasFirst = adeFirst;
        adeFirst = false;
        adeReset = false;
        if (adeClsDirty) {
            HIPT(hipMemcpyAsync(adeClsPull, adeCls, Nopen, hipMemcpyDeviceToDevice, st));
            HIPT(launchAdeKinds(nb7, adeClsPull, Nopen, Nint, adeKinds, st));
            adeClsDirty = false;
            adeBBRebuild(st);
        } else if (wasFirst && adeBBOn) {
            // the first launch handed the closed neighbours' rest populations
            // to the own slots of the buffer it read (now the output buffer)
            HIPT(launchAdeBBFill<R>(ade(), fOut(), fCurIsA ? fbbB : fbbA, st));
        }
    }
    void adeStabilize(hipStream_t st, int steps) override {
        for (int t = 0; t < steps; ++t) fusedLaunch(st, false);
        flushTimes(st);
        adeReset = true;
    }
    void adeSteps(hipStream_t st, int n) override {
        for (int t = 0; t < n; ++t) transportLaunch(st, true);
        flushTimes(st);
    }
    void equilibrium(hipStream_t st) override { equilibriumQ(st, nullptr); }

    // ---- the processors one at a time
    void adeCollideStream(hipStream_t st, int n) override {
        for (int t = 0; t < n; ++t) fusedLaunch(st, false);
        flushTimes(st);
    }
    void kinetics(hipStream_t st, int which) override {
        KinDev k = kin(true);
        if (!(which & CLB_KIN_BIOTIC)) k.bio = 0;
        if (!(which & CLB_KIN_ABIOTIC)) { k.nrxn = 0; k.dissol = 0; k.symAb = nullptr; }
        if (!k.bio && k.nrxn == 0 && !k.dissol && !k.symAb) k.on = 0;
        if (k.dissol) {
            // the biotic family first, on its own, then park + abiotic + gather
            if (k.bio) { KinDev kb = k; kb.nrxn = 0; kb.dissol = 0; HIPT(launchKinetics<R>(ade(), kb, st)); }
            k.bio = 0;
            HIPT(launchDissolPark<R>(ade(), k, st));
        }
        HIPT(launchKinetics<R>(ade(), k, st));
    }
    Status dissolConvert(hipStream_t st, int *reopenedOut) override {
        unsigned int n = 0;
        if (reopenedOut) *reopenedOut = 0;
        if (!p.dissol_on) return {CLB_OK, ""};
        DissolConv dc;
        dc.n = p.dissol_n_phases;
        for (int q = 0; q < CLB_DEV_MAXP; ++q) {
            dc.slot[q] = q < dc.n ? slot[p.dissol_species[q]] : 0;
            dc.threshold[q] = q < dc.n ? p.dissol_reopen_frac * p.dissol_full[q] : 0.0;
        }
        HIPT(hipMemsetAsync(ucnt + 3, 0, sizeof(unsigned int), st));
        HIPT(launchDissolConvert<R>(ade(), phase, dc, ucnt + 3, st));
        n = readU(st, ucnt + 3, 1)[0];
        if (n) nearDirty = true;
        if (reopenedOut) *reopenedOut = (int) n;
        return {CLB_OK, ""};
    }
    void maskTotalNow(hipStream_t st, double *gmax, long long *msum) override { maskTotal(st, gmax, msum); }
    void caSweep(hipStream_t st) override {
        if (p.n_microbes <= 0) return;
        HIPT(launchCaPushGather<R>(ca(), caSweep_++, st, false));
        cnt[6]++;
    }
    Status ageUpdate(hipStream_t st, bool toFixedPoint) override { return ageRelax(st, toFixedPoint); }
    void soluteDynamics(hipStream_t st) override { HIPT(launchOmegaSwitch(mask, list, adeCls, Nopen, st)); }
    void soluteSolid(hipStream_t st, int *reopenedOut) override {
        unsigned int nre = 0;
        HIPT(hipMemsetAsync(ucnt + 1, 0, sizeof(unsigned int), st));
        HIPT(launchSolidSwitch(mask, list, adeCls, reopened, Nopen, ucnt + 1, st));
        nre = readU(st, ucnt + 1, 1)[0];
        if (nre) { HIPT(launchAdeReopen<R>(ade(), reopened, st)); adeBBRebuild(st); }
        adeClsDirty = true;
        if (reopenedOut) *reopenedOut = (int) nre;
    }
    void precipConvert(hipStream_t st, int *changed) override {
        unsigned int ch = 0;
        if (p.precip_species >= 0 && p.precip_species < p.n_species && !p.mobile[p.precip_species]) {
            HIPT(hipMemsetAsync(ucnt, 0, sizeof(unsigned int), st));
            HIPT(launchPrecip<R>(nx, ny, N, mask, imm + (size_t) slot[p.precip_species] * N, p.precip_max_rho, p.precip_phase_id, phase, ucnt, st));
            ch = readU(st, ucnt, 1)[0];
            if (ch) nearDirty = true;
        }
        if (changed) *changed = ch ? 1 : 0;
    }
    void equilibriumQ(hipStream_t st, const int *halt) {
        if (!p.eq_on || p.eq_nc <= 0) return;
        HIPT(launchEquilibrium<R>(ade(halt), eqdev(), ucnt + 2, st));
    }

    // mask/total update + global max of the total + mask sum
    void maskTotal(hipStream_t st, double *gmax, long long *msum) {
        HIPT(launchCaMaskTotal<R>(ca(), 0, st));
        readCtl(st, gmax, msum);
    }
    // the max total and mask sum the last final reduction published
    void readCtl(hipStream_t st, double *gmax, long long *msum) {
        HIPT(hipMemcpyAsync(ctlHost, ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, st));
        waitHost(st);
        *gmax = ctlHost->gmax;
        *msum = ctlHost->msum;
    }
    // updateAgeDistance3D: one Jacobi sweep, or sweeps to the fixed point.  To
    // the fixed point the sweeps run in place over the voxels holding biomass
    // (the only ages that can move; any order of the monotone updates reaches
    // the Jacobi fixed point, and faster).  They are queued eight at a time,
    // each skipping itself once the one before raised nothing, and the counts
    // of one batch are read back together: the first sweep that raised nothing
    // stands on the fixed point.  More than 1000 sweeps is an error.  The
    // sweeps read the compacted count on the device, so they are queued
    // behind the compaction without a round trip to the host.
    Status ageRelax(hipStream_t st, bool toFixedPoint) {
        if (!toFixedPoint) return ageOrderedSweep(st);
        ageRelaxQueue(st);
        waitHost(st);
        return ageRelaxFinish(st);
    }
    static constexpr int kAgeBatch = 8;
    // the compaction and the first batch of in-place sweeps, with the copy of
    // their counts to pinned memory queued behind them (the caller waits, for
    // this and whatever else it queued)
    void ageRelaxQueue(hipStream_t st) {
        HIPT(launchAgeCompact<R>(ca(), ageList, reinterpret_cast<unsigned int *>(partsLL), ageCnt, st));
        HIPT(hipMemsetAsync(ageCnt + 1, 0, kAgeBatch * sizeof(unsigned int), st));
        for (int k = 0; k < kAgeBatch; ++k) HIPT(launchCaAgeInplace<R>(ca(), age, ageList, ageCnt, ageCnt + 1, k, st));
        HIPT(hipMemcpyAsync(pinU, ageCnt + 1, kAgeBatch * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    }
    // after that wait: the first batch's counts, then further batches until a
    // sweep raised nothing
    Status ageRelaxFinish(hipStream_t st) {
        const unsigned int *raised = pinU;
        int swept = 0;
        while (true) {
            for (int k = 0; k < kAgeBatch; ++k) {
                ++swept;
                if (raised[k] == 0) return {CLB_OK, ""};
                if (swept > 1000) return {CLB_E_AGE_SWEEPS, "age relaxation did not settle"};
            }
            HIPT(hipMemsetAsync(ageCnt + 1, 0, kAgeBatch * sizeof(unsigned int), st));
            for (int k = 0; k < kAgeBatch; ++k) HIPT(launchCaAgeInplace<R>(ca(), age, ageList, ageCnt, ageCnt + 1, k, st));
            raised = readU(st, ageCnt + 1, kAgeBatch);
        }
    }
    // the single in-place sweep after a mask change (complab.cpp:3079), solved
    // as its triangular system over the voxels holding biomass (k_ca_age_ordered):
    // relaxation launches queued eight at a time until one changes nothing;
    // the planes x + y + z = k are exact after k + 1 launches, which bounds them
    Status ageOrderedSweep(hipStream_t st) {
        const int K = 8;
        HIPT(launchAgeCompact<R>(ca(), ageList, reinterpret_cast<unsigned int *>(partsLL), ageCnt, st));
        HIPT(hipMemcpyAsync(age2, age, (size_t) N * sizeof(int), hipMemcpyDeviceToDevice, st));
        unsigned int *chg = ageCnt + 16;
        const long long cap = (long long) nx + ny + nz + 2;
        long long launches = 0;
        while (true) {
            HIPT(hipMemsetAsync(chg, 0, K * sizeof(unsigned int), st));
            for (int k = 0; k < K; ++k) HIPT(launchCaAgeOrdered<R>(ca(), age, age2, ageList, ageCnt, chg, k, st));
            const unsigned int *changed = readU(st, chg, K);
            for (int k = 0; k < K; ++k) {
                ++launches;
                if (changed[k] == 0) { std::swap(age, age2); return {CLB_OK, ""}; }
                if (launches > cap) return {CLB_E_AGE_SWEEPS, "ordered age sweep did not settle"};
            }
        }
    }
    Status caBlock(hipStream_t st, int *sweepsOut) override {
        if (sweepsOut) *sweepsOut = 0;
        if (p.n_microbes <= 0) return {CLB_OK, ""};
        double gmax; long long msum;
        maskTotal(st, &gmax, &msum);
        return caFrom(st, gmax, msum, sweepsOut);
    }
    // the CA loop (complab.cpp:2793-3055) from a mask/total reduction already made
    Status caFrom(hipStream_t st, double gmax, long long msum, int *sweepsOut) {
        int sweeps = 0;
        if (sweepsOut) *sweepsOut = 0;
        if (!std::isfinite(gmax)) return {CLB_E_NONFINITE, "non-finite biomass"};
        if (gmax - p.bmax > 1e-12) {
            double prev = gmax; int stall = 0, wc = 0;
            while (gmax - p.bmax > 1e-12) {
                HIPT(launchCaPushGather<R>(ca(), caSweep_++, st, true));     // the sweep and the mask/total update
                // the age relaxation of every 50th sweep does not depend on
                // the sweep's reduction: it is queued before the wait for it
                const bool ageNow = wc % 50 == 0;
                if (ageNow) ageRelaxQueue(st);
                readCtl(st, &gmax, &msum);
                ++sweeps; cnt[6]++;
                if (!std::isfinite(gmax)) { if (sweepsOut) *sweepsOut = sweeps; return {CLB_E_NONFINITE, "non-finite biomass"}; }
                if (gmax > prev - 1e-12) { if (++stall >= 100) { if (sweepsOut) *sweepsOut = sweeps; return {CLB_E_CA_STALL, "CA spill stalled"}; } }
                else stall = 0;
                prev = gmax;
                if (ageNow) { Status s = ageRelaxFinish(st); if (s.code) return s; }
                if (wc > 1000) { if (sweepsOut) *sweepsOut = sweeps; return {CLB_E_CA_SWEEPS, "CA spill did not settle"}; }
                ++wc;
            }
        }
        lastMaskSum = msum;
        if (sweepsOut) *sweepsOut = sweeps;
        return {CLB_OK, ""};
    }
    long long lastMaskSum = 0;

    // an in-step flow re-solve (complab.cpp:3087-3093, 3162-3168, 3221-3226)
    // with the main loop's own tracer (ns_convg2, created empty before the
    // loop and never reset, complab.cpp:2166), the re-solve tolerance and cap
    Status resolveSolve(hipStream_t st, int *cv) {
        int it = 0;
        Status s = nsSolve(st, p.ns_max_iter, p.ns_conv_rule, p.ns_conv_eps, resolveFresh ? 1 : 0, &it, cv);
        resolveFresh = false;
        return s;
    }
    // re-solve after a mask change (maskPath, complab.cpp:3085-3130) or a
    // precipitation seal (3159-3178).  A re-solve that runs out of iterations
    // with no flow through the outlet plane x = nx-2 clogs the case
    // (percolationFlag); so does one whose energy is not finite on the
    // mask-change path, where the velocity is then not coupled (3119-3121)
    Status resolve(hipStream_t st, double perm, bool maskPath) {
        nsUpdateDynamics(st, perm);
        int cv = 0;
        Status s = resolveSolve(st, &cv);
        if (s.code == CLB_E_NONFINITE && maskPath) { cnt[4]++; clogged = 1; return {CLB_OK, ""}; }
        if (s.code) return s;
        cnt[4]++;
        if (!cv) {
            nsSaturate = 1;
            double stt[4];
            nsStats(st, stt);
            if (stt[2] > 1e-12) nsSaturate = 0;            // thrd, complab3d_processors.hh:60
            else clogged = 1;
        }
        if (!(maskPath && clogged)) nsCouple(st);
        return {CLB_OK, ""};
    }

    bool precipDue(long long it) const {
        return p.precip_on && p.precip_interval > 0 && it % p.precip_interval == 0 && p.precip_species >= 0;
    }
    bool dissolDue(long long it) const {
        return p.dissol_on && p.dissol_interval > 0 && it % p.dissol_interval == 0;
    }
    bool hostDue(long long it) const { return precipDue(it) || dissolDue(it); }
    // the host part of a step after its transport, equilibrium and (CA cases)
    // mask/total reduction: CA spill, mask-change consequences, precipitation
    Status stepTail(hipStream_t st, double gmax, long long msum) {
        if (p.n_microbes > 0) {
            int sw = 0;
            Status s = caFrom(st, gmax, msum, &sw);
            if (s.code) return s;
            if (lastMaskSum != oldTotMask) {
                oldTotMask = lastMaskSum;
                cnt[5]++;
                // the solute switch does not depend on the ages: queued first,
                // it runs before the age sweep's wait
                if (p.ade_update_interval > 0 && iT % p.ade_update_interval == 0) HIPT(launchOmegaSwitch(mask, list, adeCls, Nopen, st));
                Status a = ageRelax(st, false);
                if (a.code) return a;
                if (p.ns_update_interval > 0 && iT % p.ns_update_interval == 0 && p.flow_on && !nsSaturate) {
                    Status r = resolve(st, p.perm_ratio, true);
                    if (r.code) return r;
                }
            }
        }
        if (precipDue(iT)) {
            HIPT(hipMemsetAsync(ucnt, 0, 2 * sizeof(unsigned int), st));
            HIPT(launchPrecip<R>(nx, ny, N, mask, imm + (size_t) slot[p.precip_species] * N, p.precip_max_rho, p.precip_phase_id, phase, ucnt, st));
            const unsigned int ch = readU(st, ucnt, 1)[0];
            if (ch) {
                cnt[5]++;
                nearDirty = true;
                if (p.flow_on && !nsSaturate) { Status r = resolve(st, p.precip_perm_ratio, false); if (r.code) return r; }
                HIPT(hipMemsetAsync(ucnt + 1, 0, sizeof(unsigned int), st));
                HIPT(launchSolidSwitch(mask, list, adeCls, reopened, Nopen, ucnt + 1, st));
                const unsigned int nre = readU(st, ucnt + 1, 1)[0];
                if (nre) { HIPT(launchAdeReopen<R>(ade(), reopened, st)); adeBBRebuild(st); }
                adeClsDirty = true;
            }
        }
        if (dissolDue(iT)) {                                  // complab.cpp:3203-3247
            int n = 0;
            Status d = dissolConvert(st, &n);
            if (d.code) return d;
            if (n > 0) {
                cnt[5]++;
                cnt[11] += n;
                if (p.flow_on) {
                    nsUpdateDynamics(st, p.perm_ratio);
                    nsSaturate = 0;
                    int cv = 0;
                    Status r = resolveSolve(st, &cv);
                    if (r.code) return r;
                    cnt[4]++;
                    nsCouple(st);
                    clogged = 0;                                 // reopening can undo a clog (complab.cpp:3234)
                }
                soluteSolid(st, nullptr);
            }
        }
        ++iT; cnt[7]++;
        if (clogged) return {CLB_CLOGGED, "the pore space no longer percolates"};   // complab.cpp:3261
        return {CLB_OK, ""};
    }

    // one coupled step with the host in the loop
    Status coupledStep(hipStream_t st) {
        transportLaunch(st, true);
        if (p.eq_on) equilibrium(st);
        double gmax = 0; long long msum = 0;
        if (p.n_microbes > 0) maskTotal(st, &gmax, &msum);
        return stepTail(st, gmax, msum);
    }

    // CA cases: steps are queued in batches.  Each step's mask/total reduction
    // decides on the device whether the step needs the host; the first that
    // does halts the batch (every later kernel of the batch returns at once)
    // and the host finishes that step.  A precipitation step always ends a
    // batch and always halts.  The batch length adapts: it doubles after a
    // batch that ran through and shrinks to the halting position otherwise.
    Status runBatch(hipStream_t st, int want, int *finished) {
        *finished = 0;
        int len = want;
        for (int j = 0; j < len; ++j) if (hostDue(iT + j)) { len = j + 1; break; }
        if (evUsed + (size_t) len > kEvPool) flushTimes(st);
        HIPT(launchStepCtlReset(ctl, oldTotMask, st));
        const long long l0 = launchNo;
        // the transport kernels make the step's mask and total update and its
        // block partials (interior, then planes) when every open voxel passes
        // through them; the final reduction takes the step's decision
        const bool fused = Sm > 0 && Nint > 0 && !symOn && !symAb && !rateOn;
        const int nparts = (int) (nblk(Nint, 256) + nblk(Nopen - Nint, 256));
        for (int j = 0; j < len; ++j) {
            transportLaunch(st, true, &ctl->halt, fused);
            if (p.eq_on) equilibriumQ(st, &ctl->halt);
            if (fused) HIPT(launchCaMaskFinal<R>(ca(), nparts, hostDue(iT + j) ? 2 : 1, st));
            else HIPT(launchCaMaskTotal<R>(ca(), hostDue(iT + j) ? 2 : 1, st));
        }
        HIPT(hipMemcpyAsync(ctlHost, ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, st));
        waitHost(st);
        const int fin = ctlHost->done;
        const bool halt = ctlHost->halt != 0;
        const int ran = fin + (halt ? 1 : 0);               // transport launches that did work
        if (fin > len || ran > len) return {CLB_E_HIP, "step batch bookkeeping failed"};
        if ((len - ran) & 1) fCurIsA = !fCurIsA;            // launches that returned at once swapped nothing
        while (evUsed > 0 && evLaunch[evUsed - 1] >= l0 + ran) --evUsed;   // and are not timed
        iT += fin; cnt[7] += fin;
        *finished = fin;
        batchLen = halt ? std::max(1, fin + 1) : std::min(kBatchMax, 2 * batchLen);
        if (halt) {
            Status s = stepTail(st, ctlHost->gmax, ctlHost->msum);
            if (s.code == CLB_CLOGGED) *finished = fin + 1;      // the step ran to its end
            if (s.code) return s;
            *finished = fin + 1;
        }
        return {CLB_OK, ""};
    }

    Status runSteps(hipStream_t st, int n, int *done) override {
        int k = 0;
        Status s{CLB_OK, ""};
        if (clogged) { if (done) *done = 0; return {CLB_CLOGGED, "the pore space no longer percolates"}; }
        if (p.n_microbes > 0) {
            while (k < n) {
                int fin = 0;
                s = runBatch(st, std::min(n - k, batchLen), &fin);
                k += fin;
                if (s.code) break;
            }
        } else {
            for (; k < n; ++k) { s = coupledStep(st); if (s.code) { if (s.code == CLB_CLOGGED) ++k; break; } }
        }
        flushTimes(st);
        if (done) *done = k;
        return s;
    }

    // ------------------------------------------------------------------ fields
    Status readField(hipStream_t st, int field, int index, double *dst) override {
        switch (field) {
        case CLB_FIELD_CONC:
            if (index < 0 || index >= p.n_species) return {CLB_E_ARG, "species index out of range"};
            if (p.mobile[index]) {
                AdeDev<R> a = ade();
                HIPT(hipMemsetAsync(stage, 0, N * sizeof(double), st));
                HIPT(launchAdeDensity<R>(a, slot[index], stage, st));
            } else HIPT(launchToDouble<R>(imm + (size_t) slot[index] * N, stage, N, st));
            break;
        case CLB_FIELD_BIOMASS:
            if (index < 0 || index >= p.n_microbes) return {CLB_E_ARG, "microbe index out of range"};
            HIPT(launchToDouble<R>(B + (size_t) index * N, stage, N, st));
            break;
        case CLB_FIELD_BTOTAL: HIPT(launchToDouble<R>(Btot, stage, N, st)); break;
        case CLB_FIELD_MASK: HIPT(launchI8ToDouble(mask, stage, N, st)); break;
        case CLB_FIELD_VEL:
            if (index < 0 || index > 2) return {CLB_E_ARG, "velocity component out of range"};
            HIPT(hipMemsetAsync(stage, 0, N * sizeof(double), st));
            HIPT(launchScatterCompact<R>(u + (size_t) index * Nopen, list, Nopen, stage, st));
            break;
        case CLB_FIELD_AGE: HIPT(launchI32ToDouble(age, stage, N, st)); break;
        case CLB_FIELD_PHASE: HIPT(launchI8ToDouble(phase, stage, N, st)); break;
        case CLB_FIELD_DG:
        case CLB_FIELD_FT:                                   // computeEnergyField3D, complab3d_energyfield.hh:88-150
            if (!p.thm_on) return {CLB_E_STATE, "the case has no thermodynamic block"};
            if (index < 0 || index >= p.n_microbes || (p.thm_microbe >= 0 && index != p.thm_microbe))
                return {CLB_E_ARG, "that microbe has no thermodynamic block"};
            HIPT(launchFillD(stage, (size_t) N, 0.0, st));
            HIPT(launchEnergyField<R>(ade(), kin(true), stage, field == CLB_FIELD_FT ? 1 : 0, st));
            break;
        case CLB_FIELD_RATE:
            if (!rateOn) return {CLB_E_STATE, "rate capture is off"};
            if (index < 0 || index >= p.n_species) return {CLB_E_ARG, "species index out of range"};
            HIPT(hipMemcpyAsync(stage, rateBuf + (size_t) index * N, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            break;
        case CLB_FIELD_NSRHO:
            if (!p.flow_on) return {CLB_E_STATE, "the case has no flow lattice"};
            HIPT(launchNsDensity<RN>(ns(), stage, st));
            break;
        default: return {CLB_E_ARG, "unknown field"};
        }
        HIPT(hipMemcpyAsync(dst, stage, N * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
        return {CLB_OK, ""};
    }
    Status writeField(hipStream_t st, int field, int index, const double *src) override {
        HIPT(hipMemcpyAsync(stage, src, N * sizeof(double), hipMemcpyHostToDevice, st));
        if (field == CLB_FIELD_CONC) {
            if (index < 0 || index >= p.n_species) return {CLB_E_ARG, "species index out of range"};
            if (p.mobile[index]) {
                if (!adeFirst) return {CLB_E_STATE, "a concentration field can only be written before the first step"};
                HIPT(launchAdeSet<R>(ade(), slot[index], stage, st));
                adeBBRebuild(st);
            } else HIPT(launchFromDouble<R>(stage, imm + (size_t) slot[index] * N, N, st));
        } else if (field == CLB_FIELD_BIOMASS) {
            if (index < 0 || index >= p.n_microbes) return {CLB_E_ARG, "microbe index out of range"};
            HIPT(launchFromDouble<R>(stage, B + (size_t) index * N, N, st));
            std::vector<R> tot(N, (R) 0), one(N);
            std::vector<double> acc(N, 0.0);
            for (int m = 0; m < p.n_microbes; ++m) {
                HIPT(hipMemcpyAsync(one.data(), B + (size_t) m * N, N * sizeof(R), hipMemcpyDeviceToHost, st));
                HIPT(hipStreamSynchronize(st));
                for (long long i = 0; i < N; ++i) acc[i] += (double) one[i];
            }
            for (long long i = 0; i < N; ++i) tot[i] = (R) acc[i];
            HIPT(hipMemcpyAsync(Btot, tot.data(), N * sizeof(R), hipMemcpyHostToDevice, st));
        } else if (field == CLB_FIELD_VEL) {
            if (index < 0 || index > 2) return {CLB_E_ARG, "velocity component out of range"};
            HIPT(launchGatherCompact<R>(stage, list, Nopen, u + (size_t) index * Nopen, st));
        } else if (field == CLB_FIELD_AGE) {
            HIPT(launchDoubleToI32(stage, age, N, st));
            HIPT(hipMemcpyAsync(age2, age, (size_t) N * sizeof(int), hipMemcpyDeviceToDevice, st));
        } else return {CLB_E_ARG, "field cannot be written"};
        HIPT(hipStreamSynchronize(st));
        return {CLB_OK, ""};
    }
    void speciesTotals(hipStream_t st, double *out) override { speciesTotalsWith(st, parts, scal, out); }
    // the reductions write every block partial before the final sums read
    // them, so neither buffer needs clearing
    void speciesTotalsWith(hipStream_t st, double *partBuf, double *sums, double *out) {
        for (int s = 0; s < p.n_species; ++s) {
            if (p.mobile[s]) {
                HIPT(launchAdeTotal<R>(ade(), slot[s], partBuf, st));
                HIPT(launchFinalSum(partBuf, nblk(Nint, 256), 1, 0, 0, sums + s, st));
            } else {
                HIPT(launchScalarStats<R>(nx, ny, N, imm + (size_t) slot[s] * N, partBuf, st));
                HIPT(launchFinalSum(partBuf, nblk(N, 256), 4, 0, 0, sums + s, st));
            }
        }
        HIPT(hipMemcpyAsync(out, sums, p.n_species * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
    }
    // doubles of partials + sums the totals need in a caller's buffer
    size_t totalsScratchDoubles() const {
        size_t need = 0;
        for (int s = 0; s < p.n_species; ++s)
            need = std::max(need, p.mobile[s] ? (size_t) nblk(Nint, 256) : (size_t) 4 * nblk(N, 256));
        return need + CLB_MAX_SPECIES;
    }
    Status speciesTotalsScratch(hipStream_t st, double *buf, size_t cap, double *out) override {
        const size_t need = totalsScratchDoubles();
        if (need > cap)
            return {CLB_E_ARG, "the case's reductions need " + std::to_string(need * sizeof(double)) +
                               " bytes of scratch; the static scratch holds " + std::to_string(cap * sizeof(double))};
        speciesTotalsWith(st, buf + CLB_MAX_SPECIES, buf, out);
        return {CLB_OK, ""};
    }
    // per species, then per microbe: total, minimum, maximum, held
    void fieldStats(hipStream_t st, double *out) override {
        const int S = p.n_species;
        for (int r = 0; r < S + p.n_microbes; ++r) {
            double v[4] = {0, 0, 0, 0};
            if (r < S && p.mobile[r]) {
                HIPT(launchAdeTotal<R>(ade(), slot[r], parts, st));
                HIPT(launchFinalSum(parts, nblk(Nint, 256), 1, 0, 0, scal, st));
                const long long g = nblk(Nopen, 256);
                HIPT(launchAdeHeld<R>(ade(), slot[r], parts, st));
                HIPT(launchFinalSum(parts, g, 3, 1, 1, scal + 1, st));
                HIPT(launchFinalSum(parts, g, 3, 2, 1, scal + 2, st));
                HIPT(launchFinalSum(parts, g, 3, 0, 0, scal + 3, st));
            } else {
                const R *f = r < S ? imm + (size_t) slot[r] * N : B + (size_t) (r - S) * N;
                const long long g = nblk(N, 256);
                HIPT(launchScalarStats<R>(nx, ny, N, f, parts, st));
                for (int k = 0; k < 4; ++k) HIPT(launchFinalSum(parts, g, 4, k, k == 1 || k == 2, scal + k, st));
            }
            HIPT(hipMemcpyAsync(v, scal, sizeof(v), hipMemcpyDeviceToHost, st));
            HIPT(hipStreamSynchronize(st));
            const bool none = v[2] < -1e299;                   // no voxel counted
            out[4 * r + 0] = v[0];
            out[4 * r + 1] = none ? 0.0 : -v[1];
            out[4 * r + 2] = none ? 0.0 : v[2];
            out[4 * r + 3] = (r < S && !p.mobile[r]) ? 0.0 : v[3];   // nothing of an immobile species streams
        }
    }
    void counters(hipStream_t st, int64_t out[CLB_NCOUNTERS]) override {
        unsigned long long cc = 0, dc[3] = {0, 0, 0};
        unsigned int u3[4];
        std::vector<int8_t> m(N);
        HIPT(hipMemcpyAsync(&cc, clampCount, sizeof(cc), hipMemcpyDeviceToHost, st));
        HIPT(hipMemcpyAsync(u3, ucnt, sizeof(u3), hipMemcpyDeviceToHost, st));
        HIPT(hipMemcpyAsync(dc, dcount, sizeof(dc), hipMemcpyDeviceToHost, st));
        HIPT(hipMemcpyAsync(m.data(), mask, N, hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
        for (int k = 0; k < CLB_NCOUNTERS; ++k) out[k] = cnt[k];
        // open and biofilm voxels of the reaction box on the live mask (MaskedScalarCounts3D)
        int64_t open = 0, bio = 0;
        for (long long i = 0; i < N; ++i) {
            const int x = (int) (i % nx);
            if (x < 1 || x > nx - 2) continue;
            if (m[i] >= CLB_PORE) ++open;
            if (m[i] >= CLB_BIOFILM) ++bio;
        }
        out[0] = open; out[1] = bio;
        out[2] = (int64_t) cc;
        out[3] = (int64_t) u3[2];
        out[8] = (int64_t) dc[0]; out[9] = (int64_t) dc[1]; out[10] = (int64_t) dc[2];
    }
    void setTimeStride(int stride) override { timeStride = stride; }
    Status setTuning(int knob, int value) override {
        if (knob == CLB_TUNE_NS_SHAPE) {
            if (value < 0) return {CLB_E_ARG, "flow launch shape must be >= 0"};
            nsShape = value;
            return {CLB_OK, ""};
        }
        if (knob == CLB_TUNE_ADE_BB) {
            if (value != 0 && value != 1) return {CLB_E_ARG, "the transport bounce-back store knob takes 0 or 1"};
            if (launchNo > 0) return {CLB_E_STATE, "the transport bounce-back store is chosen before the first transport launch"};
            if (value == 1 && p.eq_on) return {CLB_E_ARG, "the equilibrium kernel does not keep the transport bounce-back store in step"};
            if (value == 1 && Sm > 0 && Nint > 0 && !fbbA) {
                fbbA = dalloc<R>(allocs, (size_t) Sm * adeBBStride());
                fbbB = dalloc<R>(allocs, (size_t) Sm * adeBBStride());
            }
            adeBBOn = value == 1 && Sm > 0 && Nint > 0;
            return {CLB_OK, ""};
        }
        return {CLB_E_ARG, "unknown tuning knob"};
    }
    void kernelTime(hipStream_t, int reset, double *ms, int64_t *n) override {
        if (ms) *ms = klaunch ? kms / (double) klaunch : 0.0;
        if (n) *n = klaunch;
        if (reset) { kms = 0; klaunch = 0; }
    }
    void flowKernelTime(int reset, double *ms, int64_t *n) override {
        if (ms) *ms = nsKlaunch ? nsKms / (double) nsKlaunch : 0.0;
        if (n) *n = nsKlaunch;
        if (reset) { nsKms = 0; nsKlaunch = 0; }
    }
};

// ------------------------------------------------------------------ the batched form
// Many cases stepped together.  Each kernel of a flow iteration or of a
// transport step covers every member of the batch, through a table of member
// descriptors and a map of the launch's workgroups to (member, workgroup of
// the member's own launch), so each member computes exactly what its own
// launches compute: the same voxel arithmetic, the same block partials summed
// in the same order, the same step decisions on the device.  What a step
// needs the host for (a CA spill, a mask change and its flow re-solve,
// precipitation) runs on the member's own path between runs of steps.  Runs
// of steps and of flow iterations are captured once in HIP graphs and
// replayed; the transport workgroup size is the autotuner's pick.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void) hipFree(p); }
    // room for n elements (the contents are not kept); true when the buffer moved
    bool ensure(size_t n) {
        if (p && n <= cap) return false;
        if (p) { HIPT(hipFree(p)); p = nullptr; cap = 0; }
        const size_t c = std::max<size_t>(n, 1);
        HIPT(hipMalloc(&p, c * sizeof(T)));
        cap = c;
        return true;
    }
};
template <typename T>
struct PinBuf {
    T *p = nullptr;
    size_t cap = 0;
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { if (p) (void) hipHostFree(p); }
    void ensure(size_t n) {
        if (p && n <= cap) return;
        if (p) { HIPT(hipHostFree(p)); p = nullptr; cap = 0; }
        const size_t c = std::max<size_t>(n, 1);
        void *h = nullptr;
        HIPT(hipHostMalloc(&h, c * sizeof(T), hipHostMallocDefault));
        p = static_cast<T *>(h);
        cap = c;
    }
};
// an instantiated graph and what its kernel nodes were captured with (device
// pointers, grid sizes): replayed while that stays the same
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> key;
    GraphExec() = default;
    GraphExec(const GraphExec &) = delete;
    GraphExec &operator=(const GraphExec &) = delete;
    ~GraphExec() { reset(); }
    void reset() {
        if (exec) (void) hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    template <class F>
    void ensure(hipStream_t st, const std::vector<uintptr_t> &k, F &&body) {
        if (exec && key == k) return;
        reset();
        HIPT(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        hipGraph_t g = nullptr;
        try {
            body();
        } catch (...) {
            (void) hipStreamEndCapture(st, &g);
            if (g) (void) hipGraphDestroy(g);
            throw;
        }
        HIPT(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void) hipGraphDestroy(g);
        if (e != hipSuccess) { exec = nullptr; throw HipFail{e, "hipGraphInstantiate"}; }
        key = k;
    }
    void launch(hipStream_t st) { HIPT(hipGraphLaunch(exec, st)); }
};

struct BatchBase {
    std::vector<int> ids;
    std::vector<CaseBase *> members;
    virtual ~BatchBase() { for (CaseBase *c : members) c->inBatches--; }
    virtual bool accepts(CaseBase *c) const = 0;
    virtual void attach(CaseBase *c) = 0;
    virtual void detach(size_t k) = 0;
    virtual Status nsSolve(hipStream_t, int maxIter, int rule, double eps, int reset, int *iters, int *conv, int *status) = 0;
    virtual Status stabilize(hipStream_t, int steps, int *status) = 0;
    virtual Status runSteps(hipStream_t, int n, int *done, int *status) = 0;
    // the transport workgroup sizes the autotuner tries, and per mode (0 the
    // coupled steps, 1 the stabilization) the pick (-1: not tuned yet) and the
    // measured time per step of each candidate
    static constexpr int kTiles[3] = {256, 128, 64};
    int pick[2] = {-1, -1};
    bool fixed[2] = {false, false};
    double tuneUs[2][3] = {{0, 0, 0}, {0, 0, 0}};
};
constexpr int BatchBase::kTiles[3];

template <typename R, typename RN>
struct Batch : BatchBase {
    using C = Case<R, RN>;
    static constexpr int kFlowRun = 100;        // flow iterations per replay (even: members keep their parity)
    static constexpr int kSegSteps = 64;        // coupled steps per run between host checks
    static constexpr int kSegStab = 512;        // stabilization steps per run
    std::vector<C *> m;
    // flow: descriptors [parity][member], workgroup map, convergence read-back
    PinBuf<NsMember<RN>> nsHost;
    DevBuf<NsMember<RN>> nsTab;
    PinBuf<int2> nsMapHost;
    DevBuf<int2> nsMap;
    DevBuf<long long> nsStat;
    PinBuf<long long> nsStatHost;
    GraphExec nsGraph;
    // transport, per candidate workgroup size: descriptors [parity][member] and
    // the map (interior workgroups by species class, then the plane workgroups)
    struct Plan {
        struct Range { int cls, off, cnt; };
        std::vector<Range> bulk;
        int plnOff = 0, plnCnt = 0, members = 0;
        bool anyCa = false;
    };
    PinBuf<AdeMember<R>> adeHost[3];
    DevBuf<AdeMember<R>> adeTab[3];
    PinBuf<int2> mapHost[3];
    DevBuf<int2> mapDev[3];
    Plan plan[3];
    GraphExec pair[3];                          // two steps (both parities) per candidate
    DevBuf<StepCtl> ctlDev;
    PinBuf<StepCtl> ctlHost;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ~Batch() override {
        if (ev0) (void) hipEventDestroy(ev0);
        if (ev1) (void) hipEventDestroy(ev1);
    }
    bool accepts(CaseBase *c) const override { return dynamic_cast<C *>(c) != nullptr; }
    void attach(CaseBase *c) override { m.push_back(static_cast<C *>(c)); members.push_back(c); c->inBatches++; }
    void detach(size_t k) override {
        members[k]->inBatches--;
        members.erase(members.begin() + (long) k);
        m.erase(m.begin() + (long) k);
        ids.erase(ids.begin() + (long) k);
    }

    // ---- flow
    static int nsPlanes(const C *c) { return (int) nblk(c->Nopen - c->Nint, 256); }
    static int nsInterior(const C *c) { return (int) nblk(c->Nint, 256); }
    // descriptors and map of the members `act` (parity cur[j] now); returns the workgroups
    int nsBuild(hipStream_t st, const std::vector<int> &act, const std::vector<char> &cur, int rule, double eps) {
        HIPT(hipStreamSynchronize(st));                 // the staging buffers are free
        const int n = (int) act.size();
        nsHost.ensure(2 * (size_t) n);
        nsTab.ensure(2 * (size_t) n);
        size_t nb = 0;
        for (int k = 0; k < n; ++k) {
            C *c = m[act[k]];
            for (int t = 0; t < 2; ++t) {
                const bool a = (t == 0) == (cur[act[k]] != 0);
                NsMember<RN> &e = nsHost.p[(size_t) t * n + k];
                memset(&e, 0, sizeof(e));
                e.a = c->ns();
                e.a.cur = a ? c->gA : c->gB;
                e.a.out = a ? c->gB : c->gA;
                e.a.bbCur = a ? c->bbA : c->bbB;
                e.a.bbOut = a ? c->bbB : c->bbA;
                e.a.first = 0;
                e.a.shape = 0;
                e.gB = nsInterior(c);
                e.gP = nsPlanes(c);
                e.partE = c->parts;
                e.partN = c->parts + (e.gB + e.gP);
                e.tracer = c->tracer;
                e.rule = rule;
                e.eps = eps;
                e.active = 1;
            }
            nb += (size_t) nsInterior(c) + nsPlanes(c);
        }
        if (nb > (size_t) INT_MAX) throw HipFail{hipErrorInvalidValue, "batch flow grid too large"};
        nsMapHost.ensure(nb);
        nsMap.ensure(nb);
        size_t q = 0;
        for (int k = 0; k < n; ++k) {
            const int g = nsInterior(m[act[k]]) + nsPlanes(m[act[k]]);
            for (int b = 0; b < g; ++b) nsMapHost.p[q++] = make_int2(k, b);
        }
        HIPT(hipMemcpyAsync(nsTab.p, nsHost.p, 2 * (size_t) n * sizeof(NsMember<RN>), hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(nsMap.p, nsMapHost.p, nb * sizeof(int2), hipMemcpyHostToDevice, st));
        return (int) nb;
    }
    void nsReadStatus(hipStream_t st, int n) {
        nsStat.ensure(4 * (size_t) n);
        nsStatHost.ensure(4 * (size_t) n);
        HIPT(launchNsGroupStatus<RN>(nsTab.p, n, nsStat.p, st));
        HIPT(hipMemcpyAsync(nsStatHost.p, nsStat.p, 4 * (size_t) n * sizeof(long long), hipMemcpyDeviceToHost, st));
        HIPT(hipStreamSynchronize(st));
    }
    // Every member's flow solve as its own nsSolve does it: the first iteration
    // on the member's own path (the hand-over after an initialisation, the
    // refresh of changed classes), the rest in launches over every member
    // still iterating, each stopping on its own tracer.
    Status nsSolve(hipStream_t st, int maxIter, int rule, double eps, int reset, int *iters, int *conv, int *status) override {
        const int n = (int) m.size();
        std::vector<int> act;
        std::vector<char> cur0(n, 1), cur(n, 1);
        for (int j = 0; j < n; ++j) {
            C *c = m[j];
            int it = 0, cv = 0;
            Status s = c->nsSolve(st, std::min(maxIter, 1), rule, eps, reset, &it, &cv);
            iters[j] = it; conv[j] = cv; status[j] = s.code;
            if (s.code || !c->p.flow_on || maxIter <= 1 || it < 1 || cv) continue;
            cur0[j] = cur[j] = c->gCurIsA ? 1 : 0;
            act.push_back(j);
        }
        if (act.empty()) return {CLB_OK, ""};
        const long long want = (long long) maxIter - 1;
        int na = (int) act.size();
        int nb = nsBuild(st, act, cur, rule, eps);
        nsReadStatus(st, na);
        std::vector<long long> base(n, 0);
        for (int k = 0; k < na; ++k) base[act[k]] = nsStatHost.p[4 * k];
        auto finish = [&](int j, long long eff, const long long *sv) {
            m[j]->gCurIsA = (eff & 1) ? !cur0[j] : (cur0[j] != 0);
            iters[j] = 1 + (int) eff;
            conv[j] = sv[1] ? 1 : 0;
            if (sv[2]) status[j] = CLB_E_NONFINITE;
        };
        long long ran = 0;
        std::vector<long long> last(4 * (size_t) n, 0);
        while (ran < want && na > 0) {
            const int run = (int) std::min<long long>(kFlowRun, want - ran);
            if (run == kFlowRun) {
                const std::vector<uintptr_t> key = {(uintptr_t) nsTab.p, (uintptr_t) nsMap.p, (uintptr_t) nb, (uintptr_t) na};
                nsGraph.ensure(st, key, [&] {
                    for (int k = 0; k < kFlowRun; ++k)
                        HIPT(launchNsStepGroup<RN>(nsTab.p + (size_t) (k & 1) * na, na, nsMap.p, nb, st));
                });
                nsGraph.launch(st);
            } else {
                for (int k = 0; k < run; ++k) HIPT(launchNsStepGroup<RN>(nsTab.p + (size_t) (k & 1) * na, na, nsMap.p, nb, st));
            }
            nsReadStatus(st, na);
            std::vector<int> still;
            for (int k = 0; k < na; ++k) {
                const int j = act[k];
                const long long *sv = nsStatHost.p + 4 * k;
                memcpy(&last[4 * (size_t) j], sv, 4 * sizeof(long long));
                const long long eff = sv[0] - base[j];
                if (!sv[3] && eff < ran + run) return {CLB_E_HIP, "flow iterations were lost"};
                if (sv[3]) finish(j, eff, sv);
                else still.push_back(j);
            }
            ran += run;
            if ((int) still.size() != na) {
                act.swap(still);
                na = (int) act.size();
                if (na > 0) {
                    // tables at each member's parity now; the count and the
                    // parity at the end still run from the start of the solve
                    for (int j : act) cur[j] = (ran & 1) ? !cur0[j] : cur0[j];
                    nb = nsBuild(st, act, cur, rule, eps);
                }
            }
        }
        for (int j : act) finish(j, ran, &last[4 * (size_t) j]);
        return {CLB_OK, ""};
    }

    // ---- transport
    static int sClass(int Sm) { return Sm <= 1 ? 1 : Sm <= 2 ? 2 : Sm <= 3 ? 3 : Sm <= 4 ? 4 : Sm <= 6 ? 6 : 8; }
    static bool stepsGroupable(const C *c) {
        return c->Sm > 0 && c->Nint > 0 && !c->symOn && !c->symAb && !c->rateOn && !c->p.eq_on && !c->p.dissol_on;
    }
    static bool stabGroupable(const C *c) { return c->Sm > 0 && c->Nint > 0; }
    // descriptors and map of group G for candidate ti: the coupled step
    // (kinetics, CA decision) or the stabilization step (transport alone)
    void adeBuild(hipStream_t st, const std::vector<int> &G, int ti, bool steps) {
        HIPT(hipStreamSynchronize(st));
        const int T = kTiles[ti], n = (int) G.size();
        adeHost[ti].ensure(2 * (size_t) n);
        adeTab[ti].ensure(2 * (size_t) n);
        Plan &P = plan[ti];
        P = Plan();
        P.members = n;
        size_t nInt = 0, nPln = 0;
        for (int k = 0; k < n; ++k) {
            C *c = m[G[k]];
            const bool ca = steps && c->p.n_microbes > 0;
            P.anyCa = P.anyCa || ca;
            for (int t = 0; t < 2; ++t) {
                const bool a = (t == 0) == c->fCurIsA;
                AdeMember<R> &e = adeHost[ti].p[(size_t) t * n + k];
                memset(&e, 0, sizeof(e));
                e.a = c->ade(ca ? &c->ctl->halt : nullptr);
                e.a.cur = a ? c->fA : c->fB;
                e.a.out = a ? c->fB : c->fA;
                e.a.bbCur = a ? c->fbbA : c->fbbB;
                e.a.bbOut = a ? c->fbbB : c->fbbA;
                e.a.first = 0;
                e.a.reset = 0;
                e.a.maskTotal = ca ? 1 : 0;
                e.k = c->kin(steps);
                e.k.dissol = 0;
                e.k.nearPhase = 0;
                if (!e.k.bio && e.k.nrxn == 0) e.k.on = 0;
                e.c = c->ca();
                e.gInt = (int) nblk(c->Nint, T);
                e.gPln = (int) nblk(c->Nopen - c->Nint, T);
                e.ca = ca ? 1 : 0;
                e.iT0 = c->iT;
                e.oldMask = c->oldTotMask;
                e.precipEvery = c->p.precip_on && c->p.precip_interval > 0 && c->p.precip_species >= 0 ? c->p.precip_interval : 0;
                e.dissolEvery = c->p.dissol_on && c->p.dissol_interval > 0 ? c->p.dissol_interval : 0;
            }
            const AdeMember<R> &e = adeHost[ti].p[k];
            if (e.gPln >= 65536) throw HipFail{hipErrorInvalidValue, "too many plane workgroups for the batch map"};
            nInt += (size_t) e.gInt;
            nPln += (size_t) e.gPln * (size_t) c->Sm;
        }
        if (nInt + nPln > (size_t) INT_MAX) throw HipFail{hipErrorInvalidValue, "batch transport grid too large"};
        mapHost[ti].ensure(nInt + nPln);
        mapDev[ti].ensure(nInt + nPln);
        int q = 0;
        const int classes[6] = {1, 2, 3, 4, 6, 8};
        for (int cl : classes) {
            const int off = q;
            for (int k = 0; k < n; ++k) {
                if (sClass(m[G[k]]->Sm) != cl) continue;
                const int g = adeHost[ti].p[k].gInt;
                for (int b = 0; b < g; ++b) mapHost[ti].p[q++] = make_int2(k, b);
            }
            if (q > off) P.bulk.push_back({cl, off, q - off});
        }
        P.plnOff = q;
        for (int k = 0; k < n; ++k) {
            const int g = adeHost[ti].p[k].gPln, sm = m[G[k]]->Sm;
            for (int ms = 0; ms < sm; ++ms)
                for (int b = 0; b < g; ++b) mapHost[ti].p[q++] = make_int2(k, b | (ms << 16));
        }
        P.plnCnt = q - P.plnOff;
        HIPT(hipMemcpyAsync(adeTab[ti].p, adeHost[ti].p, 2 * (size_t) n * sizeof(AdeMember<R>), hipMemcpyHostToDevice, st));
        HIPT(hipMemcpyAsync(mapDev[ti].p, mapHost[ti].p, (size_t) q * sizeof(int2), hipMemcpyHostToDevice, st));
    }
    // one step of every member at parity t with candidate ti
    void adeStep(hipStream_t st, int ti, int t) {
        const Plan &P = plan[ti];
        const AdeMember<R> *E = adeTab[ti].p + (size_t) t * P.members;
        const int T = kTiles[ti];
        for (const auto &b : P.bulk) HIPT(launchAdeStepGroup<R>(E, mapDev[ti].p + b.off, b.cnt, nullptr, 0, b.cls, T, st));
        if (P.plnCnt) HIPT(launchAdeStepGroup<R>(E, nullptr, 0, mapDev[ti].p + P.plnOff, P.plnCnt, 1, T, st));
        if (P.anyCa) HIPT(launchCaMaskFinalGroup<R>(E, P.members, st));
    }
    void adePair(hipStream_t st, int ti) {
        const Plan &P = plan[ti];
        std::vector<uintptr_t> key = {(uintptr_t) adeTab[ti].p, (uintptr_t) mapDev[ti].p, (uintptr_t) P.members,
                                      (uintptr_t) P.anyCa, (uintptr_t) P.plnOff, (uintptr_t) P.plnCnt};
        for (const auto &b : P.bulk) { key.push_back((uintptr_t) b.cls); key.push_back((uintptr_t) b.off); key.push_back((uintptr_t) b.cnt); }
        pair[ti].ensure(st, key, [&] { adeStep(st, ti, 0); adeStep(st, ti, 1); });
        pair[ti].launch(st);
    }
    // L steps of group G, pairs of steps replayed from a graph (each member
    // back at its parity after a pair).  The first run of a mode with room for
    // it captures the pair of every candidate workgroup size, replays each
    // once untimed and once timed, and keeps the fastest: every candidate
    // computes the same steps, so the pick changes only the time.  Returns
    // the candidate that ran the last steps.
    int adeRun(hipStream_t st, const std::vector<int> &G, int L, int mode) {
        const bool steps = mode == 0;
        const int ng = (int) G.size();
        int left = L;
        int ti = pick[mode] >= 0 ? pick[mode] : 0;
        if (pick[mode] < 0 && L >= 12) {
            if (!ev0) { HIPT(hipEventCreate(&ev0)); HIPT(hipEventCreate(&ev1)); }
            for (int c = 0; c < 3; ++c) adeBuild(st, G, c, steps);
            if (steps) HIPT(launchCtlResetGroup<R>(adeTab[0].p, ng, st));
            for (int c = 0; c < 3; ++c) adePair(st, c);
            for (int c = 0; c < 3; ++c) {
                HIPT(hipEventRecord(ev0, st));
                pair[c].launch(st);
                HIPT(hipEventRecord(ev1, st));
                HIPT(hipEventSynchronize(ev1));
                float ms = 0;
                HIPT(hipEventElapsedTime(&ms, ev0, ev1));
                tuneUs[mode][c] = 500.0 * ms;              // microseconds per step
            }
            left -= 12;
            ti = 0;
            for (int c = 1; c < 3; ++c) if (tuneUs[mode][c] < tuneUs[mode][ti]) ti = c;
            pick[mode] = ti;
        } else {
            adeBuild(st, G, ti, steps);
            if (steps) HIPT(launchCtlResetGroup<R>(adeTab[ti].p, ng, st));
        }
        for (int p = 0; p < left / 2; ++p) adePair(st, ti);
        if (left & 1) adeStep(st, ti, 0);
        return ti;
    }

    // Every member's stabilization as its own adeStabilize does it: a step
    // that needs the member's own path first (hand-over, reset, changed
    // classes), the rest together, then the reset mark.
    Status stabilize(hipStream_t st, int steps, int *status) override {
        const int nm = (int) m.size();
        std::vector<int> G, left(nm, 0);
        for (int j = 0; j < nm; ++j) {
            status[j] = CLB_OK;
            C *c = m[j];
            if (!stabGroupable(c) || steps <= 0) { c->adeStabilize(st, steps); continue; }
            int todo = steps;
            if (c->adeFirst || c->adeReset || c->adeClsDirty) { c->adeCollideStream(st, 1); --todo; }
            left[j] = todo;
            if (todo > 0) G.push_back(j);
            else c->adeReset = true;
        }
        while (!G.empty()) {
            int L = kSegStab;
            for (int j : G) L = std::min(L, left[j]);
            adeRun(st, G, L, 1);
            std::vector<int> next;
            for (int j : G) {
                C *c = m[j];
                if (L & 1) c->fCurIsA = !c->fCurIsA;
                left[j] -= L;
                if (left[j] > 0) next.push_back(j);
                else c->adeReset = true;
            }
            G.swap(next);
        }
        HIPT(hipStreamSynchronize(st));
        return {CLB_OK, ""};
    }

    // Every member's coupled steps as its own runSteps takes them.  Members
    // the group launches do not cover (equilibrium, dissolution, symbolic
    // laws, rate capture, no mobile species) run on their own path.  The
    // others run together in runs that end at the first step any member
    // needs the host for; within a run a CA member's own step decision stops
    // it on the device (its later launches return at once), and the host
    // then finishes that step on the member's path.
    Status runSteps(hipStream_t st, int n, int *done, int *status) override {
        const int nm = (int) m.size();
        std::vector<int> G;
        for (int j = 0; j < nm; ++j) {
            done[j] = 0;
            status[j] = CLB_OK;
            C *c = m[j];
            if (n <= 0) continue;
            if (!stepsGroupable(c) || c->clogged) {
                int d = 0;
                Status s = c->runSteps(st, n, &d);
                done[j] = d;
                status[j] = s.code;
                continue;
            }
            G.push_back(j);
        }
        while (!G.empty()) {
            std::vector<int> ready;
            for (int j : G) {
                C *c = m[j];
                if (c->adeFirst || c->adeReset || c->adeClsDirty) {
                    int d = 0;
                    Status s = c->runSteps(st, 1, &d);
                    done[j] += d;
                    if (s.code) { status[j] = s.code; continue; }
                }
                if (done[j] < n) ready.push_back(j);
            }
            G.swap(ready);
            if (G.empty()) break;
            int L = kSegSteps;
            for (int j : G) {
                C *c = m[j];
                L = std::min(L, n - done[j]);
                for (int d = 1; d <= L; ++d)
                    if (c->hostDue(c->iT + d - 1)) { L = d; break; }
            }
            const int ti = adeRun(st, G, L, 0);
            const int ng = (int) G.size();
            if (plan[ti].anyCa) {
                ctlDev.ensure(ng);
                ctlHost.ensure(ng);
                HIPT(launchCtlGatherGroup<R>(adeTab[ti].p, ng, ctlDev.p, st));
                HIPT(hipMemcpyAsync(ctlHost.p, ctlDev.p, (size_t) ng * sizeof(StepCtl), hipMemcpyDeviceToHost, st));
            }
            HIPT(hipStreamSynchronize(st));
            std::vector<int> next;
            for (int k = 0; k < ng; ++k) {
                const int j = G[k];
                C *c = m[j];
                Status s{CLB_OK, ""};
                if (c->p.n_microbes > 0) {
                    const StepCtl cs = ctlHost.p[k];
                    const int fin = cs.done;
                    const bool halt = cs.halt != 0;
                    const int ranK = fin + (halt ? 1 : 0);            // launches that did work
                    if (fin < 0 || ranK > L) return {CLB_E_HIP, "step batch bookkeeping failed"};
                    if (ranK & 1) c->fCurIsA = !c->fCurIsA;
                    c->iT += fin;
                    c->cnt[7] += fin;
                    done[j] += fin;
                    if (halt) {
                        s = c->stepTail(st, cs.gmax, cs.msum);
                        if (s.code == CLB_OK || s.code == CLB_CLOGGED) done[j] += 1;
                    }
                } else {
                    // no step of the run but its last can need the host
                    if (L & 1) c->fCurIsA = !c->fCurIsA;
                    c->iT += L - 1;
                    c->cnt[7] += L - 1;
                    done[j] += L - 1;
                    s = c->stepTail(st, 0.0, 0);
                    if (s.code == CLB_OK || s.code == CLB_CLOGGED) done[j] += 1;
                }
                if (s.code) { status[j] = s.code; continue; }
                if (done[j] < n) next.push_back(j);
            }
            G.swap(next);
        }
        return {CLB_OK, ""};
    }
};

}  // namespace

// ------------------------------------------------------------------ static scratch
// clb_species_totals_noalloc keeps its reduction partials and sums in this
// file-scope device array instead of the case's buffers: fixed size, one copy
// per GPU at that GPU's own address, held until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): clb_ctx_create looks each device's copy up
// once with hipGetSymbolAddress, with that device current, and keeps the
// pointer in kScratchDev.  Every call on one GPU shares the array.
constexpr size_t kScratchDoubles = CLB_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex kScratchMu;
static double *kScratchDev[kMaxDevices] = {};
// the device's copy of `scratch` (looked up on the first call for a device)
static hipError_t scratchOn(int device, double **out) {
    std::lock_guard<std::mutex> lk(kScratchMu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!kScratchDev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        kScratchDev[device] = static_cast<double *>(ptr);
    }
    *out = kScratchDev[device];
    return hipSuccess;
}

struct clb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::mutex mu;                              // guards the maps (cases may be created from several threads)
    std::map<int, std::unique_ptr<CaseBase>> cases;
    std::map<int, std::unique_ptr<BatchBase>> batches;
    int next = 1;
    int nextBatch = 1;
};

namespace {
int fail(clb_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}
CaseBase *getCase(clb_ctx *ctx, int id) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->cases.find(id);
    return it == ctx->cases.end() ? nullptr : it->second.get();
}
BatchBase *getBatch(clb_ctx *ctx, int id) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->batches.find(id);
    return it == ctx->batches.end() ? nullptr : it->second.get();
}

#define GUARD_BEGIN try {
#define GUARD_END                                                                              \
    }                                                                                          \
    catch (const HipFail &h) {                                                                 \
        return fail(ctx, h.e == hipErrorOutOfMemory ? CLB_E_NOMEM : CLB_E_HIP,                 \
                    h.what + ": " + hipGetErrorString(h.e));                                   \
    }                                                                                          \
    catch (const std::exception &ex) { return fail(ctx, CLB_E_HIP, ex.what()); }

#define CASE_OR_FAIL(var)                                                                      \
    if (!ctx) return CLB_E_ARG;                                                                \
    CaseBase *var = getCase(ctx, case_id);                                                     \
    if (!var) return fail(ctx, CLB_E_ARG, "unknown case id");                                  \
    HIPT(hipSetDevice(ctx->device));

int checkParams(clb_ctx *ctx, const clb_case_params *p, const int8_t *geometry) {
    if (!p || !geometry) return fail(ctx, CLB_E_ARG, "null parameters or geometry");
    Status blk = checkParamBlock(p);
    if (blk.code) return fail(ctx, blk.code, blk.msg);
    const long long N = (long long) p->nx * p->ny * p->nz;
    // a no-dynamics grain voxel must not touch open water through a face
    long long open = 0;
    for (long long i = 0; i < N; ++i) {
        int g = geometry[i];
        if (g < 0 || g > CLB_BIOFILM + CLB_MAX_MICROBES + 2) return fail(ctx, CLB_E_GEOMETRY, "unknown material code in geometry");
        if (g >= CLB_PORE) ++open;
    }
    if (open == 0) return fail(ctx, CLB_E_GEOMETRY, "the geometry has no open voxel");
    // (a dissolvable grain of solid material is meant to touch the water)
    const bool solidGrainPhase = grainPhaseOf(*p, CLB_SOLID) > 0;
    for (int z = 0; z < p->nz && !solidGrainPhase; ++z) for (int y = 0; y < p->ny; ++y) for (int x = 0; x < p->nx; ++x) {
        long long i = x + (long long) p->nx * (y + (long long) p->ny * z);
        if (geometry[i] != CLB_SOLID) continue;
        const int d[6][3] = {{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
        for (int k = 0; k < 6; ++k) {
            int X = x + d[k][0], Y = y + d[k][1], Z = z + d[k][2];
            if (X < 0 || Y < 0 || Z < 0 || X >= p->nx || Y >= p->ny || Z >= p->nz) continue;
            if (geometry[X + (long long) p->nx * (Y + (long long) p->ny * Z)] >= CLB_PORE)
                return fail(ctx, CLB_E_GEOMETRY, "a no-dynamics (0) voxel touches open water; grains need a bounce-back (1) shell");
        }
    }
    return CLB_OK;
}
}  // namespace

extern "C" {

int clb_abi_version(void) { return CLB_ABI_VERSION; }

int clb_ctx_create(int device, clb_ctx **out) {
    if (!out) return CLB_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CLB_E_HIP;
    if (device < 0 || device >= n) return CLB_E_ARG;
    clb_ctx *c = new clb_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) { delete c; return CLB_E_HIP; }
    double *scr = nullptr;                       // this device's static scratch, once per device
    if (scratchOn(device, &scr) != hipSuccess) { (void) hipStreamDestroy(c->stream); delete c; return CLB_E_HIP; }
    *out = c;
    return CLB_OK;
}
int clb_ctx_destroy(clb_ctx *ctx) {
    if (!ctx) return CLB_E_ARG;
    (void) hipSetDevice(ctx->device);
    ctx->batches.clear();
    ctx->cases.clear();
    if (ctx->stream) (void) hipStreamDestroy(ctx->stream);
    delete ctx;
    return CLB_OK;
}
const char *clb_last_error(const clb_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int clb_device_name(clb_ctx *ctx, char *dst, size_t n) {
    if (!ctx || !dst || n == 0) return CLB_E_ARG;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, ctx->device) != hipSuccess) return fail(ctx, CLB_E_HIP, "hipGetDeviceProperties");
    snprintf(dst, n, "%s (%s, %d CUs)", pr.name, pr.gcnArchName, pr.multiProcessorCount);
    return CLB_OK;
}
int clb_sync(clb_ctx *ctx) {
    if (!ctx) return CLB_E_ARG;
    return hipStreamSynchronize(ctx->stream) == hipSuccess ? CLB_OK : fail(ctx, CLB_E_HIP, "hipStreamSynchronize");
}

size_t clb_case_footprint(const clb_case_params *p) {
    if (!p) return 0;
    const size_t N = (size_t) p->nx * p->ny * p->nz;
    const size_t r = p->precision == CLB_FP64 ? 8 : 4;
    const size_t rn = p->precision == CLB_FP32 ? 4 : 8;
    int sm = 0, si = 0;
    for (int s = 0; s < p->n_species && s < CLB_MAX_SPECIES; ++s) (p->mobile[s] ? sm : si)++;
    size_t b = 0;
    const size_t NP = (size_t) popPad((long long) N);  // tiled population lattices
    b += N * 9;                                         // mask, geom, dist, phase, classes, reopened
    b += 2 * (size_t) std::max(sm, 1) * 7 * NP * r;     // ADE populations
    b += 2 * (size_t) std::max(sm, 1) * 6 * std::max(NP, (size_t) 64) * r;   // their bounce-back stores
    b += N * (sizeof(int) + 6 * sizeof(int) + sizeof(uint16_t));   // voxel list, D3Q7 table, pull kinds
    b += (size_t) std::max(si, 1) * N * r;
    b += (size_t) std::max(p->n_microbes, 1) * N * r + N * r;
    if (p->n_microbes > 0) b += (size_t) p->n_microbes * 6 * N * r + N;
    b += 3 * N * sizeof(int);                           // ages (two) and the biomass voxel list
    if (p->flow_on) b += 2 * 19 * NP * rn + 2 * 18 * NP * rn + N * (18 * sizeof(int) + sizeof(NsNbRecord) + sizeof(unsigned long long));
    if (p->dissol_on) b += (size_t) std::max(p->n_species, 1) * N * sizeof(double) + N * sizeof(int);
    b += 3 * N * r;
    b += N * sizeof(double);                            // staging
    b += ((N + 127) / 128) * (5 * sizeof(double) + sizeof(long long));
    b += ((N + 63) / 64 + 2) * (sizeof(double) + sizeof(long long));   // mask/total partials
    return b;
}

int clb_case_create(clb_ctx *ctx, const clb_case_params *p, const int8_t *geometry, const double *biomass0, int *case_id) {
    if (!ctx || !case_id) return CLB_E_ARG;
    int rc = checkParams(ctx, p, geometry);
    if (rc) return rc;
    GUARD_BEGIN
    HIPT(hipSetDevice(ctx->device));
    std::unique_ptr<CaseBase> c;
    if (p->precision == CLB_FP32) { auto *k = new Case<float>(); c.reset(k); k->create(*p, geometry, biomass0, ctx->stream); }
    else if (p->precision == CLB_FP32_ADE) { auto *k = new Case<float, double>(); c.reset(k); k->create(*p, geometry, biomass0, ctx->stream); }
    else { auto *k = new Case<double>(); c.reset(k); k->create(*p, geometry, biomass0, ctx->stream); }
    std::lock_guard<std::mutex> lk(ctx->mu);
    int id = ctx->next++;
    ctx->cases[id] = std::move(c);
    *case_id = id;
    return CLB_OK;
    GUARD_END
}
int clb_case_destroy(clb_ctx *ctx, int case_id) {
    if (!ctx) return CLB_E_ARG;
    std::unique_ptr<CaseBase> gone;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->cases.find(case_id);
        if (it == ctx->cases.end()) return fail(ctx, CLB_E_ARG, "unknown case id");
        if (it->second->inBatches > 0) return fail(ctx, CLB_E_STATE, "the case belongs to a batch: remove it first");
        gone = std::move(it->second);
        ctx->cases.erase(it);
    }
    (void) hipSetDevice(ctx->device);
    gone.reset();
    return CLB_OK;
}

int clb_case_get_params(clb_ctx *ctx, int case_id, clb_case_params *out) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    *out = c->p;
    return CLB_OK;
    GUARD_END
}
int clb_set_ade_relaxation(clb_ctx *ctx, int case_id, const double *omp, const double *omb, double dt) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!omp || !omb) return fail(ctx, CLB_E_ARG, "null relaxation arrays");
    c->setRelaxation(omp, omb, dt);
    return CLB_OK;
    GUARD_END
}

int clb_ns_init(clb_ctx *ctx, int case_id, double delta_p) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!c->p.flow_on) return fail(ctx, CLB_E_STATE, "case was created with flow_on = 0");
    c->nsInit(ctx->stream, delta_p);
    return CLB_OK;
    GUARD_END
}
int clb_ns_solve(clb_ctx *ctx, int case_id, int max_iter, int rule, double eps, int reset_tracer, int *iters, int *converged) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!c->p.flow_on) return fail(ctx, CLB_E_STATE, "case was created with flow_on = 0");
    if (max_iter < 0 || (rule != CLB_CONV_TRACER && rule != CLB_CONV_RELDELTA)) return fail(ctx, CLB_E_ARG, "bad solve arguments");
    Status s = c->nsSolve(ctx->stream, max_iter, rule, eps, reset_tracer, iters, converged);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_ns_couple(clb_ctx *ctx, int case_id) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->nsCouple(ctx->stream);
    HIPT(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
    GUARD_END
}
int clb_ns_stats(clb_ctx *ctx, int case_id, double out[4]) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    c->nsStats(ctx->stream, out);
    return CLB_OK;
    GUARD_END
}
int clb_ns_update_dynamics(clb_ctx *ctx, int case_id, double perm_ratio) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->nsUpdateDynamics(ctx->stream, perm_ratio);
    return CLB_OK;
    GUARD_END
}
int clb_ade_stabilize(clb_ctx *ctx, int case_id, int steps) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (steps < 0) return fail(ctx, CLB_E_ARG, "negative step count");
    c->adeStabilize(ctx->stream, steps);
    return CLB_OK;
    GUARD_END
}
int clb_run_steps(clb_ctx *ctx, int case_id, int nsteps, int *done) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (nsteps < 0) return fail(ctx, CLB_E_ARG, "negative step count");
    Status s = c->runSteps(ctx->stream, nsteps, done);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_ade_steps(clb_ctx *ctx, int case_id, int nsteps) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (nsteps < 0) return fail(ctx, CLB_E_ARG, "negative step count");
    c->adeSteps(ctx->stream, nsteps);
    return CLB_OK;
    GUARD_END
}
int clb_ade_collide_stream(clb_ctx *ctx, int case_id, int nsteps) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (nsteps < 0) return fail(ctx, CLB_E_ARG, "negative step count");
    c->adeCollideStream(ctx->stream, nsteps);
    return CLB_OK;
    GUARD_END
}
int clb_kinetics(clb_ctx *ctx, int case_id, int which) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (which < 0 || which > (CLB_KIN_BIOTIC | CLB_KIN_ABIOTIC)) return fail(ctx, CLB_E_ARG, "bad kinetics selector");
    c->kinetics(ctx->stream, which);
    HIPT(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
    GUARD_END
}
int clb_mask_total(clb_ctx *ctx, int case_id, double *max_total, int64_t *mask_sum) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (c->p.n_microbes <= 0) return fail(ctx, CLB_E_STATE, "case has no CA biomass");
    double g = 0; long long m = 0;
    c->maskTotalNow(ctx->stream, &g, &m);
    if (max_total) *max_total = g;
    if (mask_sum) *mask_sum = m;
    return CLB_OK;
    GUARD_END
}
int clb_ca_sweep(clb_ctx *ctx, int case_id) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (c->p.n_microbes <= 0) return fail(ctx, CLB_E_STATE, "case has no CA biomass");
    c->caSweep(ctx->stream);
    HIPT(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
    GUARD_END
}
int clb_age_update(clb_ctx *ctx, int case_id, int to_fixed_point) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    Status s = c->ageUpdate(ctx->stream, to_fixed_point != 0);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_update_solute_dynamics(clb_ctx *ctx, int case_id) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->soluteDynamics(ctx->stream);
    HIPT(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
    GUARD_END
}
int clb_update_solute_solid(clb_ctx *ctx, int case_id, int *reopened) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->soluteSolid(ctx->stream, reopened);
    return CLB_OK;
    GUARD_END
}
int clb_dissol_convert(clb_ctx *ctx, int case_id, int *reopened) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!c->p.dissol_on) return fail(ctx, CLB_E_STATE, "case was created with dissol_on = 0");
    Status s = c->dissolConvert(ctx->stream, reopened);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_precip_convert(clb_ctx *ctx, int case_id, int *changed) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!c->p.precip_on) return fail(ctx, CLB_E_STATE, "case was created with precip_on = 0");
    c->precipConvert(ctx->stream, changed);
    return CLB_OK;
    GUARD_END
}
int clb_equilibrium(clb_ctx *ctx, int case_id) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->equilibrium(ctx->stream);
    HIPT(hipStreamSynchronize(ctx->stream));
    return CLB_OK;
    GUARD_END
}
int clb_ca_spill(clb_ctx *ctx, int case_id, int *sweeps) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    Status s = c->caBlock(ctx->stream, sweeps);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_read_field(clb_ctx *ctx, int case_id, int field, int index, double *dst) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!dst) return fail(ctx, CLB_E_ARG, "null destination");
    Status s = c->readField(ctx->stream, field, index, dst);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_write_field(clb_ctx *ctx, int case_id, int field, int index, const double *src) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!src) return fail(ctx, CLB_E_ARG, "null source");
    Status s = c->writeField(ctx->stream, field, index, src);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_set_tuning(clb_ctx *ctx, int case_id, int knob, int value) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    Status r = c->setTuning(knob, value);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_species_totals(clb_ctx *ctx, int case_id, double *out) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    c->speciesTotals(ctx->stream, out);
    return CLB_OK;
    GUARD_END
}
int clb_species_totals_noalloc(clb_ctx *ctx, int case_id, double *out) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    double *scr = nullptr;
    HIPT(scratchOn(ctx->device, &scr));
    Status r = c->speciesTotalsScratch(ctx->stream, scr, kScratchDoubles, out);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_field_stats(clb_ctx *ctx, int case_id, double *out) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    c->fieldStats(ctx->stream, out);
    return CLB_OK;
    GUARD_END
}
int clb_counters(clb_ctx *ctx, int case_id, int64_t out[CLB_NCOUNTERS]) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!out) return fail(ctx, CLB_E_ARG, "null output");
    c->counters(ctx->stream, out);
    return CLB_OK;
    GUARD_END
}
int clb_stream_copy(clb_ctx *ctx, size_t bytes, int reps, double *gbps, double *msOut) {
    if (!ctx || bytes < 16 || reps < 1) return CLB_E_ARG;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CLB_OK;
    try {
        HIPT(hipSetDevice(ctx->device));
        const long long n16 = (long long) (bytes / 16);
        HIPT(hipMalloc(&a, (size_t) n16 * 16));
        HIPT(hipMalloc(&b, (size_t) n16 * 16));
        HIPT(hipMemsetAsync(a, 1, (size_t) n16 * 16, ctx->stream));
        HIPT(hipEventCreate(&e0));
        HIPT(hipEventCreate(&e1));
        HIPT(launchCopy16(a, b, n16, ctx->stream));
        double best = 1e300;
        for (int r = 0; r < reps; ++r) {
            HIPT(hipEventRecord(e0, ctx->stream));
            HIPT(launchCopy16(a, b, n16, ctx->stream));
            HIPT(hipEventRecord(e1, ctx->stream));
            HIPT(hipEventSynchronize(e1));
            float ms = 0;
            HIPT(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, (double) ms);
        }
        if (gbps) *gbps = 2.0 * (double) n16 * 16 / (best * 1e-3) / 1e9;
        if (msOut) *msOut = best;
    } catch (const HipFail &h) {
        rc = fail(ctx, CLB_E_HIP, h.what + ": " + hipGetErrorString(h.e));
    }
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    if (a) (void) hipFree(a);
    if (b) (void) hipFree(b);
    return rc;
}

int clb_set_symbolic(clb_ctx *ctx, int case_id, int microbe, const clb_sym_program *prog) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    Status r = c->setSymbolic(microbe, prog);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_checkpoint_size(clb_ctx *ctx, int case_id, size_t *bytes) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!bytes) return fail(ctx, CLB_E_ARG, "null output");
    *bytes = c->checkpointBytes();
    return CLB_OK;
    GUARD_END
}
int clb_checkpoint_save(clb_ctx *ctx, int case_id, void *dst, size_t bytes) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!dst) return fail(ctx, CLB_E_ARG, "null buffer");
    Status r = c->checkpointSave(ctx->stream, (uint8_t *) dst, bytes);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_checkpoint_load(clb_ctx *ctx, int case_id, const void *src, size_t bytes) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (!src) return fail(ctx, CLB_E_ARG, "null buffer");
    Status r = c->checkpointLoad(ctx->stream, (const uint8_t *) src, bytes);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_rate_capture(clb_ctx *ctx, int case_id, int on) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    Status r = c->rateCapture(ctx->stream, on != 0);
    if (r.code) return fail(ctx, r.code, r.msg);
    return CLB_OK;
    GUARD_END
}
int clb_set_timing_stride(clb_ctx *ctx, int case_id, int stride) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    if (stride < 1) return fail(ctx, CLB_E_ARG, "timing stride below 1");
    c->setTimeStride(stride);
    return CLB_OK;
    GUARD_END
}
int clb_kernel_time(clb_ctx *ctx, int case_id, int reset, double *avg_ms, int64_t *launches) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->kernelTime(ctx->stream, reset, avg_ms, launches);
    return CLB_OK;
    GUARD_END
}
int clb_flow_kernel_time(clb_ctx *ctx, int case_id, int reset, double *avg_ms, int64_t *launches) {
    GUARD_BEGIN
    CASE_OR_FAIL(c)
    c->flowKernelTime(reset, avg_ms, launches);
    return CLB_OK;
    GUARD_END
}

// ---- the batched form
#define BATCH_OR_FAIL(var)                                                                     \
    if (!ctx) return CLB_E_ARG;                                                                \
    BatchBase *var = getBatch(ctx, batch_id);                                                  \
    if (!var) return fail(ctx, CLB_E_ARG, "unknown batch id");                                 \
    HIPT(hipSetDevice(ctx->device));

int clb_batch_create(clb_ctx *ctx, int precision, const int *case_ids, int n, int *batch_id) {
    if (!ctx || !batch_id || n < 0 || (n > 0 && !case_ids)) return CLB_E_ARG;
    GUARD_BEGIN
    std::unique_ptr<BatchBase> b;
    if (precision == CLB_FP64) b.reset(new Batch<double, double>());
    else if (precision == CLB_FP32) b.reset(new Batch<float, float>());
    else if (precision == CLB_FP32_ADE) b.reset(new Batch<float, double>());
    else return fail(ctx, CLB_E_ARG, "unknown precision");
    for (int k = 0; k < n; ++k) {
        CaseBase *c = getCase(ctx, case_ids[k]);
        if (!c) return fail(ctx, CLB_E_ARG, "unknown case id");
        if (!b->accepts(c)) return fail(ctx, CLB_E_ARG, "a batch holds cases of its own precision only");
        for (int q = 0; q < k; ++q)
            if (case_ids[q] == case_ids[k]) return fail(ctx, CLB_E_ARG, "a case appears twice in the batch");
    }
    for (int k = 0; k < n; ++k) { b->attach(getCase(ctx, case_ids[k])); b->ids.push_back(case_ids[k]); }
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int id = ctx->nextBatch++;
    ctx->batches[id] = std::move(b);
    *batch_id = id;
    return CLB_OK;
    GUARD_END
}
int clb_batch_destroy(clb_ctx *ctx, int batch_id) {
    if (!ctx) return CLB_E_ARG;
    std::unique_ptr<BatchBase> gone;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->batches.find(batch_id);
        if (it == ctx->batches.end()) return fail(ctx, CLB_E_ARG, "unknown batch id");
        gone = std::move(it->second);
        ctx->batches.erase(it);
    }
    (void) hipSetDevice(ctx->device);
    gone.reset();
    return CLB_OK;
}
int clb_batch_add(clb_ctx *ctx, int batch_id, int case_id) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    CaseBase *c = getCase(ctx, case_id);
    if (!c) return fail(ctx, CLB_E_ARG, "unknown case id");
    if (!b->accepts(c)) return fail(ctx, CLB_E_ARG, "a batch holds cases of its own precision only");
    for (int id : b->ids) if (id == case_id) return fail(ctx, CLB_E_ARG, "the case is in the batch already");
    b->attach(c);
    b->ids.push_back(case_id);
    return CLB_OK;
    GUARD_END
}
int clb_batch_remove(clb_ctx *ctx, int batch_id, int case_id) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    for (size_t k = 0; k < b->ids.size(); ++k)
        if (b->ids[k] == case_id) { b->detach(k); return CLB_OK; }
    return fail(ctx, CLB_E_ARG, "the case is not in the batch");
    GUARD_END
}
int clb_batch_members(clb_ctx *ctx, int batch_id, int *n, int *case_ids, int cap) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (!n) return fail(ctx, CLB_E_ARG, "null output");
    *n = (int) b->ids.size();
    if (case_ids) for (int k = 0; k < *n && k < cap; ++k) case_ids[k] = b->ids[k];
    return CLB_OK;
    GUARD_END
}
int clb_batch_ns_solve(clb_ctx *ctx, int batch_id, int max_iter, int rule, double eps, int reset_tracer, int *iters,
                       int *converged, int *status) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (!iters || !converged || !status) return fail(ctx, CLB_E_ARG, "null output arrays");
    if (rule != CLB_CONV_TRACER && rule != CLB_CONV_RELDELTA) return fail(ctx, CLB_E_ARG, "unknown convergence rule");
    Status s = b->nsSolve(ctx->stream, max_iter, rule, eps, reset_tracer, iters, converged, status);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_batch_ade_stabilize(clb_ctx *ctx, int batch_id, int steps, int *status) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (!status) return fail(ctx, CLB_E_ARG, "null output array");
    Status s = b->stabilize(ctx->stream, steps, status);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_batch_run_steps(clb_ctx *ctx, int batch_id, int nsteps, int *done, int *status) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (!done || !status) return fail(ctx, CLB_E_ARG, "null output arrays");
    Status s = b->runSteps(ctx->stream, nsteps, done, status);
    return s.code ? fail(ctx, s.code, s.msg) : CLB_OK;
    GUARD_END
}
int clb_batch_set_tile(clb_ctx *ctx, int batch_id, int mode, int tile) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (mode != CLB_BATCH_STEPS && mode != CLB_BATCH_STABILIZE) return fail(ctx, CLB_E_ARG, "unknown batch mode");
    if (tile == 0) { b->pick[mode] = -1; return CLB_OK; }
    for (int c = 0; c < 3; ++c)
        if (BatchBase::kTiles[c] == tile) { b->pick[mode] = c; return CLB_OK; }
    return fail(ctx, CLB_E_ARG, "workgroup size must be 0 (tune), 64, 128 or 256");
    GUARD_END
}
int clb_batch_tuning(clb_ctx *ctx, int batch_id, int mode, int *tile, double us_per_step[3]) {
    GUARD_BEGIN
    BATCH_OR_FAIL(b)
    if (mode != CLB_BATCH_STEPS && mode != CLB_BATCH_STABILIZE) return fail(ctx, CLB_E_ARG, "unknown batch mode");
    if (tile) *tile = b->pick[mode] >= 0 ? BatchBase::kTiles[b->pick[mode]] : 0;
    if (us_per_step) for (int c = 0; c < 3; ++c) us_per_step[c] = b->tuneUs[mode][c];
    return CLB_OK;
    GUARD_END
}

}  // extern "C"

// clb_common.h -- lattice constants and index helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>

namespace clb {

// D3Q7: rest, +x, -x, +y, -y, +z, -z; weights 1/4, 1/8 (cs^2 = 1/4)
static __device__ __constant__ int D7C[7][3] = {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1}};
// D3Q19: rest, 6 faces, 12 edges, opposite pairs adjacent; weights 1/3, 1/18, 1/36
static __device__ __constant__ int D19C[19][3] = {
    {0,0,0},
    {1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
    {1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0},
    {1,0,1},{-1,0,-1},{1,0,-1},{-1,0,1},
    {0,1,1},{0,-1,-1},{0,1,-1},{0,-1,1}};

#define W7R 0.25
#define W7F 0.125
__device__ __forceinline__ double w7(int q) { return q == 0 ? W7R : W7F; }
// NaN-propagating max (a NaN total must reach the host's finiteness check)
__device__ __forceinline__ double nanMax(double a, double b) {
    return (isnan(a) || isnan(b)) ? __builtin_nan("") : fmax(a, b);
}
// a queued step batch was halted by an earlier step: the kernel does nothing
// (the flag is written by an earlier kernel of the stream, never during this one)
__device__ __forceinline__ bool halted(const int *h) {
    return h && *h;
}
__device__ __forceinline__ int opp7(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

// a voxel's first slot in the transport bounce-back store follows the closed
// links of its tile's earlier voxels: counted here from the kinds (any lane
// order, for the kernels that change populations between launches; the
// transport kernel takes a wave prefix sum instead)
// (the kinds array is padded to whole tiles: the tile's 64 entries are one
// 128-byte line, read as eight independent 16-byte loads)
__device__ inline long long adeBBSlotOf(const uint16_t *kinds, int i) {
    const uint4 *kv = reinterpret_cast<const uint4 *>(kinds + (i & ~63));
    const int lane = i & 63;
    uint4 w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = kv[k];
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const unsigned int h[4] = {w[k].x, w[k].y, w[k].z, w[k].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const unsigned int kd = (h[e >> 1] >> (16 * (e & 1))) & 0xffffu;
            if (8 * k + e < lane) c += __popc(adeKind2Mask(kd));
        }
    }
    return (long long) (i >> 6) * (64 * 6) + c;
}
// the same as a prefix sum over the wave (every lane of the wave takes part;
// the wave's voxels are the 64 of one tile)
__device__ __forceinline__ long long adeBBBase(int i, unsigned int kd, bool valid) {
    const int c = valid ? __popc(adeKind2Mask(kd)) : 0;
    int x = c;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return (long long) (i >> 6) * (64 * 6) + (x - c);
}
// voxel i's new own populations f of mobile slot ms into the store paired with
// cur, for its closed links (first slot j0)
template <typename R, typename T>
__device__ inline void adeBBPut(const AdeDev<R> &a, unsigned int kd, long long j0, int ms, const T f[7]) {
    if (!adeKind2Mask(kd)) return;
    R *b = a.bbCur + (long long) ms * a.bbStride;
    long long j = j0;
#pragma unroll
    for (int q = 1; q < 7; ++q)
        if (((kd >> (2 * (q - 1))) & 3u) == 2u) b[j++] = (R) f[opp7(q)];
}

__device__ __forceinline__ int opp19(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__device__ __forceinline__ double w19(int q) { return q == 0 ? (1.0 / 3.0) : (q < 7 ? (1.0 / 18.0) : (1.0 / 36.0)); }

__device__ __forceinline__ void coords(long long idx, int nx, int ny, int &x, int &y, int &z) {
    const long long t = idx / nx;
    x = (int) (idx - t * nx);
    z = (int) (t / ny);
    y = (int) (t - (long long) z * ny);
}

}  // namespace clb

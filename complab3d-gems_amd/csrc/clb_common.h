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
__device__ __forceinline__ int opp19(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
__device__ __forceinline__ double w19(int q) { return q == 0 ? (1.0 / 3.0) : (q < 7 ? (1.0 / 18.0) : (1.0 / 36.0)); }

__device__ __forceinline__ void coords(long long idx, int nx, int ny, int &x, int &y, int &z) {
    const long long t = idx / nx;
    x = (int) (idx - t * nx);
    z = (int) (t / ny);
    y = (int) (t - (long long) z * ny);
}

}  // namespace clb

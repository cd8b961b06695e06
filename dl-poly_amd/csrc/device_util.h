// Small device helpers shared by the kernels: wavefront/block reductions and
// minimum-image arithmetic for orthorhombic cells.
#pragma once
#include <hip/hip_runtime.h>

namespace dlpx {

// Sum across the 64 lanes of a wavefront.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum across groups of G consecutive lanes (G a power of two <= 64); every
// lane of the group receives the group total.
template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of NV values per thread; the totals land in out[0..NV) of
// thread 0.  lds must hold NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds, double *out)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[k * nw + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < nw; w++) s += lds[threadIdx.x * nw + w];
        out[threadIdx.x] = s;
    }
}

// floor(x / d) and the remainder x - d floor(x / d) in [0, d), for |x| < 2^20,
// from a float reciprocal inv = 1/d: the estimate is within one of the
// truth and one correction step fixes it (integer division by a runtime
// value costs tens of instructions on this hardware).
__device__ __forceinline__ int fdiv_floor(int x, int d, float inv, int &r)
{
    int q = (int)floorf((float)x * inv);
    r = x - q * d;
    if (r < 0) {
        r += d;
        q--;
    } else if (r >= d) {
        r -= d;
        q++;
    }
    return q;
}

// r_ij = x_i - (x_j + s L) with s = rint((x_i - x_j)/L) per axis.
__device__ __forceinline__ double min_image_1(double xi, double xj, double L, double invL)
{
    const double s = rint((xi - xj) * invL);
    return xi - (xj + s * L);
}

}  // namespace dlpx

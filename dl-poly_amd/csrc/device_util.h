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

// Block-wide sum of NV values per thread; total k lands in out[k], written
// by thread k, which also returns it (other threads return 0).  lds must hold
// NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ double block_sum(double (&v)[NV], double *lds, double *out)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[k * nw + wid] = s;
    }
    __syncthreads();
    double total = 0.0;
    if (threadIdx.x < NV) {
        for (int w = 0; w < nw; w++) total += lds[threadIdx.x * nw + w];
        out[threadIdx.x] = total;
    }
    return total;
}

// Wait until this thread's memory operations, the fire-and-forget atomics
// included, have completed, so that an atomic issued after it (a workgroup
// counter) cannot overtake them.  Cross-workgroup results here travel only
// through device-scope atomics, which need no cache writeback; an
// agent-scope fence would write back the XCD's L2 in every workgroup.
__device__ __forceinline__ void wait_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Coherent read of a value other workgroups of the same launch built with
// device-scope atomics.
template <typename T>
__device__ __forceinline__ T load_coherent(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of column col over rows 0..rows-1 of a row-major buffer of partial
// sums (stride doubles per row), in a fixed order, by a whole workgroup of NT
// threads (a power of two); lds holds NT doubles.  Every thread gets the
// total.
template <int NT>
__device__ __forceinline__ double column_sum(const double *red, int rows, int stride, int col, double *lds)
{
    double s = 0.0;
    for (int r = threadIdx.x; r < rows; r += NT) s += red[(size_t)r * stride + col];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
        __syncthreads();
    }
    const double t = lds[0];
    __syncthreads();
    return t;
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

// Velocity-Verlet NVE integration on the device (nve.F90:127-215) and the
// host-order <-> cell-order transfers of per-atom data.
#include <cmath>
#include <cstring>

#include "common.h"
#include "device_util.h"

namespace dlpx {

// Stage 1: v += dt/2 F/m ; x += dt v  (atoms of mass <= 1e-6 stay frozen),
// fused with the displacement test of vnl_check (neighbours.F90:150-178):
// the largest minimum-image |x - x_ref|^2 goes to *dmax (as the bit pattern
// of a non-negative double), and the last workgroup to finish copies it to
// host memory and resets *dmax and *done for the next step.
struct Box {
    double L[3], invL[3];
};
// 1024-thread workgroups: the maximum is reduced in LDS first, so only one
// atomic per workgroup (63 at 64k atoms) meets the single device word.
constexpr int kVv1Threads = 1024;
__global__ void __launch_bounds__(kVv1Threads) k_vv1(double4 *pos, double4 *vel, const double4 *frc, int n,
                                                     double dt, const double4 *__restrict__ xref, Box b,
                                                     unsigned long long *dmax, unsigned int *done,
                                                     unsigned long long *host_dmax)
{
    __shared__ double wmax[kVv1Threads / 64];
    __shared__ bool last;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (i < n) {
        double4 v = vel[i];
        double4 p = pos[i];
        if (v.w > 1.0e-6) {
            const double4 f = frc[i];
            const double tmp = (0.5 * dt) / v.w;
            v.x += tmp * f.x;
            v.y += tmp * f.y;
            v.z += tmp * f.z;
            p.x += dt * v.x;
            p.y += dt * v.y;
            p.z += dt * v.z;
            vel[i] = v;
            pos[i] = p;
        }
        const double4 q = xref[i];
        const double dx = min_image_1(p.x, q.x, b.L[0], b.invL[0]);
        const double dy = min_image_1(p.y, q.y, b.L[1], b.invL[1]);
        const double dz = min_image_1(p.z, q.z, b.L[2], b.invL[2]);
        r2 = dx * dx + dy * dy + dz * dz;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = r2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wmax[w]);
        atomicMax(dmax, (unsigned long long)__double_as_longlong(m));
        __threadfence();
        last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        *host_dmax = atomicAdd(dmax, 0ull);
        __threadfence_system();
        *dmax = 0ull;
        *done = 0u;
    }
}

// Stage 2: v += dt/2 F/m, and sum m v^2 per block.  With FIN the
// reciprocal forces are first added to the total force as k_finish does:
// frc += 2 scale (frc_recip - mean frc_recip) (ewald_general.F90:857-864).
// Frozen atoms (frz[gid[i]] /= 0; frz may be null) get zero force and
// velocity, as freeze_atoms does after each force evaluation
// (configuration.F90:3240-3258), so they neither move nor add kinetic energy.
template <bool FIN>
__global__ void __launch_bounds__(256) k_vv2(double4 *vel, double4 *frc, int n, double dt, double *red,
                                             const double4 *__restrict__ frc_recip, double two_scale,
                                             const double *__restrict__ red_out, double inv_n,
                                             const int32_t *__restrict__ gid, const int32_t *__restrict__ frz)
{
    __shared__ double lred[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (i < n) {
        double4 v = vel[i];
        double4 f = frc[i];
        if (FIN) {
            const double mx = red_out[R_FNETX] * inv_n, my = red_out[R_FNETY] * inv_n, mz = red_out[R_FNETZ] * inv_n;
            const double4 r = frc_recip[i];
            f.x += (r.x - mx) * two_scale;
            f.y += (r.y - my) * two_scale;
            f.z += (r.z - mz) * two_scale;
            frc[i] = f;
        }
        if (frz && frz[gid[i]] != 0) {
            f.x = f.y = f.z = 0.0;
            v.x = v.y = v.z = 0.0;
            frc[i] = f;
            vel[i] = v;
        }
        if (v.w > 1.0e-6) {
            const double tmp = (0.5 * dt) / v.w;
            v.x += tmp * f.x;
            v.y += tmp * f.y;
            v.z += tmp * f.z;
            vel[i] = v;
            acc[0] = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
        }
    }
    block_sum<1>(acc, lred, red + (size_t)blockIdx.x * kReduceSlots + R_KIN);
}

// freeze_atoms (configuration.F90:3240-3258): zero force and velocity of
// frozen atoms (frz[gid[i]] /= 0)
__global__ void k_freeze(double4 *vel, double4 *frc, int n, const int32_t *__restrict__ gid,
                         const int32_t *__restrict__ frz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || frz[gid[i]] == 0) return;
    double4 v = vel[i], f = frc[i];
    v.x = v.y = v.z = 0.0;
    f.x = f.y = f.z = 0.0;
    vel[i] = v;
    frc[i] = f;
}

int freeze_atoms(Ctx *c)
{
    if (c->nfrozen == 0) return 0;
    hipLaunchKernelGGL(k_freeze, dim3(ceil_div(c->natms, 256)), dim3(256), 0, c->stream, c->vel, c->frc, c->natms,
                       c->gid, c->frz);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

// host-order staging (8 doubles per atom: x y z q vx vy vz m) -> cell order
__global__ void k_load(const double *__restrict__ st, const int32_t *__restrict__ st_t, int n, int cap,
                       const int32_t *__restrict__ slot, bool identity, bool with_vel, double4 *pos, int32_t *typ,
                       int32_t *gid, double4 *vel)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n) return;
    const int s = identity ? h : slot[h];
    const double *r = st + (size_t)8 * h;
    pos[s] = make_double4(r[0], r[1], r[2], r[3]);
    typ[s] = st_t[h] - 1;
    if (identity) gid[s] = h;
    if (with_vel) vel[s] = make_double4(r[4], r[5], r[6], r[7]);
}

// cell order -> host-order staging: x y z | vx vy vz | fx fy fz
__global__ void k_store(double *__restrict__ st, int n, const int32_t *__restrict__ gid, const double4 *pos,
                        const double4 *vel, const double4 *frc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int h = gid[s];
    double *r = st + (size_t)9 * h;
    if (pos) {
        const double4 p = pos[s];
        r[0] = p.x; r[1] = p.y; r[2] = p.z;
    }
    if (vel) {
        const double4 v = vel[s];
        r[3] = v.x; r[4] = v.y; r[5] = v.z;
    }
    const double4 f = frc[s];
    r[6] = f.x; r[7] = f.y; r[8] = f.z;
}

int nve_first(Ctx *c, double dt)
{
    Box b;
    for (int d = 0; d < 3; d++) {
        b.L[d] = c->L[d];
        b.invL[d] = c->invL[d];
    }
    // device words flags[4..5] (maximum) and flags[6] (workgroup count) are
    // zero between launches; the result lands in host word h_flags[8..9]
    hipLaunchKernelGGL(k_vv1, dim3(ceil_div(c->natms, kVv1Threads)), dim3(kVv1Threads), 0, c->stream, c->pos,
                       c->vel, c->frc,
                       c->natms, dt, c->xref, b, (unsigned long long *)(c->flags + 4), (unsigned int *)(c->flags + 6),
                       (unsigned long long *)(c->h_flags_dev + 8));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

double nve_first_displacement(Ctx *c)
{
    double r2;
    memcpy(&r2, c->h_flags + 8, sizeof(r2));
    return std::sqrt(r2);
}

int nve_second(Ctx *c, double dt, bool /*kinetic*/)
{
    const int nb = ceil_div(c->natms, 256);
    int r;
    if ((r = ensure_red_rows(c, nb))) return r;
    const int32_t *frz = c->nfrozen > 0 ? c->frz : nullptr;
    if (c->finish_in_vv2)
        hipLaunchKernelGGL(k_vv2<true>, dim3(nb), dim3(256), 0, c->stream, c->vel, c->frc, c->natms, dt, c->red,
                           c->frc_recip, 2.0 * recip_scale(c), c->red_out, 1.0 / (double)c->natms, c->gid, frz);
    else
        hipLaunchKernelGGL(k_vv2<false>, dim3(nb), dim3(256), 0, c->stream, c->vel, c->frc, c->natms, dt, c->red,
                           nullptr, 0.0, nullptr, 0.0, c->gid, frz);
    c->finish_in_vv2 = false;
    c->red_rows[R_KIN] = nb;
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int load_atoms(Ctx *c, const double *st_dev, const int32_t *st_t_dev, bool identity, bool with_vel)
{
    hipLaunchKernelGGL(k_load, dim3(ceil_div(c->natms, 256)), dim3(256), 0, c->stream, st_dev, st_t_dev, c->natms,
                       c->cap, c->slot, identity, with_vel, c->pos, c->typ, c->gid, c->vel);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int store_atoms(Ctx *c, double *st_dev, bool with_pos, bool with_vel)
{
    hipLaunchKernelGGL(k_store, dim3(ceil_div(c->natms, 256)), dim3(256), 0, c->stream, st_dev, c->natms, c->gid,
                       with_pos ? c->pos : nullptr, with_vel ? c->vel : nullptr, c->frc);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

}  // namespace dlpx

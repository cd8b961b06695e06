// Velocity-Verlet NVE integration on the device (nve.F90:127-215) and the
// host-order <-> cell-order transfers of per-atom data.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace dlpx {

// Stage 1: v += dt/2 F/m ; x += dt v  (atoms of mass <= 1e-6 stay frozen),
// fused with the displacement test of vnl_check (neighbours.F90:150-178):
// the largest minimum-image |x - x_ref|^2 goes to *dmax (as the bit pattern
// of a non-negative double).  The last workgroup to finish resets *dmax and
// *done for the next step and publishes the result: to host memory
// (host_dmax, the host then decides), and/or as the device's own rebuild
// decision (upd: 1 when max |dx| >= Nearest(0.5,-1) padding, neighbours.F90:
// 157-182, or always with no padding), which the list-build kernels read;
// a decided rebuild also clears the build's counters and counts itself.
struct Box {
    double L[3], invL[3];
};
struct Vv1Args {
    double4 *pos, *vel;
    const double4 *frc, *xref;
    int n;
    double dt;
    Box b;
    unsigned long long *dmax;
    unsigned int *done;
    unsigned long long *host_dmax;  // may be null
    int32_t *flags;                 // F_UPD and the build counters; null: no device decision
    double half_pad;                // Nearest(0.5,-1) * padding (0: rebuild every step)
};
// 1024-thread workgroups: the maximum is reduced in LDS first, so only one
// atomic per workgroup (63 at 64k atoms) meets the single device word.
constexpr int kVv1Threads = 1024;
__device__ __forceinline__ void vv1_body(const Vv1Args &a, int bid, int nb)
{
    __shared__ double wmax[kVv1Threads / 64];
    __shared__ bool last;
    const int i = bid * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (i < a.n) {
        double4 v = a.vel[i];
        double4 p = a.pos[i];
        if (v.w > 1.0e-6) {
            const double4 f = a.frc[i];
            const double tmp = (0.5 * a.dt) / v.w;
            v.x += tmp * f.x;
            v.y += tmp * f.y;
            v.z += tmp * f.z;
            p.x += a.dt * v.x;
            p.y += a.dt * v.y;
            p.z += a.dt * v.z;
            a.vel[i] = v;
            a.pos[i] = p;
        }
        const double4 q = a.xref[i];
        const double dx = min_image_1(p.x, q.x, a.b.L[0], a.b.invL[0]);
        const double dy = min_image_1(p.y, q.y, a.b.L[1], a.b.invL[1]);
        const double dz = min_image_1(p.z, q.z, a.b.L[2], a.b.invL[2]);
        r2 = dx * dx + dy * dy + dz * dz;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = r2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wmax[w]);
        atomicMax(a.dmax, (unsigned long long)__double_as_longlong(m));
        __threadfence();
        last = atomicAdd(a.done, 1u) == (unsigned)nb - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        const unsigned long long bits = atomicAdd(a.dmax, 0ull);
        if (a.host_dmax) {
            *a.host_dmax = bits;
            __threadfence_system();
        }
        if (a.flags) {
            const int upd = (a.half_pad <= 0.0 || sqrt(__longlong_as_double((long long)bits)) >= a.half_pad) ? 1 : 0;
            a.flags[F_UPD] = upd;
            if (upd) {
                a.flags[F_OVF] = 0;
                a.flags[F_IBIG] = 0;
                a.flags[F_PAIRS] = 0;
                a.flags[F_NREB] += 1;
            }
        }
        *a.dmax = 0ull;
        *a.done = 0u;
    }
}
__global__ void __launch_bounds__(kVv1Threads) k_vv1(Vv1Args a) { vv1_body(a, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(kVv1Threads) k_vv1_batch(const Vv1Args *__restrict__ av,
                                                           const int32_t *__restrict__ boff, int nsys)
{
    const int lb = blockIdx.x;  // stage 1 reads atoms in order: no XCD renumbering
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    vv1_body(av[s], lb - boff[s], boff[s + 1] - boff[s]);
}

// Stage 2: v += dt/2 F/m, and sum m v^2 per block.  With fin the
// reciprocal forces are first added to the total force as k_finish does:
// frc += 2 scale (frc_recip - mean frc_recip) (ewald_general.F90:857-864).
// Frozen atoms (frz[gid[i]] /= 0; frz may be null) get zero force and
// velocity, as freeze_atoms does after each force evaluation
// (configuration.F90:3240-3258), so they neither move nor add kinetic energy.
struct Vv2Args {
    double4 *vel, *frc;
    int n;
    double dt;
    double *red;
    const double4 *frc_recip;  // fin only
    double two_scale;
    const double *red_out;
    double inv_n;
    const int32_t *gid, *frz;
    int fin;
};
__device__ __forceinline__ void vv2_body(const Vv2Args &a, int bid)
{
    __shared__ double lred[4];
    const int i = bid * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (i < a.n) {
        double4 v = a.vel[i];
        double4 f = a.frc[i];
        if (a.fin) {
            const double mx = a.red_out[R_FNETX] * a.inv_n, my = a.red_out[R_FNETY] * a.inv_n,
                         mz = a.red_out[R_FNETZ] * a.inv_n;
            const double4 r = a.frc_recip[i];
            f.x += (r.x - mx) * a.two_scale;
            f.y += (r.y - my) * a.two_scale;
            f.z += (r.z - mz) * a.two_scale;
            a.frc[i] = f;
        }
        if (a.frz && a.frz[a.gid[i]] != 0) {
            f.x = f.y = f.z = 0.0;
            v.x = v.y = v.z = 0.0;
            a.frc[i] = f;
            a.vel[i] = v;
        }
        if (v.w > 1.0e-6) {
            const double tmp = (0.5 * a.dt) / v.w;
            v.x += tmp * f.x;
            v.y += tmp * f.y;
            v.z += tmp * f.z;
            a.vel[i] = v;
            acc[0] = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
        }
    }
    block_sum<1>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_KIN);
}
__global__ void __launch_bounds__(256) k_vv2(Vv2Args a) { vv2_body(a, blockIdx.x); }
__global__ void __launch_bounds__(256) k_vv2_batch(const Vv2Args *__restrict__ av, const int32_t *__restrict__ boff,
                                                   int nsys)
{
    const int lb = blockIdx.x;
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    vv2_body(av[s], lb - boff[s]);
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

Vv1Args vv1_args(Ctx *c, double dt, bool device_decides)
{
    Vv1Args a;
    a.pos = c->pos;
    a.vel = c->vel;
    a.frc = c->frc;
    a.xref = c->xref;
    a.n = c->natms;
    a.dt = dt;
    for (int d = 0; d < 3; d++) {
        a.b.L[d] = c->L[d];
        a.b.invL[d] = c->invL[d];
    }
    // device words F_DMAX (maximum) and F_DONE (workgroup count) are zero
    // between launches
    a.dmax = (unsigned long long *)(c->flags + F_DMAX);
    a.done = (unsigned int *)(c->flags + F_DONE);
    a.host_dmax = device_decides ? nullptr : (unsigned long long *)(c->h_flags_dev + 8);
    a.flags = device_decides ? c->flags : nullptr;
    a.half_pad = kHalfMinus * c->prm.padding;
    return a;
}
int vv1_blocks(const Ctx *c) { return ceil_div(c->natms, kVv1Threads); }

int nve_first(Ctx *c, double dt)
{
    // the result lands in host word h_flags[8..9]
    hipLaunchKernelGGL(k_vv1, dim3(vv1_blocks(c)), dim3(kVv1Threads), 0, c->stream, vv1_args(c, dt, false));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

double nve_first_displacement(Ctx *c)
{
    double r2;
    memcpy(&r2, c->h_flags + 8, sizeof(r2));
    return std::sqrt(r2);
}

Vv2Args vv2_args(Ctx *c, double dt, bool fin)
{
    Vv2Args a;
    a.vel = c->vel;
    a.frc = c->frc;
    a.n = c->natms;
    a.dt = dt;
    a.red = c->red;
    a.fin = fin ? 1 : 0;
    a.frc_recip = fin ? c->frc_recip : nullptr;
    a.two_scale = fin ? 2.0 * recip_scale(c) : 0.0;
    a.red_out = fin ? c->red_out : nullptr;
    a.inv_n = 1.0 / (double)c->natms;
    a.gid = c->gid;
    a.frz = c->nfrozen > 0 ? c->frz : nullptr;
    return a;
}
int vv2_blocks(const Ctx *c) { return ceil_div(c->natms, 256); }

int nve_second(Ctx *c, double dt, bool /*kinetic*/)
{
    const int nb = vv2_blocks(c);
    int r;
    if ((r = ensure_red_rows(c, nb))) return r;
    hipLaunchKernelGGL(k_vv2, dim3(nb), dim3(256), 0, c->stream, vv2_args(c, dt, c->finish_in_vv2));
    c->finish_in_vv2 = false;
    c->red_rows[R_KIN] = nb;
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

// Batched stages 1 and 2 over the systems of an ensemble (each with its own
// time step, Ctx::run_dt); stage 1 leaves the rebuild decision on the device.
int vv1_prepare(const std::vector<Ctx *> &cs, BatchLaunch &bl)
{
    std::vector<Vv1Args> a;
    std::vector<int> nb;
    for (Ctx *c : cs) {
        a.push_back(vv1_args(c, c->run_dt, true));
        nb.push_back(vv1_blocks(c));
    }
    return batch_upload(cs[0], bl, a.data(), sizeof(Vv1Args), nb, false);
}
int vv1_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl)
{
    hipLaunchKernelGGL(k_vv1_batch, dim3(bl.total), dim3(kVv1Threads), 0, s, (const Vv1Args *)bl.args, bl.boff,
                       bl.nsys);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}
int vv2_prepare(const std::vector<Ctx *> &cs, BatchLaunch &bl)
{
    std::vector<Vv2Args> a;
    std::vector<int> nb;
    for (Ctx *c : cs) {
        a.push_back(vv2_args(c, c->run_dt, c->prm.electrostatics != 0));
        nb.push_back(vv2_blocks(c));
    }
    return batch_upload(cs[0], bl, a.data(), sizeof(Vv2Args), nb, false);
}
int vv2_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl)
{
    hipLaunchKernelGGL(k_vv2_batch, dim3(bl.total), dim3(256), 0, s, (const Vv2Args *)bl.args, bl.boff, bl.nsys);
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

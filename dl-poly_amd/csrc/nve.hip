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
// of a non-negative double).  The last workgroup to finish copies it to host
// memory (host_dmax: the host decides the rebuild) and resets *dmax and
// *done for the next step.  Host-decided runs only; a batched run decides on
// the device (k_step_batch).
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
    unsigned long long *host_dmax;
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
        *a.host_dmax = atomicAdd(a.dmax, 0ull);
        __threadfence_system();
        *a.dmax = 0ull;
        *a.done = 0u;
    }
}
__global__ void __launch_bounds__(kVv1Threads) k_vv1(Vv1Args a) { vv1_body(a, blockIdx.x, gridDim.x); }
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

static Vv1Args vv1_args(Ctx *c, double dt)
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
    a.host_dmax = (unsigned long long *)(c->h_flags_dev + 8);
    return a;
}
int vv1_blocks(const Ctx *c) { return ceil_div(c->natms, kVv1Threads); }

int nve_first(Ctx *c, double dt)
{
    // the result lands in host word h_flags[8..9]
    hipLaunchKernelGGL(k_vv1, dim3(vv1_blocks(c)), dim3(kVv1Threads), 0, c->stream, vv1_args(c, dt));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

double nve_first_displacement(Ctx *c)
{
    double r2;
    memcpy(&r2, c->h_flags + 8, sizeof(r2));
    return std::sqrt(r2);
}

static Vv2Args vv2_args(Ctx *c, double dt, bool fin)
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

// One launch per step of a resident batched run (every system of the batch):
// S2, stage 2 of the step just evaluated: the reciprocal finish (frc +=
// 2 scale (frc_recip - mean), the mean from the gather's atomic totals in
// acc), frozen atoms, the second half-kick and m v^2; the step's partial
// sums reduced into red_out (workgroup b takes columns b, b + nb, ...; the
// kinetic energy is added atomically, the reciprocal energy and net force
// come from the gather's slots), and the step's record when records are
// kept; and
// S1, stage 1 of the next step: half-kick, drift, and the displacement test,
// whose largest value the last workgroup to finish turns into the rebuild
// decision (as k_vv1).  Workgroups exchange data only through device-scope
// atomics; the last one also clears the atomic totals for the next step.
struct StepArgs {
    double4 *pos, *vel, *frc;
    const double4 *frc_recip;  // null: no reciprocal part
    const double4 *xref;
    const int32_t *gid, *frz;  // frz may be null
    int n;
    double dt, two_scale, inv_n, half_pad;
    Box b;
    double *red, *red_out;
    double *acc;               // the gather's kAccSlots x 4 totals (null: no reciprocal part)
    int rows[kReduceSlots];    // partial rows per column of the step
    double *energies;          // per-step records (kReduceSlots each) or null
    int ecap;
    int32_t *flags;
};
__device__ __forceinline__ bool atomic_column(int col) { return col == R_KIN || (col >= R_ERECIP && col <= R_FNETZ); }
// 256 threads: one atom per thread in workgroups small enough to spread a
// 64k-atom system over every CU (the partial-sum columns go to the first
// workgroups, one column each)
constexpr int kStepThreads = 256;
template <bool S2, bool S1>
__device__ __forceinline__ void step_body(const StepArgs &a, int bid, int nb)
{
    __shared__ double lsum[kStepThreads];
    __shared__ double wmax[kStepThreads / 64];
    __shared__ double lred[kStepThreads / 64];
    __shared__ double recip[4];  // reciprocal energy and net force (unscaled)
    __shared__ bool last;
    const int t = threadIdx.x;
    if (S2) {
        // the gather's slot totals, summed in a fixed order by the first wavefront
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (a.acc && t < kAccSlots)
            for (int k = 0; k < 4; k++) v[k] = a.acc[4 * t + k];
        if (t < 64)
            for (int k = 0; k < 4; k++) v[k] = wave_sum(v[k]);
        if (t == 0)
            for (int k = 0; k < 4; k++) recip[k] = v[k];
        __syncthreads();
    }
    // the step's record index (the last workgroup advances it after every
    // workgroup has read it)
    const int rec = (S2 && a.energies) ? a.flags[F_STEP] : a.ecap;
    double *record = rec < a.ecap ? a.energies + (size_t)rec * kReduceSlots : nullptr;
    if (S2) {
        // the columns the step's kernels filled with partial rows
        for (int col = bid; col < kReduceSlots; col += nb) {
            if (atomic_column(col)) continue;
            const double v = column_sum<kStepThreads>(a.red, a.rows[col], kReduceSlots, col, lsum);
            if (t == 0) {
                a.red_out[col] = v;
                if (record) record[col] = v;
            }
        }
    }
    const int i = bid * blockDim.x + t;
    double ke = 0.0, r2 = 0.0;
    if (i < a.n) {
        double4 v = a.vel[i];
        double4 f = a.frc[i];
        if (S2) {
            if (a.frc_recip) {
                const double mx = recip[1] * a.inv_n, my = recip[2] * a.inv_n, mz = recip[3] * a.inv_n;
                const double4 r = a.frc_recip[i];
                f.x += (r.x - mx) * a.two_scale;
                f.y += (r.y - my) * a.two_scale;
                f.z += (r.z - mz) * a.two_scale;
            }
            if (a.frz && a.frz[a.gid[i]] != 0) {
                f.x = f.y = f.z = 0.0;
                v.x = v.y = v.z = 0.0;
            }
            a.frc[i] = f;
            if (v.w > 1.0e-6) {
                const double tmp = (0.5 * a.dt) / v.w;
                v.x += tmp * f.x;
                v.y += tmp * f.y;
                v.z += tmp * f.z;
                ke = v.w * (v.x * v.x + v.y * v.y + v.z * v.z);
            }
        }
        if (S1) {
            double4 p = a.pos[i];
            if (v.w > 1.0e-6) {
                const double tmp = (0.5 * a.dt) / v.w;
                v.x += tmp * f.x;
                v.y += tmp * f.y;
                v.z += tmp * f.z;
                p.x += a.dt * v.x;
                p.y += a.dt * v.y;
                p.z += a.dt * v.z;
                a.pos[i] = p;
            }
            const double4 q = a.xref[i];
            const double dx = min_image_1(p.x, q.x, a.b.L[0], a.b.invL[0]);
            const double dy = min_image_1(p.y, q.y, a.b.L[1], a.b.invL[1]);
            const double dz = min_image_1(p.z, q.z, a.b.L[2], a.b.invL[2]);
            r2 = dx * dx + dy * dy + dz * dz;
        }
        a.vel[i] = v;
    }
    if (S2) {
        double acc[1] = {ke};
        const double tot = block_sum<1>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_KIN);
        if (t == 0) atomicAdd(&a.red_out[R_KIN], tot);
    }
    if (S1) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
        if ((t & 63) == 0) wmax[t >> 6] = r2;
    }
    __syncthreads();
    if (t == 0) {
        if (S1) {
            double m = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wmax[w]);
            atomicMax((unsigned long long *)(a.flags + F_DMAX), (unsigned long long)__double_as_longlong(m));
        }
        wait_memory();
        last = atomicAdd((unsigned int *)(a.flags + F_DONE), 1u) == (unsigned)nb - 1;
    }
    __syncthreads();
    if (!last) return;
    if (S2) {
        // the gather's slots cleared for the next step (every workgroup has
        // read them)
        if (a.acc && t < 4 * kAccSlots)
            __hip_atomic_store(&a.acc[t], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (t != 0) return;
    if (S2) {
        // the atomic totals into red_out and the record; the kinetic one is
        // cleared for the next step
        const double kin = load_coherent(&a.red_out[R_KIN]);
        __hip_atomic_store(&a.red_out[R_KIN], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; k < 4; k++) a.red_out[R_ERECIP + k] = recip[k];
        if (record) {
            record[R_KIN] = kin;
            for (int k = 0; k < 4; k++) record[R_ERECIP + k] = recip[k];
        }
        if (a.energies) a.flags[F_STEP] = rec + 1;
    }
    if (S1) {
        const unsigned long long bits = load_coherent((unsigned long long *)(a.flags + F_DMAX));
        const int upd = (a.half_pad <= 0.0 || sqrt(__longlong_as_double((long long)bits)) >= a.half_pad) ? 1 : 0;
        a.flags[F_UPD] = upd;
        if (upd) {
            a.flags[F_OVF] = 0;
            a.flags[F_IBIG] = 0;
            a.flags[F_PAIRS] = 0;
            a.flags[F_NREB] += 1;
        }
        __hip_atomic_store((unsigned long long *)(a.flags + F_DMAX), 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store((unsigned int *)(a.flags + F_DONE), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool S2, bool S1>
__global__ void __launch_bounds__(kStepThreads) k_step_batch(const StepArgs *__restrict__ av,
                                                            const int32_t *__restrict__ boff, int nsys)
{
    const int lb = blockIdx.x;  // atoms in order: no XCD renumbering
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    step_body<S2, S1>(av[s], lb - boff[s], boff[s + 1] - boff[s]);
}

int step_blocks(const Ctx *c) { return ceil_div(c->natms, kStepThreads); }

int step_prepare(const std::vector<Ctx *> &cs, const std::vector<RedRowsHost> &rows, bool energies,
                 BatchLaunch &bl)
{
    std::vector<StepArgs> av;
    std::vector<int> nb;
    for (size_t k = 0; k < cs.size(); k++) {
        Ctx *c = cs[k];
        StepArgs a{};
        a.pos = c->pos;
        a.vel = c->vel;
        a.frc = c->frc;
        a.frc_recip = c->prm.electrostatics ? c->frc_recip : nullptr;
        a.xref = c->xref;
        a.gid = c->gid;
        a.frz = c->nfrozen > 0 ? c->frz : nullptr;
        a.n = c->natms;
        a.dt = c->run_dt;
        a.two_scale = c->prm.electrostatics ? 2.0 * recip_scale(c) : 0.0;
        a.inv_n = 1.0 / (double)c->natms;
        a.half_pad = kHalfMinus * c->prm.padding;
        for (int d = 0; d < 3; d++) {
            a.b.L[d] = c->L[d];
            a.b.invL[d] = c->invL[d];
        }
        a.red = c->red;
        a.red_out = c->red_out;
        a.acc = c->prm.electrostatics ? c->acc : nullptr;
        for (int q = 0; q < kReduceSlots; q++) a.rows[q] = rows[k].rows[q];
        a.energies = energies ? c->energies_dev : nullptr;
        a.ecap = c->energies_cap;
        a.flags = c->flags;
        av.push_back(a);
        nb.push_back(step_blocks(c));
    }
    return batch_upload(cs[0], bl, av.data(), sizeof(StepArgs), nb, false);
}

int step_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl, bool stage2, bool stage1)
{
    const dim3 g(bl.total), b(kStepThreads);
    const StepArgs *av = (const StepArgs *)bl.args;
    if (stage2 && stage1) hipLaunchKernelGGL((k_step_batch<true, true>), g, b, 0, s, av, bl.boff, bl.nsys);
    else if (stage2) hipLaunchKernelGGL((k_step_batch<true, false>), g, b, 0, s, av, bl.boff, bl.nsys);
    else if (stage1) hipLaunchKernelGGL((k_step_batch<false, true>), g, b, 0, s, av, bl.boff, bl.nsys);
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

// The C ABI (include/dlpx.h): context life cycle, tables, and the
// link_cell_pairs / vnl_check / two_body_forces / resident-MD entry points.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "common.h"

namespace dlpx {

int load_atoms(Ctx *c, const double *st_dev, const int32_t *st_t_dev, bool identity, bool with_vel);
struct Ensemble;
uint64_t next_serial();
void ensemble_free(void *e);
int ensemble_run(const std::vector<Ctx *> &cs, Ensemble *&E, int nsteps, bool energies, std::vector<int> &rebuilds);
int store_atoms(Ctx *c, double *st_dev, bool with_pos, bool with_vel);
int spme_finish(Ctx *c, double two_scale);

int fail(Ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <typename T>
static int dalloc(Ctx *c, T **p, size_t count)
{
    DLPX_HIP(c, hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1)));
    return 0;
}

static int set_cell(Ctx *c, const double *cell)
{
    if (!cell) return fail(c, DLPX_ERR_BAD_ARG, "cell is NULL");
    if (cell[1] != 0.0 || cell[2] != 0.0 || cell[3] != 0.0 || cell[5] != 0.0 || cell[6] != 0.0 || cell[7] != 0.0)
        return fail(c, DLPX_ERR_UNSUPPORTED, "only orthorhombic cells (imcon 1, 2) are supported");
    double det;
    invert_cell(cell, c->rcell, &det);
    if (std::fabs(det) < 1.0e-6) return fail(c, DLPX_ERR_SINGULAR_CELL, "singular cell");
    std::memcpy(c->cell, cell, sizeof(double) * 9);
    for (int d = 0; d < 3; d++) {
        c->L[d] = cell[4 * d];
        if (c->L[d] <= 0.0) return fail(c, DLPX_ERR_BAD_ARG, "cell lengths must be positive");
        c->invL[d] = 1.0 / c->L[d];
    }
    double p[10];
    dcell(cell, p);
    c->volm = p[9];
    return 0;
}

int upload_tables(Ctx *c)
{
    if (!c->tables_dirty) return c->comb_dirty && !c->noalloc_now ? upload_comb(c) : 0;
    const int nt = c->prm.ntypes;
    if (c->nkeys != nt * (nt + 1) / 2) return fail(c, DLPX_ERR_STATE, "vdW tables not set for every type pair");
    if (c->prm.electrostatics && c->nsamples <= 0) return fail(c, DLPX_ERR_STATE, "erfc tables not set");
    const bool mixed = c->prm.precision == DLPX_MIXED;
    const size_t rs = mixed ? sizeof(float) : sizeof(double);
    const int mg1 = c->max_grid + 1, ns1 = c->nsamples + 1;
    // pair keys whose interactions are identical (same tables and, for direct
    // evaluation, the same form and parameters) share one device row, so many
    // atom types of few distinct interactions still fit the pair kernel's LDS
    std::vector<int> row_of(c->nkeys, -1), key_of_row;
    for (int k = 0; k < c->nkeys; k++) {
        if (c->ltp[k] < 0) continue;
        for (int q = 0; q < (int)key_of_row.size() && row_of[k] < 0; q++) {
            const int k2 = key_of_row[q];
            bool same = c->ltp[k2] == c->ltp[k] &&
                        !std::memcmp(&c->vdw_frc[(size_t)k * mg1], &c->vdw_frc[(size_t)k2 * mg1], sizeof(double) * mg1) &&
                        !std::memcmp(&c->vdw_pot[(size_t)k * mg1], &c->vdw_pot[(size_t)k2 * mg1], sizeof(double) * mg1);
            if (same && c->vdw_direct)
                same = !std::memcmp(&c->gen_params[kVdwParams * (size_t)k], &c->gen_params[kVdwParams * (size_t)k2],
                                    sizeof(double) * kVdwParams);
            if (same) row_of[k] = q;
        }
        if (row_of[k] < 0) {
            row_of[k] = (int)key_of_row.size();
            key_of_row.push_back(k);
        }
    }
    c->nrows = (int)key_of_row.size();
    const int nrows_alloc = std::max(c->nrows, 1);
    std::vector<double> v((size_t)2 * nrows_alloc * mg1, 0.0), e((size_t)2 * std::max(ns1, 1), 0.0);
    for (int q = 0; q < c->nrows; q++) {
        const int k = key_of_row[q];
        for (int l = 0; l < mg1; l++) {
            v[2 * ((size_t)q * mg1 + l)] = c->vdw_frc[(size_t)k * mg1 + l];
            v[2 * ((size_t)q * mg1 + l) + 1] = c->vdw_pot[(size_t)k * mg1 + l];
        }
    }
    for (int l = 1; l < ns1; l++) {  // sample l <-> r = l * spacing
        e[2 * l] = c->erfc_d[l - 1];
        e[2 * l + 1] = c->erfc_t[l - 1];
    }
    if (c->d_vdw) (void)hipFree(c->d_vdw);
    if (c->d_erfc) (void)hipFree(c->d_erfc);
    c->d_vdw = c->d_erfc = nullptr;
    DLPX_HIP(c, hipMalloc(&c->d_vdw, rs * v.size()));
    DLPX_HIP(c, hipMalloc(&c->d_erfc, rs * e.size()));
    if (mixed) {
        std::vector<float> vf(v.begin(), v.end()), ef(e.begin(), e.end());
        for (auto &x : vf)
            if (!std::isfinite(x)) x = 3.0e38f;  // slot 0 holds Huge
        DLPX_HIP(c, dev_upload(c->stream, c->d_vdw, vf.data(), rs * vf.size()));
        DLPX_HIP(c, dev_upload(c->stream, c->d_erfc, ef.data(), rs * ef.size()));
    } else {
        DLPX_HIP(c, dev_upload(c->stream, c->d_vdw, v.data(), rs * v.size()));
        DLPX_HIP(c, dev_upload(c->stream, c->d_erfc, e.data(), rs * e.size()));
    }
    std::vector<int32_t> kt((size_t)nt * nt);
    for (int a = 1; a <= nt; a++)
        for (int b = 1; b <= nt; b++) {
            const int hi = std::max(a, b), lo = std::min(a, b);
            const int key = hi * (hi - 1) / 2 + lo - 1;
            // vdw.F90:1883: an undefined pair (tab_potential(0,k) ~ 0) is skipped
            kt[(size_t)(a - 1) * nt + (b - 1)] = (c->ltp[key] < 0) ? -1 : row_of[key];
        }
    if (c->d_keytab) (void)hipFree(c->d_keytab);
    c->d_keytab = nullptr;
    DLPX_HIP(c, hipMalloc((void **)&c->d_keytab, sizeof(int32_t) * kt.size()));
    DLPX_HIP(c, dev_upload(c->stream, c->d_keytab, kt.data(), sizeof(int32_t) * kt.size()));
    if (c->vdw_direct) {
        // vdw_method direct: per row {form, prepared parameters, afs, bfs}; afs,
        // bfs follow vdw_direct_fs_generate (vdw.F90:969-1048) under force-shifting
        constexpr int S = kDirectStride;
        std::vector<double> d((size_t)S * nrows_alloc, 0.0);
        for (int row = 0; row < c->nrows; row++) {
            const int k = key_of_row[row];
            const int kind = c->ltp[k];
            if (!vdw_kind_supported(kind)) continue;
            double q[kVdwParams];
            vdw_prepare(kind, c->gen_params.data() + kVdwParams * (size_t)k, c->prm.vdw_cutoff, q);
            d[S * row] = (double)kind;
            for (int j = 0; j < kVdwParams; j++) d[S * row + 1 + j] = q[j];
            // none for n-m shift, DPD, nDPD, LJF and 12-6-MDF (vdw.F90:1015-1040)
            if (c->vdw_force_shift && kind != 7 && kind != 10 && kind != 23 && kind != 21 && kind != 20) {
                double z, dz;
                vdw_form(kind, q, c->prm.vdw_cutoff, &z, &dz);
                d[S * row + S - 2] = dz / c->prm.vdw_cutoff;
                d[S * row + S - 1] = -z - dz;
            }
        }
        if (c->d_vdw_direct) (void)hipFree(c->d_vdw_direct);
        c->d_vdw_direct = nullptr;
        DLPX_HIP(c, hipMalloc(&c->d_vdw_direct, rs * d.size()));
        if (mixed) {
            std::vector<float> df(d.begin(), d.end());
            DLPX_HIP(c, dev_upload(c->stream, c->d_vdw_direct, df.data(), rs * df.size()));
        } else {
            DLPX_HIP(c, dev_upload(c->stream, c->d_vdw_direct, d.data(), rs * d.size()));
        }
    }
    c->tables_dirty = false;
    c->comb_dirty = true;
    return c->noalloc_now ? 0 : upload_comb(c);
}

// host arrays (strided) -> pinned staging -> device staging -> cell order
static int upload(Ctx *c, const dlpx_atoms *a, bool identity, const double *vx, const double *vy, const double *vz,
                  const double *mass, int vstride = 0)
{
    if (!a || !a->x || !a->y || !a->z || !a->chge || !a->ltype) return fail(c, DLPX_ERR_BAD_ARG, "atom arrays");
    if (a->natms <= 0 || a->natms > c->cap) return fail(c, DLPX_ERR_BAD_ARG, "natms outside 1..natms_max");
    if (!identity && a->natms != c->natms) return fail(c, DLPX_ERR_STATE, "natms changed since the list was built");
    const int n = a->natms, st = a->stride, vst = vstride > 0 ? vstride : a->stride;
    auto at = [st](const double *p, int i) { return *(const double *)((const char *)p + (size_t)st * i); };
    auto atv = [vst](const double *p, int i) { return *(const double *)((const char *)p + (size_t)vst * i); };
    for (int i = 0; i < n; i++) {
        double *r = c->h_st + (size_t)8 * i;
        r[0] = at(a->x, i);
        r[1] = at(a->y, i);
        r[2] = at(a->z, i);
        r[3] = at(a->chge, i);
        if (vx) {
            r[4] = atv(vx, i);
            r[5] = atv(vy, i);
            r[6] = atv(vz, i);
            r[7] = atv(mass, i);
        }
        const int t = a->ltype[i];
        if (t < 1 || t > c->prm.ntypes) return fail(c, DLPX_ERR_BAD_ARG, "ltype outside 1..ntypes");
        c->h_t[i] = t;
    }
    c->natms = n;
    {
        // charge per atom type, for the pair kernel's combined-table path
        std::vector<double> tq(c->prm.ntypes, 0.0);
        std::vector<char> seen(c->prm.ntypes, 0);
        bool per_type = true;
        for (int i = 0; i < n && per_type; i++) {
            const int t = c->h_t[i] - 1;
            const double q = c->h_st[(size_t)8 * i + 3];
            if (!seen[t]) {
                seen[t] = 1;
                tq[t] = q;
            } else if (q != tq[t]) {
                per_type = false;
            }
        }
        if (per_type != c->q_per_type || (per_type && tq != c->type_q)) c->comb_dirty = true;
        c->q_per_type = per_type;
        c->type_q = tq;
    }
    DLPX_HIP(c, hipMemcpyAsync(c->stage, c->h_st, sizeof(double) * 8 * n, hipMemcpyHostToDevice, c->stream));
    DLPX_HIP(c, hipMemcpyAsync(c->stage_i, c->h_t, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    return load_atoms(c, c->stage, c->stage_i, identity, vx != nullptr);
}

static void charges_summary(Ctx *c, const dlpx_atoms *a)
{
    // spme.F90:159-231 self interaction, two_body.F90:689-696 net charge
    double sq = 0.0, sum = 0.0;
    for (int i = 0; i < a->natms; i++) {
        const double q = *(const double *)((const char *)a->chge + (size_t)a->stride * i);
        sq += q * q;
        sum += q;
    }
    const double scaling = kR4pie0 / c->prm.epsilon;
    c->self_energy = -sq * (scaling * c->prm.alpha * (1.0 / std::sqrt(kPi)) / 1.0);
    c->sumchg = sum;
    c->self_done = true;
}

double recip_scale(const Ctx *c)
{
    return kPi * std::sqrt(kPi) * std::pow(c->prm.alpha, 1 - 3) * (0.5 * (1.0 / c->volm)) * (kR4pie0 / c->prm.epsilon);
}

// pair + SPME forces into c->frc (cell order); partial sums reduced to
// red_out.  With defer_finish the reciprocal forces are left in frc_recip for
// the velocity stage to add (nve_second), saving a pass over the atoms.
static int evaluate(Ctx *c, bool defer_finish = false)
{
    int r;
    if ((r = upload_tables(c))) return r;
    // one capacity for every stage's partial rows, so no stage reallocates the
    // buffer after an earlier stage of this evaluation has written to it
    if ((r = ensure_red_rows(c, std::max({pair_blocks(c), gather_blocks(c), conv_blocks(c), ceil_div(c->natms, 256), 2048}))))
        return r;
    for (int k = 0; k < kReduceSlots; k++) c->red_rows[k] = 0;
    if (c->timing) {
        // stages one after another on one stream, each between HIP events
        c->sp = c->stream;
        DLPX_HIP(c, hipEventRecord(c->ev[0], c->stream));
        if ((r = pair_forces(c))) return r;
        if ((r = excl_forces(c))) return r;
        DLPX_HIP(c, hipEventRecord(c->ev[1], c->stream));
        if (c->prm.electrostatics) {
            if ((r = spme_forces(c))) return r;  // records ev[2..4]
        } else {
            for (int k = 2; k <= 4; k++) DLPX_HIP(c, hipEventRecord(c->ev[k], c->stream));
        }
    } else if (c->prm.electrostatics) {
        // the reciprocal-space chain and the real-space pair kernel read the
        // same positions and write disjoint outputs (frc_recip and its
        // reduction columns; frc and its columns): run them concurrently
        c->sp = c->recip_stream;
        DLPX_HIP(c, hipEventRecord(c->fork_ev, c->stream));
        DLPX_HIP(c, hipStreamWaitEvent(c->recip_stream, c->fork_ev, 0));
        if ((r = spme_forces(c))) return r;
        if ((r = pair_forces(c))) return r;
        if ((r = excl_forces(c))) return r;
        DLPX_HIP(c, hipEventRecord(c->join_ev, c->recip_stream));
        DLPX_HIP(c, hipStreamWaitEvent(c->stream, c->join_ev, 0));
        c->sp = c->stream;
    } else {
        if ((r = pair_forces(c))) return r;
    }
    for (int k = R_ENGVDW; k <= R_S33; k++) c->red_rows[k] = pair_blocks(c);
    if (c->excl_total > 0 && c->prm.electrostatics)
        for (int k = R_ENGEX; k <= R_XS33; k++) c->red_rows[k] = excl_blocks(c);
    if ((r = reduce_partials(c))) return r;
    c->finish_in_vv2 = false;
    if (c->prm.electrostatics) {
        if (defer_finish)
            c->finish_in_vv2 = true;
        else if ((r = spme_finish(c, 2.0 * recip_scale(c))))
            return r;
    }
    if (c->timing) {
        DLPX_HIP(c, hipEventRecord(c->ev[5], c->stream));
        c->timed_evals++;
        c->timing_pending = true;
    }
    return 0;
}

// accumulate the per-stage event times of the last evaluation (ms):
// [0] pair, [1] spread, [2] FFT + influence, [3] gather, [4] reductions/finish
static int collect_timing(Ctx *c)
{
    if (!c->timing || !c->timing_pending) return 0;
    DLPX_HIP(c, hipEventSynchronize(c->ev[5]));
    float ms;
    for (int k = 0; k < 5; k++) {
        DLPX_HIP(c, hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
        c->stage_ms[k] += ms;
    }
    c->timing_pending = false;
    return 0;
}

// two_body.F90:708-790 bookkeeping from the reduced sums
static void assemble(Ctx *c, const double *red, dlpx_two_body_out *o)
{
    std::memset(o, 0, sizeof(*o));
    const double engvdw = red[R_ENGVDW], virvdw = red[R_VIRVDW];
    // the real-space stress trace is sum gamma r^2 = -(virvdw + virrl)
    // (vdw.F90:1985-2022, ewald_spole.F90:189-197), so the pair kernel does not
    // accumulate the Coulomb virial separately; with no electrostatics it is
    // exactly zero rather than the rounding difference of two vdW sums
    const double engrl = red[R_ENGRL];
    const double virrl = c->prm.electrostatics ? -(red[R_S11] + red[R_S22] + red[R_S33]) - virvdw : 0.0;
    double st[9] = {red[R_S11], red[R_S12], red[R_S13], red[R_S12], red[R_S22],
                    red[R_S23], red[R_S13], red[R_S23], red[R_S33]};
    double engrc = 0.0, virrc = 0.0, engnz = 0.0, virnz = 0.0, engex = 0.0, virex = 0.0;
    if (c->prm.electrostatics) {
        // Ewald exclusion correction: energy, and virial = -trace of its stress
        // (ewald_spole.F90:615-630)
        const double xs[9] = {red[R_XS11], red[R_XS12], red[R_XS13], red[R_XS12], red[R_XS22],
                              red[R_XS23], red[R_XS13], red[R_XS23], red[R_XS33]};
        engex = red[R_ENGEX];
        virex = -(xs[0] + xs[4] + xs[8]);
        for (int k = 0; k < 9; k++) st[k] += xs[k];
        const double scale = kPi * std::sqrt(kPi) * std::pow(c->prm.alpha, 1 - 3) * (0.5 * (1.0 / c->volm)) *
                             (kR4pie0 / c->prm.epsilon);
        const double eng = red[R_ERECIP] * scale;
        double rs[9] = {red[R_RS11], red[R_RS12], red[R_RS13], red[R_RS12], red[R_RS22],
                        red[R_RS23], red[R_RS13], red[R_RS23], red[R_RS33]};
        for (int k = 0; k < 9; k++) rs[k] *= scale;
        rs[0] += eng;
        rs[4] += eng;
        rs[8] += eng;
        for (int k = 0; k < 9; k++) st[k] += rs[k];
        virrc = -(rs[0] + rs[4] + rs[8]);
        engrc = eng + c->self_energy;
        if (std::fabs(c->sumchg) > 1.0e-6) {
            const double factor_nz = -0.5 * (kPi * kR4pie0 / c->prm.epsilon) * std::pow(c->sumchg / c->prm.alpha, 2);
            engnz = factor_nz / c->volm;
            virnz = -3.0 * engnz;
        }
    }
    o->buffer[7] = engvdw;
    o->buffer[8] = virvdw;
    o->buffer[9] = engrc;
    o->buffer[10] = virrc;
    o->buffer[11] = engrl;
    o->buffer[12] = virrl;
    o->buffer[15] = engex;
    o->buffer[16] = virex;
    o->engcpe = engrc + engrl + engex + engnz;
    o->vircpe = virrc + virrl + virex + virnz;
    o->engsrp = engvdw + c->elrc;
    o->virsrp = virvdw + c->vlrc;
    const double dnz = -virnz / 3.0, dlrc = -c->vlrc / 3.0;
    st[0] += dnz + dlrc;
    st[4] += dnz + dlrc;
    st[8] += dnz + dlrc;
    std::memcpy(o->stress, st, sizeof(st));
}

static int fetch_red(Ctx *c)
{
    DLPX_HIP(c, hipMemcpyAsync(c->h_red, c->red_out, sizeof(double) * kReduceSlots, hipMemcpyDeviceToHost, c->stream));
    DLPX_HIP(c, hipStreamSynchronize(c->stream));
    return collect_timing(c);
}

// What a captured step graph bakes in (buffers, sizes, launch shapes, the
// time step): a change of any of them forces a new capture.
static uint64_t bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, sizeof(u));
    return u;
}

std::vector<uint64_t> step_graph_key(const Ctx *c, double tstep)
{
    uint64_t t;
    memcpy(&t, &tstep, sizeof(t));
    return {(uint64_t)c->pos,      (uint64_t)c->vel,     (uint64_t)c->frc,     (uint64_t)c->frc_recip,
            (uint64_t)c->xref,     (uint64_t)c->typ,     (uint64_t)c->nbr,     (uint64_t)c->cnt,
            (uint64_t)c->cell_start, (uint64_t)c->red,   (uint64_t)c->red_out, (uint64_t)c->red_cap,
            (uint64_t)c->Q,        (uint64_t)c->Phi,     (uint64_t)c->H,       (uint64_t)c->d_vdw,   (uint64_t)c->d_erfc,
            (uint64_t)c->d_keytab, (uint64_t)c->d_vdw_direct, (uint64_t)c->cap_lane, (uint64_t)c->ncell,  (uint64_t)c->natms,
            (uint64_t)c->nc[0],    (uint64_t)c->nc[1],   (uint64_t)c->nc[2],   (uint64_t)c->gid,
            (uint64_t)c->slot,     (uint64_t)c->excl_off, (uint64_t)c->excl_ids, (uint64_t)c->excl_total,
            (uint64_t)c->frz,      (uint64_t)c->nfrozen, t, (uint64_t)c->theta,
            // the cell: pair, spreading and influence arguments carry it
            bits(c->cell[0]), bits(c->cell[4]), bits(c->cell[8])};
}

// Start of a resident run: clear the counters of the fused stage-1 kernel
// and do stage 1 of the first step.
static int md_begin(Ctx *c, int nsteps, double tstep)
{
    // displacement maximum and workgroup counter of the fused stage-1 kernel
    DLPX_HIP(c, hipMemsetAsync(c->flags + 4, 0, sizeof(int32_t) * 4, c->stream));
    return nsteps > 0 ? nve_first(c, tstep) : 0;
}

// Step s of a host-decided resident run (nve.F90:127-215 with
// refresh_mappings/vnl_check and the force call, drivers.F90:2105-2199): wait
// for this context's stage 1, rebuild the list if an atom moved by padding/2,
// then forces, stage 2 and the next step's stage 1.  Returns with the step's
// work queued on c->stream.
static int md_step(Ctx *c, int s, int nsteps, double tstep, bool energies, int &rebuilds)
{
    int r;
    int update = 1;
    if (c->prm.padding > 0.0) {
        DLPX_HIP(c, hipStreamSynchronize(c->stream));
        update = nve_first_displacement(c) >= kHalfMinus * c->prm.padding ? 1 : 0;
    }
    if (update) {
        if ((r = build_list(c))) return r;
        rebuilds++;
    }
    const bool last = s == nsteps - 1;
    if ((r = evaluate(c, true))) return r;
    if ((r = nve_second(c, tstep, true))) return r;
    if (c->timing && (r = collect_timing(c))) return r;
    if (energies) {
        if ((r = reduce_partials(c))) return r;
        DLPX_HIP(c, hipMemcpyAsync(c->energies_dev + (size_t)kReduceSlots * s, c->red_out,
                                   sizeof(double) * kReduceSlots, hipMemcpyDeviceToDevice, c->stream));
    }
    if (!last && (r = nve_first(c, tstep))) return r;
    return 0;
}

// A resident run with host-decided rebuilds (the list regrows when it
// overflows), from the current positions and velocities: the list is built
// and the forces evaluated first, as at dlpx_md_load.  Per-step records go to
// energies_dev when energies is set.
int md_fallback(Ctx *c, int nsteps, double tstep, bool energies, int &rebuilds)
{
    int r;
    if ((r = build_list(c))) return r;
    if ((r = evaluate(c))) return r;
    if ((r = freeze_atoms(c))) return r;
    rebuilds = 0;
    if ((r = md_begin(c, nsteps, tstep))) return r;
    for (int s = 0; s < nsteps; s++)
        if ((r = md_step(c, s, nsteps, tstep, energies, rebuilds))) return r;
    return 0;
}

// Scratch of dlpx_two_body_forces_noalloc: the partial-sum rows of one force
// evaluation (one row of kReduceSlots doubles per workgroup of a reducing
// kernel) followed by their kReduceSlots column totals.  kScratchRows rows
// (7.3 MB) cover up to about a million atoms at spline order 8, where the
// gather has one workgroup per 32 atoms.  Each device holds its own copy
// until the process ends; every call on a device shares it.  Host code uses
// the address hipGetSymbolAddress gives for the device, never the array's
// name (a host-side stand-in).
constexpr int kScratchRows = 32768;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(size_t)kScratchRows * kReduceSlots + kReduceSlots];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (dlpx_create runs on it).
static int scratch_address(Ctx *c)
{
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, DLPX_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[c->device]) return 0;
    void *p = nullptr;
    DLPX_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[c->device] = (double *)p;
    return 0;
}

static double *scratch_of(const Ctx *c)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
}

}  // namespace dlpx

using namespace dlpx;

extern "C" {

int dlpx_version(void) { return 1; }

const char *dlpx_last_error(const dlpx_ctx *ctx)
{
    return ctx ? reinterpret_cast<const Ctx *>(ctx)->err.c_str() : "no context";
}

int dlpx_create(const dlpx_params *p, dlpx_ctx **out)
{
    if (!p || !out) return DLPX_ERR_BAD_ARG;
    *out = nullptr;
    if (p->natms_max <= 0 || p->ntypes <= 0 || p->cutoff <= 0.0 || p->padding < 0.0 || p->vdw_cutoff <= 0.0 ||
        (p->precision != DLPX_FP64 && p->precision != DLPX_MIXED) || p->epsilon <= 0.0)
        return DLPX_ERR_BAD_ARG;
    if (p->electrostatics && p->alpha <= 0.0) return DLPX_ERR_BAD_ARG;
    Ctx *c = new (std::nothrow) Ctx();
    if (!c) return DLPX_ERR_BAD_ARG;
    c->prm = *p;
    if (p->natms_max >= kMaxAtoms) {
        delete c;
        return DLPX_ERR_UNSUPPORTED;
    }
    c->cap = p->natms_max;
    int r = 0;
    auto bail = [&](int code) {
        dlpx_destroy(reinterpret_cast<dlpx_ctx *>(c));
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->recip_stream, hipStreamNonBlocking) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess)
        return bail(DLPX_ERR_HIP);
    c->sp = c->stream;
    // lanes per atom of the pair kernel; DLPX_PAIR_LANES (4, 8, 16) overrides
    c->pair_lanes = 8;
    // resident steps replay a captured graph unless DLPX_GRAPHS=0
    if (const char *e = std::getenv("DLPX_GRAPHS")) c->use_graphs = std::atoi(e) != 0;
    // the list build bins by half-edge subcells unless DLPX_LIST_SUBCELLS=0
    if (const char *e = std::getenv("DLPX_LIST_SUBCELLS")) c->list_subcells = std::atoi(e) != 0;
    // DLPX_FFT_LINES=0: the generic Stockham FFT kernels for every length
    if (const char *e = std::getenv("DLPX_FFT_LINES")) c->fft_lines = std::atoi(e) != 0;
    // DLPX_REL_PAIRS=0: the mixed pair kernel forms separations from FP64 positions
    // DLPX_FFT_PLANES=0: separate x and y line kernels where the plane kernels would run
    if (const char *e = std::getenv("DLPX_FFT_PLANES")) c->fft_planes = std::atoi(e) != 0;
    if (const char *e = std::getenv("DLPX_REL_PAIRS")) c->rel_pairs = std::atoi(e) != 0;
    if (const char *e = std::getenv("DLPX_COMB_THREADS")) c->comb_threads = std::atoi(e) == 512 ? 512 : 1024;
    if (const char *e = std::getenv("DLPX_GRAPH_STEPS")) c->graph_steps = std::max(1, std::min(64, std::atoi(e)));
    if (const char *e = std::getenv("DLPX_OVERLAP")) c->overlap = std::max(0, std::min(2, std::atoi(e)));
    if (const char *e = std::getenv("DLPX_PAIR_LANES")) {
        const int v = std::atoi(e);
        if (v == 4 || v == 8 || v == 16) c->pair_lanes = v;
    }
    for (int k = 0; k < 8; k++)
        if (hipEventCreate(&c->ev[k]) != hipSuccess) return bail(DLPX_ERR_HIP);
    const size_t n = c->cap;
    if ((r = dalloc(c, &c->pos, n)) || (r = dalloc(c, &c->pos_alt, n)) || (r = dalloc(c, &c->vel, n)) ||
        (r = dalloc(c, &c->vel_alt, n)) || (r = dalloc(c, &c->frc, n)) || (r = dalloc(c, &c->frc_recip, n)) ||
        (r = dalloc(c, &c->xref, n)) || (r = dalloc(c, &c->typ, n)) || (r = dalloc(c, &c->typ_alt, n)) ||
        (r = dalloc(c, &c->gid, n)) || (r = dalloc(c, &c->gid_alt, n)) || (r = dalloc(c, &c->slot, n)) ||
        (r = dalloc(c, &c->stage, 9 * n)) || (r = dalloc(c, &c->stage_i, n)) || (r = dalloc(c, &c->cell_of, n)) ||
        (r = dalloc(c, &c->perm, n)) || (r = dalloc(c, &c->cnt, n * kLanesPerAtom)) || (r = dalloc(c, &c->flags, F_WORDS)) ||
        (r = dalloc(c, &c->red_out, kReduceSlots)) || (r = dalloc(c, &c->acc, 4 * kAccSlots)) ||
        (r = dalloc(c, &c->sub_of, n)) || (r = dalloc(c, &c->prel, n)))
        return bail(r);
    if (dev_zero(c->stream, c->acc, sizeof(double) * 4 * kAccSlots) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (dev_zero(c->stream, c->vel, sizeof(double4) * n) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (dev_zero(c->stream, c->flags, sizeof(int32_t) * F_WORDS) != hipSuccess) return bail(DLPX_ERR_HIP);
    c->serial = next_serial();
    // host words the GPU writes directly (the displacement maximum of each step)
    if (hipHostMalloc((void **)&c->h_flags, 64, hipHostMallocMapped) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (hipHostGetDevicePointer((void **)&c->h_flags_dev, c->h_flags, 0) != hipSuccess) return bail(DLPX_ERR_HIP);
    if (hipHostMalloc((void **)&c->h_red, sizeof(double) * kReduceSlots, hipHostMallocDefault) != hipSuccess)
        return bail(DLPX_ERR_HIP);
    if (hipHostMalloc((void **)&c->h_st, sizeof(double) * 9 * n, hipHostMallocDefault) != hipSuccess)
        return bail(DLPX_ERR_HIP);
    if (hipHostMalloc((void **)&c->h_t, sizeof(int32_t) * n, hipHostMallocDefault) != hipSuccess)
        return bail(DLPX_ERR_HIP);
    if (p->electrostatics && (r = spme_setup(c))) return bail(r);
    if ((r = scratch_address(c))) return bail(r);
    *out = reinterpret_cast<dlpx_ctx *>(c);
    return 0;
}

int dlpx_destroy(dlpx_ctx *ctx)
{
    if (!ctx) return 0;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->recip_stream) (void)hipStreamSynchronize(c->recip_stream);
    if (c->ens) ensemble_free(c->ens);
    if (c->ens_multi) ensemble_free(c->ens_multi);
    void *dev[] = {c->pos, c->pos_alt, c->vel, c->vel_alt, c->frc, c->frc_recip, c->xref, c->typ, c->typ_alt,
                   c->gid, c->gid_alt, c->slot, c->stage, c->stage_i, c->cell_of, c->perm, c->cell_count,
                   c->cell_start, c->cell_fill, c->key_start, c->nbr, c->cnt, c->flags, c->d_vdw, c->d_erfc, c->d_keytab,
                   c->d_vdw_direct, c->d_comb, c->d_comb_row, c->sub_of, c->prel, c->excl_off, c->excl_ids, c->frz, c->pp, c->Q2,
                   c->Q, c->Phi, c->H, c->theta, c->tw, c->bsp_n2, c->red, c->red_out, c->acc, c->energies_dev,
                   c->snap_d, c->snap_i};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (c->h_st) (void)hipHostFree(c->h_st);
    if (c->h_t) (void)hipHostFree(c->h_t);
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->h_red) (void)hipHostFree(c->h_red);
    for (int k = 0; k < 8; k++)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->recip_stream) (void)hipStreamDestroy(c->recip_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int dlpx_set_vdw_tables(dlpx_ctx *ctx, int32_t nkeys, int32_t max_grid, const double *tab_potential,
                        const double *tab_force, const int32_t *ltp)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    const int nt = c->prm.ntypes;
    if (nkeys != nt * (nt + 1) / 2 || max_grid < 10 || !tab_potential || !tab_force || !ltp)
        return fail(c, DLPX_ERR_BAD_ARG, "vdW tables: need ntypes*(ntypes+1)/2 rows of max_grid+1 values");
    c->nkeys = nkeys;
    c->max_grid = max_grid;
    const size_t m = (size_t)nkeys * (max_grid + 1);
    c->vdw_pot.assign(tab_potential, tab_potential + m);
    c->vdw_frc.assign(tab_force, tab_force + m);
    c->ltp.assign(ltp, ltp + nkeys);
    c->gen_params.clear();
    c->vdw_direct = false;  // host tables carry no analytic forms
    c->tables_dirty = true;
    return 0;
}

int dlpx_set_erfc_tables(dlpx_ctx *ctx, int32_t nsamples, const double *erfc, const double *erfc_deriv)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    if (nsamples < 10 || !erfc || !erfc_deriv) return fail(c, DLPX_ERR_BAD_ARG, "erfc tables");
    c->nsamples = nsamples;
    c->erfc_t.assign(erfc, erfc + nsamples);
    c->erfc_d.assign(erfc_deriv, erfc_deriv + nsamples);
    c->tables_dirty = true;
    return 0;
}

int dlpx_generate_vdw_tables(dlpx_ctx *ctx, int32_t nkeys, const int32_t *kind, const double *params)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    const int nt = c->prm.ntypes;
    if (nkeys != nt * (nt + 1) / 2 || !kind || !params) return fail(c, DLPX_ERR_BAD_ARG, "vdW parameters");
    const int mg = table_size(c->prm.vdw_cutoff);
    std::vector<double> pot((size_t)nkeys * (mg + 1), 0.0), frc((size_t)nkeys * (mg + 1), 0.0);
    for (int k = 0; k < nkeys; k++) {
        if (vdw_kind_supported(kind[k])) {
            vdw_table(kind[k], params + kVdwParams * (size_t)k, c->prm.vdw_cutoff, mg,
                      pot.data() + (size_t)k * (mg + 1), frc.data() + (size_t)k * (mg + 1));
        } else if (kind[k] >= 0) {
            return fail(c, DLPX_ERR_UNSUPPORTED,
                        "vdW kind " + std::to_string(kind[k]) + " is not generated (1-24 are; 0 is tabulated input)");
        }
    }
    std::vector<int32_t> ltp(kind, kind + nkeys);
    const bool direct = c->vdw_direct;
    const int r = dlpx_set_vdw_tables(ctx, nkeys, mg, pot.data(), frc.data(), ltp.data());
    if (r == 0) {
        c->gen_params.assign(params, params + kVdwParams * (size_t)nkeys);
        c->vdw_direct = direct;  // the forms are known again
    }
    return r;
}

int dlpx_generate_erfc_tables(dlpx_ctx *ctx)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    const int ns = table_size(c->prm.cutoff);
    std::vector<double> t(ns), d(ns);
    erfc_tables(c->prm.cutoff, c->prm.alpha, ns, t.data(), d.data());
    return dlpx_set_erfc_tables(ctx, ns, t.data(), d.data());
}

int dlpx_set_lrc(dlpx_ctx *ctx, double elrc, double vlrc)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    c->elrc = elrc;
    c->vlrc = vlrc;
    return 0;
}

int dlpx_compute_lrc(dlpx_ctx *ctx, const double cell[9], const double *num_type)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !num_type) return DLPX_ERR_BAD_ARG;
    int r;
    if ((r = set_cell(c, cell))) return r;
    if (c->gen_params.empty())
        return fail(c, DLPX_ERR_STATE, "tables were not generated here: pass the host's vdws%elrc/vlrc to dlpx_set_lrc");
    vdw_lrc(c->prm.ntypes, num_type, c->ltp.data(), c->gen_params.data(), c->prm.vdw_cutoff, c->volm, &c->elrc,
            &c->vlrc);
    if (c->vdw_force_shift) c->elrc = c->vlrc = 0.0;  // vdw.F90:662: no correction with force-shifting
    return 0;
}

int dlpx_set_vdw_force_shift(dlpx_ctx *ctx, int32_t on)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || (on != 0 && on != 1)) return DLPX_ERR_BAD_ARG;
    c->vdw_force_shift = on != 0;
    if (c->vdw_force_shift) c->elrc = c->vlrc = 0.0;
    c->tables_dirty = true;  // the direct shift constants depend on it
    return 0;
}

int dlpx_set_vdw_direct(dlpx_ctx *ctx, int32_t on)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || (on != 0 && on != 1)) return DLPX_ERR_BAD_ARG;
    if (on && c->gen_params.empty())
        return fail(c, DLPX_ERR_STATE, "vdw_method direct needs the forms: call dlpx_generate_vdw_tables first");
    c->vdw_direct = on != 0;
    c->tables_dirty = true;
    return 0;
}

int dlpx_set_exclusions(dlpx_ctx *ctx, int32_t natms, int32_t mxexcl, const int32_t *list_excl)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    if (c->resident) return fail(c, DLPX_ERR_STATE, "set the exclusions before dlpx_md_load");
    if (natms < 0 || natms > c->cap || mxexcl < 0 || (natms > 0 && !list_excl))
        return fail(c, DLPX_ERR_BAD_ARG, "exclusions: natms outside 0..natms_max, negative mxexcl or no array");
    // rows as sorted, duplicate-free host indices (build_excl.F90 add_exclusion keeps one copy)
    std::vector<int32_t> off(natms + 1, 0), ids;
    const size_t ld = (size_t)mxexcl + 1;
    for (int i = 0; i < natms; i++) {
        const int32_t *row = list_excl + ld * i;
        const int m = row[0];
        if (m < 0 || m > mxexcl) return fail(c, DLPX_ERR_BAD_ARG, "exclusions: count outside 0..mxexcl");
        std::vector<int32_t> r;
        for (int k = 1; k <= m; k++) {
            const int j = row[k] - 1;
            if (j < 0 || j >= natms || j == i)
                return fail(c, DLPX_ERR_BAD_ARG, "exclusions: atom " + std::to_string(i + 1) + " lists id " +
                                                     std::to_string(row[k]));
            r.push_back(j);
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        ids.insert(ids.end(), r.begin(), r.end());
        off[i + 1] = (int32_t)ids.size();
    }
    // a pair is dropped from whichever end holds it, so both ends must list it
    for (int i = 0; i < natms; i++)
        for (int k = off[i]; k < off[i + 1]; k++) {
            const int j = ids[k];
            if (!std::binary_search(ids.begin() + off[j], ids.begin() + off[j + 1], i))
                return fail(c, DLPX_ERR_BAD_ARG, "exclusions are not symmetric: " + std::to_string(i + 1) +
                                                     " excludes " + std::to_string(j + 1) + " but not the reverse");
        }
    if (c->excl_off) (void)hipFree(c->excl_off);
    if (c->excl_ids) (void)hipFree(c->excl_ids);
    c->excl_off = c->excl_ids = nullptr;
    c->excl_natms = c->excl_total = 0;
    c->have_list = false;  // the list was built under the old exclusions
    if (ids.empty()) return 0;
    int r;
    if ((r = dalloc(c, &c->excl_off, off.size())) || (r = dalloc(c, &c->excl_ids, ids.size()))) return r;
    DLPX_HIP(c, dev_upload(c->stream, c->excl_off, off.data(), sizeof(int32_t) * off.size()));
    DLPX_HIP(c, dev_upload(c->stream, c->excl_ids, ids.data(), sizeof(int32_t) * ids.size()));
    c->excl_natms = natms;
    c->excl_total = (int)ids.size();
    return 0;
}

int dlpx_set_frozen(dlpx_ctx *ctx, int32_t natms, const int32_t *lfrzn)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    if (c->resident) return fail(c, DLPX_ERR_STATE, "set the frozen atoms before dlpx_md_load");
    if (natms < 0 || natms > c->cap || (natms > 0 && !lfrzn))
        return fail(c, DLPX_ERR_BAD_ARG, "frozen flags: natms outside 0..natms_max or no array");
    std::vector<int32_t> f(natms);
    int nf = 0;
    for (int i = 0; i < natms; i++) {
        f[i] = lfrzn[i] != 0 ? 1 : 0;
        nf += f[i];
    }
    if (c->frz) (void)hipFree(c->frz);
    c->frz = nullptr;
    c->frz_natms = c->nfrozen = 0;
    c->have_list = false;  // the list was built without these flags
    if (nf == 0) return 0;
    int r;
    if ((r = dalloc(c, &c->frz, (size_t)natms))) return r;
    DLPX_HIP(c, dev_upload(c->stream, c->frz, f.data(), sizeof(int32_t) * natms));
    c->frz_natms = natms;
    c->nfrozen = nf;
    return 0;
}

int dlpx_compute_lrc_frozen(dlpx_ctx *ctx, const double cell[9], const double *num_type, const double *num_frozen)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !num_type || !num_frozen) return DLPX_ERR_BAD_ARG;
    int r;
    if ((r = set_cell(c, cell))) return r;
    if (c->gen_params.empty())
        return fail(c, DLPX_ERR_STATE, "tables were not generated here: pass the host's vdws%elrc/vlrc to dlpx_set_lrc");
    vdw_lrc(c->prm.ntypes, num_type, c->ltp.data(), c->gen_params.data(), c->prm.vdw_cutoff, c->volm, &c->elrc,
            &c->vlrc, num_frozen);
    if (c->vdw_force_shift) c->elrc = c->vlrc = 0.0;  // vdw.F90:662
    return 0;
}

int dlpx_set_per_particle(dlpx_ctx *ctx, int32_t on)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || (on != 0 && on != 1)) return DLPX_ERR_BAD_ARG;
    if (on && c->pair_lanes != 8)
        return fail(c, DLPX_ERR_UNSUPPORTED, "per-particle output runs with 8 pair lanes (DLPX_PAIR_LANES)");
    c->per_particle = on != 0;
    c->pp_valid = false;
    return 0;
}

int dlpx_add_per_particle(dlpx_ctx *ctx, double *energy, double *stress)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    if (!c->pp_valid)
        return fail(c, DLPX_ERR_STATE, "no per-particle data: dlpx_set_per_particle(ctx, 1), then dlpx_two_body_forces");
    const int n = c->natms;
    std::vector<double> pp((size_t)n * kPPStride);
    std::vector<double4> fr(c->prm.electrostatics ? n : 0);
    std::vector<int32_t> gid(n);
    DLPX_HIP(c, hipMemcpy(pp.data(), c->pp, sizeof(double) * pp.size(), hipMemcpyDeviceToHost));
    DLPX_HIP(c, hipMemcpy(gid.data(), c->gid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (c->prm.electrostatics)
        DLPX_HIP(c, hipMemcpy(fr.data(), c->frc_recip, sizeof(double4) * n, hipMemcpyDeviceToHost));
    // ewald_spole.F90:429-431,464-466 (reciprocal energy and stress scaled,
    // the self energy shared evenly) and two_body.F90:783 (the vdW long-range
    // correction shared evenly); the exclusion correction and the net-charge
    // term have no per-particle part in the reference
    const double scale = c->prm.electrostatics ? kPi * std::sqrt(kPi) * std::pow(c->prm.alpha, 1 - 3) *
                                                     (0.5 * (1.0 / c->volm)) * (kR4pie0 / c->prm.epsilon)
                                               : 0.0;
    const double share = (c->prm.electrostatics ? c->self_energy : 0.0) / n + c->elrc / n;
    const int sym[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};  // 3x3 element -> (xx, xy, xz, yy, yz, zz)
    for (int i = 0; i < n; i++) {
        const double *p = pp.data() + (size_t)i * kPPStride;
        const int h = gid[i];
        if (energy) energy[h] += p[0] + (c->prm.electrostatics ? fr[i].w * scale : 0.0) + share;
        if (stress)
            for (int k = 0; k < 9; k++)
                stress[(size_t)9 * h + k] += p[1 + sym[k]] + (c->prm.electrostatics ? p[8 + sym[k]] * scale : 0.0);
    }
    return 0;
}

int dlpx_get_lrc(dlpx_ctx *ctx, double *elrc, double *vlrc)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !elrc || !vlrc) return DLPX_ERR_BAD_ARG;
    *elrc = c->elrc;
    *vlrc = c->vlrc;
    return 0;
}

int dlpx_link_cell_pairs(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms, int64_t *npairs,
                         int32_t *ibig)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    int r;
    if ((r = set_cell(c, cell))) return r;
    c->resident = false;
    if ((r = upload(c, atoms, true, nullptr, nullptr, nullptr, nullptr))) return r;
    r = build_list(c);
    if (ibig) *ibig = c->h_flags[1];
    if (r) return r;
    // tables already given go to the device with the list, so that the force
    // calls after it need no allocation of their own
    const int nt = c->prm.ntypes;
    if ((c->tables_dirty || c->comb_dirty) && c->nkeys == nt * (nt + 1) / 2 &&
        (!c->prm.electrostatics || c->nsamples > 0) && (r = upload_tables(c)))
        return r;
    if (npairs) *npairs = (int64_t)c->h_flags[3] / 2;
    return 0;
}

int dlpx_get_pairs(dlpx_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *npairs)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    if (!c->have_list) return fail(c, DLPX_ERR_STATE, "no neighbour list");
    const int n = c->natms, G = kLanesPerAtom;
    std::vector<int32_t> nbr((size_t)c->cap_lane * c->cap * G), cnt((size_t)n * G), gid(n);
    std::vector<double4> pos(n);
    DLPX_HIP(c, hipMemcpy(nbr.data(), c->nbr, sizeof(int32_t) * nbr.size(), hipMemcpyDeviceToHost));
    DLPX_HIP(c, hipMemcpy(cnt.data(), c->cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
    DLPX_HIP(c, hipMemcpy(gid.data(), c->gid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    DLPX_HIP(c, hipMemcpy(pos.data(), c->xref, sizeof(double4) * n, hipMemcpyDeviceToHost));
    int64_t np = 0;
    for (int i = 0; i < n; i++)
        for (int g = 0; g < G; g++)
            for (int k = 0; k < std::min(cnt[(size_t)i * G + g], c->cap_lane); k++) {
                const int j = list_j(nbr[list_slot(i, k, g, c->cap_lane)]);
                const int hi = gid[i], hj = gid[j];
                if (hi >= hj) continue;
                if (np < cap && pairs) {
                    int32_t *o = pairs + 5 * np;
                    o[0] = hi;
                    o[1] = hj;
                    const double d[3] = {pos[i].x - pos[j].x, pos[i].y - pos[j].y, pos[i].z - pos[j].z};
                    for (int a = 0; a < 3; a++) o[2 + a] = (int32_t)std::rint(d[a] * c->invL[a]);
                }
                np++;
            }
    if (npairs) *npairs = np;
    return 0;
}

int dlpx_count_pairs(dlpx_ctx *ctx, double r, int64_t *npairs)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !npairs || !(r > 0.0)) return DLPX_ERR_BAD_ARG;
    if (!c->have_list) return fail(c, DLPX_ERR_STATE, "no neighbour list");
    DLPX_HIP(c, hipStreamSynchronize(c->stream));
    return count_pairs_dev(c, r, npairs);
}

int dlpx_vnl_check(dlpx_ctx *ctx, const dlpx_atoms *atoms, int32_t *update)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !update) return DLPX_ERR_BAD_ARG;
    if (!c->have_list) {
        *update = 1;
        return 0;
    }
    int r;
    if ((r = upload(c, atoms, false, nullptr, nullptr, nullptr, nullptr))) return r;
    int u = 0;
    if ((r = vnl_check_dev(c, &u))) return r;
    *update = u;
    return 0;
}

int dlpx_two_body_forces(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms, dlpx_two_body_out *out)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out) return DLPX_ERR_BAD_ARG;
    if (!c->have_list) return fail(c, DLPX_ERR_STATE, "call dlpx_link_cell_pairs first");
    if (!atoms->fx || !atoms->fy || !atoms->fz) return fail(c, DLPX_ERR_BAD_ARG, "force arrays");
    int r;
    if ((r = set_cell(c, cell))) return r;
    if ((r = upload(c, atoms, false, nullptr, nullptr, nullptr, nullptr))) return r;
    charges_summary(c, atoms);
    c->pp_valid = false;
    if (c->per_particle) {
        // per-atom records and the stress-grid copy, allocated on first use
        if (c->cap > c->pp_cap) {
            if (c->pp) (void)hipFree(c->pp);
            c->pp = nullptr;
            c->pp_cap = 0;
            if ((r = dalloc(c, &c->pp, (size_t)c->cap * kPPStride))) return r;
            c->pp_cap = c->cap;
        }
        if (c->prm.electrostatics && !c->Q2) {
            const size_t rs = c->prm.precision == DLPX_MIXED ? sizeof(float) : sizeof(double);
            DLPX_HIP(c, hipMalloc(&c->Q2, rs * c->nreal));
        }
        c->pp_now = true;
    }
    r = evaluate(c);
    c->pp_now = false;
    if (r) return r;
    c->pp_valid = c->per_particle;
    if ((r = store_atoms(c, c->stage, false, false))) return r;
    const int n = c->natms;
    DLPX_HIP(c, hipMemcpyAsync(c->h_st, c->stage, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
    if ((r = fetch_red(c))) return r;
    const int st = atoms->stride;
    for (int i = 0; i < n; i++) {
        const double *f = c->h_st + (size_t)9 * i + 6;
        *(double *)((char *)atoms->fx + (size_t)st * i) += f[0];
        *(double *)((char *)atoms->fy + (size_t)st * i) += f[1];
        *(double *)((char *)atoms->fz + (size_t)st * i) += f[2];
    }
    assemble(c, c->h_red, out);
    return 0;
}

int dlpx_two_body_forces_noalloc(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms,
                                 dlpx_two_body_out *out)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out || !atoms) return DLPX_ERR_BAD_ARG;
    if (!c->have_list) return fail(c, DLPX_ERR_STATE, "call dlpx_link_cell_pairs first");
    if (!atoms->fx || !atoms->fy || !atoms->fz) return fail(c, DLPX_ERR_BAD_ARG, "force arrays");
    if (c->per_particle)
        return fail(c, DLPX_ERR_UNSUPPORTED, "per-particle output needs buffers of its own: use dlpx_two_body_forces");
    if (c->tables_dirty)
        return fail(c, DLPX_ERR_STATE, "tables changed since the list build: call dlpx_link_cell_pairs first");
    double *scr = scratch_of(c);
    if (!scr) return fail(c, DLPX_ERR_STATE, "no scratch address for this device");
    int r;
    if ((r = set_cell(c, cell))) return r;
    if ((r = upload(c, atoms, false, nullptr, nullptr, nullptr, nullptr))) return r;
    const int rows = std::max({pair_blocks(c), gather_blocks(c), conv_blocks(c), excl_blocks(c)});
    if (rows > kScratchRows)
        return fail(c, DLPX_ERR_BAD_ARG,
                    "the evaluation needs " + std::to_string(rows) + " partial-sum rows, beyond the " +
                        std::to_string(kScratchRows) + " of the allocation-free scratch (about " +
                        std::to_string(32 * kScratchRows) + " atoms at spline order 8)");
    charges_summary(c, atoms);
    c->pp_valid = false;
    // the evaluation writes its partial sums and totals into the scratch: every
    // reducing kernel writes each row it owns before the reduction reads it,
    // so nothing needs clearing between calls
    double *const red = c->red, *const red_out = c->red_out;
    const int red_cap = c->red_cap;
    c->red = scr;
    c->red_cap = kScratchRows;
    c->red_out = scr + (size_t)kScratchRows * kReduceSlots;
    c->noalloc_now = true;  // no table upload: stale combined rows fall back to the two-table pair path
    r = evaluate(c);
    c->noalloc_now = false;
    if (!r) r = store_atoms(c, c->stage, false, false);
    if (!r) {
        const hipError_t e = hipMemcpyAsync(c->h_st, c->stage, sizeof(double) * 9 * c->natms,
                                            hipMemcpyDeviceToHost, c->stream);
        r = e != hipSuccess ? fail(c, DLPX_ERR_HIP + (int)e, std::string("hipMemcpyAsync: ") + hipGetErrorString(e))
                            : fetch_red(c);  // waits for the stream: the scratch is free again on return
    }
    c->red = red;
    c->red_cap = red_cap;
    c->red_out = red_out;
    if (r) {
        (void)hipStreamSynchronize(c->stream);
        return r;
    }
    const int st = atoms->stride;
    for (int i = 0; i < c->natms; i++) {
        const double *f = c->h_st + (size_t)9 * i + 6;
        *(double *)((char *)atoms->fx + (size_t)st * i) += f[0];
        *(double *)((char *)atoms->fy + (size_t)st * i) += f[1];
        *(double *)((char *)atoms->fz + (size_t)st * i) += f[2];
    }
    assemble(c, c->h_red, out);
    return 0;
}

int dlpx_md_load(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms, const double *vx, const double *vy,
                 const double *vz, const double *mass, int32_t vstride)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !vx || !vy || !vz || !mass || vstride < 0) return DLPX_ERR_BAD_ARG;
    int r;
    if ((r = set_cell(c, cell))) return r;
    c->resident = true;
    c->have_last_out = false;
    if ((r = upload(c, atoms, true, vx, vy, vz, mass, vstride))) return r;
    charges_summary(c, atoms);
    if ((r = build_list(c))) return r;
    // forces at the starting configuration (md_vv computes them before step 1)
    if ((r = evaluate(c))) return r;
    if ((r = freeze_atoms(c))) return r;
    if ((r = fetch_red(c))) return r;
    // the starting configuration's output, until a run with energies replaces it
    assemble(c, c->h_red, &c->last_out);
    c->have_last_out = true;
    return 0;
}

// Room for nsteps per-step records of the two-body output.
static int energies_room(Ctx *c, int nsteps)
{
    if (nsteps <= c->energies_cap) return 0;
    if (c->energies_dev) (void)hipFree(c->energies_dev);
    c->energies_dev = nullptr;
    c->energies_cap = 0;
    DLPX_HIP(c, hipMalloc((void **)&c->energies_dev, sizeof(double) * kReduceSlots * (size_t)nsteps));
    c->energies_cap = nsteps;
    return 0;
}

// The per-step records of a run as (engcpe, vircpe, engsrp, virsrp, kinetic)
// per step; the last step's output is kept for dlpx_md_last_output.
static int collect_energies(Ctx *c, int nsteps, double *energies)
{
    if (nsteps <= 0) return 0;
    std::vector<double> red((size_t)kReduceSlots * nsteps);
    DLPX_HIP(c, hipMemcpy(red.data(), c->energies_dev, sizeof(double) * red.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < nsteps; s++) {
        dlpx_two_body_out o;
        assemble(c, red.data() + (size_t)kReduceSlots * s, &o);
        double *e = energies + (size_t)5 * s;
        e[0] = o.engcpe;
        e[1] = o.vircpe;
        e[2] = o.engsrp;
        e[3] = o.virsrp;
        e[4] = 0.5 * red[(size_t)kReduceSlots * s + R_KIN];
        if (s == nsteps - 1) {
            c->last_out = o;
            c->have_last_out = true;
        }
    }
    return 0;
}

int dlpx_md_run(dlpx_ctx *ctx, int32_t nsteps, double tstep, double *energies, int32_t *nrebuilds)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || nsteps < 0) return DLPX_ERR_BAD_ARG;
    if (!c->resident || !c->have_list) return fail(c, DLPX_ERR_STATE, "call dlpx_md_load first");
    int r, rebuilds = 0;
    if (energies && (r = energies_room(c, nsteps))) return r;
    c->run_dt = tstep;
    if (c->use_graphs && !c->timing) {
        // device-decided rebuilds and captured steps: a batch of one
        std::vector<Ctx *> cs{c};
        std::vector<int> reb;
        Ensemble *E = reinterpret_cast<Ensemble *>(c->ens);
        r = ensemble_run(cs, E, nsteps, energies != nullptr, reb);
        c->ens = E;
        if (r) return r;
        rebuilds = reb.empty() ? 0 : reb[0];
    } else {
        // stage timing (or DLPX_GRAPHS=0): eager steps, host-decided rebuilds
        if ((r = md_begin(c, nsteps, tstep))) return r;
        for (int s = 0; s < nsteps; s++)
            if ((r = md_step(c, s, nsteps, tstep, energies != nullptr, rebuilds))) return r;
        DLPX_HIP(c, hipStreamSynchronize(c->stream));
    }
    if (energies && (r = collect_energies(c, nsteps, energies))) return r;
    if (nrebuilds) *nrebuilds = rebuilds;
    return 0;
}

int dlpx_md_last_output(dlpx_ctx *ctx, dlpx_two_body_out *out)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out) return DLPX_ERR_BAD_ARG;
    if (!c->have_last_out) return fail(c, DLPX_ERR_STATE, "no resident step with energies has run");
    *out = c->last_out;
    return 0;
}

int dlpx_md_run_ensemble_energies(dlpx_ctx *const *ctxs, int32_t nctx, int32_t nsteps, const double *tsteps,
                                  double *energies, int32_t *nrebuilds)
{
    if (!ctxs || nctx < 0 || nsteps < 0 || !tsteps) return DLPX_ERR_BAD_ARG;
    std::vector<Ctx *> cs(nctx);
    for (int k = 0; k < nctx; k++) {
        cs[k] = reinterpret_cast<Ctx *>(ctxs[k]);
        if (!cs[k]) return DLPX_ERR_BAD_ARG;
        if (!cs[k]->resident || !cs[k]->have_list) return fail(cs[k], DLPX_ERR_STATE, "call dlpx_md_load first");
        if (cs[k]->device != cs[0]->device) return fail(cs[k], DLPX_ERR_BAD_ARG, "ensemble contexts on different devices");
        for (int j = 0; j < k; j++)
            if (cs[j] == cs[k]) return fail(cs[k], DLPX_ERR_BAD_ARG, "a context appears twice in the ensemble");
    }
    if (nctx == 0) return 0;
    std::vector<int> rebuilds(nctx, 0);
    int r;
    bool batched = true;
    for (int k = 0; k < nctx; k++) {
        cs[k]->run_dt = tsteps[k];
        if (energies && (r = energies_room(cs[k], nsteps))) return r;
        // stage timing on the first context times the batched steps eagerly
        batched = batched && cs[k]->use_graphs && (k == 0 || !cs[k]->timing) &&
                  cs[k]->prm.precision == cs[0]->prm.precision;
    }
    if (batched) {
        // every stage one launch over all systems, steps replaying one graph
        Ensemble *E = reinterpret_cast<Ensemble *>(cs[0]->ens_multi);
        r = ensemble_run(cs, E, nsteps, energies != nullptr, rebuilds);
        cs[0]->ens_multi = E;
        if (r) return r;
    } else {
        // eager lockstep over systems with host-decided rebuilds
        for (int k = 0; k < nctx; k++)
            if ((r = md_begin(cs[k], nsteps, tsteps[k]))) return r;
        for (int s = 0; s < nsteps; s++)
            for (int k = 0; k < nctx; k++)
                if ((r = md_step(cs[k], s, nsteps, tsteps[k], energies != nullptr, rebuilds[k]))) return r;
        for (int k = 0; k < nctx; k++) DLPX_HIP(cs[k], hipStreamSynchronize(cs[k]->stream));
    }
    for (int k = 0; k < nctx; k++) {
        if (energies && (r = collect_energies(cs[k], nsteps, energies + (size_t)5 * nsteps * k))) return r;
        if (nrebuilds) nrebuilds[k] = rebuilds[k];
    }
    return 0;
}

int dlpx_md_run_ensemble(dlpx_ctx *const *ctxs, int32_t nctx, int32_t nsteps, const double *tsteps,
                         int32_t *nrebuilds)
{
    return dlpx_md_run_ensemble_energies(ctxs, nctx, nsteps, tsteps, nullptr, nrebuilds);
}

int dlpx_md_store(dlpx_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz, double *fx,
                  double *fy, double *fz, int32_t stride, int32_t vstride)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || stride < 0 || vstride < 0) return DLPX_ERR_BAD_ARG;
    if (!c->resident) return fail(c, DLPX_ERR_STATE, "call dlpx_md_load first");
    int r;
    if ((r = store_atoms(c, c->stage, true, true))) return r;
    const int n = c->natms;
    DLPX_HIP(c, hipMemcpyAsync(c->h_st, c->stage, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
    DLPX_HIP(c, hipStreamSynchronize(c->stream));
    double *dst[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
    for (int k = 0; k < 9; k++) {
        if (!dst[k]) continue;
        const size_t st = (k >= 3 && k < 6 && vstride > 0) ? vstride : stride;
        for (int i = 0; i < n; i++) *(double *)((char *)dst[k] + st * i) = c->h_st[(size_t)9 * i + k];
    }
    return 0;
}

int dlpx_set_timing(dlpx_ctx *ctx, int32_t enable)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return DLPX_ERR_BAD_ARG;
    c->timing = enable != 0;
    c->timing_pending = false;
    c->timed_evals = 0;
    for (double &v : c->stage_ms) v = 0.0;
    return 0;
}

int dlpx_get_timing(dlpx_ctx *ctx, double *ms, int32_t n)
{
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !ms) return DLPX_ERR_BAD_ARG;
    for (int k = 0; k < n && k < 8; k++) ms[k] = c->stage_ms[k];
    if (n > 7) ms[7] = (double)c->timed_evals;
    return 0;
}

}  // extern "C"

// Resident MD over a batch of systems: every stage of a step is one launch
// over all systems (per-system argument structs in device memory, see
// batch_find), the list rebuild is decided on the device, and steps replay
// one captured graph with no host round trip.  A single context's resident
// run is a batch of one.  Also the autotuner that picks the launch shapes.
//
// A step, as in md_vv (drivers.F90:2059-2290): the list build of each system
// whose stage 1 decided a rebuild (vnl_check, neighbours.F90:123-296; the
// other systems' build kernels return at once), the two-body forces (pair
// kernel beside the reciprocal-space chain on a second stream,
// two_body.F90:76-796), the partial-sum reductions, stage 2 (with the
// reciprocal finish), and stage 1 of the next step.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace dlpx {

std::vector<uint64_t> step_graph_key(const Ctx *c, double tstep);
int md_fallback(Ctx *c, int nsteps, double tstep, bool energies, int &rebuilds);

int batch_upload(Ctx *c, BatchLaunch &bl, const void *args, size_t bytes_each, const std::vector<int> &nblocks,
                 bool xcd)
{
    batch_free(bl);
    const int nsys = (int)nblocks.size();
    std::vector<int32_t> off(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) off[s + 1] = off[s] + nblocks[s];
    bl.nsys = nsys;
    bl.total = xcd ? xcd_grid(off[nsys]) : off[nsys];
    DLPX_HIP(c, hipMalloc(&bl.args, std::max<size_t>(bytes_each * nsys, 1)));
    DLPX_HIP(c, hipMalloc((void **)&bl.boff, sizeof(int32_t) * (nsys + 1)));
    DLPX_HIP(c, dev_upload(c->stream, bl.args, args, bytes_each * nsys));
    DLPX_HIP(c, dev_upload(c->stream, bl.boff, off.data(), sizeof(int32_t) * (nsys + 1)));
    return 0;
}

void batch_free(BatchLaunch &bl)
{
    if (bl.args) (void)hipFree(bl.args);
    if (bl.boff) (void)hipFree(bl.boff);
    bl.args = nullptr;
    bl.boff = nullptr;
    bl.nsys = bl.total = 0;
}

// Launch-shape choices of a batch: lanes per atom of the pair kernel and the
// largest FFT lines per workgroup.
struct Shape {
    int pair_lanes = 8;
    int fft_lpb = 16;
};

struct Ensemble {
    std::vector<Ctx *> cs;
    std::vector<uint64_t> key;
    Shape shape;
    bool energies = false;
    bool electro = false;
    BatchLaunch step, excl;
    ListBatch list;
    PairBatch pair;
    SpmeBatch spme;
    hipGraphExec_t graph = nullptr;   // one step
    hipGraphExec_t graph_k = nullptr; // kGraphSteps steps
    hipEvent_t fork = nullptr, join = nullptr, spread = nullptr;

    ~Ensemble() { release(); }
    void release()
    {
        for (BatchLaunch *b : {&step, &excl, &spme.spread, &spme.gather}) batch_free(*b);
        for (auto &g : spme.fft.groups) batch_free(g.bl);
        spme.fft.groups.clear();
        for (auto &g : pair.groups) batch_free(g.bl);
        batch_free(pair.prel);
        pair.groups.clear();
        list_batch_free(list);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_k) (void)hipGraphExecDestroy(graph_k);
        graph = graph_k = nullptr;
        for (hipEvent_t e : {fork, join, spread})
            if (e) (void)hipEventDestroy(e);
        fork = join = spread = nullptr;
    }
};

static std::atomic<uint64_t> g_serial{1};
uint64_t next_serial() { return g_serial.fetch_add(1); }

void ensemble_free(void *e) { delete reinterpret_cast<Ensemble *>(e); }

static std::vector<uint64_t> ens_key(const std::vector<Ctx *> &cs, bool energies, const Shape &sh)
{
    std::vector<uint64_t> k = {(uint64_t)cs.size(), energies ? 1ull : 0ull, (uint64_t)sh.pair_lanes,
                               (uint64_t)sh.fft_lpb};
    for (Ctx *c : cs) {
        k.push_back(c->serial);
        const std::vector<uint64_t> g = step_graph_key(c, c->run_dt);
        k.insert(k.end(), g.begin(), g.end());
        k.push_back((uint64_t)c->energies_dev);
        k.push_back((uint64_t)c->energies_cap);
        k.push_back((uint64_t)c->vdw_force_shift);
        k.push_back((uint64_t)c->vdw_direct);
        k.push_back((uint64_t)c->nrows);
    }
    return k;
}

// Argument arrays of every stage for the ensemble's systems and shape.
static int ens_prepare(Ensemble &E, bool energies, const Shape &sh)
{
    E.release();
    E.key.clear();
    E.energies = energies;
    E.shape = sh;
    const std::vector<Ctx *> &cs = E.cs;
    Ctx *c0 = cs[0];
    int r;
    E.electro = false;
    for (Ctx *c : cs) {
        if (c->prm.precision != c0->prm.precision)
            return fail(c, DLPX_ERR_UNSUPPORTED, "systems of one batch must share the precision mode");
        E.electro = E.electro || c->prm.electrostatics;
    }
    // room for every stage's partial-sum rows before any argument struct
    // records the buffer's address (upper bounds: the z pass has at most
    // K1 K2 workgroups, the pair kernel at most natms * 16 / 512)
    for (Ctx *c : cs) {
        const int ub = std::max({ceil_div((long)c->natms * 16, 512), step_blocks(c), excl_blocks(c),
                                 c->prm.electrostatics ? c->K[0] * c->K[1] : 1,
                                 c->prm.electrostatics ? ceil_div(c->natms, 16) : 1});
        if ((r = ensure_red_rows(c, ub))) return r;
    }
    std::vector<int> npair;
    if ((r = list_prepare(cs, E.list)) || (r = excl_prepare(cs, E.excl)) ||
        (r = pair_prepare(cs, E.pair, sh.pair_lanes, npair)))
        return r;
    E.spme.fft.lpb_cap = sh.fft_lpb;
    if (E.electro && (r = spme_prepare(cs, E.spme))) return r;
    std::vector<RedRowsHost> rows(cs.size());
    for (size_t k = 0; k < cs.size(); k++) {
        Ctx *c = cs[k];
        for (int q = 0; q < kReduceSlots; q++) rows[k].rows[q] = 0;
        for (int q = R_ENGVDW; q <= R_S33; q++) rows[k].rows[q] = npair[k];
        rows[k].rows[R_KIN] = step_blocks(c);
        const int nex = (c->excl_total > 0 && c->prm.electrostatics) ? excl_blocks(c) : 0;
        for (int q = R_ENGEX; q <= R_XS33; q++) rows[k].rows[q] = nex;
        if (c->prm.electrostatics) {
            for (int q = R_ERECIP; q <= R_FNETZ; q++) rows[k].rows[q] = E.spme.gather_rows[k];
            for (int q = R_RS11; q <= R_RS33; q++) rows[k].rows[q] = E.spme.fft.conv_rows[k];
        }
        for (int q = 0; q < kReduceSlots; q++)
            if (rows[k].rows[q] > c->red_cap) return fail(c, DLPX_ERR_STATE, "partial-sum rows exceed the buffer");
    }
    if ((r = step_prepare(cs, rows, energies, E.step))) return r;
    DLPX_HIP(c0, hipEventCreateWithFlags(&E.fork, hipEventDisableTiming));
    DLPX_HIP(c0, hipEventCreateWithFlags(&E.join, hipEventDisableTiming));
    DLPX_HIP(c0, hipEventCreateWithFlags(&E.spread, hipEventDisableTiming));
    E.key = ens_key(cs, energies, sh);
    return 0;
}

// Queue one step on the first context's streams: list builds (gated),
// forces, then one kernel for stage 2, the step's reductions and (with vv1)
// stage 1 of the next step.
static int ens_enqueue(Ensemble &E, bool with_vv1)
{
    Ctx *c0 = E.cs[0];
    hipStream_t s = c0->stream, s2 = c0->recip_stream;
    int r;
    if (c0->timing) {
        // stage timing (eager): every stage on one stream between the first
        // context's events, [6]-[7] list, [0]-[1] pair, [1]-[2] spreading,
        // [2]-[3] convolution, [3]-[4] gather, [4]-[5] step kernel
        DLPX_HIP(c0, hipEventRecord(c0->ev[6], s));
        if ((r = list_launch(c0, s, E.list))) return r;
        DLPX_HIP(c0, hipEventRecord(c0->ev[7], s));
        DLPX_HIP(c0, hipEventRecord(c0->ev[0], s));
        if ((r = pair_launch(c0, s, E.pair)) || (r = excl_launch(c0, s, E.excl))) return r;
        DLPX_HIP(c0, hipEventRecord(c0->ev[1], s));
        if (E.electro) {
            if ((r = spme_launch(c0, s, E.spme, c0->ev + 2))) return r;
        } else {
            for (int k = 2; k <= 4; k++) DLPX_HIP(c0, hipEventRecord(c0->ev[k], s));
        }
        if ((r = step_launch(c0, s, E.step, true, with_vv1))) return r;
        DLPX_HIP(c0, hipEventRecord(c0->ev[5], s));
        return 0;
    }
    // the binning and sort of a due rebuild; the reciprocal-space chain needs
    // no more (positions and link cells), so the branches fork here.
    // Ctx::overlap: 2 the chain on the second stream beside the list build
    // and the pair kernel; 1 the chain on the first stream and the pair
    // kernel held until the spreading is done (both lean on LDS); 0 the chain
    // on the first stream, the pair kernel free.  Default: 2 for one system,
    // 0 for several (measured: 64k ions 0.229 ms/step with 2 against 0.236
    // and 0.241 with 1 and 0; 256 systems 4.92 ms/step with 0 against 5.19
    // with 2).  The graph executor maps branches onto its own queues, so the
    // layout changes which kernels it starts together.
    if ((r = list_launch(c0, s, E.list, 0))) return r;
    if (!E.electro) {
        if ((r = list_launch(c0, s, E.list, 1)) || (r = pair_launch(c0, s, E.pair)) || (r = excl_launch(c0, s, E.excl)))
            return r;
        return step_launch(c0, s, E.step, true, with_vv1);
    }
    DLPX_HIP(c0, hipEventRecord(E.fork, s));
    DLPX_HIP(c0, hipStreamWaitEvent(s2, E.fork, 0));
    const int ov = c0->overlap >= 0 ? c0->overlap : (E.cs.size() == 1 ? 2 : 0);
    hipStream_t sr = ov == 2 ? s2 : s, sl = ov == 2 ? s : s2;  // reciprocal chain, list + pair
    if (ov == 2 && (r = spme_launch(c0, sr, E.spme))) return r;
    if ((r = list_launch(c0, sl, E.list, 1))) return r;
    if (ov == 1) {
        hipEvent_t marks[3] = {E.spread, nullptr, nullptr};
        if ((r = spme_launch(c0, sr, E.spme, marks))) return r;
        DLPX_HIP(c0, hipStreamWaitEvent(sl, E.spread, 0));
    }
    if ((r = pair_launch(c0, sl, E.pair)) || (r = excl_launch(c0, sl, E.excl))) return r;
    if (ov == 0 && (r = spme_launch(c0, sr, E.spme))) return r;
    // the second stream joins the first
    DLPX_HIP(c0, hipEventRecord(E.join, s2));
    DLPX_HIP(c0, hipStreamWaitEvent(s, E.join, 0));
    return step_launch(c0, s, E.step, true, with_vv1);
}


// Capture nsteps consecutive steps (each with the next step's stage 1) into
// one graph.
static int ens_capture(Ensemble &E, hipGraphExec_t &out, int nsteps)
{
    Ctx *c0 = E.cs[0];
    if (out) (void)hipGraphExecDestroy(out);
    out = nullptr;
    DLPX_HIP(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
    int r = 0;
    for (int k = 0; k < nsteps && !r; k++) r = ens_enqueue(E, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c0->stream, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(c0, DLPX_ERR_HIP + (int)e, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        out = nullptr;
        return fail(c0, DLPX_ERR_HIP + (int)ei, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return 0;
}

// ---- autotuner -----------------------------------------------------------
// Candidate shapes are timed on real steps of a run (each candidate is a
// correct implementation, so the trajectory is unaffected beyond summation
// order) and the fastest is kept per size class for the process, and in a
// text cache file when DLPX_TUNE_CACHE names one.
static std::mutex g_tune_mu;
static std::map<std::string, Shape> g_tune;
static bool g_tune_loaded = false;

static std::string size_class(const std::vector<Ctx *> &cs)
{
    // systems by precision, electrostatics, spline order and the octave of
    // their atom count; the batch by the total atom count's half-octave
    long tot = 0;
    std::map<std::string, int> parts;
    for (Ctx *c : cs) {
        tot += c->natms;
        char b[96];
        std::snprintf(b, sizeof b, "p%d.e%d.s%d.n%d", c->prm.precision, c->prm.electrostatics, c->prm.nsplines,
                      (int)std::floor(std::log2((double)std::max(c->natms, 1))));
        parts[b]++;
    }
    std::string k = "v1.gfx950.T" + std::to_string((int)std::floor(2.0 * std::log2((double)std::max(tot, 1L))));
    for (auto &p : parts) k += "." + p.first + "x" + std::to_string(p.second > 8 ? 9 : p.second);
    return k;
}

static void tune_load()
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *f = std::getenv("DLPX_TUNE_CACHE");
    if (!f) return;
    FILE *fp = std::fopen(f, "r");
    if (!fp) return;
    char key[512];
    int a, b;
    while (std::fscanf(fp, "%511s %d %d", key, &a, &b) == 3) g_tune[key] = Shape{a, b};
    std::fclose(fp);
}

static void tune_store(const std::string &k, const Shape &sh)
{
    g_tune[k] = sh;
    const char *f = std::getenv("DLPX_TUNE_CACHE");
    if (!f) return;
    FILE *fp = std::fopen(f, "w");
    if (!fp) return;
    for (auto &e : g_tune) std::fprintf(fp, "%s %d %d\n", e.first.c_str(), e.second.pair_lanes, e.second.fft_lpb);
    std::fclose(fp);
}

static bool tune_enabled()
{
    const char *e = std::getenv("DLPX_AUTOTUNE");
    return !e || std::atoi(e) != 0;
}

// ---- runs ----------------------------------------------------------------
// Device copies of what a run changes, restored if a device-decided list
// build overflows (the run is then repeated with host-decided rebuilds,
// which regrow the list).
static int snapshot(Ctx *c, bool restore)
{
    const size_t n = c->natms;
    if (!c->snap_d) {
        DLPX_HIP(c, hipMalloc((void **)&c->snap_d, sizeof(double4) * 2 * (size_t)c->cap));
        DLPX_HIP(c, hipMalloc((void **)&c->snap_i, sizeof(int32_t) * 3 * (size_t)c->cap));
    }
    struct {
        void *live, *copy;
        size_t bytes;
    } parts[] = {{c->pos, c->snap_d, sizeof(double4) * n},
                 {c->vel, c->snap_d + c->cap, sizeof(double4) * n},
                 {c->typ, c->snap_i, sizeof(int32_t) * n},
                 {c->gid, c->snap_i + c->cap, sizeof(int32_t) * n},
                 {c->slot, c->snap_i + 2 * (size_t)c->cap, sizeof(int32_t) * n}};
    for (auto &p : parts) {
        if (restore) DLPX_HIP(c, hipMemcpyAsync(p.live, p.copy, p.bytes, hipMemcpyDeviceToDevice, c->stream));
        else DLPX_HIP(c, hipMemcpyAsync(p.copy, p.live, p.bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    return 0;
}

// nsteps resident steps of every context, all on the first context's
// streams.  Per-step energy records go to each context's energies_dev when
// energies is set.  rebuilds[k] receives the device-decided rebuild counts.
int ensemble_run(const std::vector<Ctx *> &cs, Ensemble *&E, int nsteps, bool energies, std::vector<int> &rebuilds)
{
    Ctx *c0 = cs[0];
    int r;
    rebuilds.assign(cs.size(), 0);
    if (nsteps <= 0) return 0;
    // the other contexts' queued work finishes before the shared streams take over
    for (Ctx *c : cs)
        if (c != c0) DLPX_HIP(c, hipStreamSynchronize(c->stream));
    for (Ctx *c : cs) {
        if ((r = snapshot(c, false))) return r;
        // counters of the run: sticky overflow, rebuild count, record index;
        // F_DMAX and F_DONE of stage 1 are zero between launches
        DLPX_HIP(c, hipMemsetAsync(c->flags + F_OVF_ANY, 0, sizeof(int32_t), c0->stream));
        DLPX_HIP(c, hipMemsetAsync(c->flags + F_NREB, 0, sizeof(int32_t) * 2, c0->stream));
        DLPX_HIP(c, hipMemsetAsync(c->flags + F_DMAX, 0, sizeof(int32_t) * 3, c0->stream));
        // the totals that the gather and the step kernel accumulate atomically
        DLPX_HIP(c, hipMemsetAsync(c->acc, 0, sizeof(double) * 4 * kAccSlots, c0->stream));
        DLPX_HIP(c, hipMemsetAsync(c->red_out + R_KIN, 0, sizeof(double), c0->stream));
    }
    for (Ctx *c : cs)
        if (c != c0) DLPX_HIP(c, hipStreamSynchronize(c->stream));  // the snapshots
    // shape: cached for this size class, else tuned on this run's first
    // steps when it has enough of them, else the default
    const std::string cls = size_class(cs);
    Shape sh;
    bool tune = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load();
        auto it = g_tune.find(cls);
        if (it != g_tune.end()) sh = it->second;
        // a batch of several systems is tuned on its own first steps; a single
        // system keeps the default shape (8 pair lanes per atom, 16 FFT lines
        // per workgroup), the best one measured at 4k-1.7M atoms
        else tune = tune_enabled() && !energies && !c0->timing && nsteps >= 40 && cs.size() > 1;
    }
    for (Ctx *c : cs)
        if ((r = upload_tables(c))) return r;
    if (!E) E = new Ensemble();
    E->cs = cs;
    auto ready = [&](const Shape &s) -> int {
        if (E->key.empty() || E->key != ens_key(cs, energies, s)) return ens_prepare(*E, energies, s);
        return 0;
    };
    int s = 0;
    // stage 1 of the first step (every system decides its rebuild)
    if ((r = ready(sh))) return r;
    if ((r = step_launch(c0, c0->stream, E->step, false, true))) return r;
    if (tune) {
        // each candidate: one eager step (kernel attributes), a capture, then
        // timed replays; the fastest median replay wins
        const Shape cands[] = {{4, 16}, {8, 16}, {16, 16}, {8, 8}, {16, 8}, {4, 8}};
        double best = 1e30;
        Shape best_sh = sh;
        hipEvent_t e0, e1;
        DLPX_HIP(c0, hipEventCreate(&e0));
        DLPX_HIP(c0, hipEventCreate(&e1));
        for (const Shape &cand : cands) {
            if (s + 6 > nsteps - 2) break;
            if ((r = ready(cand))) return r;
            if ((r = ens_enqueue(*E, true))) return r;
            s++;
            if ((r = ens_capture(*E, E->graph, 1))) return r;
            std::vector<float> t;
            for (int k = 0; k < 5; k++, s++) {
                DLPX_HIP(c0, hipEventRecord(e0, c0->stream));
                DLPX_HIP(c0, hipGraphLaunch(E->graph, c0->stream));
                DLPX_HIP(c0, hipEventRecord(e1, c0->stream));
                DLPX_HIP(c0, hipEventSynchronize(e1));
                float ms;
                DLPX_HIP(c0, hipEventElapsedTime(&ms, e0, e1));
                t.push_back(ms);
            }
            std::sort(t.begin(), t.end());
            if (t[2] < best) {
                best = t[2];
                best_sh = cand;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        sh = best_sh;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            tune_store(cls, sh);
        }
        if ((r = ready(sh))) return r;
    }
    if (c0->timing) {
        // eager timed steps: each one's stage times into the first context's
        // stage_ms (as an evaluation's; [5] the list chain, per step)
        for (; s < nsteps - 1; s++) {
            if ((r = ens_enqueue(*E, true))) return r;
            DLPX_HIP(c0, hipEventSynchronize(c0->ev[5]));
            float ms;
            for (int k = 0; k < 5; k++) {
                DLPX_HIP(c0, hipEventElapsedTime(&ms, c0->ev[k], c0->ev[k + 1]));
                c0->stage_ms[k] += ms;
            }
            DLPX_HIP(c0, hipEventElapsedTime(&ms, c0->ev[6], c0->ev[7]));
            c0->stage_ms[5] += ms;
            c0->timed_evals++;
        }
    } else if (s < nsteps - 1) {
        // one eager step (every kernel's attributes set), then the graph
        if ((r = ens_enqueue(*E, true))) return r;
        s++;
        // several steps per graph (Ctx::graph_steps) save the launch gap
        // between replays
        const int gk = c0->graph_steps;
        if (gk > 1 && s + gk <= nsteps - 1) {
            if (!E->graph_k && (r = ens_capture(*E, E->graph_k, gk))) return r;
            for (; s + gk <= nsteps - 1; s += gk) DLPX_HIP(c0, hipGraphLaunch(E->graph_k, c0->stream));
        }
        if (s < nsteps - 1) {
            if (!E->graph && (r = ens_capture(*E, E->graph, 1))) return r;
            for (; s < nsteps - 1; s++) DLPX_HIP(c0, hipGraphLaunch(E->graph, c0->stream));
        }
    }
    // the last step: no stage 1 after it
    if ((r = ens_enqueue(*E, false))) return r;
    DLPX_HIP(c0, hipStreamSynchronize(c0->stream));
    for (size_t k = 0; k < cs.size(); k++) {
        Ctx *c = cs[k];
        int32_t fl[F_WORDS];
        DLPX_HIP(c, hipMemcpy(fl, c->flags, sizeof(fl), hipMemcpyDeviceToHost));
        rebuilds[k] = fl[F_NREB];
        if (fl[F_ERR] != 0) {
            c->have_list = false;
            return fail(c, DLPX_ERR_STATE, "internal error: a list build met an atom index out of range");
        }
        c->have_list = true;
        c->finish_in_vv2 = false;
        if (fl[F_OVF_ANY] != 0) {
            // a list build overflowed: repeat this system's run from the
            // snapshot with host-decided rebuilds (the list regrows)
            if ((r = snapshot(c, true))) return r;
            int reb = 0;
            if ((r = md_fallback(c, nsteps, c->run_dt, energies, reb))) return r;
            rebuilds[k] = reb;
            DLPX_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return 0;
}

}  // namespace dlpx

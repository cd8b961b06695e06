// Internal definitions shared by the dlpx HIP sources (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "dlpx.h"

namespace dlpx {

constexpr double kPi = 3.14159265358979323846;
constexpr double kR4pie0 = 138935.4835;                   // constants.F90:100
constexpr double kZeroPlus = 2.2250738585072014e-308;     // Tiny(1.0_wp)
constexpr double kHalfMinus = 0.49999999999999994;        // Nearest(0.5_wp, -1.0_wp)
constexpr int kLanesPerAtom = 4;                          // lanes sharing one atom's list
constexpr int kMaxSplines = 16;
constexpr int kReduceSlots = 28;
// per-atom record of the per-particle outputs: [0] pair energy share,
// [1..6] pair stress share (xx, xy, xz, yy, yz, zz), [8..13] reciprocal
// stress (same order, before the SPME scale)
constexpr int kPPStride = 16;
// slots of the gather's atomic energy and net-force totals in batched runs:
// workgroup b adds into slot b % kAccSlots, so no address takes more than a
// few dozen atomics
constexpr int kAccSlots = 64;

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of an
// MI355X (workgroup b on XCD b % 8), each with its own L2.  Kernels whose
// workgroups read spatially local data (atoms in cell order, grid windows)
// renumber them so that each XCD takes one contiguous range of the work:
// launch xcd_grid(nb) workgroups and use xcd_block(blockIdx.x, gridDim.x) as
// the logical index (values >= nb do no work).  The mapping is a bijection
// whatever the dispatch order, so only the cache locality depends on it.
constexpr int kXcds = 8;
inline int xcd_grid(int nb) { return (nb + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_block(int b, int nb_launched)
{
    return (b % kXcds) * (nb_launched / kXcds) + b / kXcds;
}

// Batched launches: the workgroups of one launch serve many independent
// systems.  boff (device memory) holds nsys+1 non-decreasing workgroup
// offsets; logical workgroup lb belongs to the system s with
// boff[s] <= lb < boff[s+1] and is that system's workgroup lb - boff[s].  The
// search runs on wavefront-uniform values, so it is scalar code.  A kernel
// body written for one system takes its arguments as a struct plus its
// logical workgroup index and count; a single-system kernel passes its own
// kernel arguments, a batched one the struct of the system found here.
__device__ __forceinline__ int batch_find(const int32_t *__restrict__ boff, int nsys, int lb)
{
    int lo = 0, hi = nsys;  // boff[lo] <= lb < boff[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (boff[mid] <= lb) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The (system, workgroup, workgroup count) of a batched launch of
// xcd_grid(boff[nsys]) workgroups, XCD-aware over the whole launch; false for
// the padding workgroups past the last system.
struct BatchPos {
    int s, b, nb;
};
__device__ __forceinline__ bool batch_pos(const int32_t *__restrict__ boff, int nsys, BatchPos &bp)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    if (lb >= boff[nsys]) return false;
    bp.s = batch_find(boff, nsys, lb);
    bp.b = lb - boff[bp.s];
    bp.nb = boff[bp.s + 1] - boff[bp.s];
    return true;
}

// Words of Ctx::flags (int32, device).
enum Flag : int {
    F_OVF = 0,      // largest per-lane count of an overflowing list build (0: none)
    F_IBIG = 1,     // largest per-atom count of the last build
    F_OVF_ANY = 2,  // sticky overflow of any device-decided build in a resident run
    F_PAIRS = 3,    // stored entries of the last build
    F_DMAX = 4,     // [4..5] displacement maximum (double bits), stage 1
    F_DONE = 6,     // workgroups of stage 1 finished
    F_UPD = 7,      // 1: the list is rebuilt before the coming force evaluation
    F_NREB = 8,     // rebuilds decided on the device since the run began
    F_STEP = 9,     // per-step energy records written since the run began
    F_ERR = 10,     // sticky: a list-build kernel met an index outside the atoms (internal error)
    F_WDONE = 12,   // workgroups of the list binning finished (zero between launches)
    F_WORDS = 16
};

// Verlet list layout: atom-major, the G lanes of one atom interleaved, so the
// entries of one atom (and of one workgroup's atoms) are contiguous in HBM.
// Lane g of atom i holds its c-th neighbour at this index.
__host__ __device__ inline size_t list_slot(int i, int c, int lane, int cap_lane)
{
    return ((size_t)i * cap_lane + c) * kLanesPerAtom + lane;
}
// A list entry is j | code << 25: j the neighbour's sorted index (below
// 2^25, the largest natms_max), code the subcell offset of j from i in the
// subcell build, (dz + 2) 25 + (dy + 2) 5 + (dx + 2) for offsets -2..2 (0 from
// the link-cell build).
constexpr int kListCodeShift = 25;
constexpr uint32_t kListJMask = (1u << kListCodeShift) - 1u;
constexpr int kMaxAtoms = 1 << kListCodeShift;
__host__ __device__ inline int list_j(int32_t e) { return (int)((uint32_t)e & kListJMask); }
__host__ __device__ inline int list_code(int32_t e) { return (int)((uint32_t)e >> kListCodeShift); }

// Columns of the per-evaluation partial-sum buffer (FP64, one row per block;
// each column is written by exactly one kernel, which records its row count
// in Ctx::red_rows, and is summed in fixed order).
enum Red : int {
    R_ENGVDW = 0, R_VIRVDW, R_ENGRL, R_VIRRL,
    R_S11, R_S12, R_S13, R_S22, R_S23, R_S33,   // real-space stress
    R_ERECIP,                                    // sum_i q_i sum M Phi (before scale)
    R_FNETX, R_FNETY, R_FNETZ,                   // sum of unscaled reciprocal forces
    R_RS11, R_RS12, R_RS13, R_RS22, R_RS23, R_RS33,  // reciprocal stress (before scale)
    R_KIN,                                       // kinetic energy (2 x)
    R_ENGEX,                                     // Ewald exclusion correction energy
    R_XS11, R_XS12, R_XS13, R_XS22, R_XS23, R_XS33   // its stress
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Ctx {
    dlpx_params prm{};
    std::string err;
    hipStream_t stream = nullptr;
    // SPME runs on recip_stream, concurrently with the pair kernel on stream,
    // between a fork and a join event; sp is the stream the SPME kernels of
    // the current evaluation use (stream itself when stages are timed).
    hipStream_t recip_stream = nullptr;
    int pair_lanes = kLanesPerAtom;  // lanes per atom in the pair kernel: 4, 8 or 16
    hipStream_t sp = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // resident runs replay one captured step graph (batch.cpp) unless
    // DLPX_GRAPHS=0 or stage timing asks for eager steps
    bool use_graphs = true;
    bool finish_in_vv2 = false;      // the last evaluation left the reciprocal finish to stage 2
    bool vdw_force_shift = false;    // vdws%l_force_shift
    bool vdw_direct = false;         // vdws%l_direct: analytic forms per pair
    bool list_subcells = true;       // list build over half-edge subcells (needs >= 3 link cells per axis)
    bool fft_lines = true;           // compile-time-length line kernels for the FFT lengths that have them
    bool rel_pairs = true;           // mixed combined-table pairs from subcell-relative FP32 positions
    int comb_threads = 1024;         // workgroup size of the combined-table pair kernel (512 or 1024)
    bool fft_planes = true;          // x and y passes fused per z-plane where the mesh allows
    int graph_steps = 8;             // resident steps per replayed graph (1..64)
    int overlap = -1;                // resident steps' branch layout 0..2 (batch.cpp, ens_enqueue); -1 by batch size
    int device = 0;

    // box (orthorhombic)
    double cell[9]{};
    double L[3]{}, invL[3]{};
    double rcell[9]{};
    double volm = 0.0;

    // atoms
    int natms = 0;
    int cap = 0;            // allocated atoms
    bool have_list = false;
    bool resident = false;
    bool have_types = false;

    // host tables
    int nkeys = 0, max_grid = 0;
    int nrows = 0;                         // distinct table rows on the device (keys with equal tables share one)
    std::vector<double> vdw_pot, vdw_frc;  // nkeys x (max_grid+1)
    std::vector<int32_t> ltp;
    std::vector<double> gen_params;        // kVdwParams per key when the tables were generated here
    int nsamples = 0;
    std::vector<double> erfc_t, erfc_d;    // nsamples
    double elrc = 0.0, vlrc = 0.0;
    bool tables_dirty = true;

    // frozen atoms (config%lfrzn), per host index
    int frz_natms = 0;               // atoms the flags were given for (0: none)
    int nfrozen = 0;
    int32_t *frz = nullptr;          // 1 = frozen

    // excluded pairs (neigh%list_excl), symmetric, as CSR over host indices
    int excl_natms = 0;              // atoms the lists were given for (0: none)
    int excl_total = 0;              // entries (each pair twice)
    int32_t *excl_off = nullptr;     // excl_natms+1 row starts
    int32_t *excl_ids = nullptr;     // excl_total host indices, ascending per row

    // per-atom device arrays (sorted order unless noted)
    double4 *pos = nullptr, *pos_alt = nullptr;     // x, y, z, q
    double4 *vel = nullptr, *vel_alt = nullptr;     // vx, vy, vz, mass
    double4 *frc = nullptr;                          // total force
    double4 *frc_recip = nullptr;                    // unscaled reciprocal force
    double4 *xref = nullptr;                         // positions at list build
    int32_t *typ = nullptr, *typ_alt = nullptr;      // 0-based type
    int32_t *gid = nullptr, *gid_alt = nullptr;      // sorted -> host index
    int32_t *slot = nullptr;                         // host index -> sorted
    double *stage = nullptr;                         // host-order staging, 9 doubles per atom
    int32_t *stage_i = nullptr;                      // host-order int staging
    double *h_st = nullptr;                          // pinned host staging, 9 doubles per atom
    int32_t *h_t = nullptr;                          // pinned host staging, 1 int per atom
    int32_t *cell_of = nullptr, *perm = nullptr;

    // link cells
    int nc[3]{1, 1, 1};
    int ncell = 0, ncell_cap = 0;
    int32_t *cell_count = nullptr, *cell_start = nullptr, *cell_fill = nullptr;
    // binning keys: link cells, or subcells (2 x 2 x 2 per link cell, numbered
    // cell-major so each link cell's atoms stay contiguous)
    int nkey_cap = 0;
    int32_t *key_start = nullptr;

    // Verlet list: lane-interleaved chunks; element (c, i, g) at (c*cap + i)*G + g
    int cap_lane = 0;
    bool list_coded = false;         // entries carry subcell offsets (the subcell build)
    int32_t *sub_of = nullptr;       // per sorted atom: its subcell at the build, sx | sy << 10 | sz << 20
    float4 *prel = nullptr;          // per sorted atom: position relative to that subcell (FP32), type bits
    int32_t *nbr = nullptr;
    int32_t *cnt = nullptr;          // cap*G per-lane counts
    int32_t *flags = nullptr;        // [0] overflow lane max, [1] ibig, [3] pairs, [4..5] displacement
                                     // maximum (double bits), [6] workgroups done
    int32_t *h_flags = nullptr;      // pinned, mapped host words: [0..3] list flags, [4..5] vnl_check,
                                     // [8..9] displacement maximum of the last md step
    int32_t *h_flags_dev = nullptr;  // device address of h_flags

    // tables on device
    void *d_vdw = nullptr;           // nkeys x (max_grid+1) x {frc, pot}, FP64 or FP32
    void *d_erfc = nullptr;          // nsamples+1 x {deriv, erfc} (index 0 virtual)
    int32_t *d_keytab = nullptr;     // ntypes*ntypes -> key or -1
    // mixed precision, charges fixed per atom type and one sample grid for the
    // vdW and erfc tables: one combined row per type pair (pair.hip)
    std::vector<double> type_q;      // charge per type of the last atom upload
    bool q_per_type = false;         // every atom of a type has that type's charge
    bool comb_dirty = true;          // the combined rows do not match tables and charges
    bool comb_valid = false;         // d_comb holds the combined rows
    bool noalloc_now = false;        // the evaluation under way may not allocate
    void *d_comb = nullptr;          // ntypes*(ntypes+1)/2 rows x (max_grid+1) float4 samples
    int32_t *d_comb_row = nullptr;   // ntypes*ntypes -> combined row
    size_t comb_bytes = 0;
    void *d_vdw_direct = nullptr;    // vdw_method direct: kDirectStride per row, FP64 or FP32

    // SPME
    int K[3]{0, 0, 0};
    int K1h = 0, K1hp = 0;           // K1/2+1 and the padded row stride of H
    size_t nreal = 0, nhalf = 0;
    void *Q = nullptr;               // real charge grid K3 x K2 x K1; all zero between evaluations (the
                                     // forward x pass clears each line it has read)
    void *Phi = nullptr;             // real potential grid (the convolution's output)
    void *H = nullptr;               // complex half grid K3 x K2 x K1hp (x < K1h used)
    void *tw = nullptr;              // twiddles per axis: 3 x max(K) complex
    double *bsp_n2 = nullptr;        // |b|^2 per axis, K1 + K2 + K3 (FP64)
    void *theta = nullptr;           // {theta(m), its stress weight} on the half spectrum (layout of H), grid precision
    double theta_key[13]{};          // K1..K3, rcell, alpha it was computed for
    bool theta_valid = false;
    void *Q2 = nullptr;              // per-particle stress grid (a copy of the charge grid, then convolved)
    double self_energy = 0.0;
    double sumchg = 0.0;
    bool self_done = false;

    // reductions
    double *red = nullptr;           // red_cap rows x kReduceSlots partials
    int red_cap = 0;
    int red_rows[kReduceSlots]{};    // rows written per column by the last evaluation
    double *red_out = nullptr;       // kReduceSlots totals (device)
    double *acc = nullptr;           // kAccSlots x 4 atomic totals of the gather (batched runs), zero between steps
    double *h_red = nullptr;         // pinned mirror of red_out
    double *energies_dev = nullptr;  // per-step energies (resident runs)
    int energies_cap = 0;
    dlpx_two_body_out last_out{};    // the last step of the last resident run with energies
    bool have_last_out = false;

    // per-particle energies and stresses (stats%collect_pp) of dlpx_two_body_forces
    bool per_particle = false;       // requested (dlpx_set_per_particle)
    bool pp_now = false;             // the evaluation under way computes them
    bool pp_valid = false;           // the last dlpx_two_body_forces computed them
    double *pp = nullptr;            // kPPStride doubles per atom (sorted order), see pair.hip
    int pp_cap = 0;

    // resident runs
    double run_dt = 0.0;             // time step of the run under way
    uint64_t serial = 0;             // unique per context in the process (ensemble cache keys)
    void *ens = nullptr;             // this context's own batched step (an Ensemble of one)
    void *ens_multi = nullptr;       // the last ensemble run led by this context
    int32_t *snap_i = nullptr;       // device copies of the state at the start of a run
    double4 *snap_d = nullptr;       // (restored if a device-decided list build overflows)

    // timing
    bool timing = false;
    bool timing_pending = false;
    int timed_evals = 0;
    hipEvent_t ev[8]{};
    double stage_ms[8]{};      // pair, spread, fft+influence, gather, finish, list build
};

// One batched launch (see batch_find): a device array of nsys per-system
// argument structs, the systems' workgroup offsets, the launch's workgroup
// count (rounded up to a multiple of the XCD count when xcd) and its dynamic
// LDS bytes.  Owned by the ensemble that prepared it.
struct BatchLaunch {
    void *args = nullptr;
    int32_t *boff = nullptr;
    int nsys = 0, total = 0;
    size_t lds = 0;
};
// the batched kernels of one list build (neighbours.hip)
struct ListBatch {
    BatchLaunch wrap, scat, sort, bsub, blink, excl;
    int32_t *gate = nullptr;  // one system: its F_UPD word, else null (per-system gates only)
    // several systems: the compacted chain's plan (k_list_plan), its inputs
    // and the fixed grid of each kernel
    int32_t *plan = nullptr, *plan_nb = nullptr;
    int32_t **plan_gates = nullptr;
    int plan_grid[6] = {0, 0, 0, 0, 0, 0};
    int nsys = 0;
};
// the batched passes of one SPME convolution (fft.hip) and of the
// reciprocal-space stage (spme.hip).  Each pass is one launch per group of
// systems: a group per line length with compiled line kernels
// (fft_lines.hip), one group for the other systems (generic kernels).
struct FftGroup {
    BatchLaunch bl;
    int kind = 0;  // 0 forward x, 1 forward y, 2 z with the influence function, 3 inverse y, 4 inverse x,
                   // 5 forward x and y per plane, 6 inverse y and x per plane
    int n = 0;     // line length of a line-kernel group; 0: generic kernels
    int lpb = 0;   // lines (x-line pairs) per workgroup
};
struct FftBatch {
    std::vector<FftGroup> groups;  // in launch order
    int lpb_cap = 16;              // largest generic lines per workgroup to use (tuning knob)
    std::vector<int> conv_rows;    // z-pass workgroups (reduction rows) per system
};
struct SpmeBatch {
    BatchLaunch spread, gather;
    FftBatch fft;
    int S = 0;                     // spline order (0: no system has electrostatics)
    std::vector<int> gather_rows;  // gather workgroups per system
};
// the pair kernel's launches, one per template variant present
struct PairBatch {
    struct Group {
        BatchLaunch bl;
        int gp = 8;
        bool fs = false, direct = false, pp = false, gt = false, comb = false, rel = false;
        int nt = 512;
    };
    std::vector<Group> groups;
    BatchLaunch prel;  // the relative positions of the systems whose pair kernel reads them
};
struct RedRowsHost {
    int rows[kReduceSlots];
};
int batch_upload(Ctx *c, BatchLaunch &bl, const void *args, size_t bytes_each, const std::vector<int> &nblocks,
                 bool xcd);
void batch_free(BatchLaunch &bl);

// Device memory clears and host-to-device copies for set-up code: issued on
// the given stream and waited for before returning.  The context streams are
// non-blocking, so the null stream's hipMemset/hipMemcpy would not be ordered
// with the kernels queued on them.  Never used inside a stream capture.
inline hipError_t dev_zero(hipStream_t s, void *p, size_t bytes)
{
    const hipError_t e = hipMemsetAsync(p, 0, bytes, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}
inline hipError_t dev_upload(hipStream_t s, void *dst, const void *src, size_t bytes)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

// error helpers
int fail(Ctx *c, int code, const std::string &msg);
#define DLPX_HIP(c, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::dlpx::fail((c), DLPX_ERR_HIP + (int)e_,                           \
                                std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Let a kernel use up to the full 160 KB of dynamic LDS, once per device and
// process (each launch still gets only what it asks for).  Done at a kernel's
// first launch on a device, which is never inside a stream capture (captured
// steps follow an eager one).
#define DLPX_LDS_MAX(c, fn)                                                                                     \
    do {                                                                                                        \
        static std::atomic<uint64_t> done_{0};                                                                 \
        const uint64_t bit_ = 1ull << ((c)->device & 63);                                                      \
        if (!(done_.load(std::memory_order_acquire) & bit_)) {                                                 \
            DLPX_HIP((c), hipFuncSetAttribute((const void *)(fn), hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                              160 * 1024));                                                    \
            done_.fetch_or(bit_, std::memory_order_acq_rel);                                                   \
        }                                                                                                       \
    } while (0)

// host table generation (tables.cpp)
// vdW parameters per pair key: vdws%param(1:max_param, k), max_param = 7
// (bounds.F90:666), in internal units as the FIELD reader leaves them
constexpr int kVdwParams = 7;
// device row of vdw_method direct: {form, prepared params[7], afs, bfs}
constexpr int kDirectStride = 3 + kVdwParams;
int table_size(double cut);
bool vdw_kind_supported(int kind);
// params of one key -> the constants the evaluators take (vdw_prepare)
void vdw_prepare(int kind, const double *p, double rvdw, double *q);
void vdw_form(int kind, const double *q, double r, double *e, double *g);  // U(r), gamma = -r dU/dr
void vdw_table(int kind, const double *p, double rvdw, int max_grid, double *pot, double *frc);
void erfc_tables(double rc, double alpha, int n, double *erfc_t, double *erfc_d);
void bspline_norm2(int K, int S, double *norm2);
void invert_cell(const double *a, double *b, double *det);
void dcell(const double *a, double *b);
void vdw_lrc(int ntypes, const double *num_type, const int *kinds, const double *params, double rvdw,
             double volm, double *elrc, double *vlrc, const double *num_frz = nullptr);

// device stages (launch wrappers, one per .hip file)
int upload_tables(Ctx *c);
int upload_comb(Ctx *c);                     // the combined rows when the conditions hold (pair.hip)
int build_list(Ctx *c);                      // from c->pos (current order), sorts & permutes
int vnl_check_dev(Ctx *c, int *update);
int count_pairs_dev(Ctx *c, double r, int64_t *npairs);
int pair_forces(Ctx *c);                     // writes c->frc and reduction slots
int excl_forces(Ctx *c);                     // adds the exclusion correction into c->frc
int excl_blocks(const Ctx *c);
int spme_forces(Ctx *c);                     // adds into c->frc and reduction slots
int spme_setup(Ctx *c);
int ensure_red_rows(Ctx *c, int rows);
int reduce_partials(Ctx *c);                 // red -> red_out (device)
int pair_blocks(const Ctx *c);
int gather_blocks(const Ctx *c);
int nve_first(Ctx *c, double dt);            // also the vnl_check displacement (flags[4..6])
double nve_first_displacement(Ctx *c);       // after the stream has reached nve_first
int nve_second(Ctx *c, double dt, bool kinetic);  // also does a deferred reciprocal finish
int freeze_atoms(Ctx *c);                          // zero force and velocity of frozen atoms
double recip_scale(const Ctx *c);                   // the SPME energy scale (ewald_spole.F90:429)

// Influence function of the SPME convolution over the half spectrum
// (x index 0..K1/2), ewald_spole.F90:1296-1380.
struct InflParams {
    int K[3], K1h;
    double rcell[9];
    double cut2, test_fac, rcf;  // cut_off^2, (1e-6/rcf)^2, pi/alpha
    double alpha;
    // 0: the potential's influence function theta(m), with the reciprocal
    // stress partials in reduction slots R_RS11..R_RS33; 1: the per-particle
    // stress kernel B(m) f'(x) pi/(x alpha) (ewald_general.F90:1033-1057), no
    // reductions
    int mode;
};
// out <- FFT-(kernel . FFT+(in)) for real grids in and out (may be the same);
// with clear_in the forward pass leaves in all zero.
int fft_convolve(Ctx *c, const InflParams &ip, void *in, void *out, bool clear_in);
// theta(m) of ip into c->theta, launched on s when the mesh, cell or alpha
// differ from those of the table's last computation
int theta_update(Ctx *c, const InflParams &ip, hipStream_t s);
int conv_blocks(const Ctx *c);  // workgroups of the z pass (reduction rows)

// batched stages (one launch per stage over the systems of an ensemble)
int list_prepare(const std::vector<Ctx *> &cs, ListBatch &lb);
int list_launch(Ctx *c, hipStream_t s, const ListBatch &lb, int part = -1);
void list_batch_free(ListBatch &lb);
int pair_prepare(const std::vector<Ctx *> &cs, PairBatch &pb, int lanes, std::vector<int> &rows);
int pair_launch(Ctx *c, hipStream_t s, const PairBatch &pb);
int pair_blocks_exact(const Ctx *c);
int excl_prepare(const std::vector<Ctx *> &cs, BatchLaunch &bl);
int excl_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl);
InflParams infl_params(const Ctx *c);
int fft_prepare(const std::vector<Ctx *> &cs, const std::vector<InflParams> &ips, FftBatch &fb);
int fft_launch(Ctx *c, hipStream_t s, const FftBatch &fb);
int spme_prepare(const std::vector<Ctx *> &cs, SpmeBatch &sb);
// marks (timed runs): events recorded after the spreading, the convolution
// and the gather
int spme_launch(Ctx *c, hipStream_t s, const SpmeBatch &sb, hipEvent_t *marks = nullptr);
int vv2_blocks(const Ctx *c);
int excl_blocks(const Ctx *c);
// one kernel per step of a batched run: stage 2 of the step, its reductions
// and per-step record, stage 1 of the next step (nve.hip)
int step_blocks(const Ctx *c);
int step_prepare(const std::vector<Ctx *> &cs, const std::vector<RedRowsHost> &rows, bool energies,
                 BatchLaunch &bl);
int step_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl, bool stage2, bool stage1);

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace dlpx

// Internal definitions shared by the host code and the HIP kernels of libdebye.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <new>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/debye.h"

namespace dbx {

// ---------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define DBX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::dbx::Error(e_ == hipErrorOutOfMemory ? DBX_ERR_NOMEM : DBX_ERR_HIP, \
                               std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define DBX_REQUIRE(cond, code, msg)                   \
    do {                                               \
        if (!(cond)) throw ::dbx::Error((code), (msg)); \
    } while (0)

// ---------------------------------------------------------------------------------
// device layout of a model
// ---------------------------------------------------------------------------------
// Atoms are stored per subgroup as float4 (x, y, z, 0) in a spatial order: every
// subgroup is cut by recursive median bisection into tiles of at most 64 atoms (all
// tiles of a subgroup but its last are full).  Coordinates are the caller's float
// values bit for bit; only their order changes, which leaves every histogram unchanged.
constexpr int kTileAtoms = 64;

struct Tile {          // 32 bytes
    float cx, cy, cz;  // centre (midpoint of the tile's bounding box)
    float rad;         // >= max distance of a tile atom from the centre (rounded up)
    int32_t start;     // first atom (index into the float4 position array)
    int32_t count;     // atoms in the tile (1..64)
    int32_t group;
    int32_t parent;    // the tile's own index (pairs within one subgroup: B tiles from here on)
};

// One pass-1 work item: the pairs of a list of A tiles (subgroup row, 64-atom tiles) with the
// B atoms (subgroup col, 64-atom tiles [b0, b1)) whose distance from the A tile's centre lies
// in shell `shell` of pair `pair`, counted in one LDS window starting at bin k_lo: kP1Bins bins
// of 32-bit counters, or 2 kP1Bins bins of 16-bit counters (kP1Win16).  A shell wider than one
// window is covered by several jobs with consecutive windows: each counts only the pairs
// inside its own window, so every pair is counted by exactly one of them.  Degraded jobs
// (window narrower than the geometry, no siblings) send out-of-window pairs straight to the
// global histogram.
struct P1Job {
    int32_t pair;
    int32_t a_off, a_cnt;  // A tiles: alist[a_off .. a_off + a_cnt)
    int32_t b0, b1;
    int32_t same;   // 1: one subgroup, count b >= a (upper triangle incl. self)
    int32_t shell;  // index into the pair's shell-edge table
    int32_t k_lo;   // first bin of the LDS window
    int32_t e2_off; // offset of the pair's shell-edge table in the edge array
    int32_t degraded;  // 1: the window cannot hold the job's geometry (checked kernel)
    int32_t win;       // kP1WinFirst: no window below this one; kP1WinLast: none above;
                       // kP1Win16: a window of 2 kP1Bins 16-bit counters
    int32_t pad;
};
constexpr int kP1WinFirst = 1, kP1WinLast = 2, kP1WinOnly = 3;
constexpr int kP1Win16 = 4;
// planner: a 64-atom A tile whose shells would be thinner than kP1SingleShell windows in one
// 32-bit window counts them in 16-bit windows, one when that leaves a shell of kP1SingleShell
// windows, else several with shells kP1MultiShell windows wide.  Measured on the 20 nm
// rocksalt sphere at 0.0005 A and the spinel sphere at 1e-4 A against variants (32- or
// 16-atom A tiles with thinner shells: +14-60%; other shell fractions: up to +35%).
constexpr double kP1MultiShell = 0.75;
constexpr double kP1SingleShell = 0.25;

constexpr int kP1Threads = 1024;   // 16 waves of 64
constexpr int kP1Bins = 35328;     // LDS window: 138 KiB of uint32 counters (+ guard and sentinel)
constexpr int kP1Cand = 1024;      // LDS candidate list (A tile, B tile), 4 KiB
// + a 1 KiB B-atom stage per wave (16 KiB)

// Host-managed suballocator over one device allocation (the serving budget): allocation and
// release are host bookkeeping only, so releasing memory never synchronises the device.
// First fit over a free list ordered by offset, 256-byte granules, neighbours coalesced.
class DevArena {
public:
    DevArena(char *base, size_t bytes);
    void *alloc(size_t bytes);  // nullptr when no free range is large enough
    void free(void *p);
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }
    size_t largest_free() const;

private:
    char *base_;
    size_t cap_, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    std::map<size_t, size_t> live_;  // offset -> size
};

struct ModelDev {
    DevArena *arena = nullptr;  // device memory from this arena (else hipMalloc)
    int n_groups = 0;
    int64_t n_atoms = 0;
    std::vector<int64_t> group_n;
    std::vector<int32_t> group_tile0, group_tile1;  // tile range per subgroup
    std::vector<Tile> tiles_h;                      // host copy (job planning)
    std::vector<float> bbox;                        // per subgroup: min xyz, max xyz
    float4 *pos = nullptr;                          // device, spatially ordered
    Tile *tiles = nullptr;                          // device
    // pass-1 work plan for one bin width, built on first use and kept with the model
    double plan_rstep = -1.0;
    int plan_cu = 0;
    int plan_jobs = 0, plan_fast = 0;
    P1Job *d_jobs = nullptr;
    float *d_edge2 = nullptr;
    int32_t *d_alist = nullptr;
    int4 *d_list = nullptr;          // launch list of this model alone: (0, job, piece, pieces)
    int plan_list_n = 0, plan_list_fast = 0;
    std::vector<double> plan_cost;   // estimated cost of each job (launch order)
    std::vector<P1Job> plan_h;       // host copy of the jobs
};

// One model of a pass-1 launch, as the kernel sees it (64 bytes).
struct P1Ref {
    const float4 *pos;
    const Tile *tiles;
    const P1Job *jobs;
    const int32_t *alist;
    const float *edge2;
    unsigned long long *hist;  // this model's [P][n_bins] block
    int64_t n_bins;
    float rp;
    int32_t n_rows;            // P, the rows of hist
    unsigned long long *oor;   // this model's count of pairs at or beyond n_bins
};

// device memory of a model: from its arena when it has one (throws DBX_ERR_BUDGET when the
// arena is full), else hipMalloc
void *model_alloc(ModelDev &m, size_t bytes);
void model_free(ModelDev &m, void *p);

}  // namespace dbx

struct dbx_model {
    dbx::ModelDev d;
};

struct dbx_ctx {
    int device = 0;
    size_t budget = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::recursive_mutex mu;
    std::set<dbx_model *> models;  // live models, released by dbx_close if still there
    // grow-only device scratch, one buffer per slot (dbx::kSlot*)
    void *slot_ptr[15] = {};
    size_t slot_bytes[15] = {};
    void *pin_ptr[6] = {};
    size_t pin_bytes[6] = {};
    // buffers that growths replaced while enqueued work may still use them (release_retired)
    std::vector<void *> retired_dev, retired_pin;
    bool no_alloc = false;  // inside dbx_debye_noalloc: waits leave the retired buffers alone
    // pinned staging of uploads (dbx::stage): a grow-only buffer handed out in consecutive regions
    char *stage_buf = nullptr;
    size_t stage_cap = 0, stage_used = 0;
    // the q grid last uploaded to slot kSlotQin (host copy and device address)
    std::vector<double> q_host;
    const void *q_dev = nullptr;
    // [0] out-of-range pairs, [1] pairs outside the LDS window (checked kernel),
    // [2] fast jobs queued for the checked kernel, [3] pair evaluations; then the retry list
    uint64_t *counters = nullptr;
    uint64_t *counters_h = nullptr;  // pinned host copy of counters[0..3]
    bool p1_timed = false;
    int cu_count = 0;
    int64_t last_p1_jobs = 0;
    uint64_t last_p1_misses = 0;
    uint64_t last_p1_retried = 0;
    uint64_t last_p1_evals = 0;   // pair evaluations of the last pass 1 (in- and out-of-window)
    float last_p1_kernel_ms = 0.f;  // hipEvent time of the last pass-1 histogram kernel
    float last_p2_ms = 0.f;         // hipEvent time of the last pass-2 structure computation
    hipEvent_t ev[4] = {};
    bool p2_timed = false;
    bool p2_refining = false;  // an FP64 recomputation: the pass-2 time keeps the first pass's start
    int pass2_mode = DBX_PASS2_MFMA;  // DBX_PASS2_MFMA or DBX_PASS2_FP64
    double rel_err_target = 1e-6;     // models whose I bound exceeds it are redone in FP64 (0: never)
    int last_p2_refined = 0;          // models redone in FP64 by the last call
    int last_p2_bins = 0;             // compacted bins of the last pass 2 (known after sync_pass2)
    uint64_t *p2_readback_h = nullptr;  // kReadbackBytes of pinned host memory for pass 2, from dbx_open
    // the compacted-bin count and largest row digit class of each pass 2 enqueued since the last
    // sync_pass2 (pinned, 2 words each, from dbx_open), and the mode of each
    uint64_t *p2_status_h = nullptr;
    int p2_pending = 0;
    int p2_pending_mode[64] = {};
    // profiling of the matrix-core contraction (dbx_debug_set_profiling)
    bool profiling = false;
    hipEvent_t ev_mf[2] = {};
    bool mf_timed = false;
    double mf_flops = 0, mf_useful = 0;
    int mf_digits = 0, mf_bins = 0;  // the profiled contraction's largest digit count, compacted bins
    double mf_mean_digits = 0;      // and its digit count averaged over the row tiles
};

namespace dbx {

enum { kSlotPass1 = 0, kSlotPass2 = 1, kSlotPass2b = 2, kSlotHist = 3, kSlotQ = 4, kSlotDigits = 5,
       kSlotRtab = 6, kSlotBatch = 7, kSlotRows = 8, kSlotPairs = 9, kSlotBatchHist = 10,
       kSlotOor = 11, kSlotQin = 12, kSlotRefine = 13, kSlotProf = 14 };
// device buffer of at least `bytes` for `slot`; contents are not preserved when it grows (the
// old buffer stays valid for work already enqueued until release_retired)
void *scratch(dbx_ctx *ctx, int slot, size_t bytes);
// pinned host buffer of at least `bytes` for `slot`: destinations of downloads, read after the
// caller's wait (uploads take their staging from stage()); grows like scratch
enum { kPinMisc = 0, kPinOor = 1, kPinStepI = 2, kPinI = 3, kPinRel = 4, kPinStepRel = 5 };
void *pinned(dbx_ctx *ctx, int slot, size_t bytes);
// frees the buffers that growths of scratch and pinned replaced; the stream must be idle
void release_retired(dbx_ctx *ctx);
// Pinned host memory for one upload enqueued on the context's stream, a region of its own:
// uploads enqueued one after another with no wait between them never share staging memory.
// Regions are handed out again once the stream has been waited for (stage_reset after a
// synchronisation; stage itself waits when the buffer is full, and may then move it): fill a
// region and enqueue its copy before the next stage call on the context.
void *stage(dbx_ctx *ctx, size_t bytes);
void stage_reset(dbx_ctx *ctx);
// the q grid on the device (slot kSlotQin), uploaded through pinned memory unless the slot
// already holds the same values
const double *device_q(dbx_ctx *ctx, const double *q, int nq);

// pass 1 (pass1.hip): histograms of all subgroup pairs into d_hist[P][n_bins]
// (zeroed here).  Returns the number of pairs with bin >= n_bins.
struct P1Stats {
    uint64_t out_of_range = 0;
    uint64_t window_misses = 0;
    uint64_t retried = 0;
    int n_jobs = 0;
    int n_checked = 0;
};
P1Stats pass1_histograms(dbx_ctx *ctx, ModelDev &m, double rstep, int64_t n_bins,
                         uint64_t *d_hist, bool sync_stats = true);
// several models in one launch pair (fast + checked), all with bin width rstep
struct P1Item {
    ModelDev *m;
    int64_t n_bins;
    uint64_t *d_hist;  // [P][n_bins] of this model (zeroed by the call)
    uint64_t *d_oor = nullptr;  // its out-of-range counter (zeroed by pass1_batch); null: the context's
};
P1Stats pass1_batch(dbx_ctx *ctx, const std::vector<P1Item> &items, double rstep, bool sync_stats);
// reads the pass-1 counters of the last launch (after the stream is synchronised)
void pass1_read_stats(dbx_ctx *ctx, P1Stats &st);
// host planning of pass 1: work items (largest first) and the shell-edge tables
void plan_pass1(const ModelDev &m, double rstep, std::vector<P1Job> &jobs, std::vector<float> &edge2,
                std::vector<int32_t> &alist, int n_cu = 256, std::vector<double> *cost = nullptr);

// the pass-1 plan on the host (make_plan_host, no device) and its installation on a model
struct PlanHost {
    double rstep = 0;
    int cu = 0;
    std::vector<P1Job> jobs;
    std::vector<float> edge2;
    std::vector<int32_t> alist;
    std::vector<double> cost;
    std::vector<int4> list;
    int n_fast = 0, list_fast = 0;
};
PlanHost make_plan_host(const ModelDev &m, double rstep, int n_cu);
size_t plan_device_bytes(const PlanHost &ph);
void install_plan(dbx_ctx *ctx, ModelDev &m, PlanHost &&ph);

// models (api.cpp): host tiling, device upload, release
void tile_model_host(const float *x, const float *y, const float *z, const int64_t *group_n, int S,
                     ModelDev &m, std::vector<float4> &pos_h);
size_t model_device_bytes(const ModelDev &m, size_t n_pos);
void upload_model(dbx_ctx *ctx, ModelDev &m, const std::vector<float4> &pos_h);
void free_model(ModelDev &m);

// pass 2 (pass2.hip)
// one row of pass 2: a partial histogram at hist_base + h_off with n_bins bins
struct RowDesc {
    int64_t h_off;
    int64_t n_bins;
    int32_t first_n;  // weight rule: the first first_n nonzero bins at r <= 0.0005 count once
    int32_t pad;
};
// the scattering parameters of one subgroup pair (row) for the prefactor
struct GroupPair {
    double occ_a, occ_b, beq_a, beq_b, ca, cb;
    double aa[5], ba[5], ab[5], bb[5];
};
GroupPair make_group_pair(const dbx_group_params &ga, const dbx_group_params &gb);
void prefactor_dev(dbx_ctx *ctx, const dbx_group_params &ga, const dbx_group_params &gb,
                   const double *d_q, int nq, double *d_ff);
// ff[r][q] for every row r
void prefactor_rows(dbx_ctx *ctx, const std::vector<GroupPair> &rows, const double *d_q, int nq,
                    double *d_ff);
// Where pass 2's device buffers and its read-back words come from: the context's grow-only
// slots (ctx_pass2_mem), or a fixed device region (the allocation-free entry point).  dev()
// returns a buffer of at least `bytes` that no other dev() call of the same pass 2 overlaps.
// pinned host memory that dbx_open allocates for the allocation-free pass 2's read-backs
constexpr size_t kReadbackBytes = 64 << 10;
// pass-2 status words a context holds (passes 2 enqueued between two sync_pass2)
constexpr int kP2StatusMax = 64;
struct P2Mem {
    virtual ~P2Mem() = default;
    virtual void *dev(int slot, size_t bytes) = 0;
    // bytes that dev() can still hand out in this pass 2 (SIZE_MAX: as many as the device holds)
    virtual size_t room() const { return SIZE_MAX; }
    // pinned host memory for uploads, or null when this provider has less
    virtual void *staging(size_t bytes) = 0;
};
// Pass 2 leaves its compacted-bin count and largest row digit class in pinned status words
// (pass2_status_slot, at most kP2StatusMax between waits); sync_pass2 waits for the context's
// stream and checks them: a count too large for the matrix cores' digits throws DBX_ERR_RANGE.
// Every caller waits through sync_pass2 before it uses pass 2's results.
unsigned long long *pass2_status_slot(dbx_ctx *ctx, int mode);
void sync_pass2(dbx_ctx *ctx);
// marks an FP64 recomputation of pass 2 for its lifetime
struct P2Refining {
    dbx_ctx *ctx;
    explicit P2Refining(dbx_ctx *c) : ctx(c) { ctx->p2_refining = true; }
    ~P2Refining() { ctx->p2_refining = false; }
};
P2Mem &ctx_pass2_mem(dbx_ctx *ctx);
// S[r][q] for every row r (rows of any models sharing rstep, first_n and the q grid)
// d_B (optional) receives each row's absolute error bound of S (the same for every q); mode:
// DBX_PASS2_MFMA or DBX_PASS2_FP64, -1 for the context's
void debye_structure_rows(dbx_ctx *ctx, const uint64_t *d_hist_base, const std::vector<RowDesc> &rows,
                          double rstep, const double *d_q, int nq, double *d_S, double *d_B = nullptr,
                          int mode = -1);
// The rows of several models again with FP64 arithmetic, in one pass 2: model m is rows
// [first, first + count) of rows (longest first) given as models[m] = (first, count); its sums go
// to S_dst[m] (count x nq) and its bounds to B_dst[m] (count).
void structure_fp64_models(dbx_ctx *ctx, const uint64_t *d_hist_base, const std::vector<RowDesc> &rows,
                           const std::vector<int2> &models, double rstep, const double *d_q, int nq,
                           const std::vector<double *> &S_dst, const std::vector<double *> &B_dst);
// S[p][q] for the P partials of one model, d_hist[P][n_bins], buffers from mem (the row
// descriptors are written on the device: no host staging)
void debye_structure_mem(dbx_ctx *ctx, P2Mem &mem, const uint64_t *d_hist, int P, int64_t n_bins, double rstep,
                         int first_n, const double *d_q, int nq, double *d_S, double *d_B = nullptr, int mode = -1);
// I[mrows[m].z][q] = sum over rows [mrows[m].x, mrows[m].x + mrows[m].y) of ff S; with the row
// bounds d_B, d_rel[mrows[m].z][q] (optional) receives the relative error bound of each I
void combine_models(dbx_ctx *ctx, const double *d_ff, const double *d_S, const double *d_B,
                    const std::vector<int4> &mrows, int nq, double *d_I, double *d_rel);
// S_p(q) for all p into d_S[P][nq] from device histograms d_hist[P][n_bins].
void debye_structure(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep,
                     int first_n, const double *d_q, int nq, double *d_S, double *d_B = nullptr, int mode = -1);
// I[q] = sum_p ff[p][q] S[p][q]
void launch_prefactor(hipStream_t st, const GroupPair *d_pairs, int n, const double *d_q, int nq, double *d_ff);
// d_mrows null: one model, rows m0.x .. m0.x + m0.y - 1 into I[m0.z]
void launch_combine(hipStream_t st, const double *d_ff, const double *d_S, const double *d_B, const int4 *d_mrows,
                    int nm, int nq, double *d_I, double *d_rel, int4 m0 = make_int4(0, 0, 0, 0));
void combine_intensity(dbx_ctx *ctx, const double *d_ff, const double *d_S, const double *d_B, int P, int nq,
                       double *d_I, double *d_rel);
// *d_out = the bit pattern of the largest of the non-negative doubles d_v[0..n)
void launch_max_bits(hipStream_t st, const double *d_v, int n, unsigned long long *d_out);


// An entry point's body under the context's lock, with its device current: C++ errors become
// the returned status and the context's message (no exception crosses the C ABI).
template <class F>
int guarded(dbx_ctx *ctx, F &&f)
{
    if (!ctx) return DBX_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    try {
        DBX_HIP(hipSetDevice(ctx->device));
        ctx->p2_pending = 0;  // status words of a call that failed before its wait are stale
        f();
        ctx->err.clear();
        return DBX_OK;
    } catch (const Error &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return DBX_ERR_NOMEM;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return DBX_ERR_ARG;
    }
}

// true when an intensity bound rel (relative) does not meet target (> 0); NaN does not
inline bool exceeds_target(double rel, double target) { return target > 0 && !(rel <= target); }

// the allocation-free pass 2 (noalloc.hip): the device address of this device's fixed scratch
// region, looked up once per device (called by dbx_open with the device current)
void noalloc_init(int device);

}  // namespace dbx

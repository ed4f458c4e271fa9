// Serving: a stream of structure models of uneven size through continuous batching.
//
// No reference counterpart in-process: the reference runs one process per model
// (Matlab/multisim.m:42-58 launching cuDebye.exe, cuDebye-v15.cu:599-1003).
//
//   submit -> [prep threads: tiling + pass-1 plan on the host] -> ready queue
//          -> admission: every device buffer of the request from the budget arena
//          -> [lane threads, one stream each: one step = the admitted requests that fit the
//              step caps; pass 1 over all of them in one launch pair, then per q grid one
//              prefactor + compaction + Debye contraction + combine] -> done
// A request's device memory comes from one arena sized by the HBM budget; a request waits
// in the ready queue until its buffers fit.  Prepared requests are admitted in submission
// order, except that one may pass earlier requests still being prepared or waiting for
// memory while the oldest waiting request is younger than kAdmitHoldMs and the arena keeps
// room for those earlier requests beside it; past that, nothing passes the oldest, so the
// memory that finishing steps return goes to it.  Memory is returned at
// completion (or at dbx_release for DBX_REQ_KEEP requests, whose histograms and S_p(q) stay
// resident).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace dbx {

enum { kQueued = 0, kReady = 1, kAdmitted = 2, kRunning = 3, kDone = 4, kFailed = 5 };

// Pinned host memory for the requests' upload images, which the prep threads write: an upload
// from pinned memory is one DMA that the lane's stream runs on its own, where one from pageable
// memory is staged through the runtime's buffers on the calling thread (measured ~16 us per
// request on the host per step, most of a step's host time).  Chunks of kPinChunk bytes (or one
// image's size) are allocated on demand up to kPinCap in all, each handed out by an offset
// allocator; an image that finds no room stays in pageable memory.  Thread-safe.
class PinnedPool {
public:
    static constexpr size_t kPinChunk = (size_t)64 << 20;
    static constexpr size_t kPinCap = (size_t)1 << 30;
    ~PinnedPool()
    {
        for (Chunk &c : chunks_) (void)hipHostFree(c.base);
    }
    char *alloc(size_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (Chunk &c : chunks_)
            if (void *p = c.arena->alloc(bytes)) return (char *)p;
        const size_t want = std::max(kPinChunk, (bytes + 255) & ~(size_t)255);
        if (total_ + want > kPinCap) return nullptr;
        void *b = nullptr;
        if (hipHostMalloc(&b, want, hipHostMallocDefault) != hipSuccess) return nullptr;
        chunks_.push_back(Chunk{(char *)b, want, std::unique_ptr<DevArena>(new DevArena((char *)b, want))});
        total_ += want;
        return (char *)chunks_.back().arena->alloc(bytes);
    }
    void free(char *p)
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (Chunk &c : chunks_)
            if (p >= c.base && p < c.base + c.bytes) {
                c.arena->free(p);
                return;
            }
    }

private:
    struct Chunk {
        char *base;
        size_t bytes;
        std::unique_ptr<DevArena> arena;
    };
    std::mutex mu_;
    std::vector<Chunk> chunks_;
    size_t total_ = 0;
};

// a request's upload image: pinned from the pool when it has room, else pageable
struct Image {
    PinnedPool *pool = nullptr;
    char *pinned = nullptr;
    std::vector<char> pageable;
    size_t bytes = 0;
    char *data() { return pinned ? pinned : pageable.data(); }
    void assign(PinnedPool *p, size_t n)
    {
        release();
        bytes = n;
        pool = p;
        pinned = p ? p->alloc(n) : nullptr;
        if (!pinned) pageable.assign(n, 0);
    }
    void release()
    {
        if (pinned) pool->free(pinned);
        pinned = nullptr;
        std::vector<char>().swap(pageable);
        bytes = 0;
    }
    ~Image() { release(); }
};

// how long the oldest waiting request may be passed by later ones at admission
constexpr double kAdmitHoldMs = 50.0;

static double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    std::vector<float> x, y, z;
    std::vector<int64_t> group_n;
    std::vector<dbx_group_params> params;
    std::vector<double> q;
    double rstep = 0;
    int64_t n_bins = 0;
    int first_n = 0;
    int flags = 0;
    int P = 0;
    int64_t n_atoms = 0;
    // host preparation
    ModelDev m;
    std::vector<float4> pos_h;
    PlanHost plan;
    // everything uploaded, in its device layout: one copy per request
    Image image;
    size_t off[7] = {};  // pos, tiles, jobs, edge2, list, alist, q
    // device buffers (arena)
    std::vector<void *> dev;
    uint64_t *d_hist = nullptr;
    // DBX_REQ_KEEP requests only (reweighting): sums, prefactors, I, bounds, pair parameters
    double *d_S = nullptr;  // [P][nq]
    double *d_q = nullptr;
    double *d_ff = nullptr;  // [P][nq]: the step's prefactors, then the last reweight's
    double *d_I = nullptr;   // [nq]
    double *d_B = nullptr;   // [P] error bound of each S_p
    double *d_rel = nullptr; // [nq] relative error bound of d_I
    GroupPair *d_pairs = nullptr;
    int4 *d_mrows = nullptr;
    size_t bytes = 0;
    size_t est_bytes = 0;  // lower bound of bytes, known at submission (before preparation)
    // results
    std::vector<double> I, rel;
    uint64_t oor = 0;
    int state = kQueued;
    int err = DBX_OK;
    std::string msg;
    double t_submit = 0, t_admit = 0, t_done = 0;
    // reweight: captured graph, pinned staging
    hipGraph_t rw_graph = nullptr;
    hipGraphExec_t rw_exec = nullptr;
    GroupPair *h_pairs = nullptr;
    double *h_I = nullptr;   // [2 nq]: I, then its relative error bound
};

}  // namespace dbx

struct dbx_server {
    int device = 0;
    dbx_server_config cfg{};
    void *arena_mem = nullptr;
    std::unique_ptr<dbx::DevArena> arena;
    std::mutex mu;
    std::condition_variable cv_prep, cv_sched, cv_done;
    dbx::PinnedPool pins;  // upload images (outlives every request below)
    std::deque<dbx::Request *> prep_q, admitted_q;
    std::map<uint64_t, dbx::Request *> ready;  // prepared, by submission order
    std::set<uint64_t> preparing;              // submitted, not yet prepared
    std::unordered_map<uint64_t, std::unique_ptr<dbx::Request>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> threads;
    dbx_ctx *lane[2] = {nullptr, nullptr};
    dbx_ctx *rw_ctx = nullptr;
    std::mutex rw_mu;
    std::string err;
    dbx_server_stats st{};
    int running_steps = 0;  // lanes inside a step
    int in_api = 0;         // callers inside an entry point (dbx_server_close waits for none)
    std::condition_variable cv_api;
};

namespace dbx {

static size_t al256(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }

// the request's uploads in their device layout (positions, tiles, plan, q grid), 256-byte aligned
static void build_image(Request &r, PinnedPool *pool)
{
    const void *src[7] = {r.pos_h.data(), r.m.tiles_h.data(), r.plan.jobs.data(), r.plan.edge2.data(),
                          r.plan.list.data(), r.plan.alist.data(), r.q.data()};
    const size_t len[7] = {sizeof(float4) * r.pos_h.size(), sizeof(Tile) * r.m.tiles_h.size(),
                           sizeof(P1Job) * r.plan.jobs.size(), sizeof(float) * r.plan.edge2.size(),
                           sizeof(int4) * r.plan.list.size(), sizeof(int32_t) * r.plan.alist.size(),
                           sizeof(double) * r.q.size()};
    size_t off = 0;
    for (int i = 0; i < 7; i++) {
        r.off[i] = off;
        off += al256(len[i]);
    }
    r.image.assign(pool, off);
    for (int i = 0; i < 7; i++)
        if (len[i]) std::memcpy(r.image.data() + r.off[i], src[i], len[i]);
    // the tile-ordered positions now live in the image only
    std::vector<float4>().swap(r.pos_h);
}

// device bytes of a prepared request while resident
static size_t request_bytes(const Request &r)
{
    const size_t nq = r.q.size(), P = (size_t)r.P;
    const size_t kept = (r.flags & DBX_REQ_KEEP) ? 2 * al256(sizeof(double) * P * nq) + 2 * al256(sizeof(double) * nq) +
                                                       al256(sizeof(double) * P) + al256(sizeof(GroupPair) * P) +
                                                       al256(sizeof(int4))
                                                 : 0;
    return al256(r.image.bytes) + al256(sizeof(uint64_t) * P * (size_t)r.n_bins) + kept;
}

// a lower bound of request_bytes from what submission knows: positions, histograms, q grid
// and the kept buffers (tiles and plan are sized by preparation)
static size_t request_bytes_lower(const Request &r)
{
    Request t;
    t.q.resize(r.q.size());
    t.P = r.P;
    t.n_bins = r.n_bins;
    t.flags = r.flags;
    t.image.bytes = sizeof(float4) * (size_t)r.n_atoms + sizeof(double) * r.q.size();
    return request_bytes(t);
}

// every device buffer of the request from the arena; false (nothing kept) when one does not fit
static bool alloc_request(DevArena &A, Request &r)
{
    const size_t nq = r.q.size(), P = (size_t)r.P;
    auto take = [&](size_t b) -> void * {
        void *p = A.alloc(b);
        if (p) r.dev.push_back(p);
        return p;
    };
    bool ok = true;
    auto need = [&](void *p) { ok = ok && p != nullptr; return p; };
    char *img = (char *)need(take(r.image.bytes));
    if (img) {
        r.m.pos = (float4 *)(img + r.off[0]);
        r.m.tiles = (Tile *)(img + r.off[1]);
        r.m.d_jobs = (P1Job *)(img + r.off[2]);
        r.m.d_edge2 = (float *)(img + r.off[3]);
        r.m.d_list = (int4 *)(img + r.off[4]);
        r.m.d_alist = (int32_t *)(img + r.off[5]);
        r.d_q = (double *)(img + r.off[6]);
    }
    r.d_hist = (uint64_t *)need(take(sizeof(uint64_t) * P * (size_t)r.n_bins));
    if (r.flags & DBX_REQ_KEEP) {
        r.d_S = (double *)need(take(sizeof(double) * P * nq));
        r.d_ff = (double *)need(take(sizeof(double) * P * nq));
        r.d_I = (double *)need(take(sizeof(double) * nq));
        r.d_rel = (double *)need(take(sizeof(double) * nq));
        r.d_B = (double *)need(take(sizeof(double) * P));
        r.d_pairs = (GroupPair *)need(take(sizeof(GroupPair) * P));
        r.d_mrows = (int4 *)need(take(sizeof(int4)));
    }
    if (!ok) {
        for (void *p : r.dev) A.free(p);
        r.dev.clear();
        r.m.pos = nullptr;
        r.m.tiles = nullptr;
        r.m.d_jobs = nullptr;
        r.m.d_edge2 = nullptr;
        r.m.d_list = nullptr;
        r.m.d_alist = nullptr;
    }
    return ok;
}

static void free_request_device(DevArena &A, Request &r)
{
    for (void *p : r.dev) A.free(p);
    r.dev.clear();
    r.m.pos = nullptr;
    r.m.tiles = nullptr;
    r.m.d_jobs = nullptr;
    r.m.d_edge2 = nullptr;
    r.m.d_list = nullptr;
    r.m.d_alist = nullptr;
    r.d_hist = nullptr;
    r.d_S = nullptr;
    r.d_B = nullptr;
    if (r.rw_exec) (void)hipGraphExecDestroy(r.rw_exec);
    if (r.rw_graph) (void)hipGraphDestroy(r.rw_graph);
    r.rw_exec = nullptr;
    r.rw_graph = nullptr;
}

static void fail(Request &r, int code, const std::string &msg)
{
    r.state = kFailed;
    r.err = code;
    r.msg = msg;
}

static void prep_loop(dbx_server *s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_prep.wait(lk, [&] { return s->stop || !s->prep_q.empty(); });
            if (s->stop) return;
            r = s->prep_q.front();
            s->prep_q.pop_front();
        }
        int code = DBX_OK;
        std::string msg;
        try {
            tile_model_host(r->x.data(), r->y.data(), r->z.data(), r->group_n.data(), (int)r->group_n.size(), r->m,
                            r->pos_h);
            r->plan = make_plan_host(r->m, r->rstep, s->lane[0]->cu_count);
            // the caller's coordinates are no longer needed
            std::vector<float>().swap(r->x);
            std::vector<float>().swap(r->y);
            std::vector<float>().swap(r->z);
            build_image(*r, &s->pins);
            r->bytes = request_bytes(*r);
        } catch (const Error &e) {
            code = e.code;
            msg = e.what();
        } catch (const std::exception &e) {
            code = DBX_ERR_NOMEM;
            msg = e.what();
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->preparing.erase(r->id);
        if (code == DBX_OK && r->bytes > s->arena->capacity()) {
            code = DBX_ERR_BUDGET;
            msg = "request needs " + std::to_string(r->bytes) + " device bytes, more than the serving budget";
        }
        if (code != DBX_OK) {
            fail(*r, code, msg);
            r->t_done = now_ms();
            s->st.failed++;
            s->cv_done.notify_all();
            s->cv_sched.notify_all();  // a later request may have been waiting for this one
            continue;
        }
        r->state = kReady;
        s->ready[r->id] = r;
        s->cv_sched.notify_all();
        s->cv_done.notify_all();  // a waiter looks again whether its request can still be admitted
    }
}

// admission (caller holds s->mu): prepared requests in submission order while their buffers
// fit.  While the oldest waiting request (being prepared, or prepared and waiting for memory)
// is younger than kAdmitHoldMs, a later prepared request may pass the earlier waiting ones
// if the arena keeps room for all of them beside it (their sizes, or for those still being
// prepared a lower bound known at submission): passing never takes memory that an earlier
// request needs, so under memory pressure (DBX_REQ_KEEP requests resident) requests are
// admitted in submission order.  Past kAdmitHoldMs nothing passes the oldest.  Returns the
// number admitted.
static int admit(dbx_server *s)
{
    const double now = now_ms();
    int n = 0;
    for (;;) {
        uint64_t oldest = UINT64_MAX;
        if (!s->preparing.empty()) oldest = *s->preparing.begin();
        if (!s->ready.empty()) oldest = std::min(oldest, s->ready.begin()->first);
        bool hold = false;
        if (oldest != UINT64_MAX) {
            auto it = s->reqs.find(oldest);
            hold = it != s->reqs.end() && now - it->second->t_submit > kAdmitHoldMs;
        }
        bool held_in = false;
        auto prep = s->preparing.begin();
        size_t owed = 0;  // bytes of the earlier requests still waiting
        for (auto it = s->ready.begin(); it != s->ready.end();) {
            // a request held in line: nothing after it is admitted before it
            if (hold && it->first != oldest) break;
            Request *r = it->second;
            for (; prep != s->preparing.end() && *prep < it->first; ++prep) {
                auto pr = s->reqs.find(*prep);
                if (pr != s->reqs.end()) owed += pr->second->est_bytes;
            }
            const size_t left = s->arena->capacity() - s->arena->used();
            if ((owed > 0 && (r->bytes > left || left - r->bytes < owed)) || !alloc_request(*s->arena, *r)) {
                if (hold) break;
                owed += r->bytes;
                ++it;
                continue;
            }
            it = s->ready.erase(it);
            r->state = kAdmitted;
            r->t_admit = now;
            s->admitted_q.push_back(r);
            n++;
            if (hold) {
                held_in = true;  // the line moves on: look at the next oldest
                break;
            }
        }
        if (!held_in) break;
    }
    s->st.resident_bytes = s->arena->used();
    return n;
}

// true when request r waits for memory that nothing but a dbx_release can free: it is
// prepared but not admitted, and no step runs, none is admitted and none is being prepared
static bool admission_stalled(dbx_server *s, const Request *r)
{
    return r->state == kReady && s->running_steps == 0 && s->admitted_q.empty() && s->preparing.empty();
}

// one step on a lane: uploads, pass 1 over all requests, pass 2 per q grid, results.  The
// host waits for the device once, at the end: pass 2 sizes its launches from upper bounds and
// its kernels read the compacted-bin counts from device memory (see structure_core).
static void run_step(dbx_server *s, dbx_ctx *ctx, std::vector<Request *> &batch)
{
    const double t_start = now_ms();
    ctx->p2_pending = 0;  // status words of a step that failed before its wait are stale
    const double rstep = batch[0]->rstep;
    for (Request *r : batch) {
        // the model's device arrays, its plan and its q grid: one copy of the request's image
        ModelDev &m = r->m;
        PlanHost &ph = r->plan;
        DBX_HIP(hipMemcpyAsync(m.pos, r->image.data(), r->image.bytes, hipMemcpyHostToDevice, ctx->stream));
        m.plan_h = std::move(ph.jobs);
        m.plan_cost = std::move(ph.cost);
        m.plan_jobs = (int)m.plan_h.size();
        m.plan_fast = ph.n_fast;
        m.plan_list_n = (int)ph.list.size();
        m.plan_list_fast = ph.list_fast;
        m.plan_cu = ph.cu;
        m.plan_rstep = ph.rstep;
    }
    // the requests' out-of-range counts: one array for the step, zeroed by pass 1's first launch
    uint64_t *d_oor = (uint64_t *)scratch(ctx, kSlotOor, sizeof(uint64_t) * batch.size());
    std::vector<P1Item> items;
    for (size_t i = 0; i < batch.size(); i++) {
        Request *r = batch[i];
        P1Item it{&r->m, r->n_bins, r->d_hist};
        it.d_oor = d_oor + i;
        items.push_back(it);
    }
    pass1_batch(ctx, items, rstep, false);
    // the launch list is on the device: the host copies of the plan and tiles are not needed
    // any more
    for (Request *r : batch) {
        r->plan = PlanHost();
        std::vector<P1Job>().swap(r->m.plan_h);
        std::vector<double>().swap(r->m.plan_cost);
        std::vector<Tile>().swap(r->m.tiles_h);
    }
    // pass 2: one contraction per q grid, rows longest first
    std::vector<size_t> order(batch.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (batch[a]->q.size() != batch[b]->q.size()) return batch[a]->q.size() < batch[b]->q.size();
        return batch[a]->n_bins > batch[b]->n_bins;
    });
    // each group (requests of one q grid, longest first) and its rows
    struct Group {
        std::vector<size_t> req;  // batch indices
        std::vector<RowDesc> rows;
        std::vector<GroupPair> pairs;
        std::vector<int4> mrows;  // (first row, rows, dest, -)
        uint64_t *base;
        int nq;
        size_t buf_off, I_base;
    };
    std::vector<Group> groups;
    size_t buf_total = 0, nI = 0;
    for (size_t g0 = 0; g0 < order.size();) {
        size_t g1 = g0 + 1;
        while (g1 < order.size() && batch[order[g1]]->q == batch[order[g0]]->q) g1++;
        Group G;
        for (size_t i = g0; i < g1; i++) G.req.push_back(order[i]);
        G.nq = (int)batch[G.req[0]]->q.size();
        G.base = batch[G.req[0]]->d_hist;
        for (size_t i : G.req) G.base = std::min(G.base, batch[i]->d_hist);
        for (size_t gi = 0; gi < G.req.size(); gi++) {
            Request *r = batch[G.req[gi]];
            const int S = (int)r->group_n.size();
            G.mrows.push_back(make_int4((int)G.rows.size(), r->P, (int)gi, 0));
            int p = 0;
            for (int a = 0; a < S; a++)
                for (int b = a; b < S; b++, p++) {
                    G.rows.push_back(RowDesc{(int64_t)(r->d_hist - G.base) + (int64_t)p * r->n_bins, r->n_bins,
                                             r->first_n, 0});
                    G.pairs.push_back(make_group_pair(r->params[a], r->params[b]));
                }
        }
        const size_t R = G.rows.size();
        G.buf_off = buf_total;
        buf_total += (size_t)G.nq * 2 * R + R;  // prefactors, sums, row bounds
        G.I_base = nI;
        nI += G.req.size() * (size_t)G.nq;
        groups.push_back(std::move(G));
        g0 = g1;
    }
    // every group's intensities and bounds land in pinned host memory straight from its combine
    // (row gi of a group at I_base + gi nq), the out-of-range counts in one copy; all read after
    // the step's synchronisation
    double *hI = (double *)pinned(ctx, kPinStepI, sizeof(double) * nI);
    double *hrel = (double *)pinned(ctx, kPinStepRel, sizeof(double) * nI);
    uint64_t *hoor = (uint64_t *)pinned(ctx, kPinOor, sizeof(uint64_t) * batch.size());
    DBX_HIP(hipMemcpyAsync(hoor, d_oor, sizeof(uint64_t) * batch.size(), hipMemcpyDeviceToHost, ctx->stream));
    double *buf = (double *)scratch(ctx, kSlotQ, sizeof(double) * buf_total);
    for (Group &G : groups) {
        const size_t R = G.rows.size(), nq = (size_t)G.nq;
        double *dff = buf + G.buf_off, *dS = dff + R * nq, *dB = dS + R * nq;
        const double *dq = batch[G.req[0]]->d_q;
        prefactor_rows(ctx, G.pairs, dq, G.nq, dff);
        debye_structure_rows(ctx, G.base, G.rows, rstep, dq, G.nq, dS, dB);
        combine_models(ctx, dff, dS, dB, G.mrows, G.nq, hI + G.I_base, hrel + G.I_base);
        // a kept request keeps its own S_p, prefactors and bounds (reweighting)
        for (size_t gi = 0; gi < G.req.size(); gi++) {
            Request *r = batch[G.req[gi]];
            if (!(r->flags & DBX_REQ_KEEP)) continue;
            const size_t rn = (size_t)r->P * nq, r0 = (size_t)G.mrows[gi].x;
            DBX_HIP(hipMemcpyAsync(r->d_S, dS + r0 * nq, sizeof(double) * rn, hipMemcpyDeviceToDevice, ctx->stream));
            DBX_HIP(hipMemcpyAsync(r->d_ff, dff + r0 * nq, sizeof(double) * rn, hipMemcpyDeviceToDevice, ctx->stream));
            DBX_HIP(hipMemcpyAsync(r->d_B, dB + r0, sizeof(double) * r->P, hipMemcpyDeviceToDevice, ctx->stream));
        }
    }
    const double t_issued = now_ms();
    sync_pass2(ctx);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->st.step_enqueue_ms += t_issued - t_start;
        s->st.step_wait_ms += now_ms() - t_issued;
    }
    // requests whose bound misses the target: their partials again with FP64 arithmetic, one
    // pass 2 per q grid, over their rows of the group's sums and bounds, then combined again
    bool refined = false;
    if (ctx->pass2_mode == DBX_PASS2_MFMA) {
        for (Group &G : groups) {
            const size_t R = G.rows.size(), nq = (size_t)G.nq;
            double *dff = buf + G.buf_off, *dS = dff + R * nq, *dB = dS + R * nq;
            std::vector<int2> redo;
            std::vector<double *> Sd, Bd;
            std::vector<size_t> who;
            for (size_t gi = 0; gi < G.req.size(); gi++) {
                const size_t i = G.req[gi];
                Request *r = batch[i];
                const double *rel = hrel + G.I_base + gi * nq;
                bool bad = false;
                for (size_t k = 0; k < nq && !bad; k++) bad = exceeds_target(rel[k], ctx->rel_err_target);
                if (!bad || hoor[i]) continue;
                const int r0 = G.mrows[gi].x;
                redo.push_back(make_int2(r0, r->P));
                Sd.push_back(dS + (size_t)r0 * nq);
                Bd.push_back(dB + r0);
                who.push_back(gi);
            }
            if (redo.empty()) continue;
            structure_fp64_models(ctx, G.base, G.rows, redo, rstep, batch[G.req[0]]->d_q, G.nq, Sd, Bd);
            for (size_t gi : who) {
                Request *r = batch[G.req[gi]];
                const int r0 = G.mrows[gi].x;
                launch_combine(ctx->stream, dff, dS, dB, nullptr, 1, G.nq, hI + G.I_base + gi * nq,
                               hrel + G.I_base + gi * nq, make_int4(r0, r->P, 0, 0));
                if (r->flags & DBX_REQ_KEEP) {
                    DBX_HIP(hipMemcpyAsync(r->d_S, dS + (size_t)r0 * nq, sizeof(double) * r->P * nq,
                                           hipMemcpyDeviceToDevice, ctx->stream));
                    DBX_HIP(hipMemcpyAsync(r->d_B, dB + r0, sizeof(double) * r->P, hipMemcpyDeviceToDevice,
                                           ctx->stream));
                }
            }
            refined = true;
        }
        if (refined) sync_pass2(ctx);
    }
    for (Group &G : groups)
        for (size_t gi = 0; gi < G.req.size(); gi++) {
            const size_t i = G.req[gi];
            Request *r = batch[i];
            const double *I = hI + G.I_base + gi * (size_t)G.nq, *rel = hrel + G.I_base + gi * (size_t)G.nq;
            r->I.assign(I, I + G.nq);
            r->rel.assign(rel, rel + G.nq);
            r->oor = hoor[i];
            r->image.release();  // its upload is complete
        }
}

// pair distances of a request: N(N+1)/2
static double request_pairs(const Request *r)
{
    const double n = (double)r->n_atoms;
    return 0.5 * n * (n + 1);
}

static void lane_loop(dbx_server *s, int li)
{
    (void)hipSetDevice(s->device);
    dbx_ctx *ctx = s->lane[li];
    for (;;) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                admit(s);
                if (!s->admitted_q.empty()) {
                    // batching window: while requests are still being prepared and the step is
                    // not full, wait for them up to batch_window_ms after the first admission
                    const double left = s->admitted_q.front()->t_admit + s->cfg.batch_window_ms - now_ms();
                    double pairs = 0;
                    for (Request *r : s->admitted_q) pairs += request_pairs(r);
                    const bool full = (int)s->admitted_q.size() >= s->cfg.max_batch || pairs >= s->cfg.max_step_pairs;
                    if (left > 0 && !full && !s->preparing.empty()) {
                        s->cv_sched.wait_for(lk, std::chrono::duration<double, std::milli>(left));
                        continue;
                    }
                    break;
                }
                s->cv_sched.wait(lk);
            }
            // the step: admitted requests with the head's bin width, within the caps
            const double rstep = s->admitted_q.front()->rstep;
            double pairs = 0;
            for (auto it = s->admitted_q.begin(); it != s->admitted_q.end();) {
                Request *r = *it;
                const double pr = request_pairs(r);
                if (r->rstep != rstep || (!batch.empty() && (pairs + pr > s->cfg.max_step_pairs ||
                                                             (int)batch.size() >= s->cfg.max_batch))) {
                    ++it;
                    continue;
                }
                batch.push_back(r);
                pairs += pr;
                r->state = kRunning;
                it = s->admitted_q.erase(it);
            }
            // distinct q grids of the step (one contraction each)
            std::vector<const std::vector<double> *> grids;
            for (Request *r : batch) {
                bool seen = false;
                for (auto *g : grids) seen = seen || *g == r->q;
                if (!seen) grids.push_back(&r->q);
            }
            s->st.steps++;
            s->st.max_step_requests = std::max(s->st.max_step_requests, (double)batch.size());
            s->st.max_step_qgrids = std::max(s->st.max_step_qgrids, (double)grids.size());
            s->st.pairs += pairs;
            s->running_steps++;
        }
        int code = DBX_OK;
        std::string msg;
        try {
            run_step(s, ctx, batch);
        } catch (const Error &e) {
            code = e.code;
            msg = e.what();
        } catch (const std::exception &e) {
            code = DBX_ERR_HIP;
            msg = e.what();
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->running_steps--;
        const double t = now_ms();
        for (Request *r : batch) {
            r->t_done = t;
            if (code != DBX_OK) {
                fail(*r, code, msg);
            } else if (r->oor) {
                fail(*r, DBX_ERR_RANGE,
                     std::to_string(r->oor) + " pair distances fall at or beyond the last bin (rmax too small)");
            } else {
                r->state = kDone;
            }
            if (r->state == kDone) s->st.completed++;
            else s->st.failed++;
            if (!(r->flags & DBX_REQ_KEEP) || r->state != kDone) free_request_device(*s->arena, *r);
        }
        s->st.resident_bytes = s->arena->used();
        s->cv_done.notify_all();
        s->cv_sched.notify_all();
    }
}

// A caller inside an entry point of a server: refused once dbx_server_close has begun;
// dbx_server_close waits until every admitted caller has left before it frees the server.
struct ApiCall {
    dbx_server *s;
    bool ok;
    explicit ApiCall(dbx_server *sv) : s(sv)
    {
        std::lock_guard<std::mutex> lk(s->mu);
        ok = !s->stop;
        if (ok) s->in_api++;
    }
    ~ApiCall()
    {
        if (!ok) return;
        std::lock_guard<std::mutex> lk(s->mu);
        if (--s->in_api == 0) s->cv_api.notify_all();
    }
    ApiCall(const ApiCall &) = delete;
    ApiCall &operator=(const ApiCall &) = delete;
};

}  // namespace dbx

using namespace dbx;

extern "C" {

int dbx_server_open(int device, const dbx_server_config *cfg, dbx_server **out)
{
    if (!out) return DBX_ERR_ARG;
    *out = nullptr;
    dbx_server *s = new (std::nothrow) dbx_server;
    if (!s) return DBX_ERR_NOMEM;
    s->device = device;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 256;
    if (!(s->cfg.max_step_pairs > 0)) s->cfg.max_step_pairs = 2e10;
    if (s->cfg.prep_threads <= 0) {
        // the host's cores less the two lanes, the caller and one spare, at most 12
        const int hw = (int)std::thread::hardware_concurrency();
        s->cfg.prep_threads = hw > 0 ? std::max(2, std::min(12, hw - 4)) : 6;
    }
    int rc = DBX_OK;
    for (int i = 0; i < 2 && rc == DBX_OK; i++) rc = dbx_open(device, 0, &s->lane[i]);
    if (rc == DBX_OK) rc = dbx_open(device, 0, &s->rw_ctx);
    if (rc == DBX_OK) {
        // 0: the default target; negative: never recompute
        const double target = s->cfg.rel_err_target == 0.0 ? 1e-6 : std::max(0.0, s->cfg.rel_err_target);
        for (int i = 0; i < 2; i++) {
            s->lane[i]->pass2_mode = s->cfg.pass2_mode;
            s->lane[i]->rel_err_target = target;
        }
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) rc = DBX_ERR_HIP;
        size_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : (size_t)(0.75 * (double)fr);
        budget &= ~(size_t)255;
        if (rc == DBX_OK && hipMalloc(&s->arena_mem, budget) != hipSuccess) rc = DBX_ERR_NOMEM;
        if (rc == DBX_OK) {
            s->arena.reset(new DevArena((char *)s->arena_mem, budget));
            s->st.budget_bytes = budget;
        }
    }
    if (rc != DBX_OK) {
        for (dbx_ctx *c : {s->lane[0], s->lane[1], s->rw_ctx})
            if (c) dbx_close(c);
        if (s->arena_mem) (void)hipFree(s->arena_mem);
        delete s;
        return rc;
    }
    try {
        for (int i = 0; i < s->cfg.prep_threads; i++) s->threads.emplace_back(prep_loop, s);
        for (int i = 0; i < 2; i++) s->threads.emplace_back(lane_loop, s, i);
    } catch (const std::exception &) {
        // a thread that could not start: stop and join those that did, release everything
        dbx_server_close(s);
        return DBX_ERR_NOMEM;
    }
    *out = s;
    return DBX_OK;
}

int dbx_server_close(dbx_server *s)
{
    if (!s) return DBX_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_prep.notify_all();
    s->cv_sched.notify_all();
    s->cv_done.notify_all();
    {
        // callers inside an entry point leave (waiters return DBX_ERR_STATE) before anything
        // they may touch is freed
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_api.wait(lk, [&] { return s->in_api == 0; });
    }
    for (auto &t : s->threads) t.join();
    (void)hipSetDevice(s->device);
    for (auto &kv : s->reqs) {
        Request &r = *kv.second;
        free_request_device(*s->arena, r);
        if (r.h_pairs) (void)hipHostFree(r.h_pairs);
        if (r.h_I) (void)hipHostFree(r.h_I);
    }
    for (dbx_ctx *c : {s->lane[0], s->lane[1], s->rw_ctx})
        if (c) dbx_close(c);
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    delete s;
    return DBX_OK;
}

const char *dbx_server_last_error(dbx_server *s)
{
    if (!s) return "null server";
    ApiCall call(s);
    if (!call.ok) return "server closing";
    // a copy per calling thread: the lanes and other callers may replace the message
    thread_local std::string msg;
    std::lock_guard<std::mutex> lk(s->mu);
    msg = s->err;
    return msg.c_str();
}

int dbx_submit(dbx_server *s, const dbx_request *q, uint64_t *ticket)
{
    if (!s || !q || !ticket) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    if (q->n_groups <= 0 || !q->group_n || !q->params || !q->q || q->nq <= 0 || q->n_bins <= 0 ||
        !(q->rstep > 0) || q->first_n < 0)
        return DBX_ERR_ARG;
    int64_t n = 0;
    for (int g = 0; g < q->n_groups; g++) {
        if (q->group_n[g] < 0) return DBX_ERR_ARG;
        n += q->group_n[g];
    }
    if (n > 0 && (!q->x || !q->y || !q->z)) return DBX_ERR_ARG;
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(q->x[i]) || !std::isfinite(q->y[i]) || !std::isfinite(q->z[i])) return DBX_ERR_ARG;
    std::unique_ptr<Request> r;
    try {
        r.reset(new Request);
        r->x.assign(q->x, q->x + n);
        r->y.assign(q->y, q->y + n);
        r->z.assign(q->z, q->z + n);
        r->group_n.assign(q->group_n, q->group_n + q->n_groups);
        r->params.assign(q->params, q->params + q->n_groups);
        r->q.assign(q->q, q->q + q->nq);
    } catch (const std::bad_alloc &) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "host allocation failed";
        return DBX_ERR_NOMEM;
    }
    r->rstep = q->rstep;
    r->n_bins = q->n_bins;
    r->first_n = q->first_n;
    r->flags = q->flags;
    r->P = q->n_groups * (q->n_groups + 1) / 2;
    r->n_atoms = n;
    r->est_bytes = request_bytes_lower(*r);
    // histograms alone beyond the budget: refuse now
    if (sizeof(uint64_t) * (size_t)r->P * (size_t)r->n_bins + 16 * (size_t)n > s->arena->capacity()) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request larger than the serving budget";
        return DBX_ERR_BUDGET;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t id = s->next_id;
    Request *rp = r.get();
    rp->id = id;
    rp->t_submit = now_ms();
    // registered in all three places or in none
    try {
        s->reqs[id] = std::move(r);
        s->preparing.insert(id);
        s->prep_q.push_back(rp);
    } catch (const std::bad_alloc &) {
        s->preparing.erase(id);
        s->reqs.erase(id);
        s->err = "host allocation failed";
        return DBX_ERR_NOMEM;
    }
    s->next_id++;
    *ticket = id;
    s->st.submitted++;
    s->cv_prep.notify_one();
    return DBX_OK;
}

int dbx_poll(dbx_server *s, uint64_t ticket, int *state)
{
    if (!s || !state) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return DBX_ERR_ARG;
    const int st = it->second->state;
    *state = st == kDone ? DBX_REQ_DONE : st == kFailed ? DBX_REQ_FAILED : st == kRunning ? DBX_REQ_RUNNING
                                                                                         : DBX_REQ_QUEUED;
    return DBX_OK;
}

int dbx_wait(dbx_server *s, uint64_t ticket, double *I_out, double *latency_ms, double *rel_err_out)
{
    if (!s) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(ticket) == s->reqs.end()) return DBX_ERR_ARG;
    // looked up again at every wake-up: another thread may release the request meanwhile
    Request *r = nullptr;
    bool stalled = false;
    s->cv_done.wait(lk, [&] {
        auto it = s->reqs.find(ticket);
        r = it == s->reqs.end() ? nullptr : it->second.get();
        if (!r || r->state == kDone || r->state == kFailed || s->stop) return true;
        if (r->state == kReady) {
            if (admit(s) > 0) s->cv_sched.notify_all();
            stalled = admission_stalled(s, r);
        }
        return stalled;
    });
    if (!r) return DBX_ERR_ARG;
    if (stalled) {
        // nothing runs or waits that could free memory: only a dbx_release can; the request
        // stays queued and a later wait returns its result
        s->err = "request " + std::to_string(ticket) +
                 " waits for device memory held by finished DBX_REQ_KEEP requests: release some first";
        return DBX_ERR_BUDGET;
    }
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (r->state != kDone && r->state != kFailed) return DBX_ERR_STATE;
    if (I_out && !r->I.empty()) std::memcpy(I_out, r->I.data(), sizeof(double) * r->I.size());
    if (rel_err_out && !r->rel.empty()) std::memcpy(rel_err_out, r->rel.data(), sizeof(double) * r->rel.size());
    if (r->state == kFailed) {
        s->err = r->msg;
        return r->err;
    }
    return DBX_OK;
}

int dbx_reweight(dbx_server *s, uint64_t ticket, const dbx_group_params *params, double *I_out,
                 double *rel_err_out)
{
    if (!s || !params || !I_out) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    // the decode lane is held from the lookup to the last use of the request: dbx_release
    // takes it too before it frees one
    std::lock_guard<std::mutex> rl(s->rw_mu);
    Request *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return DBX_ERR_ARG;
        r = it->second.get();
        if (r->state != kDone || !(r->flags & DBX_REQ_KEEP) || !r->d_S) {
            s->err = "reweight needs a finished request submitted with DBX_REQ_KEEP";
            return DBX_ERR_STATE;
        }
    }
    try {
        (void)hipSetDevice(s->device);
        const int nq = (int)r->q.size();
        hipStream_t st = s->rw_ctx->stream;
        if (!r->h_pairs) {
            DBX_HIP(hipHostMalloc((void **)&r->h_pairs, sizeof(GroupPair) * r->P, hipHostMallocDefault));
            DBX_HIP(hipHostMalloc((void **)&r->h_I, sizeof(double) * 2 * nq, hipHostMallocDefault));
        }
        const int S = (int)r->group_n.size();
        int p = 0;
        for (int a = 0; a < S; a++)
            for (int b = a; b < S; b++, p++) r->h_pairs[p] = make_group_pair(params[a], params[b]);
        if (!r->rw_exec) {
            // decode of a resident request: new form factors -> prefactors -> I = sum_p ff_p S_p,
            // captured once and replayed for every reweight
            const int4 mr = make_int4(0, r->P, 0, 0);
            DBX_HIP(hipMemcpyAsync(r->d_mrows, &mr, sizeof(int4), hipMemcpyHostToDevice, st));
            DBX_HIP(hipStreamSynchronize(st));
            DBX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            DBX_HIP(hipMemcpyAsync(r->d_pairs, r->h_pairs, sizeof(GroupPair) * r->P, hipMemcpyHostToDevice, st));
            launch_prefactor(st, r->d_pairs, r->P, r->d_q, nq, r->d_ff);
            launch_combine(st, r->d_ff, r->d_S, r->d_B, r->d_mrows, 1, nq, r->d_I, r->d_rel);
            DBX_HIP(hipMemcpyAsync(r->h_I, r->d_I, sizeof(double) * nq, hipMemcpyDeviceToHost, st));
            DBX_HIP(hipMemcpyAsync(r->h_I + nq, r->d_rel, sizeof(double) * nq, hipMemcpyDeviceToHost, st));
            DBX_HIP(hipStreamEndCapture(st, &r->rw_graph));
            DBX_HIP(hipGraphInstantiate(&r->rw_exec, r->rw_graph, nullptr, nullptr, 0));
        }
        DBX_HIP(hipGraphLaunch(r->rw_exec, st));
        DBX_HIP(hipStreamSynchronize(st));
        std::memcpy(I_out, r->h_I, sizeof(double) * nq);
        if (rel_err_out) std::memcpy(rel_err_out, r->h_I + nq, sizeof(double) * nq);
    } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "host allocation failed";
        return DBX_ERR_NOMEM;
    }
    return DBX_OK;
}

int dbx_request_histogram(dbx_server *s, uint64_t ticket, uint64_t *hist_out)
{
    if (!s || !hist_out) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    std::lock_guard<std::mutex> rl(s->rw_mu);  // as in dbx_reweight
    Request *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return DBX_ERR_ARG;
        r = it->second.get();
        if (r->state != kDone || !r->d_hist) {
            s->err = "histograms are kept only for finished DBX_REQ_KEEP requests";
            return DBX_ERR_STATE;
        }
    }
    (void)hipSetDevice(s->device);
    if (hipMemcpyAsync(hist_out, r->d_hist, sizeof(uint64_t) * (size_t)r->P * r->n_bins, hipMemcpyDeviceToHost,
                       s->rw_ctx->stream) != hipSuccess ||
        hipStreamSynchronize(s->rw_ctx->stream) != hipSuccess)
        return DBX_ERR_HIP;
    return DBX_OK;
}

int dbx_release(dbx_server *s, uint64_t ticket)
{
    if (!s) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return DBX_ERR_ARG;
        if (it->second->state != kDone && it->second->state != kFailed) {
            s->err = "release of a request that has not finished";
            return DBX_ERR_STATE;
        }
        r = std::move(it->second);
        s->reqs.erase(it);
    }
    {
        // a reweight of this request may still hold the decode lane
        std::lock_guard<std::mutex> rl(s->rw_mu);
        std::lock_guard<std::mutex> lk(s->mu);
        free_request_device(*s->arena, *r);
        s->st.resident_bytes = s->arena->used();
    }
    if (r->h_pairs) (void)hipHostFree(r->h_pairs);
    if (r->h_I) (void)hipHostFree(r->h_I);
    s->cv_sched.notify_all();
    return DBX_OK;
}

int dbx_server_stats_get(dbx_server *s, dbx_server_stats *out)
{
    if (!s || !out) return DBX_ERR_ARG;
    ApiCall call(s);
    if (!call.ok) return DBX_ERR_STATE;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return DBX_OK;
}

}  // extern "C"

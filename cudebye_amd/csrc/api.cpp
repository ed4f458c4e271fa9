// C ABI of libdebye (include/debye.h): context, models, pass 1, prefactor, pass 2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

#include "fastmath.h"
#include "mfdigits.h"
#include "internal.h"

namespace dbx {

// Slots start at kSlotMin bytes and grow at least twofold.  A growth does not wait: work
// already enqueued on the stream may still use the old buffer, which is retired (freed by
// release_retired once the stream is idle) rather than freed, so a growth inside a server step
// costs one allocation, not a wait for that step's earlier kernels (and a device-wide one in
// hipFree).  Only when the device has no room for the new buffer beside the old ones does it
// wait, free them and allocate the size asked for.
constexpr size_t kSlotMin = (size_t)64 << 10;

void *scratch(dbx_ctx *ctx, int slot, size_t bytes)
{
    if (ctx->slot_bytes[slot] < bytes || !ctx->slot_ptr[slot]) {
        const size_t old = ctx->slot_bytes[slot];
        size_t want = std::max({bytes + bytes / 4, 2 * old, kSlotMin});
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            DBX_HIP(hipStreamSynchronize(ctx->stream));
            release_retired(ctx);
            if (ctx->slot_ptr[slot]) DBX_HIP(hipFree(ctx->slot_ptr[slot]));
            ctx->slot_ptr[slot] = nullptr;
            ctx->slot_bytes[slot] = 0;
            want = std::max(bytes + bytes / 4, kSlotMin);
            if (hipMalloc(&p, want) != hipSuccess) {
                (void)hipGetLastError();
                want = std::max<size_t>(bytes, 256);
                DBX_HIP(hipMalloc(&p, want));
            }
        }
        if (ctx->slot_ptr[slot]) ctx->retired_dev.push_back(ctx->slot_ptr[slot]);
        ctx->slot_ptr[slot] = p;
        ctx->slot_bytes[slot] = want;
        // a new buffer holds no q grid, even at an address an earlier one had
        if (slot == kSlotQin) ctx->q_dev = nullptr;
    }
    return ctx->slot_ptr[slot];
}

void *pinned(dbx_ctx *ctx, int slot, size_t bytes)
{
    if (ctx->pin_bytes[slot] < bytes || !ctx->pin_ptr[slot]) {
        // as scratch: the old buffer may be the destination of a download still enqueued
        const size_t want = std::max({bytes + bytes / 2, 2 * ctx->pin_bytes[slot], kSlotMin});
        void *p = nullptr;
        DBX_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
        if (ctx->pin_ptr[slot]) ctx->retired_pin.push_back(ctx->pin_ptr[slot]);
        ctx->pin_ptr[slot] = p;
        ctx->pin_bytes[slot] = want;
    }
    return ctx->pin_ptr[slot];
}

void release_retired(dbx_ctx *ctx)
{
    for (void *p : ctx->retired_dev) (void)hipFree(p);
    for (void *p : ctx->retired_pin) (void)hipHostFree(p);
    ctx->retired_dev.clear();
    ctx->retired_pin.clear();
}

void *stage(dbx_ctx *ctx, size_t bytes)
{
    const size_t a = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    if (ctx->stage_used + a > ctx->stage_cap) {
        // every earlier region's upload has run once the stream is idle
        DBX_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stage_used = 0;
        if (a > ctx->stage_cap) {
            if (ctx->stage_buf) DBX_HIP(hipHostFree(ctx->stage_buf));
            ctx->stage_buf = nullptr;
            ctx->stage_cap = 0;
            const size_t want = std::max<size_t>(a + a / 2, (size_t)1 << 20);
            void *p = nullptr;
            DBX_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
            ctx->stage_buf = (char *)p;
            ctx->stage_cap = want;
        }
    }
    char *p = ctx->stage_buf + ctx->stage_used;
    ctx->stage_used += a;
    return p;
}

void stage_reset(dbx_ctx *ctx) { ctx->stage_used = 0; }

// host copy of a call's results into the caller's (pageable) memory: large ones split over
// four threads (one thread copies ~10 GB/s, so a 256-model batch's 8.4 MB of I took 0.7 ms
// with the device idle)
static void copy_out(void *dst, const void *src, size_t bytes)
{
    constexpr size_t kSplit = (size_t)4 << 20;
    if (bytes < kSplit) {
        std::memcpy(dst, src, bytes);
        return;
    }
    constexpr int kThreads = 4;
    const size_t part = (bytes / kThreads + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (int t = 1; t < kThreads; t++) {
        const size_t o = part * t;
        if (o >= bytes) break;
        auto piece = [=] { std::memcpy((char *)dst + o, (const char *)src + o, std::min(part, bytes - o)); };
        try {
            th.emplace_back(piece);
        } catch (const std::exception &) {
            piece();  // no thread to spare: this one copies it
        }
    }
    std::memcpy(dst, src, std::min(part, bytes));
    for (auto &t : th) t.join();
}

const double *device_q(dbx_ctx *ctx, const double *q, int nq)
{
    const size_t b = sizeof(double) * (size_t)nq;
    double *d = (double *)scratch(ctx, kSlotQin, b);
    if (d == ctx->q_dev && ctx->q_host.size() == (size_t)nq && std::memcmp(ctx->q_host.data(), q, b) == 0)
        return d;
    double *h = (double *)stage(ctx, b);
    std::memcpy(h, q, b);
    DBX_HIP(hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, ctx->stream));
    ctx->q_host.assign(q, q + nq);
    ctx->q_dev = d;
    return d;
}

// ---------------------------------------------------------------------------------
// spatial tiling of a model (host)
// ---------------------------------------------------------------------------------
namespace {

struct Pt {  // 16 bytes: the median splits move these records
    float x, y, z;
    int32_t id;  // atom index (models hold fewer than 2^31 atoms)
};

// One median split of pts[lo, hi) (more than `unit` atoms) along its longest bounding-box
// axis, ties broken by atom index; returns the split point.  The left part receives a whole
// number of full leaves, so only the last leaf of a range can be short.
int64_t split_range(std::vector<Pt> &pts, int64_t lo, int64_t hi, int64_t unit)
{
    const int64_t n = hi - lo;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i < hi; i++) {
        const float v[3] = {pts[i].x, pts[i].y, pts[i].z};
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], v[a]);
            mx[a] = std::max(mx[a], v[a]);
        }
    }
    int ax = 0;
    for (int a = 1; a < 3; a++)
        if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
    const int64_t nt = (n + unit - 1) / unit;
    const int64_t nl = (nt / 2) * unit;
    // one comparator per axis (no per-comparison axis switch); ties broken by atom index
    auto split = [&](auto key) {
        std::nth_element(pts.begin() + lo, pts.begin() + lo + nl, pts.begin() + hi,
                         [&](const Pt &a, const Pt &b) {
                             const float ka = key(a), kb = key(b);
                             return ka < kb || (ka == kb && a.id < b.id);
                         });
    };
    if (ax == 0)
        split([](const Pt &p) { return p.x; });
    else if (ax == 1)
        split([](const Pt &p) { return p.y; });
    else
        split([](const Pt &p) { return p.z; });
    return lo + nl;
}

// Recursive median bisection of pts[lo, hi) into leaves of at most `unit` atoms, appended
// to `out` depth first, left before right.
void kd_split(std::vector<Pt> &pts, int64_t lo, int64_t hi, int64_t unit,
              std::vector<std::pair<int64_t, int64_t>> &out)
{
    std::vector<std::pair<int64_t, int64_t>> stack{{lo, hi}};
    while (!stack.empty()) {
        const auto r = stack.back();
        stack.pop_back();
        if (r.second - r.first <= unit) {
            out.push_back(r);
            continue;
        }
        const int64_t m = split_range(pts, r.first, r.second, unit);
        stack.push_back({m, r.second});
        stack.push_back({r.first, m});
    }
}

// The same leaves for a large range with the work on several threads: the range is split
// here, its right half goes to a new thread while this one takes the left, `depth` levels
// deep (2^depth threads at most); the leaves are joined in the same depth-first order.
void kd_split_parallel(std::vector<Pt> &pts, int64_t lo, int64_t hi, int64_t unit, int depth,
                       std::vector<std::pair<int64_t, int64_t>> &out)
{
    constexpr int64_t kParallelAtoms = 32768;
    if (depth == 0 || hi - lo < kParallelAtoms) {
        kd_split(pts, lo, hi, unit, out);
        return;
    }
    const int64_t m = split_range(pts, lo, hi, unit);
    std::vector<std::pair<int64_t, int64_t>> right;
    std::exception_ptr err;
    std::thread t([&] {
        try {
            kd_split_parallel(pts, m, hi, unit, depth - 1, right);
        } catch (...) {
            err = std::current_exception();
        }
    });
    try {
        kd_split_parallel(pts, lo, m, unit, depth - 1, out);
    } catch (...) {
        t.join();
        throw;
    }
    t.join();
    if (err) std::rethrow_exception(err);
    out.insert(out.end(), right.begin(), right.end());
}

}  // namespace

// metadata of the tile pts[lo, hi): bounding-box centre and a radius bounding every atom
static Tile make_tile(const std::vector<Pt> &pts, int64_t lo, int64_t hi, int g, int64_t pos_start,
                      int32_t parent)
{
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i < hi; i++) {
        const double v[3] = {pts[i].x, pts[i].y, pts[i].z};
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], v[a]);
            mx[a] = std::max(mx[a], v[a]);
        }
    }
    Tile t;
    t.cx = (float)(0.5 * (mn[0] + mx[0]));
    t.cy = (float)(0.5 * (mn[1] + mx[1]));
    t.cz = (float)(0.5 * (mn[2] + mx[2]));
    double r2 = 0;
    for (int64_t i = lo; i < hi; i++) {
        const double dx = pts[i].x - (double)t.cx, dy = pts[i].y - (double)t.cy, dz = pts[i].z - (double)t.cz;
        r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
    }
    t.rad = (float)(std::sqrt(r2) * (1 + 1e-6) + 1e-5);
    t.start = (int32_t)pos_start;
    t.count = (int32_t)(hi - lo);
    t.group = g;
    t.parent = parent;
    return t;
}

// Host part of a model: validation, tiling, tile metadata and the tile-ordered positions.
// Each subgroup is cut into blocks of 64 atoms by median bisection; blocks are stored in
// ascending size of their bounding box, each a contiguous run of positions.
void tile_model_host(const float *x, const float *y, const float *z, const int64_t *group_n,
                            int S, ModelDev &m, std::vector<float4> &pos_h)
{
    DBX_REQUIRE(S > 0 && S <= 64, DBX_ERR_ARG, "n_groups must be in [1, 64]");
    int64_t N = 0;
    for (int g = 0; g < S; g++) {
        DBX_REQUIRE(group_n[g] >= 0, DBX_ERR_ARG, "negative subgroup size");
        N += group_n[g];
    }
    DBX_REQUIRE(N < ((int64_t)1 << 31) - kTileAtoms * 64, DBX_ERR_ARG, "too many atoms");
    for (int64_t i = 0; i < N; i++)
        DBX_REQUIRE(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]), DBX_ERR_ARG,
                    "non-finite coordinate at atom " + std::to_string(i));
    m.n_groups = S;
    m.n_atoms = N;
    m.group_n.assign(group_n, group_n + S);
    m.bbox.assign(6 * S, 0.f);
    m.group_tile0.assign(S, 0);
    m.group_tile1.assign(S, 0);
    m.tiles_h.clear();
    pos_h.clear();
    pos_h.reserve(N);
    // the subgroups' bisections are independent: those of 4096 atoms or more run on threads of
    // their own (up to eight at once, each subgroup's top levels on further threads when few
    // run); the leaves are the same whatever runs where
    std::vector<std::vector<Pt>> gpts(S);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> gblocks(S);
    std::vector<int64_t> gstart(S);
    for (int g = 0, st = 0; g < S; st += (int)group_n[g], g++) gstart[g] = st;
    auto bisect = [&](int g, int depth) {
        const int64_t n = group_n[g], start = gstart[g];
        std::vector<Pt> &pts = gpts[g];
        pts.resize(n);
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = 0; i < n; i++) {
            pts[i] = {x[start + i], y[start + i], z[start + i], (int32_t)(start + i)};
            const float v[3] = {pts[i].x, pts[i].y, pts[i].z};
            for (int a = 0; a < 3; a++) {
                mn[a] = std::min(mn[a], v[a]);
                mx[a] = std::max(mx[a], v[a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            m.bbox[6 * g + a] = n ? mn[a] : 0.f;
            m.bbox[6 * g + 3 + a] = n ? mx[a] : 0.f;
        }
        if (n) kd_split_parallel(pts, 0, n, kTileAtoms, depth, gblocks[g]);
    };
    {
        std::vector<int> big;
        for (int g = 0; g < S; g++)
            if (group_n[g] >= 4096) big.push_back(g);
        const int depth = big.size() <= 1 ? 3 : big.size() <= 2 ? 2 : big.size() <= 4 ? 1 : 0;
        std::vector<std::exception_ptr> errs(S);
        for (size_t b0 = 0; b0 < big.size(); b0 += 8) {
            std::vector<std::thread> th;
            for (size_t b = b0; b < std::min(big.size(), b0 + 8); b++) {
                const int g = big[b];
                th.emplace_back([&, g] {
                    try {
                        bisect(g, depth);
                    } catch (...) {
                        errs[g] = std::current_exception();
                    }
                });
            }
            for (auto &t : th) t.join();
        }
        for (int g = 0; g < S; g++)
            if (group_n[g] < 4096) bisect(g, 0);
        for (auto &e : errs)
            if (e) std::rethrow_exception(e);
    }
    for (int g = 0; g < S; g++) {
        const std::vector<Pt> &pts = gpts[g];
        const std::vector<std::pair<int64_t, int64_t>> &blocks = gblocks[g];
        // blocks in ascending bounding-box diagonal
        std::vector<std::pair<double, size_t>> key(blocks.size());
        for (size_t b = 0; b < blocks.size(); b++) {
            Tile t = make_tile(pts, blocks[b].first, blocks[b].second, g, 0, 0);
            key[b] = {(double)t.rad, b};
        }
        std::stable_sort(key.begin(), key.end());
        m.group_tile0[g] = (int32_t)m.tiles_h.size();
        for (auto &kb : key) {
            const int64_t lo = blocks[kb.second].first, hi = blocks[kb.second].second;
            const int32_t self = (int32_t)m.tiles_h.size();
            m.tiles_h.push_back(make_tile(pts, lo, hi, g, (int64_t)pos_h.size(), self));
            for (int64_t i = lo; i < hi; i++) pos_h.push_back(make_float4(pts[i].x, pts[i].y, pts[i].z, 0.f));
        }
        m.group_tile1[g] = (int32_t)m.tiles_h.size();
        std::vector<Pt>().swap(gpts[g]);
    }
}

DevArena::DevArena(char *base, size_t bytes) : base_(base), cap_(bytes & ~(size_t)255)
{
    if (cap_) free_[0] = cap_;
}

void *DevArena::alloc(size_t bytes)
{
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > bytes) free_[off + bytes] = sz - bytes;
        live_[off] = bytes;
        used_ += bytes;
        return base_ + off;
    }
    return nullptr;
}

void DevArena::free(void *p)
{
    if (!p) return;
    const size_t off = (size_t)((char *)p - base_);
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t o = off, sz = lv->second;
    live_.erase(lv);
    used_ -= sz;
    auto nx = free_.lower_bound(o);
    if (nx != free_.end() && nx->first == o + sz) {
        sz += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == o) {
            o = pv->first;
            sz += pv->second;
            free_.erase(pv);
        }
    }
    free_[o] = sz;
}

size_t DevArena::largest_free() const
{
    size_t b = 0;
    for (auto &f : free_) b = std::max(b, f.second);
    return b;
}

void *model_alloc(ModelDev &m, size_t bytes)
{
    if (m.arena) {
        void *p = m.arena->alloc(bytes);
        DBX_REQUIRE(p != nullptr, DBX_ERR_BUDGET, "serving memory budget exhausted");
        return p;
    }
    void *p = nullptr;
    DBX_HIP(hipMalloc(&p, bytes));
    return p;
}

void model_free(ModelDev &m, void *p)
{
    if (!p) return;
    if (m.arena)
        m.arena->free(p);
    else
        (void)hipFree(p);
}

size_t model_device_bytes(const ModelDev &m, size_t n_pos)
{
    auto al = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; };
    return al(sizeof(float4) * n_pos) + al(sizeof(Tile) * m.tiles_h.size());
}

void upload_model(dbx_ctx *ctx, ModelDev &m, const std::vector<float4> &pos_h)
{
    m.pos = (float4 *)model_alloc(m, sizeof(float4) * std::max<size_t>(1, pos_h.size()));
    m.tiles = (Tile *)model_alloc(m, sizeof(Tile) * std::max<size_t>(1, m.tiles_h.size()));
    if (!pos_h.empty())
        DBX_HIP(hipMemcpyAsync(m.pos, pos_h.data(), sizeof(float4) * pos_h.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!m.tiles_h.empty())
        DBX_HIP(hipMemcpyAsync(m.tiles, m.tiles_h.data(), sizeof(Tile) * m.tiles_h.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    DBX_HIP(hipStreamSynchronize(ctx->stream));
}

static void build_model(dbx_ctx *ctx, const float *x, const float *y, const float *z,
                        const int64_t *group_n, int S, ModelDev &m)
{
    std::vector<float4> pos_h;
    tile_model_host(x, y, z, group_n, S, m, pos_h);
    upload_model(ctx, m, pos_h);
}

void free_model(ModelDev &m)
{
    for (void *p : {(void *)m.pos, (void *)m.tiles, (void *)m.d_jobs, (void *)m.d_edge2, (void *)m.d_list,
                    (void *)m.d_alist})
        model_free(m, p);
    m.pos = nullptr;
    m.tiles = nullptr;
    m.d_jobs = nullptr;
    m.d_edge2 = nullptr;
    m.d_list = nullptr;
    m.d_alist = nullptr;
}

// The I(q) of one model (rows [row0, row0 + P) of d_S, d_B; destination dest) and its bound to
// the host (hI, hrel: nq each).  When the bound exceeds the context's target at some q and the
// rows were summed on the matrix cores, they are recomputed with FP64 arithmetic from the
// model's histograms d_hist[P][n_bins] and combined again.  Synchronises the stream.  dI == hI
// (and drel == hrel): the combine writes pinned host memory directly, no copy.
static void finish_model(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep, int first_n,
                         const double *dq, int nq, const double *dff, double *dS, double *dB, int row0, int dest,
                         double *dI, double *drel, double *hI, double *hrel)
{
    const size_t nqd = (size_t)nq;
    auto fetch = [&] {
        if (dI != hI) {
            DBX_HIP(hipMemcpyAsync(hI, dI + dest * nqd, sizeof(double) * nqd, hipMemcpyDeviceToHost, ctx->stream));
            DBX_HIP(hipMemcpyAsync(hrel, drel + dest * nqd, sizeof(double) * nqd, hipMemcpyDeviceToHost,
                                   ctx->stream));
        }
        sync_pass2(ctx);
    };
    fetch();
    if (ctx->pass2_mode != DBX_PASS2_MFMA) return;
    bool bad = false;
    for (int i = 0; i < nq && !bad; i++) bad = exceeds_target(hrel[i], ctx->rel_err_target);
    if (!bad) return;
    P2Refining mark(ctx);
    debye_structure(ctx, d_hist, P, n_bins, rstep, first_n, dq, nq, dS + (size_t)row0 * nqd, dB + row0,
                    DBX_PASS2_FP64);
    launch_combine(ctx->stream, dff, dS, dB, nullptr, 1, nq, dI, drel, make_int4(row0, P, dest, 0));
    ctx->last_p2_refined++;
    fetch();
}

}  // namespace dbx

using namespace dbx;

extern "C" {

int dbx_open(int device, size_t hbm_budget, dbx_ctx **out)
{
    if (!out) return DBX_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return DBX_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return DBX_ERR_HIP;
    dbx_ctx *c = new (std::nothrow) dbx_ctx;
    if (!c) return DBX_ERR_NOMEM;
    c->device = device;
    // on failure dbx_close releases whatever was created (members start null)
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        dbx_close(c);
        return DBX_ERR_HIP;
    }
    c->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)fr);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->counters, 64) != hipSuccess ||
        hipHostMalloc((void **)&c->counters_h, 64, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->p2_readback_h, kReadbackBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->p2_status_h, 2 * sizeof(uint64_t) * kP2StatusMax, hipHostMallocDefault) !=
            hipSuccess) {
        dbx_close(c);
        return DBX_ERR_HIP;
    }
    // the allocation-free pass 2's scratch region on this device (device current)
    try {
        noalloc_init(device);
    } catch (const Error &) {
        dbx_close(c);
        return DBX_ERR_HIP;
    }
    for (int i = 0; i < 4; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            dbx_close(c);
            return DBX_ERR_HIP;
        }
    *out = c;
    return DBX_OK;
}

int dbx_close(dbx_ctx *ctx)
{
    if (!ctx) return DBX_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (dbx_model *m : ctx->models) {
        free_model(m->d);
        delete m;
    }
    ctx->models.clear();
    release_retired(ctx);
    for (void *p : ctx->slot_ptr)
        if (p) (void)hipFree(p);
    for (void *p : ctx->pin_ptr)
        if (p) (void)hipHostFree(p);
    if (ctx->counters) (void)hipFree(ctx->counters);
    for (int i = 0; i < 2; i++)
        if (ctx->ev_mf[i]) (void)hipEventDestroy(ctx->ev_mf[i]);
    if (ctx->counters_h) (void)hipHostFree(ctx->counters_h);
    if (ctx->p2_readback_h) (void)hipHostFree(ctx->p2_readback_h);
    if (ctx->p2_status_h) (void)hipHostFree(ctx->p2_status_h);
    if (ctx->stage_buf) (void)hipHostFree(ctx->stage_buf);
    for (int i = 0; i < 4; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DBX_OK;
}

const char *dbx_last_error(const dbx_ctx *ctx)
{
    if (!ctx) return "null context";
    // a copy per calling thread: another thread's call may replace the context's message
    thread_local std::string msg;
    std::lock_guard<std::recursive_mutex> lk(const_cast<dbx_ctx *>(ctx)->mu);
    msg = ctx->err;
    return msg.c_str();
}

void *dbx_stream(dbx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int dbx_q_from_2theta(double lambda, double tt_min, double tt_max, double tt_step, double *tt,
                      double *q, int cap)
{
    const double PI = 3.14159265358979323846;  // cuDebye-v15.cu:46
    if (!(tt_step > 0) || !(lambda > 0) || !std::isfinite(tt_min) || !std::isfinite(tt_max))
        return DBX_ERR_ARG;
    const double span = (tt_max - tt_min) / tt_step;
    if (!(span >= 0) || span > 2.0e9) return DBX_ERR_ARG;
    const int n = (int)span + 1;  // cuDebye-v15.cu:690
    if (tt && q && cap >= n) {
        for (int i = 0; i < n; i++) {
            tt[i] = tt_min + (double)i * tt_step;                   // :695
            q[i] = 4 * PI * std::sin(tt[i] / 2 * PI / 180) / lambda;  // :696
        }
    }
    return n;
}

int64_t dbx_channels(double rmax, double rstep)
{
    if (!(rstep > 0) || !(rmax >= 0)) return DBX_ERR_ARG;
    const double c = rmax / rstep;
    if (c > 2.0e9) return DBX_ERR_ARG;
    return (int64_t)(int)c + 2;  // cuDebye-v15.cu:734
}

int dbx_model_create(dbx_ctx *ctx, const float *x, const float *y, const float *z,
                     const int64_t *group_n, int n_groups, dbx_model **out)
{
    if (!out || !group_n || (n_groups > 0 && (!x || !y || !z))) return DBX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        dbx_model *m = new dbx_model;
        try {
            build_model(ctx, x, y, z, group_n, n_groups, m->d);
            ctx->models.insert(m);
        } catch (...) {
            free_model(m->d);
            delete m;
            throw;
        }
        *out = m;
    });
}

int dbx_model_destroy(dbx_ctx *ctx, dbx_model *m)
{
    if (!m) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        // only a live model of this context (a second destroy is an argument error)
        DBX_REQUIRE(ctx->models.count(m) == 1, DBX_ERR_ARG, "not a live model of this context");
        // the model is released whatever the synchronisation returns (a failed one has ended
        // the stream's work); its error is reported after
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        ctx->models.erase(m);
        free_model(m->d);
        delete m;
        if (e != hipSuccess)
            throw Error(DBX_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    });
}

int dbx_set_pass2_mode(dbx_ctx *ctx, int mode)
{
    if (!ctx || (mode != DBX_PASS2_MFMA && mode != DBX_PASS2_FP64)) return DBX_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->pass2_mode = mode;
    return DBX_OK;
}

int dbx_set_rel_err_target(dbx_ctx *ctx, double target)
{
    if (!ctx || std::isnan(target)) return DBX_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->rel_err_target = target > 0 ? target : 0.0;
    return DBX_OK;
}

int dbx_last_pass1_stats(const dbx_ctx *ctx, int64_t *n_jobs, uint64_t *window_misses, uint64_t *retried)
{
    if (!ctx) return DBX_ERR_ARG;
    if (n_jobs) *n_jobs = ctx->last_p1_jobs;
    if (window_misses) *window_misses = ctx->last_p1_misses;
    if (retried) *retried = ctx->last_p1_retried;
    return DBX_OK;
}

int dbx_last_kernel_times(dbx_ctx *ctx, double *pass1_ms, double *pass2_ms)
{
    if (!ctx) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        if (pass1_ms) *pass1_ms = ctx->last_p1_kernel_ms;
        if (pass2_ms) {
            float t = 0.f;
            if (ctx->p2_timed) {
                DBX_HIP(hipEventSynchronize(ctx->ev[3]));
                DBX_HIP(hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
            }
            *pass2_ms = t;
        }
    });
}

int dbx_model_info(const dbx_model *m, int64_t *n_atoms, int *n_groups, int *n_pairs)
{
    if (!m) return DBX_ERR_ARG;
    if (n_atoms) *n_atoms = m->d.n_atoms;
    if (n_groups) *n_groups = m->d.n_groups;
    if (n_pairs) *n_pairs = m->d.n_groups * (m->d.n_groups + 1) / 2;
    return DBX_OK;
}

int dbx_histogram(dbx_ctx *ctx, dbx_model *m, double rstep, int64_t n_bins, uint64_t *out,
                  int out_is_device, uint64_t *n_out_of_range)
{
    if (!m || !out) return DBX_ERR_ARG;
    int rc = DBX_OK;
    int grc = guarded(ctx, [&] {
        const int P = m->d.n_groups * (m->d.n_groups + 1) / 2;
        const size_t bytes = sizeof(uint64_t) * (size_t)P * n_bins;
        uint64_t *d = out_is_device ? out : (uint64_t *)scratch(ctx, kSlotHist, bytes);
        P1Stats st = pass1_histograms(ctx, m->d, rstep, n_bins, d);
        if (!out_is_device) {
            DBX_HIP(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
            DBX_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (n_out_of_range) *n_out_of_range = st.out_of_range;
        if (st.out_of_range) {
            rc = DBX_ERR_RANGE;
            ctx->err = std::to_string(st.out_of_range) +
                       " pair distances fall at or beyond the last bin (rmax too small)";
        }
    });
    if (grc != DBX_OK) return grc;
    return rc;
}

int dbx_prefactor(dbx_ctx *ctx, const dbx_group_params *ga, const dbx_group_params *gb,
                  const double *q, int nq, double *ffoobb)
{
    if (!ga || !gb || !q || !ffoobb || nq <= 0) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        double *d = (double *)scratch(ctx, kSlotQ, sizeof(double) * 2 * (size_t)nq);
        DBX_HIP(hipMemcpyAsync(d, q, sizeof(double) * nq, hipMemcpyHostToDevice, ctx->stream));
        prefactor_dev(ctx, *ga, *gb, d, nq, d + nq);
        DBX_HIP(hipMemcpyAsync(ffoobb, d + nq, sizeof(double) * nq, hipMemcpyDeviceToHost, ctx->stream));
        DBX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int dbx_debye(dbx_ctx *ctx, const uint64_t *hist, int hist_is_device, int P, int64_t n_bins,
              double rstep, int first_n, const double *q, int nq, const double *ffoobb,
              double *I_out, double *S_out, double *rel_err_out)
{
    if (!hist || !q || !ffoobb || !I_out || P <= 0 || nq <= 0 || n_bins <= 0) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        const size_t hb = sizeof(uint64_t) * (size_t)P * n_bins;
        const uint64_t *dh = hist;
        if (!hist_is_device) {
            uint64_t *t = (uint64_t *)scratch(ctx, kSlotHist, hb);
            DBX_HIP(hipMemcpyAsync(t, hist, hb, hipMemcpyHostToDevice, ctx->stream));
            dh = t;
        }
        const size_t nqd = (size_t)nq;
        double *d = (double *)scratch(ctx, kSlotQ, sizeof(double) * (nqd * (2 * P + 2) + P));
        double *dff = d, *dS = dff + P * nqd, *dI = dS + P * nqd, *drel = dI + nqd, *dB = drel + nqd;
        const double *dq = device_q(ctx, q, nq);
        DBX_HIP(hipMemcpyAsync(dff, ffoobb, sizeof(double) * P * nqd, hipMemcpyHostToDevice,
                               ctx->stream));
        ctx->last_p2_refined = 0;
        double *hI = (double *)pinned(ctx, kPinI, sizeof(double) * nqd);
        double *hrel = (double *)pinned(ctx, kPinRel, sizeof(double) * nqd);
        debye_structure(ctx, dh, P, n_bins, rstep, first_n, dq, nq, dS, dB);
        combine_intensity(ctx, dff, dS, dB, P, nq, hI, hrel);  // straight into pinned host memory
        finish_model(ctx, dh, P, n_bins, rstep, first_n, dq, nq, dff, dS, dB, 0, 0, hI, hrel, hI, hrel);
        std::memcpy(I_out, hI, sizeof(double) * nqd);
        if (rel_err_out) std::memcpy(rel_err_out, hrel, sizeof(double) * nqd);
        if (S_out) {
            DBX_HIP(hipMemcpyAsync(S_out, dS, sizeof(double) * P * nqd, hipMemcpyDeviceToHost,
                                   ctx->stream));
            DBX_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int dbx_batch_intensity(dbx_ctx *ctx, const dbx_batch_item *items, int n_items, double rstep,
                        const double *q, int nq, double *I_out, uint64_t *n_out_of_range,
                        uint64_t *const *hist_out, double *rel_err_out)
{
    if (!items || n_items <= 0 || !q || nq <= 0 || !I_out) return DBX_ERR_ARG;
    for (int i = 0; i < n_items; i++)
        if (!items[i].model || !items[i].params || items[i].n_bins <= 0 || items[i].first_n < 0)
            return DBX_ERR_ARG;
    int rc = DBX_OK;
    int grc = guarded(ctx, [&] {
        // models longest first: pass 2 wants its rows ordered by length
        std::vector<int> order(n_items);
        for (int i = 0; i < n_items; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return items[a].n_bins > items[b].n_bins; });
        std::vector<int64_t> hoff(n_items);
        int64_t total = 0;
        int n_rows = 0;
        for (int i : order) {
            const int S = items[i].model->d.n_groups;
            hoff[i] = total;
            total += (int64_t)(S * (S + 1) / 2) * items[i].n_bins;
            n_rows += S * (S + 1) / 2;
        }
        uint64_t *dh = (uint64_t *)scratch(ctx, kSlotBatchHist, sizeof(uint64_t) * (size_t)total);
        uint64_t *door = (uint64_t *)scratch(ctx, kSlotOor, sizeof(uint64_t) * (size_t)n_items);
        const size_t nqd = (size_t)nq;
        double *d = (double *)scratch(ctx, kSlotQ,
                                      sizeof(double) * (nqd * (2 * (size_t)n_rows + 2 * n_items) + n_rows));
        double *dff = d, *dS = dff + n_rows * nqd, *dI = dS + n_rows * nqd, *drel = dI + n_items * nqd;
        double *dB = drel + n_items * nqd;
        const double *dq = device_q(ctx, q, nq);
        std::vector<P1Item> p1;
        std::vector<RowDesc> rows;
        std::vector<GroupPair> pairs;
        std::vector<int4> mrows;
        for (int i : order) {
            const dbx_batch_item &it = items[i];
            const int S = it.model->d.n_groups;
            P1Item pi{&it.model->d, it.n_bins, dh + hoff[i]};
            pi.d_oor = door + i;
            p1.push_back(pi);
            mrows.push_back(make_int4((int)rows.size(), S * (S + 1) / 2, i, 0));
            int p = 0;
            for (int a = 0; a < S; a++)
                for (int b = a; b < S; b++, p++) {
                    rows.push_back(RowDesc{hoff[i] + (int64_t)p * it.n_bins, it.n_bins, it.first_n, 0});
                    pairs.push_back(make_group_pair(it.params[a], it.params[b]));
                }
        }
        ctx->last_p2_refined = 0;
        pass1_batch(ctx, p1, rstep, false);
        prefactor_rows(ctx, pairs, dq, nq, dff);
        debye_structure_rows(ctx, dh, rows, rstep, dq, nq, dS, dB);
        combine_models(ctx, dff, dS, dB, mrows, nq, dI, drel);
        double *hI = (double *)pinned(ctx, kPinStepI, sizeof(double) * nqd * n_items);
        double *hrel = (double *)pinned(ctx, kPinStepRel, sizeof(double) * nqd * n_items);
        auto fetch = [&] {
            DBX_HIP(hipMemcpyAsync(hI, dI, sizeof(double) * nqd * n_items, hipMemcpyDeviceToHost, ctx->stream));
            DBX_HIP(hipMemcpyAsync(hrel, drel, sizeof(double) * nqd * n_items, hipMemcpyDeviceToHost, ctx->stream));
        };
        fetch();
        if (hist_out)
            for (int i = 0; i < n_items; i++)
                if (hist_out[i]) {
                    const int S = items[i].model->d.n_groups;
                    DBX_HIP(hipMemcpyAsync(hist_out[i], dh + hoff[i],
                                           sizeof(uint64_t) * (size_t)(S * (S + 1) / 2) * items[i].n_bins,
                                           hipMemcpyDeviceToHost, ctx->stream));
                }
        uint64_t *hoor = (uint64_t *)pinned(ctx, kPinOor, sizeof(uint64_t) * (size_t)n_items);
        DBX_HIP(hipMemcpyAsync(hoor, door, sizeof(uint64_t) * (size_t)n_items, hipMemcpyDeviceToHost, ctx->stream));
        sync_pass2(ctx);
        // models whose bound misses the target: their partials again with FP64 arithmetic
        // (all of them in one FP64 pass 2, written over their rows of dS and dB)
        if (ctx->pass2_mode == DBX_PASS2_MFMA) {
            std::vector<int4> redo;
            std::vector<int2> redo_rows;
            std::vector<double *> Sd, Bd;
            for (const int4 &mr : mrows) {
                const int i = mr.z;
                bool bad = false;
                for (int k = 0; k < nq && !bad; k++) bad = exceeds_target(hrel[i * nqd + k], ctx->rel_err_target);
                if (!bad) continue;
                redo.push_back(mr);
                redo_rows.push_back(make_int2(mr.x, mr.y));
                Sd.push_back(dS + (size_t)mr.x * nqd);
                Bd.push_back(dB + mr.x);
            }
            if (!redo.empty()) {
                structure_fp64_models(ctx, dh, rows, redo_rows, rstep, dq, nq, Sd, Bd);
                combine_models(ctx, dff, dS, dB, redo, nq, dI, drel);
                ctx->last_p2_refined = (int)redo.size();
                fetch();
                sync_pass2(ctx);
            }
        }
        copy_out(I_out, hI, sizeof(double) * nqd * n_items);
        if (rel_err_out) copy_out(rel_err_out, hrel, sizeof(double) * nqd * n_items);
        P1Stats st;
        pass1_read_stats(ctx, st);
        int bad = 0;
        for (int i = 0; i < n_items; i++) {
            if (n_out_of_range) n_out_of_range[i] = hoor[i];
            bad += hoor[i] != 0;
        }
        if (bad) {
            rc = DBX_ERR_RANGE;
            ctx->err = std::to_string(bad) + " model(s) have pair distances at or beyond their last bin";
        }
    });
    if (grc != DBX_OK) return grc;
    return rc;
}

int dbx_intensity(dbx_ctx *ctx, dbx_model *m, const dbx_group_params *params, double rstep,
                  int64_t n_bins, int first_n, const double *q, int nq, double *I_out,
                  uint64_t *hist_out, double *rel_err_out)
{
    if (!m || !params || !q || !I_out || nq <= 0) return DBX_ERR_ARG;
    int rc = DBX_OK;
    int grc = guarded(ctx, [&] {
        const int S = m->d.n_groups;
        const int P = S * (S + 1) / 2;
        const size_t hb = sizeof(uint64_t) * (size_t)P * n_bins;
        uint64_t *dh = (uint64_t *)scratch(ctx, kSlotHist, hb);
        P1Stats st = pass1_histograms(ctx, m->d, rstep, n_bins, dh, false);
        const size_t nqd = (size_t)nq;
        double *d = (double *)scratch(ctx, kSlotQ, sizeof(double) * (nqd * (2 * P + 2) + P));
        double *dff = d, *dS = dff + P * nqd, *dI = dS + P * nqd, *drel = dI + nqd, *dB = drel + nqd;
        const double *dq = device_q(ctx, q, nq);
        std::vector<GroupPair> pairs;
        for (int a = 0; a < S; a++)
            for (int b = a; b < S; b++) pairs.push_back(make_group_pair(params[a], params[b]));
        ctx->last_p2_refined = 0;
        // I and its bound land in pinned host memory straight from the combine kernel
        double *hI = (double *)pinned(ctx, kPinI, sizeof(double) * nqd);
        double *hrel = (double *)pinned(ctx, kPinRel, sizeof(double) * nqd);
        prefactor_rows(ctx, pairs, dq, nq, dff);
        debye_structure(ctx, dh, P, n_bins, rstep, first_n, dq, nq, dS, dB);
        combine_intensity(ctx, dff, dS, dB, P, nq, hI, hrel);
        if (hist_out) DBX_HIP(hipMemcpyAsync(hist_out, dh, hb, hipMemcpyDeviceToHost, ctx->stream));
        finish_model(ctx, dh, P, n_bins, rstep, first_n, dq, nq, dff, dS, dB, 0, 0, hI, hrel, hI, hrel);
        std::memcpy(I_out, hI, sizeof(double) * nqd);
        if (rel_err_out) std::memcpy(rel_err_out, hrel, sizeof(double) * nqd);
        pass1_read_stats(ctx, st);
        if (st.out_of_range) {
            rc = DBX_ERR_RANGE;
            ctx->err = std::to_string(st.out_of_range) +
                       " pair distances fall at or beyond the last bin (rmax too small)";
        }
    });
    if (grc != DBX_OK) return grc;
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// diagnostics (include/debye_debug.h)
// ---------------------------------------------------------------------------------
#include "../../include/debye_debug.h"

extern "C" {

int dbx_debug_model_tiles(const dbx_model *m, void *out, int cap)
{
    if (!m) return DBX_ERR_ARG;
    const int n = (int)m->d.tiles_h.size();
    if (out && cap >= n) std::memcpy(out, m->d.tiles_h.data(), sizeof(Tile) * n);
    return n;
}

int64_t dbx_debug_model_positions(dbx_ctx *ctx, const dbx_model *m, float *out4, int64_t cap)
{
    if (!m || !ctx) return DBX_ERR_ARG;
    const int64_t n = m->d.n_atoms;
    if (out4 && cap >= n && n > 0) {
        int rc = guarded(ctx, [&] {
            DBX_HIP(hipMemcpy(out4, m->d.pos, sizeof(float4) * n, hipMemcpyDeviceToHost));
        });
        if (rc != DBX_OK) return rc;
    }
    return n;
}

int dbx_debug_pass1_evaluations(const dbx_ctx *ctx, uint64_t *evaluated)
{
    if (!ctx || !evaluated) return DBX_ERR_ARG;
    *evaluated = ctx->last_p1_evals;
    return DBX_OK;
}

int dbx_debug_pass1_plan(const dbx_model *m, double rstep, void *jobs, int cap_jobs, float *edge2,
                         int cap_edge2, int *n_jobs, int *n_edge2)
{
    if (!m) return DBX_ERR_ARG;
    try {
        std::vector<P1Job> j;
        std::vector<float> e;
        std::vector<int32_t> al;
        plan_pass1(m->d, rstep, j, e, al);
        if (n_jobs) *n_jobs = (int)j.size();
        if (n_edge2) *n_edge2 = (int)e.size();
        if (jobs && cap_jobs >= (int)j.size()) std::memcpy(jobs, j.data(), sizeof(P1Job) * j.size());
        if (edge2 && cap_edge2 >= (int)e.size()) std::memcpy(edge2, e.data(), sizeof(float) * e.size());
    } catch (const Error &e) {
        return e.code;
    }
    return DBX_OK;
}

// Host-only tiling and pass-1 plan of a model given as arrays (no device needed).
int dbx_debug_plan_host(const float *x, const float *y, const float *z, const int64_t *group_n,
                        int n_groups, double rstep, void *tiles, int cap_tiles, float *pos4,
                        int64_t cap_pos, void *jobs, int cap_jobs, float *edge2, int cap_edge2,
                        int32_t *alist, int cap_alist, int *n_tiles, int *n_jobs, int *n_edge2,
                        int *n_alist)
{
    if (!x || !group_n) return DBX_ERR_ARG;
    try {
        ModelDev m;
        std::vector<float4> pos_h;
        tile_model_host(x, y, z, group_n, n_groups, m, pos_h);
        std::vector<P1Job> j;
        std::vector<float> e;
        std::vector<int32_t> al;
        plan_pass1(m, rstep, j, e, al);
        if (n_tiles) *n_tiles = (int)m.tiles_h.size();
        if (n_jobs) *n_jobs = (int)j.size();
        if (n_edge2) *n_edge2 = (int)e.size();
        if (n_alist) *n_alist = (int)al.size();
        if (alist && cap_alist >= (int)al.size()) std::memcpy(alist, al.data(), sizeof(int32_t) * al.size());
        if (tiles && cap_tiles >= (int)m.tiles_h.size())
            std::memcpy(tiles, m.tiles_h.data(), sizeof(Tile) * m.tiles_h.size());
        if (pos4 && cap_pos >= (int64_t)pos_h.size())
            std::memcpy(pos4, pos_h.data(), sizeof(float4) * pos_h.size());
        if (jobs && cap_jobs >= (int)j.size()) std::memcpy(jobs, j.data(), sizeof(P1Job) * j.size());
        if (edge2 && cap_edge2 >= (int)e.size()) std::memcpy(edge2, e.data(), sizeof(float) * e.size());
    } catch (const Error &e) {
        return e.code;
    }
    return DBX_OK;
}

// the pass-2 sine evaluated on the host (same source as the device code)
int dbx_debug_set_profiling(dbx_ctx *ctx, int on)
{
    if (!ctx) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        if (on && !ctx->ev_mf[0]) {
            DBX_HIP(hipEventCreate(&ctx->ev_mf[0]));
            DBX_HIP(hipEventCreate(&ctx->ev_mf[1]));
        }
        ctx->profiling = on != 0;
        ctx->mf_timed = false;
    });
}

int dbx_debug_pass2_profile(dbx_ctx *ctx, double *kernel_ms, double *mfma_flops, double *useful_flops,
                            int *n_bins_compacted, int *count_digits, double *mean_digits)
{
    if (!ctx) return DBX_ERR_ARG;
    return guarded(ctx, [&] {
        DBX_REQUIRE(ctx->mf_timed, DBX_ERR_STATE, "no profiled matrix-core pass 2 on this context");
        float ms = 0.f;
        DBX_HIP(hipEventElapsedTime(&ms, ctx->ev_mf[0], ctx->ev_mf[1]));
        if (kernel_ms) *kernel_ms = ms;
        if (mfma_flops) *mfma_flops = ctx->mf_flops;
        if (useful_flops) *useful_flops = ctx->mf_useful;
        if (n_bins_compacted) *n_bins_compacted = ctx->mf_bins;
        if (count_digits) *count_digits = ctx->mf_digits;
        if (mean_digits) *mean_digits = ctx->mf_mean_digits;
    });
}

int dbx_debug_pass2_refined(const dbx_ctx *ctx, int *n_models)
{
    if (!ctx || !n_models) return DBX_ERR_ARG;
    *n_models = ctx->last_p2_refined;
    return DBX_OK;
}

int dbx_debug_sin_host(const double *x, int n, double *out)
{
    if (!x || !out || n < 0) return DBX_ERR_ARG;
    for (int i = 0; i < n; i++) out[i] = sin_cw(x[i]);
    return DBX_OK;
}

int dbx_debug_sincos_host(const double *x, int n, double *s_out, double *c_out)
{
    if (!x || !s_out || !c_out || n < 0) return DBX_ERR_ARG;
    for (int i = 0; i < n; i++) sincos_cw(x[i], s_out[i], c_out[i]);
    return DBX_OK;
}

int dbx_debug_sinc_digits_host(const double *q, const double *r, int n, signed char *digits)
{
    if (!q || !r || !digits || n < 0 || n % 4) return DBX_ERR_ARG;
    for (int i = 0; i < n; i += 4) {
        uint32_t lo[4], hi[4];
        for (int u = 0; u < 4; u++) {
            const double Q = q[i + u] != 0.0 ? q[i + u] : 1e-30;
            sinc_word(sin_cw(Q * r[i + u]), kMfC / Q, 1.0 / r[i + u], lo[u], hi[u]);
        }
        int dj[kMfJ];
        pack4(lo, hi, dj);
        for (int j = 0; j < kMfJ; j++)
            for (int u = 0; u < 4; u++) digits[(size_t)(i + u) * kMfJ + j] = (signed char)(dj[j] >> (8 * u));
    }
    return DBX_OK;
}

}  // extern "C"

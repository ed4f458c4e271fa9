// Pass 1: all-pairs interatomic-distance histograms per subgroup pair.
//
// Replaces splitCalcDist and the dist / dist2 / care_dist2 kernels of cuDebye-v15.cu:88-272.
// The per-pair expression is the reference's (cuDebye-v15.cu:101-115):
//     dx, dy, dz = FP32 differences;  s = (dx*dx + dy*dy) + dz*dz, every operation rounded
//     (no fused multiply-add: the reference's .debx outputs are reproduced only by that form);
//     k = (int) roundf(sqrtf(s) * rp),  rp = (float)(1/rstep),  sqrtf correctly rounded.
// The pair set is the reference's: one subgroup -> pairs col <= row including self pairs
// (cuDebye-v15.cu:137), two subgroups -> the full n_row x n_col rectangle.
//
// MI355X design.  A workgroup (16 waves) owns a window of bin counters in LDS (kP1Bins 32-bit
// counters, or 2 kP1Bins 16-bit ones).  A wave holds the atoms of a 64-atom A tile in its
// lanes and streams B atoms past them, read as broadcasts from a per-wave LDS stage: every
// 64-lane step is one pair per lane, ~21 VALU operations and one LDS atomic.  To keep a
// workgroup's bins inside its window, B atoms are taken shell by shell: a job handles the B
// atoms whose distance from the A tile's centre lies in [s*h, (s+1)*h); every pair of such an
// atom with an A atom then lies within [s*h - rad_A, (s+1)*h + rad_A], and the planner picks
// h so that this range fits the window, or a run of consecutive windows (one job each).  At
// the end of a job the nonzero LDS counters are added to the uint64 histogram in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "binrule.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace dbx {

// uint64 add to the histogram in HBM through a global (not flat) atomic: it waits on the
// vector-memory counter only, so the LDS traffic around it is not serialised behind it
typedef __attribute__((address_space(1))) unsigned long long gu64;
__device__ __forceinline__ void hbm_add(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add((gu64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A read-only table entry at a wave-uniform index (a job, a tile, an A-list or shell-edge
// entry), loaded through the constant address space so that it becomes a scalar load: whole
// 16-byte words for structs (no constructor takes a constant-space reference), bit-copied.
typedef int v4i __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ T ld_uniform(const T *p)
{
    T t;
    if constexpr (sizeof(T) % 16 == 0) {
        typedef const __attribute__((address_space(4))) v4i CV4;
        const CV4 *q = (const CV4 *)p;
#pragma unroll
        for (size_t i = 0; i < sizeof(T) / 16; i++) {
            const v4i w = q[i];
            __builtin_memcpy(reinterpret_cast<char *>(&t) + 16 * i, &w, 16);
        }
    } else {
        static_assert(sizeof(T) == 4, "a 4-byte scalar or a multiple of 16 bytes");
        typedef const __attribute__((address_space(4))) T CT;
        t = *(const CT *)p;
    }
    return t;
}

// Position i (per lane) of a run starting at a wave-uniform base, through the global address
// space: a global (not flat) load, which waits on the vector-memory counter only
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_pos(const float4 *base, uint32_t i)
{
    typedef const __attribute__((address_space(1))) v4f GV4;
    const v4f w = ((const GV4 *)base)[i];
    return make_float4(w.x, w.y, w.z, w.w);
}

// The LDS window: h[0] guard (bins below the window), h[1 + i] bin k_lo + i for
// 0 <= i < kP1Bins, h[kP1Bins + 1] sentinel (bins above it).  v = the pair's distance in bin
// units (pair_scaled).
//
// CountFast (the window holds every pair of the job): v is clamped to the guard and sentinel
// bins before the round-half-up conversion, so a bin outside the window lands in h[0] or
// h[kP1Bins + 1] (three instructions: med3, convert, address).
struct CountFast {
    uint32_t *h;
    int off;     // 1 - k_lo: h[k + off] holds bin k
    float vlo;   // k_lo - 1 and k_lo + kP1Bins, exact in FP32 (the planner keeps these jobs
    float vhi;   // below 2^24 bins)
    __device__ __forceinline__ void operator()(float v)
    {
        const int k = round_rpi(__builtin_amdgcn_fmed3f(v, vlo, vhi));
        atomicAdd(&h[k + off], 1u);
    }
    __device__ __forceinline__ void finish() {}
};

// CountWindow (one of several consecutive windows of a job): a pair outside the window belongs
// to a sibling job; its lane adds to its own slot of `trash` instead (no branch, no shared
// address); kmin / kmax record the bins seen, to check that none falls below the first window
// or above the last.
struct CountWindow {
    uint32_t *h;
    uint32_t *trash;  // this lane's discard counter
    int k_lo;
    int kmin, kmax;
    __device__ __forceinline__ void operator()(float v)
    {
        const int k = round_rpi(v);
        kmin = min(kmin, k);
        kmax = max(kmax, k);
        const unsigned kk = (unsigned)(k - k_lo);
        atomicAdd(kk < (unsigned)kP1Bins ? &h[kk + 1] : trash, 1u);
    }
    __device__ __forceinline__ void finish() {}
};

// Windows of 16-bit counters (jobs with kP1Win16): 2 kP1Bins bins, two per LDS word.  Bin
// k_lo + i is field (i + 2) & 1 (low, high) of word (i + 2) >> 1, so word 0's high field is the
// guard and word kP1Bins + 1's low field the sentinel.  The planner starts these windows at an
// even bin, so bin k is field k & 1 of word (k >> 1) + 1 - k_lo / 2, at LDS byte address
// (2 k + base) & ~3 with base = &h[1 - k_lo / 2] (a multiple of 4): two instructions and no
// subtraction of k_lo per pair; the field's shift is (k << 4) mod 32, which the shift and
// bit-field instructions take from k << 4 directly.  A field counts up to 0x7FFF: the lane
// whose increment sets its top bit (old field 0x7FFF, which exactly one lane sees per 32768
// counts) takes 0x8000 back off the field and adds 32768 to the bin in HBM.  The lane looks at
// the old value at its next pair, so that waiting for it does not stall that pair, or at
// finish(), which the pair loops call where a lane skips a pair and at the end of every
// candidate: a field that reached 0x8000 is drained within one iteration of its wave's pair
// loop, with no barrier and no wait on another wave in between.  A carry into the neighbouring
// field would need 32768 further increments of that bin by the other waves inside that one
// iteration.  Guard and sentinel overflows are recorded in *err instead.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_addr(uint32_t *p) { return (uint32_t)(uintptr_t)(lds_u32 *)p; }
__device__ __forceinline__ uint32_t lds_add(uint32_t addr, uint32_t v)
{
    return __hip_atomic_fetch_add((lds_u32 *)(uintptr_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct Count16 {
    int k_lo;       // even
    uint32_t base;  // LDS byte address of h[1 - k_lo / 2]
    unsigned long long *hrow;
    int64_t n_bins;
    unsigned long long *oor;
    int *err;
    uint32_t p_old = 0u;  // the previous pair's old word (0: nothing to look at)
    int p_k = 0;
    uint32_t p_sh = 0u;   // p_k << 4
    __device__ __forceinline__ void init(uint32_t *h, int klo)
    {
        k_lo = klo;
        base = lds_addr(h) + 4u * (uint32_t)(1 - (klo >> 1));
    }
    __device__ __forceinline__ void look()
    {
        if (__builtin_amdgcn_ubfe(p_old, p_sh, 16) == 0x7FFFu) {
            const uint32_t a = ((uint32_t)(2 * p_k) + base) & ~3u;
            lds_add(a, 0u - (0x8000u << (p_sh & 16u)));
            if ((unsigned)(p_k - k_lo) >= (unsigned)(2 * kP1Bins))
                atomicAdd(err, 1);
            else if ((int64_t)p_k < n_bins)
                hbm_add(&hrow[p_k], 32768ull);
            else
                hbm_add(oor, 32768ull);
        }
    }
    // adds one count to bin k, or to the lane's discard word `trash` (LDS byte address) when !in
    __device__ __forceinline__ void add(int k, bool in, uint32_t trash)
    {
        const uint32_t sh = (uint32_t)k << 4;
        const uint32_t a = ((uint32_t)(2 * k) + base) & ~3u;
        const uint32_t old = lds_add(in ? a : trash, 1u << (sh & 31u));
        look();
        p_old = in ? old : 0u;
        p_k = k;
        p_sh = sh;
    }
    __device__ __forceinline__ void finish()
    {
        look();
        p_old = 0u;
    }
};

// CountFast16: the window holds every pair of the job; out-of-window values are clamped onto
// the guard or sentinel field
struct CountFast16 : Count16 {
    float vlo;   // k_lo - 1 and k_lo + 2 kP1Bins, exact in FP32
    float vhi;
    __device__ __forceinline__ void operator()(float v)
    {
        const int k = round_rpi(__builtin_amdgcn_fmed3f(v, vlo, vhi));
        add(k, true, 0u);
    }
};

// CountWindow16: one of several consecutive 16-bit windows of a job, as CountWindow
struct CountWindow16 : Count16 {
    uint32_t trash;  // LDS byte address of the lane's discard word
    int kmin, kmax;
    __device__ __forceinline__ void operator()(float v)
    {
        const int k = round_rpi(v);
        kmin = min(kmin, k);
        kmax = max(kmax, k);
        add(k, (unsigned)(k - k_lo) < (unsigned)(2 * kP1Bins), trash);
    }
};

// CountChecked (degraded jobs, and the complements of fast jobs that found pairs outside their
// window): pairs in the LDS window (kP1Bins bins from k_lo) are counted there, unless
// `complement`; pairs of the job's own window beyond it (16-bit windows) and pairs outside the
// job's window that no sibling window holds (below the first window, above the last) go to the
// uint64 histogram in HBM.  A complement run counts exactly the pairs the fast run could not.
struct CountChecked {
    uint32_t *h;
    int k_lo;
    int win;
    int width;   // the job's window: kP1Bins, or 2 kP1Bins for 16-bit windows
    bool complement;
    int64_t n_bins;
    unsigned long long *hrow;
    uint32_t miss, oor;
    __device__ __forceinline__ void operator()(float v)
    {
        const int k = round_rpi(v);
        const unsigned kk = (unsigned)(k - k_lo);
        const bool mine_hbm = complement ? !(kk < (unsigned)width) && (k < k_lo ? (win & kP1WinFirst) : (win & kP1WinLast))
                                         : (kk < (unsigned)width ? kk >= (unsigned)kP1Bins
                                                                 : (k < k_lo ? (win & kP1WinFirst) : (win & kP1WinLast)) != 0);
        if (!complement && kk < (unsigned)kP1Bins) {
            atomicAdd(&h[kk + 1], 1u);
        } else if (mine_hbm) {
            ++miss;
            if (k >= 0 && (int64_t)k < n_bins)
                hbm_add(&hrow[k], 1ull);
            else
                ++oor;
        }
    }
    __device__ __forceinline__ void finish() {}
};

// One workgroup runs one job: it zeroes its LDS window, culls the job's (A tile, B tile)
// pairs into an LDS candidate list (B tiles with an atom possibly in the shell), then its
// 16 waves take candidates one at a time from that list, so work spreads evenly over the
// waves whatever the shape of the job.  For each candidate the wave loads the B tile into
// its lanes, keeps the atoms inside the shell (ballot) and streams them past its A atoms.
//
// CHECKED = false: every pair's bin is assumed to fall in the window (the planner sized the
// shells for that).  An out-of-window bin is clamped into a guard or sentinel counter; a
// job whose sentinel is nonzero flushes nothing and is queued for the checked variant.
// CHECKED = true: out-of-window pairs go straight to the uint64 histogram in HBM (jobs whose
// window is narrower than their geometry, and the queued ones).
// zeroes the histogram block and the out-of-range count of every model of a launch:
// blockIdx.y = model (refs + y0, or ref0 alone when refs is null), blocks along x stride over
// its [P][n_bins] words; the first launch also clears the four launch counters
__global__ __launch_bounds__(256) void p1_zero_kernel(const P1Ref *__restrict__ refs, P1Ref ref0, int y0,
                                                      unsigned long long *__restrict__ counters)
{
    const P1Ref R = refs ? refs[y0 + (int)blockIdx.y] : ref0;
    if (y0 == 0 && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) counters[threadIdx.x] = 0ull;
    const int64_t n = (int64_t)R.n_rows * R.n_bins;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        R.hist[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *R.oor = 0ull;
}

template <bool CHECKED>
__global__ __launch_bounds__(kP1Threads) void p1_histogram_kernel(
    const P1Ref *__restrict__ refs, P1Ref ref0, const int4 *__restrict__ list, int n_fast, int n_checked,
    unsigned long long *__restrict__ counters, int4 *__restrict__ retry)
{
    __shared__ uint32_t h[kP1Bins + 2];
    __shared__ uint32_t cand[kP1Cand];
    __shared__ float4 stage[kP1Threads / 64][64];
    __shared__ int s_ncand, s_next, s_kmin, s_kmax, s_err;
    __shared__ uint32_t trash[64];  // per-lane discard counters of multi-window jobs

    // list entry = (model, job, piece, pieces); the checked launch runs the planned checked
    // jobs, then the fast jobs the fast launch queued for a re-run
    int4 entry;
    if (!CHECKED) {
        entry = list[blockIdx.x];
    } else {
        if ((int)blockIdx.x < n_checked) {
            entry = list[n_fast + blockIdx.x];
        } else {
            const int r = blockIdx.x - n_checked;
            if (r >= (int)counters[2]) return;
            entry = retry[r];
        }
    }
    // a negative piece count marks the complement of a fast job (see CountChecked)
    const bool complement = entry.w < 0;
    if (complement) entry.w = -entry.w;
    // the launch's models (refs), or its one model as an argument (refs null)
    const P1Ref M = refs ? refs[entry.x] : ref0;
    P1Job job = ld_uniform(&M.jobs[entry.y]);
    if (entry.w > 1) {
        // piece z of w of the job's B-tile range
        const int b0 = job.b0, nb = job.b1 - job.b0;
        job.b0 = b0 + (int)((int64_t)nb * entry.z / entry.w);
        job.b1 = b0 + (int)((int64_t)nb * (entry.z + 1) / entry.w);
    }
    const float4 *__restrict__ pos = M.pos;
    const Tile *__restrict__ tiles = M.tiles;
    const int32_t *__restrict__ alist = M.alist;
    const float *__restrict__ edge2 = M.edge2;
    const float rp = M.rp;
    const int64_t n_bins = M.n_bins;
    for (int i = threadIdx.x; i < kP1Bins + 2; i += kP1Threads) h[i] = 0u;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ai = lane;  // the lane's atom of the A tile
    const float e_lo = ld_uniform(&edge2[job.e2_off + job.shell]);
    const float e_hi = ld_uniform(&edge2[job.e2_off + job.shell + 1]);
    const float r_lo = e_lo > 0.f ? sqrtf(e_lo) : 0.f;
    const float r_hi = sqrtf(e_hi);  // +inf for the last shell
    const int k_lo = job.k_lo;
    unsigned long long *hrow = M.hist + (int64_t)job.pair * n_bins;
    const int nblk = (job.b1 - job.b0 + 63) >> 6;
    const int units = job.a_cnt * nblk;
    const bool w16 = (job.win & kP1Win16) != 0;
    const int width = w16 ? 2 * kP1Bins : kP1Bins;
    if (threadIdx.x == 0) {
        s_kmin = INT_MAX;
        s_kmax = INT_MIN;
        s_err = 0;
    }

    int cur_ia = -1;
    // pair evaluations of this wave (staged B atoms times the A tile's atoms, whether their bin
    // falls in this job's window or not), summed into counters[3] at the job's end
    unsigned long long evals = 0;
    // squared distances from the A tile's centre of the B atoms that can pair in this job: the
    // shell, narrowed to the window's reach in multi-window jobs
    float s_lo2 = e_lo, s_hi2 = e_hi;
    Tile ta = {};
    bool a_ok = false;
    int idx_a = 0;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f);

    // the job's pairs, each handed to the counter `cnt`
    auto run = [&](auto &cnt) {
        for (int u0 = 0; u0 < units; u0 += kP1Cand / 64) {
            if (threadIdx.x == 0) {
                s_ncand = 0;
                s_next = 0;
            }
            __syncthreads();
            // cull: unit = (A tile, block of 64 B tiles); candidates appended to the LDS list
            const int u1 = min(units, u0 + kP1Cand / 64);
            for (int u = u0 + wave; u < u1; u += kP1Threads / 64) {
                const int ial = u / nblk;
                const Tile A = ld_uniform(&tiles[ld_uniform(&alist[job.a_off + ial])]);
                const int j = job.b0 + (u - ial * nblk) * 64 + lane;
                bool c = false;
                if (j < job.b1 && (!job.same || j >= A.parent)) {
                    const Tile tb = tiles[j];
                    const float dx = tb.cx - A.cx, dy = tb.cy - A.cy, dz = tb.cz - A.cz;
                    const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
                    const float mg = 1e-3f + 1e-5f * (d + tb.rad);
                    c = (d + tb.rad + mg >= r_lo) && (d - tb.rad - mg < r_hi);
                }
                const uint64_t cm = __ballot(c);
                if (cm) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&s_ncand, __popcll(cm));
                    base = __builtin_amdgcn_readfirstlane(base);
                    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(cm >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)cm, 0));
                    if (c) cand[base + below] = ((uint32_t)ial << 24) | (uint32_t)(j - job.b0);
                }
            }
            __syncthreads();
            const int nc = s_ncand;
            for (;;) {
                int ci = 0;
                if (lane == 0) ci = atomicAdd(&s_next, 1);
                ci = __builtin_amdgcn_readfirstlane(ci);
                if (ci >= nc) break;
                // one candidate per wave: its tiles' fields come through scalar loads
                const uint32_t e = __builtin_amdgcn_readfirstlane(cand[ci]);
                const int ia = ld_uniform(&alist[job.a_off + (int)(e >> 24)]);
                const int j2 = job.b0 + (int)(e & 0xFFFFFFu);
                if (ia != cur_ia) {
                    cur_ia = ia;
                    ta = ld_uniform(&tiles[ia]);
                    a_ok = ai < ta.count;
                    idx_a = ta.start + ai;
                    pa = ld_pos(pos + ta.start, (uint32_t)min(ai, ta.count - 1));
                    // a B atom at distance d from the A tile's centre pairs with its atoms at
                    // distances in [d - rad, d + rad]: one whose interval misses this job's
                    // window (below it when a window lies below, above it when one lies above)
                    // has no pair to count here.  Margin: 10 bins plus 1e-5 relative.
                    if ((job.win & kP1WinOnly) != kP1WinOnly) {
                        const float rs = 1.f / rp;
                        const float lo = (float)(k_lo - 10) * rs - ta.rad;
                        const float hi = (float)(k_lo + width + 10) * rs + ta.rad;
                        const float lo_m = lo - 1e-5f * fabsf(lo);
                        const float hi_m = hi + 1e-5f * hi;
                        const float w_lo2 = (job.win & kP1WinFirst) || lo_m <= 0.f ? -1.f : lo_m * lo_m;
                        const float w_hi2 = (job.win & kP1WinLast) ? INFINITY : hi_m * hi_m;
                        s_lo2 = fmaxf(e_lo, w_lo2);
                        s_hi2 = fminf(e_hi, w_hi2);
                    }
                }
                const Tile tb = ld_uniform(&tiles[j2]);
                const bool b_ok = lane < tb.count;
                const float4 pb = ld_pos(pos + tb.start, (uint32_t)min(lane, tb.count - 1));
                const float ex = pb.x - ta.cx, ey = pb.y - ta.cy, ez = pb.z - ta.cz;
                const float d2c = (ex * ex + ey * ey) + ez * ez;
                const bool in_shell = b_ok && d2c >= s_lo2 && d2c < s_hi2;
                uint64_t bm = __ballot(in_shell);
                const bool diag = job.same && (j2 == ta.parent);  // B tile holds the A tile's atoms
                if (bm) {
                    // stage the shell's B atoms (with their atom index) in the wave's LDS area; the
                    // lanes then read them back as broadcasts, R per step
                    const int nb = __popcll(bm);
                    evals += (unsigned long long)nb * (unsigned long long)ta.count;
                    if (in_shell) {
                        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0));
                        stage[wave][below] = make_float4(pb.x, pb.y, pb.z, __int_as_float(tb.start + lane));
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (!diag) {
                        if (a_ok) {
                            // eight, four, then two B atoms per iteration: independent chains
                            // in flight, the stage address advanced once per eight pairs
                            int t = 0;
                            for (; t + 7 < nb; t += 8) {
                                float v[8];
#pragma unroll
                                for (int u = 0; u < 8; u++) {
                                    const float4 b = stage[wave][t + u];
                                    v[u] = pair_scaled(pa.x, pa.y, pa.z, b.x, b.y, b.z, rp);
                                }
#pragma unroll
                                for (int u = 0; u < 8; u++) cnt(v[u]);
                            }
                            if (t + 3 < nb) {
                                float v[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    const float4 b = stage[wave][t + u];
                                    v[u] = pair_scaled(pa.x, pa.y, pa.z, b.x, b.y, b.z, rp);
                                }
#pragma unroll
                                for (int u = 0; u < 4; u++) cnt(v[u]);
                                t += 4;
                            }
                            if (t + 1 < nb) {
                                const float4 b0 = stage[wave][t];
                                const float4 b1 = stage[wave][t + 1];
                                const float v0 = pair_scaled(pa.x, pa.y, pa.z, b0.x, b0.y, b0.z, rp);
                                const float v1 = pair_scaled(pa.x, pa.y, pa.z, b1.x, b1.y, b1.z, rp);
                                cnt(v0);
                                cnt(v1);
                                t += 2;
                            }
                            if (t < nb) {
                                const float4 b = stage[wave][t];
                                cnt(pair_scaled(pa.x, pa.y, pa.z, b.x, b.y, b.z, rp));
                            }
                        }
                    } else {
                        // the B tile is the A tile: pairs with index b >= a only; a lane that
                        // skips a pair settles its previous one, as an add would
                        for (int t = 0; t < nb; t++) {
                            const float4 b = stage[wave][t];
                            if (a_ok && idx_a <= __float_as_int(b.w))
                                cnt(pair_scaled(pa.x, pa.y, pa.z, b.x, b.y, b.z, rp));
                            else
                                cnt.finish();
                        }
                    }
                    // the candidate's last pair is settled here, before the wave can wait on
                    // other waves (next candidate, chunk barrier)
                    cnt.finish();
                    // the stage is rewritten by the next candidate only after every lane's reads
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __syncthreads();
        }
        if (lane == 0 && evals) atomicAdd(&counters[3], evals);
    };
    if (CHECKED) {
        CountChecked cnt{h, k_lo, job.win, width, complement, n_bins, hrow, 0u, 0u};
        run(cnt);
        __syncthreads();  // zeroing visible to all threads even when the job had no candidates
        uint32_t miss = cnt.miss, oor = cnt.oor;
        for (int off = 32; off > 0; off >>= 1) {
            miss += __shfl_down(miss, off);
            oor += __shfl_down(oor, off);
        }
        if (lane == 0) {
            if (miss) atomicAdd(&counters[1], (unsigned long long)miss);
            if (oor) atomicAdd(M.oor, (unsigned long long)oor);
        }
        if (complement) return;  // nothing counted in the window
    } else {
        // the window keeps its counts either way; a job with pairs outside it (never, when
        // the planner's margins hold) queues its complement for the checked variant
        bool bad;
        auto kminmax = [&](int kmin, int kmax) {
            for (int off = 32; off > 0; off >>= 1) {
                kmin = min(kmin, __shfl_xor(kmin, off));
                kmax = max(kmax, __shfl_xor(kmax, off));
            }
            if (lane == 0) {
                atomicMin(&s_kmin, kmin);
                atomicMax(&s_kmax, kmax);
            }
            __syncthreads();
            // a bin below the first window or above the last one
            return ((job.win & kP1WinFirst) && s_kmin < k_lo) || ((job.win & kP1WinLast) && s_kmax >= k_lo + width);
        };
        if ((job.win & kP1WinOnly) == kP1WinOnly) {
            if (!w16) {
                CountFast cnt{h, 1 - k_lo, (float)(k_lo - 1), (float)(k_lo + kP1Bins)};
                run(cnt);
                __syncthreads();
                bad = (h[0] | h[kP1Bins + 1]) != 0u;
            } else {
                CountFast16 cnt;
                cnt.init(h, k_lo);
                cnt.hrow = hrow, cnt.n_bins = n_bins, cnt.oor = M.oor, cnt.err = &s_err;
                cnt.vlo = (float)(k_lo - 1), cnt.vhi = (float)(k_lo + 2 * kP1Bins);
                run(cnt);
                __syncthreads();
                bad = (h[0] >> 16) != 0u || (h[kP1Bins + 1] & 0xFFFFu) != 0u || s_err != 0;
            }
        } else if (!w16) {
            CountWindow cnt{h, &trash[lane], k_lo, INT_MAX, INT_MIN};
            run(cnt);
            bad = kminmax(cnt.kmin, cnt.kmax);
        } else {
            CountWindow16 cnt;
            cnt.init(h, k_lo);
            cnt.hrow = hrow, cnt.n_bins = n_bins, cnt.oor = M.oor, cnt.err = &s_err;
            cnt.trash = lds_addr(&trash[lane]), cnt.kmin = INT_MAX, cnt.kmax = INT_MIN;
            run(cnt);
            bad = kminmax(cnt.kmin, cnt.kmax) || s_err != 0;
        }
        if (bad && threadIdx.x == 0) {
            const unsigned long long r = atomicAdd(&counters[2], 1ull);
            retry[r] = make_int4(entry.x, entry.y, entry.z, -entry.w);
        }
    }
    if (!w16) {
        for (int i = threadIdx.x; i < kP1Bins; i += kP1Threads) {
            const uint32_t v = h[i + 1];
            if (v) {
                const int64_t k = (int64_t)k_lo + i;
                if (k < n_bins)
                    hbm_add(&hrow[k], (unsigned long long)v);
                else
                    hbm_add(M.oor, (unsigned long long)v);
            }
        }
    } else {
        for (int i = threadIdx.x; i < 2 * kP1Bins; i += kP1Threads) {
            const int j = i + 2;
            const uint32_t v = (h[j >> 1] >> ((j & 1) << 4)) & 0xFFFFu;
            if (v) {
                const int64_t k = (int64_t)k_lo + i;
                if (k < n_bins)
                    hbm_add(&hrow[k], (unsigned long long)v);
                else
                    hbm_add(M.oor, (unsigned long long)v);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// host: job planning
// ---------------------------------------------------------------------------------
namespace {

double point_box_min(const Tile &t, const float *bb)
{
    double d2 = 0;
    const double c[3] = {t.cx, t.cy, t.cz};
    for (int a = 0; a < 3; a++) {
        double lo = bb[a], hi = bb[3 + a];
        double e = c[a] < lo ? lo - c[a] : (c[a] > hi ? c[a] - hi : 0.0);
        d2 += e * e;
    }
    return std::sqrt(d2);
}

double point_box_max(const Tile &t, const float *bb)
{
    double d2 = 0;
    const double c[3] = {t.cx, t.cy, t.cz};
    for (int a = 0; a < 3; a++) {
        double e = std::max(std::fabs(c[a] - bb[a]), std::fabs(bb[3 + a] - c[a]));
        d2 += e * e;
    }
    return std::sqrt(d2);
}

}  // namespace

// The plan of one model and bin width.  Per subgroup pair (row ga, col gb):
//   - A tiles (the 64-atom blocks of ga): narrow ones, whose radius leaves a shell of at least
//     kP1SingleShell windows in one 32-bit window, and wide ones, counted in 16-bit windows
//     (one, or several consecutive ones per shell);
//   - each kind sorted by radius and cut into chunks; each chunk gets its own shell width
//     h = n_windows usable - 2 rad_max - margins and, per shell it can reach, one job per
//     window;
//   - with n_cu > 0, jobs far above a CU's mean share are split along their B-tile range
//     (n_cu = 0 keeps jobs whole: the launch lists cut them into pieces instead);
//   - order: exact-window jobs, largest first, then the checked ones.
void plan_pass1(const ModelDev &m, double rstep, std::vector<P1Job> &sorted,
                std::vector<float> &edge2, std::vector<int32_t> &alist, int n_cu,
                std::vector<double> *sorted_cost)
{
    DBX_REQUIRE(rstep > 0 && std::isfinite(rstep), DBX_ERR_ARG, "rstep must be positive");
    const int S = m.n_groups;
    const double usable = (kP1Bins - 6) * rstep;
    const double usable16 = (2 * kP1Bins - 6) * rstep;  // a window of 16-bit counters

    std::vector<P1Job> jobs;
    std::vector<double> cost;
    edge2.clear();
    alist.clear();
    int p = 0;
    for (int ga = 0; ga < S; ga++) {
        for (int gb = ga; gb < S; gb++, p++) {
            const int a0 = m.group_tile0[ga], a1 = m.group_tile1[ga];
            const int b0 = m.group_tile0[gb], b1 = m.group_tile1[gb];
            if (a0 == a1 || b0 == b1) continue;
            const float *bbA = &m.bbox[6 * ga];
            const float *bbB = &m.bbox[6 * gb];
            double radA = 0;
            for (int t = a0; t < a1; t++) radA = std::max(radA, (double)m.tiles_h[t].rad);
            double D2 = 0;
            for (int a = 0; a < 3; a++) {
                double e = std::max(std::fabs((double)bbB[3 + a] - bbA[a]),
                                    std::fabs((double)bbA[3 + a] - bbB[a]));
                D2 += e * e;
            }
            const double D = std::sqrt(D2) + 1e-3;  // bound on |b - c_A|
            const double mg = 2e-3 + 4e-6 * (D + radA);
            const double h_min = std::max(0.5, 0.02 * usable);
            // A tiles: "narrow" ones leave a shell of at least kP1SingleShell windows in one
            // 32-bit window; "wide" ones count in 16-bit windows
            const double h_single = std::max(h_min, kP1SingleShell * usable);
            const double rad_fit = 0.5 * (usable - h_single - 2 * mg);
            std::vector<int32_t> sel[2];
            for (int t = a0; t < a1; t++) sel[m.tiles_h[t].rad <= rad_fit ? 0 : 1].push_back(t);
            const double nB = (double)m.group_n[gb];
            const int nbt = b1 - b0;
            for (int wide = 0; wide < 2; wide++) {
                std::vector<int32_t> &A = sel[wide];
                if (A.empty()) continue;
                std::stable_sort(A.begin(), A.end(), [&](int x, int y) { return m.tiles_h[x].rad < m.tiles_h[y].rad; });
                const double rad_typ = m.tiles_h[A[A.size() / 2]].rad;
                // windows of a wide-tile chunk (16-bit counters: twice the span per window): one
                // when it leaves a shell of kP1SingleShell windows' width, else the fewest that
                // leave kP1MultiShell windows' width, at least h_min
                const double uw = wide ? usable16 : usable;
                const int wbins = wide ? 2 * kP1Bins : kP1Bins;
                auto n_windows = [&](double rad) {
                    if (!wide || uw - 2 * rad - 2 * mg >= std::max(h_min, kP1SingleShell * uw)) return 1;
                    const double hw = std::max(h_min, kP1MultiShell * uw);
                    return std::max(1, (int)std::ceil((2 * rad + 2 * mg + hw) / uw));
                };
                const double h_typ = std::max(h_min, n_windows(rad_typ) * uw - 2 * rad_typ - 2 * mg);
                const double shell_frac = std::min(1.0, h_typ / std::max(D, h_typ));
                const double per_tile = 64.0 * std::max(1.0, nB * shell_frac);
                int CA = (int)std::ceil(6.0e6 / per_tile);
                CA = std::max(1, std::min(CA, 64));
                const int nchunks = ((int)A.size() + CA - 1) / CA;
                const int stride = (int)std::max<int64_t>(1, ((int64_t)nchunks * nbt) / 2000000);
                for (int c0 = 0; c0 < (int)A.size(); c0 += CA) {
                    const int c1 = std::min((int)A.size(), c0 + CA);
                    const int a_off = (int)alist.size();
                    double dmin = 1e300, dmax = 0, rmaxc = 0, na = 0;
                    for (int i = c0; i < c1; i++) {
                        const Tile &T = m.tiles_h[A[i]];
                        alist.push_back(A[i]);
                        dmin = std::min(dmin, point_box_min(T, bbB));
                        dmax = std::max(dmax, point_box_max(T, bbB));
                        rmaxc = std::max(rmaxc, (double)T.rad);
                        na += T.count;
                    }
                    const int nw = n_windows(rmaxc);
                    double h = nw * uw - 2 * rmaxc - 2 * mg;
                    const bool degraded = h < h_min;
                    if (degraded) h = h_min;
                    const int ns = (int)(D / h) + 1;
                    const int e2_off = (int)edge2.size();
                    edge2.push_back(-1.0f);
                    for (int sh = 1; sh < ns; sh++) edge2.push_back((float)((sh * h) * (sh * h)));
                    edge2.push_back(INFINITY);
                    // cost estimate: B tiles by centre distance from a few A tiles (sampled)
                    std::vector<double> shell_cost(ns, 0.0);
                    const int astep = std::max(1, (c1 - c0) / 4);
                    for (int i = c0; i < c1; i += astep) {
                        const Tile &T0 = m.tiles_h[A[i]];
                        const int jf = (gb == ga) ? std::max(b0, T0.parent) : b0;
                        for (int j = jf; j < b1; j += stride) {
                            const Tile &T = m.tiles_h[j];
                            const double dx = T.cx - T0.cx, dy = T.cy - T0.cy, dz = T.cz - T0.cz;
                            int sh = (int)(std::sqrt(dx * dx + dy * dy + dz * dz) / h);
                            if (sh >= ns) sh = ns - 1;
                            shell_cost[sh] += (double)T.count * stride * astep;
                        }
                    }
                    for (int sh = 0; sh < ns; sh++) {
                        if (sh * h > dmax + mg + 1e-3 || (sh + 1) * h < dmin - mg - 1e-3) continue;
                        P1Job J;
                        std::memset(&J, 0, sizeof(J));
                        J.pair = p;
                        J.a_off = a_off;
                        J.a_cnt = c1 - c0;
                        J.b0 = b0;
                        J.b1 = b1;
                        J.same = (ga == gb);
                        J.shell = sh;
                        J.e2_off = e2_off;
                        double klo;
                        if (!degraded)
                            klo = std::floor((sh * h - rmaxc - mg) / rstep) - 2;
                        else
                            klo = std::floor(((sh + 0.5) * h) / rstep) - kP1Bins / 2;
                        if (klo < 0) klo = 0;
                        // 16-bit windows start at an even bin (the counters address bin k's field
                        // from k alone); the margins above leave room for one bin less
                        if (wide && !degraded) klo -= std::fmod(klo, 2.0);
                        // neighbouring shells share straddling tiles: spread the sampled cost
                        double cst = shell_cost[sh];
                        if (sh > 0) cst += 0.25 * shell_cost[sh - 1];
                        if (sh + 1 < ns) cst += 0.25 * shell_cost[sh + 1];
                        const double jc = na / (c1 - c0) * (cst + 1.0);
                        const int wn = degraded ? 1 : nw;
                        for (int w = 0; w < wn; w++) {
                            J.k_lo = (int32_t)std::min<double>(klo + (double)w * wbins, (double)INT32_MAX - 1 - wbins);
                            J.win = (w == 0 ? kP1WinFirst : 0) | (w == wn - 1 ? kP1WinLast : 0) |
                                    (wide && !degraded ? kP1Win16 : 0);
                            // the single-window fast variants' bounds must be exact in FP32
                            J.degraded = (degraded || (wn == 1 && J.k_lo + (double)wbins >= 16777216.0)) ? 1 : 0;
                            jobs.push_back(J);
                            cost.push_back(jc);
                        }
                    }
                }
            }
        }
    }

    // split jobs far above the mean share of a CU along their B-tile range (the candidates of
    // a shell are spread over the whole range: tiles are ordered by size, not by position)
    double total = 0;
    for (double c : cost) total += c;
    const double target = n_cu > 0 ? total / (double)(n_cu * 4) : 1e300;
    std::vector<P1Job> split_jobs;
    std::vector<double> split_cost;
    for (size_t i = 0; i < jobs.size(); i++) {
        const P1Job &J = jobs[i];
        const int nb = J.b1 - J.b0;
        int pieces = (int)std::min<double>(std::ceil(cost[i] / std::max(target, 1.0)), (double)nb);
        pieces = std::max(1, std::min(pieces, 64));
        for (int k = 0; k < pieces; k++) {
            P1Job K = J;
            K.b0 = J.b0 + (int)((int64_t)nb * k / pieces);
            K.b1 = J.b0 + (int)((int64_t)nb * (k + 1) / pieces);
            if (K.b1 <= K.b0) continue;
            split_jobs.push_back(K);
            split_cost.push_back(cost[i] / pieces);
        }
    }
    // launch order: exact-window jobs first, largest first; then the degraded ones
    std::vector<int> order(split_jobs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (split_jobs[x].degraded != split_jobs[y].degraded) return split_jobs[x].degraded < split_jobs[y].degraded;
        return split_cost[x] > split_cost[y];
    });
    sorted.resize(split_jobs.size());
    for (size_t i = 0; i < order.size(); i++) sorted[i] = split_jobs[order[i]];
    if (sorted_cost) {
        sorted_cost->resize(split_jobs.size());
        for (size_t i = 0; i < order.size(); i++) (*sorted_cost)[i] = split_cost[order[i]];
    }
}

// launch entries of model `mi`: every job cut into pieces of about `target` cost along its
// B-tile range (at most 64 pieces, at most one per B tile), split into exact-window and
// checked entries with their costs
typedef std::vector<std::pair<double, int4>> EntryList;
static void add_entries(const ModelDev &m, int mi, double target, EntryList &fast, EntryList &checked)
{
    for (int j = 0; j < m.plan_jobs; j++) {
        const P1Job &J = m.plan_h[j];
        const int nb = J.b1 - J.b0;
        int pieces = (int)std::min<double>(std::ceil(m.plan_cost[j] / std::max(target, 1.0)), (double)nb);
        pieces = std::max(1, std::min(pieces, 64));
        for (int k = 0; k < pieces; k++) {
            if ((int64_t)nb * (k + 1) / pieces <= (int64_t)nb * k / pieces) continue;
            (J.degraded ? checked : fast).push_back({m.plan_cost[j] / pieces, make_int4(mi, j, k, pieces)});
        }
    }
}

static double plan_total(const ModelDev &m)
{
    double t = 0;
    for (double c : m.plan_cost) t += c;
    return t;
}

// one part of a launch list, largest cost first: a counting sort on the cost's exponent and top
// five mantissa bits (buckets of 1/32 octave; a non-negative double orders like its bits), in
// insertion order within a bucket.  The order only balances the launch (the counts do not
// depend on it), and a 256-model batch's 20,000 pieces take a fraction of what a comparison
// sort of them took (0.6 ms on the host, the device idle meanwhile).
static void order_part(const EntryList &e, int4 *out)
{
    if (e.empty()) return;
    std::vector<uint32_t> key(e.size());
    uint32_t kmin = UINT32_MAX, kmax = 0;
    for (size_t i = 0; i < e.size(); i++) {
        uint64_t b;
        std::memcpy(&b, &e[i].first, sizeof(b));
        key[i] = (uint32_t)(b >> 47);
        kmin = std::min(kmin, key[i]);
        kmax = std::max(kmax, key[i]);
    }
    std::vector<size_t> at((size_t)(kmax - kmin) + 2, 0);
    for (uint32_t k : key) at[kmax - k + 1]++;
    for (size_t d = 1; d < at.size(); d++) at[d] += at[d - 1];
    for (size_t i = 0; i < e.size(); i++) out[at[kmax - key[i]]++] = e[i].second;
}

// exact-window entries then checked ones, each part largest first
static void order_entries(EntryList &fast, EntryList &checked, int4 *out)
{
    order_part(fast, out);
    order_part(checked, out + fast.size());
}

// The pass-1 plan of a model for one bin width, built on the host (no device needed): the
// jobs are kept whole and cut into pieces per launch list (alone: about 4 pieces' worth per CU).
PlanHost make_plan_host(const ModelDev &m, double rstep, int n_cu)
{
    PlanHost ph;
    ph.rstep = rstep;
    ph.cu = n_cu;
    plan_pass1(m, rstep, ph.jobs, ph.edge2, ph.alist, 0, &ph.cost);
    for (const P1Job &J : ph.jobs)
        DBX_REQUIRE(!(J.win & kP1Win16) || (J.k_lo & 1) == 0, DBX_ERR_STATE,
                    "a 16-bit window must start at an even bin");
    ph.n_fast = 0;
    while (ph.n_fast < (int)ph.jobs.size() && !ph.jobs[ph.n_fast].degraded) ph.n_fast++;
    // the launch list of the model alone
    ModelDev tmp;
    tmp.plan_h = ph.jobs;
    tmp.plan_cost = ph.cost;
    tmp.plan_jobs = (int)ph.jobs.size();
    EntryList fast, checked;
    add_entries(tmp, 0, plan_total(tmp) / (double)(std::max(1, n_cu) * 4), fast, checked);
    ph.list.resize(fast.size() + checked.size());
    order_entries(fast, checked, ph.list.data());
    ph.list_fast = (int)fast.size();
    return ph;
}

size_t plan_device_bytes(const PlanHost &ph)
{
    auto al = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; };
    return al(sizeof(P1Job) * ph.jobs.size()) + al(sizeof(float) * ph.edge2.size()) +
           al(sizeof(int4) * ph.list.size()) + al(sizeof(int32_t) * ph.alist.size());
}

// replaces the model's plan by ph (device arrays from the model's allocator, copies on the
// context's stream, synchronised)
void install_plan(dbx_ctx *ctx, ModelDev &m, PlanHost &&ph)
{
    for (void *q : {(void *)m.d_jobs, (void *)m.d_edge2, (void *)m.d_list, (void *)m.d_alist})
        if (q) model_free(m, q);
    m.d_jobs = nullptr;
    m.d_edge2 = nullptr;
    m.d_list = nullptr;
    m.d_alist = nullptr;
    m.plan_rstep = -1.0;
    m.d_jobs = (P1Job *)model_alloc(m, sizeof(P1Job) * std::max<size_t>(1, ph.jobs.size()));
    m.d_edge2 = (float *)model_alloc(m, sizeof(float) * std::max<size_t>(1, ph.edge2.size()));
    m.d_list = (int4 *)model_alloc(m, sizeof(int4) * std::max<size_t>(1, ph.list.size()));
    m.d_alist = (int32_t *)model_alloc(m, sizeof(int32_t) * std::max<size_t>(1, ph.alist.size()));
    if (!ph.jobs.empty()) {
        DBX_HIP(hipMemcpyAsync(m.d_jobs, ph.jobs.data(), sizeof(P1Job) * ph.jobs.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        DBX_HIP(hipMemcpyAsync(m.d_list, ph.list.data(), sizeof(int4) * ph.list.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    }
    if (!ph.edge2.empty())
        DBX_HIP(hipMemcpyAsync(m.d_edge2, ph.edge2.data(), sizeof(float) * ph.edge2.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!ph.alist.empty())
        DBX_HIP(hipMemcpyAsync(m.d_alist, ph.alist.data(), sizeof(int32_t) * ph.alist.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    DBX_HIP(hipStreamSynchronize(ctx->stream));
    m.plan_h = std::move(ph.jobs);
    m.plan_cost = std::move(ph.cost);
    m.plan_jobs = (int)m.plan_h.size();
    m.plan_fast = ph.n_fast;
    m.plan_list_n = (int)ph.list.size();
    m.plan_list_fast = ph.list_fast;
    m.plan_cu = ph.cu;
    m.plan_rstep = ph.rstep;
}

// builds (or reuses) the model's device copy of the pass-1 plan for rstep
static void ensure_plan(dbx_ctx *ctx, ModelDev &m, double rstep)
{
    if (m.plan_rstep == rstep && m.plan_cu == ctx->cu_count && m.d_jobs) return;
    install_plan(ctx, m, make_plan_host(m, rstep, ctx->cu_count));
}

static P1Ref make_ref(const ModelDev &m, int64_t n_bins, uint64_t *d_hist, double rstep)
{
    P1Ref r;
    r.pos = m.pos;
    r.tiles = m.tiles;
    r.jobs = m.d_jobs;
    r.alist = m.d_alist;
    r.edge2 = m.d_edge2;
    r.hist = (unsigned long long *)d_hist;
    r.n_bins = n_bins;
    r.rp = (float)(1 / rstep);  // cuDebye-v15.cu:739
    r.n_rows = m.n_groups * (m.n_groups + 1) / 2;
    r.oor = nullptr;
    return r;
}

P1Stats pass1_batch(dbx_ctx *ctx, const std::vector<P1Item> &items, double rstep, bool sync_stats)
{
    DBX_REQUIRE(rstep > 0 && std::isfinite(rstep), DBX_ERR_ARG, "rstep must be positive");
    DBX_REQUIRE(!items.empty(), DBX_ERR_ARG, "no model for pass 1");
    P1Stats st;
    std::vector<P1Ref> refs(items.size());
    int n_fast = 0, n_checked = 0;
    int64_t max_words = 0;
    for (size_t i = 0; i < items.size(); i++) {
        const P1Item &it = items[i];
        // bins are 32-bit integers on the device (a scaled distance converts with saturation)
        DBX_REQUIRE(it.n_bins > 0 && it.n_bins < (int64_t)INT32_MAX, DBX_ERR_ARG, "bad n_bins");
        ensure_plan(ctx, *it.m, rstep);
        refs[i] = make_ref(*it.m, it.n_bins, it.d_hist, rstep);
        refs[i].oor = it.d_oor ? (unsigned long long *)it.d_oor : (unsigned long long *)ctx->counters;
        max_words = std::max<int64_t>(max_words, (int64_t)refs[i].n_rows * it.n_bins);
    }
    // launch list: one model uses its own list; several get one merged list whose pieces are
    // sized for the whole batch, exact-window entries then checked ones, largest first
    const int4 *d_list = nullptr;
    EntryList fast, checked;
    if (items.size() == 1) {
        d_list = items[0].m->d_list;
        n_fast = items[0].m->plan_list_fast;
        n_checked = items[0].m->plan_list_n - n_fast;
    } else {
        double total = 0;
        for (const P1Item &it : items) total += plan_total(*it.m);
        const double target = total / (double)(std::max(1, ctx->cu_count) * 4);
        for (size_t i = 0; i < items.size(); i++) add_entries(*items[i].m, (int)i, target, fast, checked);
        n_fast = (int)fast.size();
        n_checked = (int)checked.size();
    }
    st.n_jobs = n_fast + n_checked;
    st.n_checked = n_checked;
    const size_t ref_bytes = (sizeof(P1Ref) * refs.size() + 255) & ~(size_t)255;
    const size_t list_bytes = items.size() == 1 ? 0 : (sizeof(int4) * st.n_jobs + 255) & ~(size_t)255;
    const size_t retry_bytes = sizeof(int4) * (size_t)(n_fast + 1);
    char *buf = (char *)scratch(ctx, kSlotBatch, ref_bytes + list_bytes + retry_bytes);
    P1Ref *d_refs = (P1Ref *)buf;
    int4 *d_retry = (int4 *)(buf + ref_bytes + list_bytes);
    if (items.size() != 1) {
        int4 *hl = (int4 *)stage(ctx, sizeof(int4) * std::max(1, st.n_jobs));
        order_entries(fast, checked, hl);
        d_list = (int4 *)(buf + ref_bytes);
        DBX_HIP(hipMemcpyAsync((void *)d_list, hl, sizeof(int4) * st.n_jobs, hipMemcpyHostToDevice,
                               ctx->stream));
    }
    // one model travels as a kernel argument, several as a table in device memory
    const P1Ref ref0 = refs[0];
    if (items.size() == 1) {
        d_refs = nullptr;
    } else {
        P1Ref *hr = (P1Ref *)stage(ctx, sizeof(P1Ref) * refs.size());
        std::memcpy(hr, refs.data(), sizeof(P1Ref) * refs.size());
        DBX_HIP(hipMemcpyAsync(d_refs, hr, sizeof(P1Ref) * refs.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    // every histogram and out-of-range count, and the launch counters, zeroed by one launch
    // (per 65535 models)
    unsigned long long *cnt = (unsigned long long *)ctx->counters;
    const unsigned zx = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (max_words + 4095) / 4096));
    for (int y0 = 0; y0 < (int)refs.size(); y0 += 65535) {
        const unsigned ny = (unsigned)std::min<int>(65535, (int)refs.size() - y0);
        hipLaunchKernelGGL(p1_zero_kernel, dim3(zx, ny), dim3(256), 0, ctx->stream, d_refs, ref0, y0, cnt);
    }
    ctx->p1_timed = false;
    if (st.n_jobs > 0) {
        DBX_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
        if (n_fast > 0)
            hipLaunchKernelGGL(p1_histogram_kernel<false>, dim3((unsigned)n_fast), dim3(kP1Threads), 0,
                               ctx->stream, d_refs, ref0, d_list, n_fast, n_checked, cnt, d_retry);
        // checked variant: the degraded jobs plus whatever the fast launch queued
        hipLaunchKernelGGL(p1_histogram_kernel<true>, dim3((unsigned)(n_checked + n_fast)),
                           dim3(kP1Threads), 0, ctx->stream, d_refs, ref0, d_list, n_fast, n_checked, cnt, d_retry);
        DBX_HIP(hipGetLastError());
        DBX_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
        ctx->p1_timed = true;
    }
    DBX_HIP(hipMemcpyAsync(ctx->counters_h, ctx->counters, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->last_p1_jobs = st.n_jobs;
    if (sync_stats) {
        DBX_HIP(hipStreamSynchronize(ctx->stream));
        pass1_read_stats(ctx, st);
    }
    return st;
}

P1Stats pass1_histograms(dbx_ctx *ctx, ModelDev &m, double rstep, int64_t n_bins,
                         uint64_t *d_hist, bool sync_stats)
{
    return pass1_batch(ctx, {P1Item{&m, n_bins, d_hist}}, rstep, sync_stats);
}

void pass1_read_stats(dbx_ctx *ctx, P1Stats &st)
{
    const uint64_t *c = ctx->counters_h;
    st.out_of_range = c[0];
    st.window_misses = c[1];
    st.retried = c[2];
    ctx->last_p1_evals = c[3];
    ctx->last_p1_misses = c[1];
    ctx->last_p1_retried = c[2];
    if (ctx->p1_timed) DBX_HIP(hipEventElapsedTime(&ctx->last_p1_kernel_ms, ctx->ev[0], ctx->ev[1]));
}

}  // namespace dbx

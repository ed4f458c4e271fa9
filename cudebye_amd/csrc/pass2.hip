// Pass 2: prefactors, histogram compaction and the Debye sum.
//
// Replaces reduceCheck's compaction (cuDebye-v15.cu:283-346), atomicProp/atomicScatter
// (cuDebye-v15.cu:447-479) and intensity/debyeRowSum (cuDebye-v15.cu:493-577).
//
//   S_p(q) = self_p + sum_k A_pk sinc(q r_k),   A_pk = 2 N_pk  (0 for the weight-1 bins)
//   I(q)   = sum_p ffoobb_p(q) S_p(q)
//
// Weight rule (cuDebye-v15.cu:508-522): in each pair's own list of nonzero bins, the
// first `first_n` entries (first_n = blockDim.x of the reference launch) with
// r_k <= 0.0005 A count N once with sinc 1 (self_p); every other nonzero bin counts
// 2 N sin(q r_k)/(q r_k).  r_k = double(k) * rstep (cuDebye-v15.cu:736).
//
// The bins are compacted to the union of the nonzero bins of all P partials (ascending,
// as thrust::copy_if does per pair), so all partials share one sinc evaluation per
// (bin, q).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "fastmath.h"
#include "mfdigits.h"
#include "internal.h"

namespace dbx {

// ---------------------------------------------------------------------------------
// prefactor, cuDebye-v15.cu:447-467 (FP64)
// ---------------------------------------------------------------------------------
__device__ double prefactor_one(const GroupPair &g, double Q)
{
    const double rp16pi2 = -0.006332573977646;  // -1/(16 pi^2), cuDebye-v15.cu:452
    const double nh = Q * Q * rp16pi2;
    double v = g.occ_a * g.occ_b;
    v = v * exp(nh * (g.beq_a + g.beq_b));
    double f1 = g.ca, f2 = g.cb;
    for (int t = 0; t < 5; t++) {
        f1 += g.aa[t] * exp(g.ba[t] * nh);
        f2 += g.ab[t] * exp(g.bb[t] * nh);
    }
    return v * f1 * f2;
}

// ff[r][q], grid.y over rows
__global__ void p2_prefactor_kernel(const GroupPair *__restrict__ g, const double *__restrict__ q, int nq,
                                    double *__restrict__ ff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    ff[(int64_t)blockIdx.y * nq + i] = prefactor_one(g[blockIdx.y], q[i]);
}

GroupPair make_group_pair(const dbx_group_params &ga, const dbx_group_params &gb)
{
    GroupPair g;
    g.occ_a = ga.occ; g.occ_b = gb.occ; g.beq_a = ga.beq; g.beq_b = gb.beq;
    g.ca = ga.c; g.cb = gb.c;
    for (int t = 0; t < 5; t++) {
        g.aa[t] = ga.a[t]; g.ba[t] = ga.b[t]; g.ab[t] = gb.a[t]; g.bb[t] = gb.b[t];
    }
    return g;
}

// up to kPairArgs pairs travel as a kernel argument (no upload)
constexpr int kPairArgs = 10;
struct PairArgs {
    GroupPair g[kPairArgs];
};
__global__ void p2_prefactor_args_kernel(PairArgs pa, const double *__restrict__ q, int nq, double *__restrict__ ff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    ff[(int64_t)blockIdx.y * nq + i] = prefactor_one(pa.g[blockIdx.y], q[i]);
}

void prefactor_rows(dbx_ctx *ctx, const std::vector<GroupPair> &rows, const double *d_q, int nq, double *d_ff)
{
    const int n = (int)rows.size();
    if (n == 0) return;
    if (n <= kPairArgs) {
        PairArgs pa;
        std::copy(rows.begin(), rows.end(), pa.g);
        hipLaunchKernelGGL(p2_prefactor_args_kernel, dim3((nq + 255) / 256, n), dim3(256), 0, ctx->stream, pa, d_q,
                           nq, d_ff);
        DBX_HIP(hipGetLastError());
        return;
    }
    GroupPair *hg = (GroupPair *)stage(ctx, sizeof(GroupPair) * n);
    std::copy(rows.begin(), rows.end(), hg);
    GroupPair *dg = (GroupPair *)scratch(ctx, kSlotPairs, sizeof(GroupPair) * n);
    DBX_HIP(hipMemcpyAsync(dg, hg, sizeof(GroupPair) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(p2_prefactor_kernel, dim3((nq + 255) / 256, n), dim3(256), 0, ctx->stream, dg, d_q, nq,
                       d_ff);
    DBX_HIP(hipGetLastError());
}

void prefactor_dev(dbx_ctx *ctx, const dbx_group_params &ga, const dbx_group_params &gb,
                   const double *d_q, int nq, double *d_ff)
{
    prefactor_rows(ctx, std::vector<GroupPair>{make_group_pair(ga, gb)}, d_q, nq, d_ff);
}

// ---------------------------------------------------------------------------------
// Rows.  Pass 2 works on a list of rows, each a partial histogram of its own length at
// hist_base + h_off; the rows of one model are its P subgroup pairs, the rows of a batch are
// the rows of all its models.  Rows are ordered by n_bins, longest first, so the rows that
// hold bin k are a prefix of the list and a block of 16 rows needs the bins below the
// n_bins of its first row only.
//
// Compaction: union of the nonzero bins of all rows, ascending (for one row, the list
// thrust::copy_if keeps, cuDebye-v15.cu:315-329).
// ---------------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__device__ __forceinline__ bool row_holds(const RowDesc *rows, int r, int64_t k)
{
    return k < rows[r].n_bins;
}

constexpr int kFlagRows = 32;  // rows per thread of the nonzero-flag pass

// flags[k] = 1 where some row holds a nonzero count at bin k > k_self_max; grid.y over chunks
// of kFlagRows rows (a bin written by several chunks gets the same value from each).  Also each
// row's digit class: the count digits the matrix-core sum needs for its largest count beyond the
// self bins, the smallest I with 2 N < 127 256^(I - 1) (6: more than five), from up to five
// ballots per wave and row, an LDS maximum per workgroup and row, and a global atomic only
// where the workgroup's class exceeds the one already recorded (the same few addresses are
// hit by every workgroup: unconditional atomics there serialise).
// Rows are longest first: a wave stops at the first row that ends before its first bin.
__global__ __launch_bounds__(kScanBlock) void p2_flags_kernel(
    const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows, int n_rows, int64_t nb_max,
    int k_self_max, unsigned char *__restrict__ flags, int *__restrict__ rowcls, int *__restrict__ block_counts)
{
    __shared__ int scls[kFlagRows];
    __shared__ int wcnt[kScanBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int r0 = blockIdx.y * kFlagRows;
    const int r1 = min(n_rows, r0 + kFlagRows);
    if (threadIdx.x < kFlagRows) scls[threadIdx.x] = 0;
    __syncthreads();
    const bool kin = k < nb_max && k > k_self_max;
    const int64_t kw = k - (threadIdx.x & 63);  // the wave's first bin
    bool nz = false;
    for (int r = r0; r < r1; r++) {
        const RowDesc rd = rows[r];
        if (kw >= rd.n_bins) break;  // rows are longest first: no later one reaches the wave's bins
        const unsigned long long v2 = (kin && k < rd.n_bins) ? 2ull * hist[rd.h_off + k] : 0ull;
        nz |= v2 != 0ull;
        const uint64_t c2 = __ballot(v2 >= (unsigned long long)mf_count_limit(1));
        if (!c2) continue;
        const int c = 2 + (__ballot(v2 >= (unsigned long long)mf_count_limit(2)) != 0) +
                      (__ballot(v2 >= (unsigned long long)mf_count_limit(3)) != 0) +
                      (__ballot(v2 >= (unsigned long long)mf_count_limit(4)) != 0) +
                      (__ballot(v2 >= (unsigned long long)mf_count_limit(5)) != 0);
        if ((threadIdx.x & 63) == 0) atomicMax(&scls[r - r0], c);
    }
    if (nz) flags[k] = 1;
    // one chunk of rows (block_counts given): the block's flagged bins, for the scan
    if (block_counts) {
        const uint64_t bnz = __ballot(nz);
        if ((threadIdx.x & 63) == 0) wcnt[threadIdx.x >> 6] = __popcll(bnz);
    }
    __syncthreads();
    // (a stale read only costs an atomic: the recorded class never decreases)
    const int t = threadIdx.x;
    if (t < r1 - r0) {
        const int c = scls[t];
        if (c > 1 && c > rowcls[r0 + t]) atomicMax(&rowcls[r0 + t], c);
    }
    if (block_counts && t == 0) {
        int c = 0;
        for (int w = 0; w < kScanBlock / 64; w++) c += wcnt[w];
        block_counts[blockIdx.x] = c;
    }
}

// profiling: total nonzero bins over all rows (the bins the reference's per-pair sum contracts)
__global__ void p2_row_nnz_kernel(const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows,
                                  int n_rows, int64_t nb_max, unsigned long long *__restrict__ total)
{
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const int r0 = blockIdx.y * kFlagRows;
    const int r1 = min(n_rows, r0 + kFlagRows);
    unsigned long long c = 0;
    if (k < nb_max)
        for (int r = r0; r < r1 && row_holds(rows, r, k); r++) c += hist[rows[r].h_off + k] != 0ull;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}

__global__ void p2_zero_kernel(unsigned int *__restrict__ p, int64_t n_words)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words) p[i] = 0u;
}

__global__ void p2_count_kernel(const unsigned char *__restrict__ flags, int64_t nb_max, int *__restrict__ block_counts)
{
    __shared__ int wsum[kScanBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const bool nz = k < nb_max && flags[k];
    const uint64_t b = __ballot(nz);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kScanBlock / 64; w++) s += wsum[w];
        block_counts[blockIdx.x] = s;
    }
}

// a row's digit class as the flags kernel recorded it (0: never raised), in [1, 6]
__device__ __forceinline__ int row_class(const int *__restrict__ rowcls, int r)
{
    return min(6, max(1, rowcls[r]));
}

// exclusive scan of the block counts, one workgroup; total[0] = the sum (compacted bins),
// total[2] = the largest row digit class (at least 1).  Each thread sums a run of blocks; the
// runs' totals are scanned by waves (shuffles) and across the 16 waves, with no serial loop.
// With krows given, the rows are also ordered for the matrix-core sum: by digit class, most
// digits first, each class in the caller's order (a stable counting sort):
//   krows[p] = rows[perm[p]], and bcls[b] = the class of row 16 b of that order, the largest
//   of its 16-row block (1 past the rows), for n_blocks blocks.
// A row's sum does not depend on its place in this order: it depends on its own counts and on
// the bin splits alone.
__global__ __launch_bounds__(1024) void p2_scan_kernel(int *__restrict__ block_counts, int n_blocks,
                                                       int *__restrict__ total, const int *__restrict__ rowcls,
                                                       int n_rows, const RowDesc *__restrict__ rows,
                                                       RowDesc *__restrict__ krows, int *__restrict__ perm,
                                                       int *__restrict__ bcls, int n_rblocks)
{
    __shared__ int wsum[16], wcls[16];
    __shared__ int ccount[8], cstart[8], cnext[8];
    __shared__ int wrank[16][8];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n_blocks + 1023) / 1024;
    const int b0 = min(n_blocks, t * per), b1 = min(n_blocks, b0 + per);
    int s = 0;
    for (int i = b0; i < b1; i++) s += block_counts[i];
    int cls = 1;
    for (int r = t; r < n_rows; r += 1024) cls = max(cls, rowcls[r]);
    // inclusive scan of s within the wave, maximum of cls
    int inc = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
        cls = max(cls, __shfl_xor(cls, off));
    }
    if (lane == 63) wsum[w] = inc;
    if (lane == 0) wcls[w] = cls;
    if (t < 8) ccount[t] = 0;
    __syncthreads();
    int base = 0, all = 0, cmax = 1;
    for (int i = 0; i < 16; i++) {
        if (i < w) base += wsum[i];
        all += wsum[i];
        cmax = max(cmax, wcls[i]);
    }
    int acc = base + inc - s;  // exclusive prefix of this thread's run
    for (int i = b0; i < b1; i++) {
        const int c = block_counts[i];
        block_counts[i] = acc;
        acc += c;
    }
    if (t == 0) {
        total[0] = all;
        total[2] = cmax;
    }
    if (!krows) return;
    // rows per class, the first place of each class (class 6 first, then 5, ..., 1)
    for (int r = t; r < n_rows; r += 1024) atomicAdd(&ccount[row_class(rowcls, r)], 1);
    __syncthreads();
    if (t == 0) {
        int p = 0;
        for (int c = 6; c >= 1; c--) {
            cstart[c] = p;
            cnext[c] = p;
            p += ccount[c];
        }
    }
    __syncthreads();
    // chunks of 1024 rows in order: a row's place is its class's next place plus the rows of its
    // class before it in the chunk (earlier waves, then earlier lanes of its wave)
    for (int r0 = 0; r0 < n_rows; r0 += 1024) {
        const int r = r0 + t;
        const int c = r < n_rows ? row_class(rowcls, r) : 0;
        int below = 0;
        for (int cc = 1; cc <= 6; cc++) {
            const uint64_t m = __ballot(c == cc);
            if (lane == 0) wrank[w][cc] = __popcll(m);
            if (c == cc)
                below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
        }
        __syncthreads();
        if (r < n_rows) {
            int p = cnext[c] + below;
            for (int i = 0; i < w; i++) p += wrank[i][c];
            krows[p] = rows[r];
            perm[p] = r;
        }
        __syncthreads();
        if (t >= 1 && t <= 6) {
            int n = 0;
            for (int i = 0; i < 16; i++) n += wrank[i][t];
            cnext[t] += n;
        }
        __syncthreads();
    }
    for (int b = t; b < n_rblocks; b += 1024) {
        const int p = 16 * b;
        int c = 1;
        if (p < n_rows)
            for (int cc = 6; cc >= 1; cc--)
                if (p >= cstart[cc] && p < cstart[cc] + ccount[cc]) c = cc;
        bcls[b] = c;
    }
}

// the self-pair terms of row r: the reference's weight rule for the entries at r <= 0.0005
// (bins up to k_self_max): the first first_n nonzero ones count once (self), the others twice
// with sinc (low)
__device__ __forceinline__ void self_row(const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows,
                                         int r, int k_self_max, double *__restrict__ self, double *__restrict__ low)
{
    const RowDesc rd = rows[r];
    double s = 0;
    int rank = 0;
    for (int k = 0; k <= k_self_max; k++) {
        const unsigned long long n = k < rd.n_bins ? hist[rd.h_off + k] : 0ull;
        double lw = 0;
        if (n) {
            if (rank < rd.first_n)
                s += (double)n;
            else
                lw = 2.0 * (double)n;
            rank++;
        }
        low[(int64_t)r * (k_self_max + 1) + k] = lw;
    }
    self[r] = s;
}

// blocks [0, nblk): the compacted bin list kz from the flags and the scanned block offsets;
// blocks from nblk: the self-pair terms of one row per thread
__global__ void p2_scatter_kernel(const unsigned char *__restrict__ flags, int64_t nb_max,
                                  const int *__restrict__ block_offsets, int *__restrict__ kz, int nblk,
                                  const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows,
                                  int n_rows, int k_self_max, double *__restrict__ self, double *__restrict__ low)
{
    if ((int)blockIdx.x >= nblk) {
        const int r = (blockIdx.x - nblk) * kScanBlock + threadIdx.x;
        if (r < n_rows) self_row(hist, rows, r, k_self_max, self, low);
        return;
    }
    __shared__ int woff[kScanBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const bool nz = k < nb_max && flags[k];
    const uint64_t b = __ballot(nz);
    const int lane = threadIdx.x & 63;
    if (lane == 0) woff[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kScanBlock / 64; w++) {
            int c = woff[w];
            woff[w] = s;
            s += c;
        }
    }
    __syncthreads();
    if (nz) {
        const uint64_t below = lane ? (b & ((~0ull) >> (64 - lane))) : 0ull;
        kz[block_offsets[blockIdx.x] + woff[threadIdx.x >> 6] + __popcll(below)] = (int)k;
    }
}

// kext[b] = number of compacted bins below the n_bins of row block b's longest row
__device__ __forceinline__ void kext_entry(int b, const int *__restrict__ kz, int Kc, const RowDesc *__restrict__ rows,
                                           int n_rows, int rows_per_block, int n_blocks, int *__restrict__ kext)
{
    if (b >= n_blocks) return;
    const int r0 = b * rows_per_block, r1 = min(n_rows, r0 + rows_per_block);
    if (r0 >= n_rows) {
        kext[b] = 0;
        return;
    }
    int64_t nb = 0;  // the block's longest row
    for (int r = r0; r < r1; r++) nb = max(nb, rows[r].n_bins);
    int lo = 0, hi = Kc;  // first index with kz >= nb
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (kz[mid] < nb)
            lo = mid + 1;
        else
            hi = mid;
    }
    kext[b] = lo;
}

// weights of the compacted bins, FP64: A[r][i] = 2 N_r,kz[i] at A + r ld + i; rsum[r0 + t] += the
// row's sum of them (one atomic per wave and row).  The grid covers ld >= Kc bins; Kc = dyn[0].
__global__ void p2_weights_kernel(const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows,
                                  int r0, int nr, const int *__restrict__ kz, const int *__restrict__ dyn,
                                  int64_t ld, double *__restrict__ A, unsigned long long *__restrict__ rsum)
{
    const int Kc = dyn[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= Kc) return;  // a whole block past the compacted bins
    const int64_t k = i < Kc ? kz[i] : -1;
    for (int t = 0; t < nr; t++) {
        const RowDesc rd = rows[r0 + t];
        const unsigned long long n = (k >= 0 && k < rd.n_bins) ? 2ull * hist[rd.h_off + k] : 0ull;
        if (i < Kc) A[(int64_t)t * ld + i] = (double)n;
        unsigned long long w = n;
        for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
        if ((threadIdx.x & 63) == 0 && w) atomicAdd(&rsum[r0 + t], w);
    }
}

// weight-1 bins (cuDebye-v15.cu:508-514) and the weight-2 bins at r <= 0.0005 that the
// index condition excludes; one thread per row.  self[r] = sum of weight-1 counts,
// low[r * (k_self_max+1) + k] = 2 N for the other low bins (added to S with sinc(q r_k)).
// ---------------------------------------------------------------------------------
// Debye sum, FP64: one thread per q, bins split over blockIdx.y and staged through LDS,
// one sinc per (bin, q) shared by all rows of the pass; split partials reduced in fixed order.
// ---------------------------------------------------------------------------------
constexpr int kDsThreads = 256;
constexpr int kDsPmax = 10;  // rows per pass
constexpr int kDsChunk = 256;

// Kc = dyn[0] compacted bins in nsplit splits of ceil(Kc / nsplit) (the last ones possibly
// empty); A's rows ld apart
__global__ __launch_bounds__(kDsThreads) void p2_debye_fp64_kernel(
    const int *__restrict__ kz, const int *__restrict__ dyn, int nsplit, int64_t ld, const double *__restrict__ A,
    int np, double rstep, const double *__restrict__ q, int nq, double *__restrict__ Spart)
{
    __shared__ double sr[kDsChunk];
    __shared__ double sa[kDsPmax][kDsChunk];
    const int Kc = dyn[0];
    const int k_per_split = (Kc + nsplit - 1) / nsplit;
    const int iq = blockIdx.x * kDsThreads + threadIdx.x;
    const double Q = iq < nq ? q[iq] : 1.0;
    const int kb = blockIdx.y * k_per_split;
    const int ke = min(Kc, kb + k_per_split);
    double acc[kDsPmax];
#pragma unroll
    for (int t = 0; t < kDsPmax; t++) acc[t] = 0.0;
    for (int c0 = kb; c0 < ke; c0 += kDsChunk) {
        const int n = min(kDsChunk, ke - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kDsThreads) {
            sr[i] = (double)kz[c0 + i] * rstep;
            for (int t = 0; t < np; t++) sa[t][i] = A[(int64_t)t * ld + c0 + i];
        }
        __syncthreads();
        for (int i = 0; i < n; i++) {
            const double x = Q * sr[i];
            const double sc = x != 0.0 ? sin(x) / x : 1.0;
#pragma unroll
            for (int t = 0; t < kDsPmax; t++)
                if (t < np) acc[t] += sa[t][i] * sc;
        }
    }
    if (iq < nq) {
        double *dst = Spart + (int64_t)blockIdx.y * np * nq;
#pragma unroll
        for (int t = 0; t < kDsPmax; t++)
            if (t < np) dst[(int64_t)t * nq + iq] = acc[t];
    }
}

constexpr int kMfK = 64;       // bins per step of the matrix-core sum: the K of one matrix instruction
constexpr int kMfFold = 256;   // steps between folds of its 32-bit classes

// The compacted-bin count and the largest row digit class, as the scan kernel left them in
// device memory (dyn[0], dyn[2]): kernels sized on the host from upper bounds read the actual
// values here, so no host read-back stands between the compaction and the sums.
__device__ __forceinline__ int dyn_bins(const int *dyn) { return dyn[0]; }
__device__ __forceinline__ int dyn_class(const int *dyn) { return max(1, dyn[2]); }
__host__ __device__ constexpr int pad_k(int kc) { return (kc + kMfK - 1) & ~(kMfK - 1); }

// Error per unit weight of a row's sum over its compacted bins, sum_k A_k T_k with A_k = 2 N_k:
//   matrix cores: the sinc digits round c (T - 3/8) to an integer, |dT| <= 1 / (2c); the FP64 sine
//     and its recurrence add at most 4e-15 (sin_cw alone, below 1e-15); forming c T takes 5 u:
//     c / Q and 1 / r one rounding each, and the scaling of a run's first sine and cosine by
//     c / Q, which the linear recurrence carries to every bin of the run with a gain of at most
//     |s| + |c| <= sqrt(2), below 2.9 u of T where Q r >= 0.5 (below u for a lone sine scaled
//     once); the FP64 partial sums hold at most 0.625 sum_k A_k (|T - 3/8| <= 0.625) and take
//     nsplit + classes x folds + 2 rounded additions each, and the 3/8 sum_k A_k term one
//     conversion;
//   FP64: a sinc value within 4 u (sine, quotient, product), then kper + nsplit + 2 rounded
//     additions of partial sums below sum_k A_k (the classical gamma_n bound).
// u = 2^-53.  Every term is an upper bound, whatever the signs of the errors.
__host__ __device__ inline double pass2_unit_error(int mode, int kper_bins, int nsplit, int classes)
{
    const double u = 0x1p-53;
    if (mode == DBX_PASS2_FP64) {
        const double n = (double)kper_bins + nsplit + 2;
        return 4 * u + n * u / (1.0 - n * u);
    }
    const int folds = (kper_bins / kMfK + kMfFold - 1) / kMfFold;
    return 0.5 / kMfC + 4e-15 + 5 * u + 0.625 * u * (double)(nsplit + classes * folds + 2) + 0.375 * u;
}

// the unit error of the sums of one pass 2 split nsplit ways, from its compacted-bin count and
// largest row class in dyn (the matrix-core sums' splits: ceil(Kc_pad / 64 / nsplit) steps; the
// FP64 sums': ceil(Kc / nsplit) bins)
__device__ inline double unit_error_dyn(int mode, const int *__restrict__ dyn, int nsplit)
{
    const int Kc = dyn_bins(dyn);
    if (mode == DBX_PASS2_FP64) return pass2_unit_error(mode, (Kc + nsplit - 1) / nsplit, nsplit, 1);
    const int kper = (pad_k(Kc) / kMfK + nsplit - 1) / nsplit * kMfK;
    return pass2_unit_error(mode, kper, nsplit, min(kMfImax, dyn_class(dyn)) + kMfJ - 1);
}

// Absolute error bound of each row's structure sum S[r][q], the same for every q:
//   B[r] = e rsum[r] + 5 u (self[r] + sum_k |low[r][k]|)
// rsum[r] = sum of the weights 2 N over the row's compacted bins (exact), e = the error per unit
// weight of the mode's sum over them (pass2_unit_error); the weight-1 bins are exact integers and
// each low bin's FP64 sinc term is within 5 u of its value.  rsum null: no compacted bin.
// r: the order rsum was computed in; perm[r] (perm null: r) the caller's row
__device__ __forceinline__ void row_bound(const unsigned long long *__restrict__ rsum,
                                          const double *__restrict__ self, const double *__restrict__ low, int nlow,
                                          int r, double e, const int *__restrict__ perm, double *__restrict__ B)
{
    const int o = perm ? perm[r] : r;
    double lo = self[o];
    for (int k = 0; k < nlow; k++) lo += fabs(low[(int64_t)o * nlow + k]);
    B[o] = (rsum ? e * (double)rsum[r] : 0.0) + 5.0 * 0x1p-53 * lo;
}

// Row r = r0 + i / nq in the order the sums were computed in, q index iq = i % nq (i over
// nr rows x nq), the caller's row o = perm[r] (perm null: o = r):
//   S[o][iq] = (shift rsum[r] + sum over splits of Spart[s][i], fixed order)
//            + (self[o] + sum_k low[o][k] sinc(q r_k))
// rsum (the q-independent part of the matrix-core sum) may be null; nsplit may be 0.  The
// second term is the reference's weight rule for the entries at r <= 0.0005 among the first
// first_n (weight 1, no sinc) and their neighbours below first_n (weight 2 with sinc).
// The blocks past those of the sums write the bounds B of rows r0..r0+nr (brsum: the row sums
// of the weights, with the unit error of this pass 2: unit_error_dyn of mode, dyn and nsplit),
// one row per thread; none when B is null.
constexpr int kFinishThreads = 256;
__global__ __launch_bounds__(kFinishThreads) void p2_finish_rows_kernel(
    const double *__restrict__ Spart, int nsplit, int r0, int nr, int nq, const unsigned long long *__restrict__ rsum,
    double shift, const double *__restrict__ self, const double *__restrict__ low, int nlow, double rstep,
    const double *__restrict__ q, const int *__restrict__ perm, double *__restrict__ S,
    const unsigned long long *__restrict__ brsum, int mode, const int *__restrict__ dyn, double *__restrict__ B)
{
    const int64_t n = (int64_t)nr * nq;
    const int64_t nfin = (n + kFinishThreads - 1) / kFinishThreads;
    if ((int64_t)blockIdx.x >= nfin) {
        const int r = r0 + (int)((int64_t)blockIdx.x - nfin) * kFinishThreads + (int)threadIdx.x;
        if (r < r0 + nr) row_bound(brsum, self, low, nlow, r, unit_error_dyn(mode, dyn, nsplit), perm, B);
        return;
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = r0 + (int)(i / nq);
    const int iq = (int)(i - (int64_t)(r - r0) * nq);
    const int o = perm ? perm[r] : r;
    double acc = rsum ? shift * (double)rsum[r] : 0.0;
    // eight loads in flight, the sum kept in split order
    int s = 0;
    for (; s + 8 <= nsplit; s += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = Spart[(int64_t)(s + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    for (; s < nsplit; s++) acc += Spart[(int64_t)s * n + i];
    double lo = self[o];
    for (int k = 0; k < nlow; k++) {
        const double w = low[(int64_t)o * nlow + k];
        if (w != 0.0) {
            const double rk = (double)k * rstep;
            const double x = q[iq] * rk;
            lo += rk == 0.0 ? w : w * sin(x) / x;
        }
    }
    S[(int64_t)o * nq + iq] = acc + lo;
}

// ---------------------------------------------------------------------------------
// Debye sum on the matrix cores: S[rows x q] = A[rows x bins] . T[bins x q], A = 2 N (integer),
// T = sinc(q r), as signed 8-bit digits (the encoding in mfdigits.h) with exact 32-bit integer
// accumulation (v_mfma_i32_16x16x64_i8).  A digit product is at most 2^14 in magnitude and a
// scale class s = i + j collects at most min(I, 5) of them per bin, so a 32-bit accumulator
// over 256 steps of 64 bins stays below 2^31 and is exact whatever order the matrix core adds
// in.  Every 256 steps the classes are folded into FP64, S += sum_s 2^(8s) acc_s / c, and the
// 3/8 part, (3/8) sum_k A_k, is added per row from the exact integer sum of its counts.  The
// only approximation is the rounding of c (T - 3/8) to an integer (|err| <= 1/(2c) < 2^-48 per
// bin), with T from an FP64 sine with full argument reduction.
// ---------------------------------------------------------------------------------
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));


// balanced count digits as bytes: cdig[i][row][kk] = digit i of 2 N_row,kz[kk] at
// cdig + (i rows_pad + row) ld + kk (0 past Kc, past the row's bins or past n_rows), and
// rsum[row] += the row's sum of 2 N over the compacted bins.  Block (bx, row) of
// (ld / 1024 rounded up) x rows_pad, 256 threads, four consecutive bins per thread (Kc_pad and
// ld are multiples of 64, so a thread's four bins are all inside or all outside Kc_pad); one
// atomic per block.
constexpr int kCdThreads = 256;
__device__ __forceinline__ void count_digits_block(
    int bx, int row, unsigned long long *wsum, const unsigned long long *__restrict__ hist,
    const RowDesc *__restrict__ rows, int n_rows, const int *__restrict__ kz, int Kc, int Kc_pad, int64_t ld,
    int rows_pad, int I, signed char *__restrict__ cdig, unsigned long long *__restrict__ rsum)
{
    const int kk0 = (bx * kCdThreads + threadIdx.x) * 4;
    long long v[4] = {0, 0, 0, 0};
    if (row < n_rows && kk0 < Kc) {
        const RowDesc rd = rows[row];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kk = kk0 + u;
            if (kk < Kc) {
                const int64_t k = kz[kk];
                if (k < rd.n_bins) v[u] = (long long)(2ull * hist[rd.h_off + k]);
            }
        }
    }
    unsigned long long sum = (unsigned long long)(v[0] + v[1] + v[2] + v[3]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0 && row < n_rows) {
        unsigned long long t = 0;
        for (int w = 0; w < kCdThreads / 64; w++) t += wsum[w];
        if (t != 0ull) atomicAdd(&rsum[row], t);
    }
    if (kk0 >= Kc_pad) return;
    for (int i = 0; i < I; i++) {
        uint32_t packed = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            long long e = v[u];
            if (i + 1 < I) {
                e = ((v[u] + 128) & 255) - 128;
                v[u] = (v[u] - e) >> 8;
            }
            packed |= (uint32_t)(uint8_t)(signed char)e << (8 * u);
        }
        *reinterpret_cast<uint32_t *>(cdig + ((int64_t)i * rows_pad + row) * ld + kk0) = packed;
    }
}

// bin distances and their reciprocals: rr[kk] = kz[kk] * rstep, ir[kk] = 1 / rr[kk].  Entries
// past Kc (zero counts) continue the bins after the last one, so that the table stays strictly
// increasing in whole bins and 64 entries spanning 63 bins are always 64 consecutive bins.
__device__ __forceinline__ void rtab_entry(int i, const int *__restrict__ kz, int Kc, int Kc_pad, double rstep,
                                           double *__restrict__ rr, double *__restrict__ ir)
{
    if (i >= Kc_pad) return;
    const int k = i < Kc ? kz[i] : (Kc > 0 ? kz[Kc - 1] : 0) + (i - Kc + 1);
    const double r = (double)k * rstep;
    rr[i] = r;
    ir[i] = 1.0 / r;
}

// The inputs of the contraction: blocks [0, cd_x rows_pad) write the count digits (block b:
// bins of chunk b % cd_x of row b / cd_x; as many digits as the largest row class needs, at
// most kMfImax), the next n_rt the r table, the rest the row-block bin extents.  cd_x and n_rt
// cover ld bins, an upper bound of the compacted-bin count; blocks past the actual count write
// nothing.
__global__ __launch_bounds__(kCdThreads) void p2_mf_inputs_kernel(
    const unsigned long long *__restrict__ hist, const RowDesc *__restrict__ rows, int n_rows,
    const int *__restrict__ kz, const int *__restrict__ dyn, int64_t ld, int rows_pad, int cd_x,
    signed char *__restrict__ cdig, unsigned long long *__restrict__ rsum, double rstep, double *__restrict__ rr,
    double *__restrict__ ir, int n_rt, int n_kblocks, int *__restrict__ kext)
{
    __shared__ unsigned long long wsum[kCdThreads / 64];
    const int Kc = dyn_bins(dyn), Kc_pad = pad_k(Kc);
    const int I = min(kMfImax, dyn_class(dyn));
    const int n_cd = cd_x * rows_pad;
    const int b = blockIdx.x;
    if (b < n_cd) {
        count_digits_block(b % cd_x, b / cd_x, wsum, hist, rows, n_rows, kz, Kc, Kc_pad, ld, rows_pad, I, cdig, rsum);
    } else if (b < n_cd + n_rt) {
        rtab_entry((b - n_cd) * kCdThreads + threadIdx.x, kz, Kc, Kc_pad, rstep, rr, ir);
    } else {
        kext_entry((b - n_cd - n_rt) * kCdThreads + threadIdx.x, kz, Kc, rows, n_rows, 16, n_kblocks, kext);
    }
}

// One workgroup: WQ x WR waves.  Wave (wq, wr) owns 16 q values (block wq of the workgroup's
// 16 WQ) and RB blocks of 16 rows (block wr of its WR RB 16).  Operand layout of
// v_mfma_i32_16x16x64_i8 (lane l): A[row l%16][k 16(l/16)..+15], B[k 16(l/16)..+15][col l%16],
// C[row 4(l/16)+v][col l%16].  A step is 64 compacted bins.
//   - Count digits: the I x WR x RB fragments (1 KB each) of a step are loaded by the
//     workgroup's waves together, ahead of use, into a ring of three LDS slots (lane l of a
//     fragment holds exactly the 16 bytes matrix lane l reads); one barrier per step.
//   - Sinc digits: the 16 q x 64 bin values of a step are generated by the WR waves of a q
//     block together (16 / WR values per thread), through LDS (double-buffered) when WR > 1,
//     one step ahead of the matrix cores so that both run at once.  The step's r and 1/r come
//     through the same LDS ring.  A thread whose values lie at consecutive bins takes one
//     sine-cosine pair and steps along the following bins with a second-order recurrence.
//   - The 32-bit sums are folded into FP64 partial sums in Spart every kMfFold steps.
// The workgroup's bins end at kext of its first row block: rows are ordered by length,
// longest first.
// The tile (q group, split, row group) of a workgroup, q group fastest: the dispatcher deals
// consecutive workgroups round-robin to the 8 XCDs, so every XCD gets a share of every row group
// (the longest rows first) and the count digits of a (split, row group) are read through each
// XCD's L2 by its share of the q groups.  Binding whole (split, row group)s to one XCD would read
// them once, but rows of unequal length leave the XCDs unequally loaded, and the digit traffic
// is far below the HBM bandwidth in either case.
__device__ __forceinline__ int3 mf_tile() { return make_int3(blockIdx.x, blockIdx.y, blockIdx.z); }

// The contraction of one tile with I count digits (the tile's rows need at most I); the LDS ring
// holds NFA >= I WR RB fragments per slot.  kend: the tile's last bin.
template <int RB, int WQ, int WR, int NFA, int I>
__device__ __forceinline__ void mf_tile_sum(const int3 tile, int kend, i32x4 (*sa)[NFA][64], double (*srt)[2 * kMfK],
                                            i32x4 (*sd)[WR > 1 ? WQ : 1][kMfJ][64], const double *__restrict__ rr,
                                            const double *__restrict__ ir, int Kc_pad, int64_t ld, int k_per_split,
                                            const signed char *__restrict__ cdig, int rows_pad,
                                            const double *__restrict__ q, int nq, int nrows, int srows,
                                            double rstep, double *__restrict__ Spart)
{
    constexpr int NW = WQ * WR;
    constexpr int NF = I * WR * RB;
    static_assert(NF <= NFA, "LDS ring smaller than the tile's fragments");
    constexpr int LF = (NF + NW - 1) / NW;
    constexpr int C = I + kMfJ - 1;
    constexpr int V = 16 / WR;
    constexpr int NG = V / 4;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wq = wave % WQ, wr = wave / WQ;
    const int q0 = tile.x * (16 * WQ) + wq * 16;
    const int rowt = tile.z * (WR * RB * 16);
    const int col = lane & 15, kq = lane >> 4;
    // generator role: thread g of the q block's WR waves -> fragment lane gl, values gt..gt+V-1
    const int g = wr * 64 + lane;
    const int gl = g / WR, gt = (g % WR) * V;
    const double Q0 = q[min(q0 + (gl & 15), nq - 1)];
    const double Q = Q0 != 0.0 ? Q0 : 1e-30;
    const double cQ = kMfC / Q;
    const int gk = (gl >> 4) * 16 + gt;
    // one bin further: sin(Q rstep), and lam = 2 - 2 cos(Q rstep) = 4 sin^2(Q rstep / 2), formed
    // from the half-angle sine so that it keeps full relative precision for small Q rstep
    const double sdQ = sin_cw(Q * rstep);
    const double shQ = sin_cw(0.5 * (Q * rstep));
    const double lamQ = 4.0 * (shQ * shQ);
    const int kx = (kend + kMfK - 1) & ~(kMfK - 1);
    const int kb = tile.y * k_per_split;
    const int ke = min(min(Kc_pad, kx), kb + k_per_split);
    const int nsteps = ke > kb ? (ke - kb) / kMfK : 0;
    const int qq = q0 + col;
    // The sines hold for Q r < kSinArgMax.  A tile whose largest Q r (its last bin, at most 63
    // padding bins past the last compacted one) reaches that writes NaN sums instead: a call
    // with an error target recomputes those models in FP64 (whose sines take any argument),
    // one without returns NaN.
    if (__syncthreads_or(nsteps > 0 && Q * rr[ke - 1] >= kSinArgMax)) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rowt + (wr * RB + rb) * 16 + kq * 4 + v;
                if (qq < nq && row < nrows)
                    Spart[((int64_t)tile.y * srows + row) * nq + qq] = __builtin_nan("");
            }
        return;
    }
    i32x4 acc[RB][C];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int c = 0; c < C; c++) acc[rb][c] = i32x4{0, 0, 0, 0};
    // The 32-bit sums (digit c weighs 2^(8c)) are folded into the split's FP64 sums, which live
    // in Spart rather than in registers: each (row, q) has one owner lane, the first fold
    // stores, later ones add, and the last one scales by 1 / c.
    bool first = true;
    auto fold = [&](bool last) {
        constexpr double inv_c = 1.0 / kMfC;
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rowt + (wr * RB + rb) * 16 + kq * 4 + v;
                const bool own = qq < nq && row < nrows;
                double *p = Spart + ((int64_t)tile.y * srows + row) * nq + qq;
                double a = (own && !first) ? *p : 0.0;
#pragma unroll
                for (int c = 0; c < C; c++) a += __builtin_ldexp((double)acc[rb][c][v], 8 * c);
                if (own) *p = last ? a * inv_c : a;
            }
#pragma unroll
            for (int c = 0; c < C; c++) acc[rb][c] = i32x4{0, 0, 0, 0};
        }
        first = false;
    };
    // count-digit fragment f of the step at k0: digit f / (WR RB), row block f % (WR RB).  Wave w
    // moves fragments w, w + NW, ... (clamped to the last one: a duplicate writes the same
    // bytes), and every wave moves the step's r table (same bytes again); no branches, so the
    // scheduler can interleave all of it with the matrix instructions.
    i32x4 pf[LF];
    double prr, pir;
    auto load_a = [&](int k0) {
#pragma unroll
        for (int m = 0; m < LF; m++) {
            const int f = min(wave + m * NW, NF - 1);
            const int i = f / (WR * RB), blk = f - i * (WR * RB);
            pf[m] = *reinterpret_cast<const i32x4 *>(
                cdig + ((int64_t)i * rows_pad + rowt + blk * 16 + col) * ld + k0 + kq * 16);
        }
    };
    auto store_a = [&](int slot) {
#pragma unroll
        for (int m = 0; m < LF; m++) sa[slot][min(wave + m * NW, NF - 1)][lane] = pf[m];
    };
    auto load_r = [&](int k0) {
        prr = rr[k0 + lane];
        pir = ir[k0 + lane];
    };
    auto store_r = [&](int slot) {
        srt[slot][lane] = prr;
        srt[slot][kMfK + lane] = pir;
    };
    // Sinc digits of the step whose r table is in srt[slot]: B fragments (WR == 1) or sd[buf].
    // A lane's V values are at bins gk..gk+V-1.  Every lane first takes one sine-cosine pair
    // (s0, c0) at its first bin and steps along the next ones with the second-order recurrence
    // in Reinsch's form: d1 = c0 sin(h) - s0 lam / 2, s1 = s0 + d1, then d(k+1) = d(k) -
    // lam s(k), s(k+1) = s(k) + d(k+1), with h = Q rstep and lam = 4 sin^2(h / 2) (two operations
    // per bin).  Each step rounds s once (<= 1.1e-16) and the difference d is small (|d| <= h),
    // so fifteen steps add below 2e-15 to the sine of the sine-cosine pair, and so below 4e-15
    // to sinc where Q r >= 0.5, under 1% of the 1/c grid of the sinc digits.  That holds where
    // the lane's bins are consecutive (the table is strictly increasing in whole bins, so a
    // span of V-1 bins means that) and Q r >= 0.5 at the first; other lanes redo their values
    // with one sine each (fix_gen), a branch that no wave takes where the histograms are
    // dense.
    i32x4 Bn[kMfJ];
    auto rot_ok = [&](int slot) -> bool {
        const double span = srt[slot][gk + V - 1] - srt[slot][gk];
        return span > (V - 1.5) * rstep && span < (V - 0.5) * rstep && Q * srt[slot][gk] >= 0.5;
    };
    auto gen = [&](auto rot_tag, int slot, int buf) {
        constexpr bool ROT = decltype(rot_tag)::value;
        // ROT: the recurrence runs on sines and cosines scaled by cQ (it is linear), which
        // saves the scaling of every value
        double sv = 0.0, cv = 0.0, dv = 0.0;
        if constexpr (ROT) {
            sincos_cw(Q * srt[slot][gk], sv, cv);
            sv *= cQ;
            cv *= cQ;
        }
        int dw[NG][kMfJ];
#pragma unroll
        for (int g4 = 0; g4 < NG; g4++) {
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = g4 * 4 + u;
                double s;
                if constexpr (ROT) {
                    if (t == 1) {
                        dv = fma(cv, sdQ, -(sv * (0.5 * lamQ)));
                        sv = sv + dv;
                    } else if (t > 1) {
                        dv = fma(-lamQ, sv, dv);
                        sv = sv + dv;
                    }
                    s = sv;
                } else {
                    s = sin_cw(Q * srt[slot][gk + t]) * cQ;
                }
                sinc_word_scaled(s, srt[slot][kMfK + gk + t], lo[u], hi[u]);
            }
            pack4(lo, hi, dw[g4]);
        }
        if constexpr (WR == 1) {
#pragma unroll
            for (int j = 0; j < kMfJ; j++) Bn[j] = i32x4{dw[0][j], dw[1][j], dw[2][j], dw[3][j]};
        } else {
            static_assert(NG == 2, "two waves per q block: eight bins per thread");
#pragma unroll
            for (int j = 0; j < kMfJ; j++)
                *(reinterpret_cast<i32x2 *>(&sd[buf][wq][j][gl]) + (gt >> 3)) = i32x2{dw[0][j], dw[1][j]};
        }
    };
    // Pipeline: at step s the matrix cores take A(s) (count digits, LDS ring slot s%3) and the
    // sinc digits of step s, while the vector units generate the sinc digits of step s+1 from
    // the r table of step s+1 (ring slot (s+1)%3).  A(s+1) and the r table of step s+2 are
    // written to the ring during step s; A(s+2) and the r table of step s+3 are fetched into
    // registers.  A slot is rewritten two barriers after its last read.
    if (nsteps > 0) {
        load_a(kb);
        load_r(kb);
        store_a(0);
        store_r(0);
        if (nsteps > 1) {
            load_r(kb + kMfK);
            store_r(1);
        }
    }
    __syncthreads();
    auto gen_step = [&](int slot, int buf) { gen(std::true_type{}, slot, buf); };
    auto fix_gen = [&](int slot, int buf) {
        if (!rot_ok(slot)) gen(std::false_type{}, slot, buf);
    };
    if (nsteps > 0) {
        gen_step(0, 0);
        fix_gen(0, 0);
        if (nsteps > 1) load_a(kb + kMfK);
        if (nsteps > 2) load_r(kb + 2 * kMfK);
    }
    __syncthreads();
    // one step: the matrix instructions of step s interleaved with the generation of step s+1;
    // branch-free inside (past the last step the generator and the ring take stale or clamped,
    // in-bounds data that no later step reads)
    auto step = [&](int s, int slot, int s1, int s2) {
        i32x4 B[kMfJ];
        if constexpr (WR == 1) {
#pragma unroll
            for (int j = 0; j < kMfJ; j++) B[j] = Bn[j];
        } else {
#pragma unroll
            for (int j = 0; j < kMfJ; j++) B[j] = sd[s & 1][wq][j][lane];
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
#pragma unroll
            for (int i = 0; i < I; i++) {
                const i32x4 A = sa[slot][i * (WR * RB) + wr * RB + rb][lane];
#pragma unroll
                for (int j = 0; j < kMfJ; j++)
                    acc[rb][i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B[j], acc[rb][i + j], 0, 0, 0);
            }
        }
        gen_step(s1, (s + 1) & 1);
        store_a(s1);
        store_r(s2);
        load_a(min(kb + (s + 2) * kMfK, ke - kMfK));
        load_r(min(kb + (s + 3) * kMfK, ke - kMfK));
        // interleave: one matrix instruction, then three vector instructions of the generator
#pragma unroll
        for (int x = 0; x < RB * I * kMfJ; x++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
        }
        // outside the interleaved block: lanes whose values cannot use the recurrence
        fix_gen(s1, (s + 1) & 1);
    };
    int slot = 0, nfold = 0;
    for (int s = 0; s < nsteps; s++) {
        const int s1 = slot == 2 ? 0 : slot + 1;
        const int s2 = s1 == 2 ? 0 : s1 + 1;
        step(s, slot, s1, s2);
        if (++nfold == kMfFold) {
            fold(false);
            nfold = 0;
        }
        __syncthreads();
        slot = s1;
    }
    fold(true);
}

// the tile sum with the smallest instantiated digit count I' >= Irg (Irg in [1, I])
template <int RB, int WQ, int WR, int NFA, int I, typename... A>
__device__ __forceinline__ void mf_tile_dispatch(int Irg, A... a)
{
    if constexpr (I > 1) {
        if (Irg < I) {
            mf_tile_dispatch<RB, WQ, WR, NFA, I - 1>(Irg, a...);
            return;
        }
    }
    mf_tile_sum<RB, WQ, WR, NFA, I>(a...);
}

// One workgroup per tile (q group, split, row group).  The row group's bins end at the largest
// extent of its 16-row blocks; its count digits number bcls of its first block (rows are ordered
// by digit class, most first, so the first block's is the largest).  The compacted-bin count
// and the largest class come from device memory (dyn): a launch runs only when that class lies
// in [band_lo, band_hi] (the launches of one contraction cover disjoint bands, so exactly one
// runs), and the nsplit splits share the ceil(Kc_pad / 64 / nsplit) steps of bins, the last
// ones possibly empty.  Digit plane rows are ld bytes apart.
template <int RB, int WQ, int WR, int IMAX>
__global__ __launch_bounds__(WQ * WR * 64) void p2_debye_mfma_kernel(
    const double *__restrict__ rr, const double *__restrict__ ir, const int *__restrict__ dyn, int64_t ld,
    int nsplit, int band_lo, int band_hi, const int *__restrict__ kext, const int *__restrict__ bcls,
    const signed char *__restrict__ cdig, int rows_pad, const double *__restrict__ q, int nq, int nrows, int srows,
    double rstep, double *__restrict__ Spart)
{
    constexpr int NFA = IMAX * WR * RB;
    __shared__ i32x4 sa[3][NFA][64];
    __shared__ double srt[3][2 * kMfK];
    __shared__ i32x4 sd[WR > 1 ? 2 : 1][WR > 1 ? WQ : 1][kMfJ][64];
    const int imax = dyn_class(dyn);
    if (imax < band_lo || imax > band_hi) return;
    const int Kc_pad = pad_k(dyn_bins(dyn));
    const int k_per_split = (Kc_pad / kMfK + nsplit - 1) / nsplit * kMfK;
    const int3 tile = mf_tile();  // (q group, split, row group)
    const int b0 = tile.z * (WR * RB);
    int kend = 0;
#pragma unroll
    for (int b = 0; b < WR * RB; b++) kend = max(kend, kext[b0 + b]);
    const int Irg = min(IMAX, bcls[b0]);
    mf_tile_dispatch<RB, WQ, WR, NFA, IMAX>(Irg, tile, kend, sa, srt, sd, rr, ir, Kc_pad, ld, k_per_split, cdig,
                                            rows_pad, q, nq, nrows, srows, rstep, Spart);
}

// launch shapes of the matrix-core Debye sum: (RB, WQ, WR) by row count
struct MfShape {
    int RB, WQ, WR;
    int rows() const { return RB * WR * 16; }
    int qs() const { return WQ * 16; }
    int threads() const { return WQ * WR * 64; }
};

// 32 rows per workgroup (2 row blocks per wave) or 128 (4 row blocks, two waves per q block):
// the larger tile shares each sinc digit among more rows, at about 6.6 against 7.6 us per row
// (measured, 4096 q), but pads the rows to a multiple of 128.  The larger tile takes up to
// three count digits: five overfill its LDS, and four spill its registers.
static MfShape mf_shape(int n_rows, int I)
{
    if (n_rows <= 16) return {1, 4, 1};
    const int64_t pad32 = (n_rows + 31) / 32 * 32, pad128 = (n_rows + 127) / 128 * 128;
    if (I >= 4 || pad32 * 76 <= pad128 * 66) return {2, 4, 1};
    return {4, 4, 2};
}

// Splits of the contraction along the bins, for `tiles` (q group, row group) tiles of shape sh:
// the launch holds 8 waves per CU (every shape's register use allows 8), so a round of
// resident workgroups is cu_count 8 / (waves per workgroup).  At least eight rounds, and the
// fewest splits past that whose last round is nearly full.  A half-empty last round leaves
// half the CUs idle at the end, and row groups of unequal length balance better in smaller
// pieces; measured (contraction kernel, 4096 q): the 256-model batch's 128-row tiles 4.26-4.29
// ms at 4.5 rounds, 4.04-4.07 at 6, 3.82-3.91 at 9; 192 rows on the 32-row tile 1.13 ms at 4.5
// rounds, 1.08 at 6 and at 9; 48 rows 0.44 ms at 4 rounds, 0.45 at 8; one model's 3 rows the
// same at 4 and 8.  At most max_splits.
static int mf_splits(int tiles, const MfShape &sh, int cu_count, int max_splits)
{
    const int64_t round = (int64_t)std::max(1, cu_count) * 8 / (sh.WQ * sh.WR);
    const int nmax = std::max(1, max_splits);
    const int nlo = (int)std::min<int64_t>(nmax, std::max<int64_t>(1, (8 * round + tiles - 1) / tiles));
    int best = nlo;
    double best_fill = 0;
    for (int n = nlo; n <= std::min(nmax, 4 * nlo); n++) {
        const int64_t wg = (int64_t)tiles * n;
        const double fill = (double)wg / (double)(round * ((wg + round - 1) / round));
        if (fill >= 0.97) return n;
        if (fill > best_fill) {
            best_fill = fill;
            best = n;
        }
    }
    return best;
}

// One launch of the contraction: shape, instantiated digit count I (bodies for 1..I, chosen per
// tile at run time) and the band of largest row classes for which it runs.
struct MfLaunch {
    MfShape sh;
    int I, lo, hi;
};

// The launches of a contraction of n_rows rows.  With the largest row class known (imax > 0):
// one launch of its shape.  Unknown (imax 0, the class is read from device memory): one
// instantiation for five digits where the shape does not depend on the class; otherwise the
// 128-row tile for classes up to 3 and the 32-row tile for classes 4 and 5, of which exactly
// one runs (the other exits at once: 5 us for the 32-row tile's grid on the 256-model batch).
static std::vector<MfLaunch> mf_launches(int n_rows, int imax)
{
    if (imax > 0) return {MfLaunch{mf_shape(n_rows, imax), imax, 1, kMfImax}};
    const MfShape s3 = mf_shape(n_rows, 3), s5 = mf_shape(n_rows, kMfImax);
    if (s3.RB == s5.RB) return {MfLaunch{s5, kMfImax, 1, kMfImax}};
    return {MfLaunch{s3, 3, 1, 3}, MfLaunch{s5, kMfImax, 4, kMfImax}};
}

template <int RB, int WQ, int WR>
static void mf_launch_i(int I, dim3 grid, hipStream_t st, const double *rr, const double *ir, const int *dyn,
                        int64_t ld, int nsplit, int lo, int hi, const int *kext, const int *bcls,
                        const signed char *cdig, int rows_pad, const double *q, int nq, int nrows, int srows,
                        double rstep, double *Spart)
{
    const dim3 blk(WQ * WR * 64);
#define DBX_MF(II)                                                                                               \
    hipLaunchKernelGGL((p2_debye_mfma_kernel<RB, WQ, WR, II>), grid, blk, 0, st, rr, ir, dyn, ld, nsplit, lo, hi, \
                       kext, bcls, cdig, rows_pad, q, nq, nrows, srows, rstep, Spart)
    switch (I) {
    case 1: DBX_MF(1); break;
    case 2: DBX_MF(2); break;
    case 3: DBX_MF(3); break;
    default:
        if constexpr (WR == 1) {
            if (I == 4)
                DBX_MF(4);
            else
                DBX_MF(5);
        } else {
            throw Error(DBX_ERR_ARG, "four or five count digits need a one-row-wave shape");
        }
        break;
    }
#undef DBX_MF
}

// grid: (q groups, nsplit, row groups of the launch's shape)
static void mf_launch(const MfLaunch &L, int qgroups, int nsplit, int rows_pad, hipStream_t st, const double *rr,
                      const double *ir, const int *dyn, int64_t ld, const int *kext, const int *bcls,
                      const signed char *cdig, const double *q, int nq, int nrows, int srows, double rstep,
                      double *Spart)
{
    const dim3 grid((unsigned)qgroups, (unsigned)nsplit, (unsigned)(rows_pad / L.sh.rows()));
    if (L.sh.RB == 1)
        mf_launch_i<1, 4, 1>(L.I, grid, st, rr, ir, dyn, ld, nsplit, L.lo, L.hi, kext, bcls, cdig, rows_pad, q, nq,
                             nrows, srows, rstep, Spart);
    else if (L.sh.RB == 2)
        mf_launch_i<2, 4, 1>(L.I, grid, st, rr, ir, dyn, ld, nsplit, L.lo, L.hi, kext, bcls, cdig, rows_pad, q, nq,
                             nrows, srows, rstep, Spart);
    else
        mf_launch_i<4, 4, 2>(L.I, grid, st, rr, ir, dyn, ld, nsplit, L.lo, L.hi, kext, bcls, cdig, rows_pad, q, nq,
                             nrows, srows, rstep, Spart);
}

// I[dest][q] = sum over the model's rows of ff S, FP64, rows in order;
// mrows[m] = (first row, row count, dest, -), or m0 for a single model (mrows null).
// rel (optional) receives the relative error bound of each I[dest][q] against the exact sum of
// the same histograms and prefactors:
//   (sum_r |ff_r| B_r + (P + 8) u sum_r |ff_r S_r|) / |I|
// with B_r the row bounds (row_bound); the second term covers the FP64 rounding of
// this sum, of S_r's last additions and of the prefactors (a few units each).
__global__ void p2_combine_kernel(const double *__restrict__ ff, const double *__restrict__ S,
                                  const double *__restrict__ B, const int4 *__restrict__ mrows, int4 m0, int nq,
                                  double *__restrict__ I, double *__restrict__ rel)
{
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (iq >= nq) return;
    const int4 rr = mrows ? mrows[m] : m0;
    double acc = 0.0, mag = 0.0, bnd = 0.0;
    for (int r = rr.x; r < rr.x + rr.y; r++) {
        const double f = ff[(int64_t)r * nq + iq];
        const double t = S[(int64_t)r * nq + iq] * f;
        acc += t;
        mag += fabs(t);
        if (B) bnd += fabs(f) * B[r];
    }
    I[(int64_t)rr.z * nq + iq] = acc;
    // an I of exactly 0 from terms of exactly 0 (form factors that underflow at large Q) has
    // relative error 0, not 0 / 0
    const double num = bnd + (double)(rr.y + 8) * 0x1p-53 * mag;
    if (rel) rel[(int64_t)rr.z * nq + iq] = num == 0.0 ? 0.0 : num / fabs(acc);
}

// largest k with double(k)*rstep <= 0.0005 (the reference's test r[k] <= 0.0005)
static int self_bin_max(double rstep)
{
    int k = (int)std::floor(0.0005 / rstep);
    while ((double)(k + 1) * rstep <= 0.0005) k++;
    while (k >= 0 && (double)k * rstep > 0.0005) k--;
    return k;
}

// zeroes n_words words at z and, with rows given, writes the descriptors of a single model's
// P partials: row r's histogram at r * n_bins, all of equal length
__global__ void p2_prep_kernel(unsigned int *__restrict__ z, int64_t n_words, RowDesc *__restrict__ rows,
                               int n_rows, int64_t n_bins, int first_n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words) z[i] = 0u;
    if (rows && i < n_rows) rows[i] = RowDesc{i * n_bins, n_bins, first_n, 0};
}

// the finish launch: sums of rows [r0, r0 + nr) (and their bounds when B is given)
static void launch_finish(hipStream_t st, const double *Spart, int nsplit, int r0, int nr, int nq,
                          const unsigned long long *rsum, double shift, const double *self, const double *low,
                          int nlow, double rstep, const double *q, const int *perm, double *S,
                          const unsigned long long *brsum, int mode, const int *dyn, double *B)
{
    const int64_t nfin = ((int64_t)nr * nq + kFinishThreads - 1) / kFinishThreads;
    const int64_t nb = B ? (nr + kFinishThreads - 1) / kFinishThreads : 0;
    hipLaunchKernelGGL(p2_finish_rows_kernel, dim3((unsigned)(nfin + nb)), dim3(kFinishThreads), 0, st, Spart, nsplit,
                       r0, nr, nq, rsum, shift, self, low, nlow, rstep, q, perm, S, brsum, mode, dyn, B);
}

static_assert(sizeof(((dbx_ctx *)nullptr)->p2_pending_mode) == sizeof(int) * kP2StatusMax,
              "one mode per status slot");

unsigned long long *pass2_status_slot(dbx_ctx *ctx, int mode)
{
    if (ctx->p2_pending == kP2StatusMax) sync_pass2(ctx);
    const int i = ctx->p2_pending++;
    ctx->p2_pending_mode[i] = mode;
    return (unsigned long long *)ctx->p2_status_h + 2 * i;
}

void sync_pass2(dbx_ctx *ctx)
{
    DBX_HIP(hipStreamSynchronize(ctx->stream));
    stage_reset(ctx);
    if (!ctx->no_alloc && (!ctx->retired_dev.empty() || !ctx->retired_pin.empty())) release_retired(ctx);
    const int n = ctx->p2_pending;
    ctx->p2_pending = 0;
    bool too_large = false;
    for (int i = 0; i < n; i++) {
        const unsigned long long *w = (const unsigned long long *)ctx->p2_status_h + 2 * i;
        ctx->last_p2_bins = (int)(w[0] & 0xFFFFFFFFull);
        too_large = too_large || (ctx->p2_pending_mode[i] != DBX_PASS2_FP64 && (int)(w[1] & 0xFFFFFFFFull) > kMfImax);
    }
    DBX_REQUIRE(!too_large, DBX_ERR_RANGE, "bin count too large for the matrix-core Debye sum");
}

// Profiling (dbx_debug_set_profiling): the matrix instructions the contraction issued, from
// the launch that ran, its geometry and the row blocks' bin extents and digit classes (read
// back after a wait), and the algorithmic work, 2 nq flops per nonzero bin of every row.
static void profile_contraction(dbx_ctx *ctx, const unsigned long long *h, const RowDesc *d_rows, int n_rows,
                                int64_t nb_max, int nq, int rows_pad, int qgroups, const std::vector<MfLaunch> &Ls,
                                const unsigned long long *tot, const int *kext, const int *bcls)
{
    const int nblocks = rows_pad / 16;
    unsigned long long *hv = (unsigned long long *)pinned(ctx, kPinMisc, sizeof(unsigned long long) * 3 +
                                                                          2 * sizeof(int) * nblocks);
    int *hk = (int *)(hv + 3), *hb = hk + nblocks;
    DBX_HIP(hipMemcpyAsync(hv, tot, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, ctx->stream));
    DBX_HIP(hipMemcpyAsync(hk, kext, sizeof(int) * nblocks, hipMemcpyDeviceToHost, ctx->stream));
    DBX_HIP(hipMemcpyAsync(hb, bcls, sizeof(int) * nblocks, hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long *cnt = (unsigned long long *)scratch(ctx, kSlotProf, sizeof(unsigned long long));
    hipLaunchKernelGGL(p2_zero_kernel, dim3(1), dim3(64), 0, ctx->stream, (unsigned int *)cnt, (int64_t)2);
    hipLaunchKernelGGL(p2_row_nnz_kernel,
                       dim3((unsigned)((nb_max + kScanBlock - 1) / kScanBlock), (unsigned)((n_rows + kFlagRows - 1) / kFlagRows)),
                       dim3(kScanBlock), 0, ctx->stream, h, d_rows, n_rows, nb_max, cnt);
    DBX_HIP(hipMemcpyAsync(hv + 2, cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    sync_pass2(ctx);
    const int Kc = (int)(hv[0] & 0xFFFFFFFFull);
    const int Imax = std::max(1, (int)(hv[1] & 0xFFFFFFFFull));
    const MfLaunch *run = nullptr;
    for (const MfLaunch &L : Ls)
        if (Imax >= L.lo && Imax <= L.hi) run = &L;
    if (!run) return;
    const MfShape &sh = run->sh;
    const int bpg = sh.rows() / 16;
    double issued = 0, digits = 0;
    for (int z = 0; z < rows_pad / sh.rows(); z++) {
        int ke = 0;
        for (int b = 0; b < bpg; b++) ke = std::max(ke, hk[z * bpg + b]);
        const int kx = std::min(pad_k(Kc), pad_k(ke));
        const int Iz = std::min(run->I, hb[z * bpg]);
        issued += (double)(kx / kMfK) * qgroups * sh.WQ * sh.WR * sh.RB * Iz * kMfJ * 32768.0;
        digits += (double)Iz * sh.rows();
    }
    ctx->mf_flops = issued;
    ctx->mf_useful = (double)hv[2] * (double)nq * 2.0;
    ctx->mf_digits = Imax;
    ctx->mf_mean_digits = digits / rows_pad;
    ctx->mf_bins = Kc;
    ctx->mf_timed = true;
}

// Buffers of one pass 2 sized from upper bounds stay below this; a larger problem reads its
// compacted-bin count and digit class back first and sizes them exactly.
constexpr size_t kP2BoundBytes = (size_t)8 << 30;

// The pass-2 structure sums S[r][q] of n_rows rows (longest first; nb_max = the first row's
// length), buffers from mem.  h_rows: the row descriptors on the host (uploaded through pinned
// memory), or null for the P equal rows of one model (nb_max bins each, uni_first_n), written
// on the device.
//
// No host round trip: every launch and buffer is sized before the compaction has run, from
// upper bounds (the compacted bins from the bins of the longest row past the self bins, the
// count digits from kMfImax planes), and the kernels read the actual compacted-bin count and
// largest digit class from device memory (dyn).  Only when buffers sized that way would exceed
// kP2BoundBytes or the provider's room, or cannot be allocated, does the host read the two
// values back first and size the buffers exactly.  The splits of the sums are chosen from the
// bounds either way, so both give the same sums bit for bit.  The two values are also copied
// to the context's pinned status words, which sync_pass2 checks after the caller's wait (a
// count beyond the matrix cores' digits is reported there as DBX_ERR_RANGE).
static void structure_core(dbx_ctx *ctx, P2Mem &mem, const uint64_t *d_hist_base, const RowDesc *h_rows,
                           int n_rows, int64_t nb_max, int uni_first_n, double rstep, const double *d_q, int nq,
                           double *d_S, int mode, double *d_B)
{
    const int ksm = self_bin_max(rstep);
    const int nlow = ksm + 1;
    const int64_t nblk = (nb_max + kScanBlock - 1) / kScanBlock;
    const bool mfma = mode != DBX_PASS2_FP64;
    // upper bound of the compacted bins: every bin of the longest row past the self bins
    const int kc_cap = (int)std::max<int64_t>(0, nb_max - nlow);
    // the contraction's launches for an unknown digit class; rows padded for each of their shapes
    const std::vector<MfLaunch> Lb = mf_launches(n_rows, 0);
    int tile_rows = 16;
    for (const MfLaunch &L : Lb) tile_rows = std::max(tile_rows, L.sh.rows());
    const int rows_pad = (n_rows + tile_rows - 1) / tile_rows * tile_rows;
    const int nblocks = rows_pad / 16;
    // the splits of the sums, from the bounds
    const int qgroups = (nq + Lb[0].sh.qs() - 1) / Lb[0].sh.qs();
    const int qblocks = (nq + kDsThreads - 1) / kDsThreads;
    int nsplit;
    if (mfma) {
        const int rg = rows_pad / Lb[0].sh.rows();
        nsplit = mf_splits(qgroups * rg, Lb[0].sh, ctx->cu_count, pad_k(kc_cap) / kMfK / 32);  // >= 32 steps each
    } else {
        nsplit = std::max(1, std::min((2048 + qblocks - 1) / qblocks, (kc_cap + 511) / 512));
    }
    const int npass = std::min(n_rows, kDsPmax);
    // scratch: rows (the caller's order, then the order of the sums), the permutation between
    // them, block counts, kz, self, low, kext, block digit classes, then the parts that start at
    // zero (totals, row digit classes, row sums, flags), cleared by one launch
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off = (off + b + 255) & ~(size_t)255; return o; };
    const size_t o_rows = take(sizeof(RowDesc) * n_rows);
    const size_t o_krows = take(sizeof(RowDesc) * n_rows);
    const size_t o_perm = take(sizeof(int) * n_rows);
    const size_t o_bc = take(sizeof(int) * (nblk + 1));
    const size_t o_kz = take(sizeof(int) * nb_max);
    const size_t o_self = take(sizeof(double) * n_rows);
    const size_t o_low = take(sizeof(double) * (size_t)n_rows * nlow);
    const size_t o_kext = take(sizeof(int) * nblocks);
    const size_t o_bcls = take(sizeof(int) * nblocks);
    const size_t o_tot = take(sizeof(unsigned long long) * 2);
    const size_t o_rcls = take(sizeof(int) * n_rows);
    const size_t o_rsum = take(sizeof(unsigned long long) * n_rows);
    const size_t o_flags = take((size_t)(nblk * kScanBlock));
    char *base = (char *)mem.dev(kSlotPass2, off);
    RowDesc *d_rows = (RowDesc *)(base + o_rows);
    RowDesc *d_krows = (RowDesc *)(base + o_krows);
    int *d_perm = (int *)(base + o_perm);
    int *bc = (int *)(base + o_bc);
    unsigned long long *tot = (unsigned long long *)(base + o_tot);  // [0] Kc, [1] largest row digit class
    const int *dyn = (const int *)tot;
    int *rcls = (int *)(base + o_rcls);
    int *kz = (int *)(base + o_kz);
    double *self = (double *)(base + o_self);
    double *low = (double *)(base + o_low);
    int *kext = (int *)(base + o_kext);
    int *d_bcls = (int *)(base + o_bcls);
    unsigned long long *rsum = (unsigned long long *)(base + o_rsum);
    unsigned char *flags = (unsigned char *)(base + o_flags);
    const unsigned long long *h = (const unsigned long long *)d_hist_base;

    if (!ctx->p2_refining) DBX_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    if (h_rows) {
        RowDesc *hrows = (RowDesc *)mem.staging(sizeof(RowDesc) * n_rows);
        DBX_REQUIRE(hrows != nullptr, DBX_ERR_BUDGET, "pass-2 row list beyond the staging memory");
        std::copy(h_rows, h_rows + n_rows, hrows);
        DBX_HIP(hipMemcpyAsync(d_rows, hrows, sizeof(RowDesc) * n_rows, hipMemcpyHostToDevice, ctx->stream));
    }
    // zeroing (and the rows of one model), flags and row classes (with the block counts when one
    // chunk of rows covers every row), counts otherwise, the scan with the rows ordered by digit
    // class, then the scatter with the self-pair terms in its last blocks
    const int64_t zwords = (int64_t)(off - o_tot) / 4;
    const int64_t prep_n = std::max<int64_t>(zwords, h_rows ? 0 : n_rows);
    hipLaunchKernelGGL(p2_prep_kernel, dim3((unsigned)((prep_n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (unsigned int *)tot, zwords, h_rows ? nullptr : d_rows, n_rows, nb_max, uni_first_n);
    const int nchunks = (n_rows + kFlagRows - 1) / kFlagRows;
    hipLaunchKernelGGL(p2_flags_kernel, dim3((unsigned)nblk, (unsigned)nchunks), dim3(kScanBlock), 0, ctx->stream, h,
                       d_rows, n_rows, nb_max, ksm, flags, rcls, nchunks == 1 ? bc : nullptr);
    if (nchunks > 1)
        hipLaunchKernelGGL(p2_count_kernel, dim3((unsigned)nblk), dim3(kScanBlock), 0, ctx->stream, flags, nb_max,
                           bc);
    hipLaunchKernelGGL(p2_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, bc, (int)nblk, (int *)tot, rcls, n_rows,
                       d_rows, mfma ? d_krows : nullptr, d_perm, d_bcls, nblocks);
    hipLaunchKernelGGL(p2_scatter_kernel, dim3((unsigned)(nblk + (n_rows + kScanBlock - 1) / kScanBlock)),
                       dim3(kScanBlock), 0, ctx->stream, flags, nb_max, bc, kz, (int)nblk, h, d_rows, n_rows, ksm,
                       self, low);
    // the compacted-bin count and largest class to the status words (checked after the caller's
    // wait); sized from the bounds unless those buffers would not fit
    unsigned long long *hs = pass2_status_slot(ctx, mode);
    DBX_HIP(hipMemcpyAsync(hs, tot, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, ctx->stream));
    int64_t ld = mfma ? pad_k(std::max(kc_cap, 1)) : std::max(kc_cap, 1);
    int planes = kMfImax;
    std::vector<MfLaunch> Ls = Lb;
    // the buffers sized by the compacted bins (digit planes and r table, or FP64 weights), and
    // the partial sums after them
    double *A = nullptr, *rtab = nullptr, *Spart = nullptr;
    signed char *cdig = nullptr;
    auto take_sized = [&] {
        if (mfma) {
            rtab = (double *)mem.dev(kSlotRtab, sizeof(double) * 2 * (size_t)ld);
            cdig = (signed char *)mem.dev(kSlotDigits, (size_t)planes * rows_pad * ld);
        } else {
            A = (double *)mem.dev(kSlotDigits, sizeof(double) * (size_t)npass * ld);
        }
        Spart = (double *)mem.dev(kSlotPass2b, sizeof(double) * (size_t)nsplit * (mfma ? n_rows : npass) * nq);
    };
    const size_t spart_bytes = sizeof(double) * (size_t)nsplit * (mfma ? n_rows : npass) * nq + 1024;
    const size_t bound_bytes = spart_bytes + (mfma ? (size_t)planes * rows_pad * ld + 2 * sizeof(double) * ld
                                                   : sizeof(double) * (size_t)npass * ld);
    bool sized = false;
    if (bound_bytes <= std::min(kP2BoundBytes, mem.room())) {
        try {
            take_sized();
            sized = true;
        } catch (const Error &e) {
            // the device has no room for them now: size them exactly instead
            if (e.code != DBX_ERR_HIP && e.code != DBX_ERR_BUDGET) throw;
            (void)hipGetLastError();
        }
    }
    if (!sized) {
        sync_pass2(ctx);  // throws for a count beyond the matrix cores' digits
        const int Kc = (int)(hs[0] & 0xFFFFFFFFull);
        const int Imax = std::max(1, (int)(hs[1] & 0xFFFFFFFFull));
        ld = mfma ? pad_k(std::max(Kc, 1)) : std::max(Kc, 1);
        planes = Imax;
        if (mfma) {
            Ls = mf_launches(n_rows, Imax);
            DBX_REQUIRE(rows_pad % Ls[0].sh.rows() == 0, DBX_ERR_STATE, "pass-2 rows padded for another tile");
        }
        take_sized();
    }
    if (!mfma) {
        for (int r0 = 0; r0 < n_rows; r0 += kDsPmax) {
            const int np = std::min(kDsPmax, n_rows - r0);
            hipLaunchKernelGGL(p2_weights_kernel, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, ctx->stream, h,
                               d_rows, r0, np, kz, dyn, ld, A, rsum);
            hipLaunchKernelGGL(p2_debye_fp64_kernel, dim3(qblocks, nsplit), dim3(kDsThreads), 0, ctx->stream, kz,
                               dyn, nsplit, ld, A, np, rstep, d_q, nq, Spart);
            launch_finish(ctx->stream, Spart, nsplit, r0, np, nq, nullptr, 0.0, self, low, nlow, rstep, d_q, nullptr,
                          d_S, rsum, mode, dyn, d_B);
        }
    } else {
        const int cd_x = (int)((ld + 4 * kCdThreads - 1) / (4 * kCdThreads));
        const int n_rt = (int)((ld + kCdThreads - 1) / kCdThreads);
        const int64_t n_in = (int64_t)cd_x * rows_pad + n_rt + (nblocks + kCdThreads - 1) / kCdThreads;
        DBX_REQUIRE(n_in < INT32_MAX, DBX_ERR_ARG, "pass-2 problem too large");
        hipLaunchKernelGGL(p2_mf_inputs_kernel, dim3((unsigned)n_in), dim3(kCdThreads), 0, ctx->stream, h, d_krows,
                           n_rows, kz, dyn, ld, rows_pad, cd_x, cdig, rsum, rstep, rtab, rtab + ld, n_rt, nblocks,
                           kext);
        if (ctx->profiling) DBX_HIP(hipEventRecord(ctx->ev_mf[0], ctx->stream));
        for (const MfLaunch &L : Ls)
            mf_launch(L, qgroups, nsplit, rows_pad, ctx->stream, rtab, rtab + ld, dyn, ld, kext, d_bcls, cdig, d_q,
                      nq, n_rows, n_rows, rstep, Spart);
        if (ctx->profiling) {
            DBX_HIP(hipEventRecord(ctx->ev_mf[1], ctx->stream));
            profile_contraction(ctx, h, d_rows, n_rows, nb_max, nq, rows_pad, qgroups, Ls, tot, kext, d_bcls);
        }
        launch_finish(ctx->stream, Spart, nsplit, 0, n_rows, nq, rsum, kMfShift, self, low, nlow, rstep, d_q, d_perm,
                      d_S, rsum, mode, dyn, d_B);
    }
    DBX_HIP(hipGetLastError());
    DBX_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    ctx->p2_timed = true;
}

// the context's grow-only slots and pinned words
namespace {
struct CtxMem : P2Mem {
    dbx_ctx *ctx = nullptr;
    void *dev(int slot, size_t bytes) override { return scratch(ctx, slot, bytes); }
    void *staging(size_t bytes) override { return stage(ctx, bytes); }
};
}  // namespace

P2Mem &ctx_pass2_mem(dbx_ctx *ctx)
{
    thread_local CtxMem m;
    m.ctx = ctx;
    return m;
}

void debye_structure_rows(dbx_ctx *ctx, const uint64_t *d_hist_base, const std::vector<RowDesc> &rows,
                          double rstep, const double *d_q, int nq, double *d_S, double *d_B, int mode)
{
    const int n_rows = (int)rows.size();
    DBX_REQUIRE(n_rows > 0 && nq > 0, DBX_ERR_ARG, "empty pass-2 problem");
    for (int r = 0; r < n_rows; r++) {
        DBX_REQUIRE(rows[r].n_bins > 0, DBX_ERR_ARG, "pass-2 row without bins");
        DBX_REQUIRE(r == 0 || rows[r].n_bins <= rows[r - 1].n_bins, DBX_ERR_ARG,
                    "pass-2 rows must be ordered by length, longest first");
    }
    DBX_REQUIRE(rows[0].n_bins < (int64_t)INT32_MAX, DBX_ERR_ARG, "too many bins");
    structure_core(ctx, ctx_pass2_mem(ctx), d_hist_base, rows.data(), n_rows, rows[0].n_bins, 0, rstep, d_q, nq,
                   d_S, mode < 0 ? ctx->pass2_mode : mode, d_B);
}

void structure_fp64_models(dbx_ctx *ctx, const uint64_t *d_hist_base, const std::vector<RowDesc> &rows,
                           const std::vector<int2> &models, double rstep, const double *d_q, int nq,
                           const std::vector<double *> &S_dst, const std::vector<double *> &B_dst)
{
    std::vector<RowDesc> sub;
    for (const int2 &m : models) sub.insert(sub.end(), rows.begin() + m.x, rows.begin() + m.x + m.y);
    if (sub.empty()) return;
    const size_t n = sub.size(), nqd = (size_t)nq;
    double *dS = (double *)scratch(ctx, kSlotRefine, sizeof(double) * n * (nqd + 1));
    double *dB = dS + n * nqd;
    P2Refining mark(ctx);
    debye_structure_rows(ctx, d_hist_base, sub, rstep, d_q, nq, dS, dB, DBX_PASS2_FP64);
    size_t r0 = 0;
    for (size_t m = 0; m < models.size(); m++) {
        const size_t c = (size_t)models[m].y;
        DBX_HIP(hipMemcpyAsync(S_dst[m], dS + r0 * nqd, sizeof(double) * c * nqd, hipMemcpyDeviceToDevice,
                               ctx->stream));
        DBX_HIP(hipMemcpyAsync(B_dst[m], dB + r0, sizeof(double) * c, hipMemcpyDeviceToDevice, ctx->stream));
        r0 += c;
    }
}

void debye_structure_mem(dbx_ctx *ctx, P2Mem &mem, const uint64_t *d_hist, int P, int64_t n_bins, double rstep,
                         int first_n, const double *d_q, int nq, double *d_S, double *d_B, int mode)
{
    DBX_REQUIRE(P > 0 && nq > 0 && n_bins > 0, DBX_ERR_ARG, "empty pass-2 problem");
    DBX_REQUIRE(n_bins < (int64_t)INT32_MAX, DBX_ERR_ARG, "too many bins");
    structure_core(ctx, mem, d_hist, nullptr, P, n_bins, first_n, rstep, d_q, nq, d_S,
                   mode < 0 ? ctx->pass2_mode : mode, d_B);
}

void debye_structure(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep,
                     int first_n, const double *d_q, int nq, double *d_S, double *d_B, int mode)
{
    debye_structure_mem(ctx, ctx_pass2_mem(ctx), d_hist, P, n_bins, rstep, first_n, d_q, nq, d_S, d_B, mode);
}

void combine_models(dbx_ctx *ctx, const double *d_ff, const double *d_S, const double *d_B,
                    const std::vector<int4> &mrows, int nq, double *d_I, double *d_rel)
{
    const int nm = (int)mrows.size();
    if (nm == 1) {
        // one model: its row range as an argument, no copy
        hipLaunchKernelGGL(p2_combine_kernel, dim3((nq + 255) / 256, 1), dim3(256), 0, ctx->stream, d_ff, d_S, d_B,
                           nullptr, mrows[0], nq, d_I, d_rel);
        DBX_HIP(hipGetLastError());
        return;
    }
    int4 *hm = (int4 *)stage(ctx, sizeof(int4) * nm);
    std::copy(mrows.begin(), mrows.end(), hm);
    int4 *dm = (int4 *)scratch(ctx, kSlotRows, sizeof(int4) * nm);
    DBX_HIP(hipMemcpyAsync(dm, hm, sizeof(int4) * nm, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(p2_combine_kernel, dim3((nq + 255) / 256, nm), dim3(256), 0, ctx->stream, d_ff, d_S, d_B, dm,
                       make_int4(0, 0, 0, 0), nq, d_I, d_rel);
    DBX_HIP(hipGetLastError());
}

void combine_intensity(dbx_ctx *ctx, const double *d_ff, const double *d_S, const double *d_B, int P, int nq,
                       double *d_I, double *d_rel)
{
    combine_models(ctx, d_ff, d_S, d_B, std::vector<int4>{make_int4(0, P, 0, 0)}, nq, d_I, d_rel);
}

// *out = the largest of v[0..n) as its bit pattern (non-negative doubles order like their bits;
// a NaN sorts above every number), one workgroup
__global__ __launch_bounds__(1024) void p2_max_bits_kernel(const double *__restrict__ v, int n,
                                                            unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long sm[1024 / 64];
    unsigned long long m = 0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v[i]);
        m = b > m ? b : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / 64; w++) m = sm[w] > m ? sm[w] : m;
        *out = m;
    }
}

void launch_max_bits(hipStream_t st, const double *d_v, int n, unsigned long long *d_out)
{
    hipLaunchKernelGGL(p2_max_bits_kernel, dim3(1), dim3(1024), 0, st, d_v, n, d_out);
    DBX_HIP(hipGetLastError());
}

// stream-explicit launches for callers that record them into a graph (no allocation, no copy)
void launch_prefactor(hipStream_t st, const GroupPair *d_pairs, int n, const double *d_q, int nq, double *d_ff)
{
    hipLaunchKernelGGL(p2_prefactor_kernel, dim3((nq + 255) / 256, n), dim3(256), 0, st, d_pairs, d_q, nq, d_ff);
    DBX_HIP(hipGetLastError());
}

void launch_combine(hipStream_t st, const double *d_ff, const double *d_S, const double *d_B, const int4 *d_mrows,
                    int nm, int nq, double *d_I, double *d_rel, int4 m0)
{
    hipLaunchKernelGGL(p2_combine_kernel, dim3((nq + 255) / 256, nm), dim3(256), 0, st, d_ff, d_S, d_B, d_mrows,
                       m0, nq, d_I, d_rel);
    DBX_HIP(hipGetLastError());
}

}  // namespace dbx

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
#include <cmath>
#include <vector>

#include "fastmath.h"
#include "internal.h"

namespace dbx {

// ---------------------------------------------------------------------------------
// prefactor, cuDebye-v15.cu:447-467 (FP64)
// ---------------------------------------------------------------------------------
struct GroupPair {
    double occ_a, occ_b, beq_a, beq_b, ca, cb;
    double aa[5], ba[5], ab[5], bb[5];
};

__global__ void p2_prefactor_kernel(GroupPair g, const double *__restrict__ q, int nq,
                                    double *__restrict__ ff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double rp16pi2 = -0.006332573977646;  // -1/(16 pi^2), cuDebye-v15.cu:452
    const double nh = q[i] * q[i] * rp16pi2;
    double v = g.occ_a * g.occ_b;
    v = v * exp(nh * (g.beq_a + g.beq_b));
    double f1 = g.ca, f2 = g.cb;
    for (int t = 0; t < 5; t++) {
        f1 += g.aa[t] * exp(g.ba[t] * nh);
        f2 += g.ab[t] * exp(g.bb[t] * nh);
    }
    ff[i] = v * f1 * f2;
}

void prefactor_dev(dbx_ctx *ctx, const dbx_group_params &ga, const dbx_group_params &gb,
                   const double *d_q, int nq, double *d_ff)
{
    GroupPair g;
    g.occ_a = ga.occ; g.occ_b = gb.occ; g.beq_a = ga.beq; g.beq_b = gb.beq;
    g.ca = ga.c; g.cb = gb.c;
    for (int t = 0; t < 5; t++) {
        g.aa[t] = ga.a[t]; g.ba[t] = ga.b[t]; g.ab[t] = gb.a[t]; g.bb[t] = gb.b[t];
    }
    hipLaunchKernelGGL(p2_prefactor_kernel, dim3((nq + 255) / 256), dim3(256), 0, ctx->stream, g,
                       d_q, nq, d_ff);
    DBX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------
// compaction: union of nonzero bins of P partials, ascending
// ---------------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__global__ void p2_count_kernel(const unsigned long long *__restrict__ hist, int P, int64_t n_bins,
                                int k_self_max, int *__restrict__ block_counts,
                                unsigned long long *__restrict__ max_count)
{
    __shared__ int wsum[kScanBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    bool nz = false;
    unsigned long long mx = 0;
    if (k < n_bins && k > k_self_max)
        for (int p = 0; p < P; p++) {
            const unsigned long long v = hist[(int64_t)p * n_bins + k];
            nz |= v != 0ull;
            mx = v > mx ? v : mx;
        }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(mx, off);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(max_count, mx);
    const uint64_t b = __ballot(nz);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kScanBlock / 64; w++) s += wsum[w];
        block_counts[blockIdx.x] = s;
    }
}

// exclusive scan of the block counts (one block; n_blocks small)
__global__ void p2_scan_kernel(int *__restrict__ block_counts, int n_blocks, int *__restrict__ total)
{
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < n_blocks; i++) {
            int c = block_counts[i];
            block_counts[i] = s;
            s += c;
        }
        *total = s;
    }
}

__global__ void p2_scatter_kernel(const unsigned long long *__restrict__ hist, int P, int64_t n_bins,
                                  int k_self_max, const int *__restrict__ block_offsets,
                                  int *__restrict__ kz)
{
    __shared__ int woff[kScanBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    bool nz = false;
    if (k < n_bins && k > k_self_max)
        for (int p = 0; p < P; p++) nz |= hist[(int64_t)p * n_bins + k] != 0ull;
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

// weights of the compacted bins, FP64: A[p][i] = 2 N_p,kz[i]
__global__ void p2_weights_kernel(const unsigned long long *__restrict__ hist, int P, int64_t n_bins,
                                  const int *__restrict__ kz, int Kc, double *__restrict__ A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Kc) return;
    const int64_t k = kz[i];
    for (int p = 0; p < P; p++) A[(int64_t)p * Kc + i] = 2.0 * (double)hist[(int64_t)p * n_bins + k];
}

// weight-1 bins (cuDebye-v15.cu:508-514) and the weight-2 bins at r <= 0.0005 that the
// index condition excludes; one thread per partial.  self[p] = sum of weight-1 counts,
// low[p * (k_self_max+1) + k] = 2 N for the other low bins (added to S with sinc(q r_k)).
__global__ void p2_self_kernel(const unsigned long long *__restrict__ hist, int P, int64_t n_bins,
                               int k_self_max, int first_n, double *__restrict__ self,
                               double *__restrict__ low)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double s = 0;
    int rank = 0;
    for (int k = 0; k <= k_self_max; k++) {
        const unsigned long long n = k < n_bins ? hist[(int64_t)p * n_bins + k] : 0ull;
        double lw = 0;
        if (n) {
            if (rank < first_n)
                s += (double)n;
            else
                lw = 2.0 * (double)n;
            rank++;
        }
        low[p * (k_self_max + 1) + k] = lw;
    }
    self[p] = s;
}

// ---------------------------------------------------------------------------------
// Debye sum, FP64: one thread per q, bins split over blockIdx.y and staged through LDS,
// one sinc per (bin, q) shared by all partials; split partials reduced in fixed order.
// ---------------------------------------------------------------------------------
constexpr int kDsThreads = 256;
constexpr int kDsPmax = 10;  // partials per pass (S <= 4 subgroups in one pass)
constexpr int kDsChunk = 256;

__global__ __launch_bounds__(kDsThreads) void p2_debye_fp64_kernel(
    const int *__restrict__ kz, int Kc, int k_per_split, const double *__restrict__ A, int p0,
    int np, double rstep, const double *__restrict__ q, int nq, double *__restrict__ Spart)
{
    __shared__ double sr[kDsChunk];
    __shared__ double sa[kDsPmax][kDsChunk];
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
            for (int t = 0; t < np; t++) sa[t][i] = A[(int64_t)(p0 + t) * Kc + c0 + i];
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

__global__ void p2_reduce_splits_kernel(const double *__restrict__ Spart, int nsplit, int p0, int np,
                                        int nq, double *__restrict__ S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np * nq) return;
    double acc = 0.0;
    for (int s = 0; s < nsplit; s++) acc += Spart[(int64_t)s * np * nq + i];
    S[(int64_t)p0 * nq + i] = acc;
}

// ---------------------------------------------------------------------------------
// Debye sum on the matrix cores: S[rows x q] = A[rows x bins] . T[bins x q],
// A = 2 N (integer), T = sinc(q r).  Split-BF16 limbs, exact FP32 accumulation:
//   A = sum_i e_i 256^i,              e_i in [0, 255]           (count digits, I of them)
//   T ~ V / 2^40,  V = rint(T 2^40) = sum_j d_j 256^j,  d_j in [-128, 128], d_4 in [-56, 256]
// Every digit is an integer that BF16 holds exactly and every product e_i d_j is an integer
// below 2^16 in magnitude, so an FP32 accumulator that collects the products of one scale
// class s = i + j (at most 5 digit pairs) over 32 bins stays below 2^24 and is exact,
// whatever order the matrix core adds in.  After every 32 bins the classes are folded into
// FP64: S += sum_s 2^(8s - 40) acc_s.  The only approximation is the rounding of T to 2^-40
// (|error| <= 2^-41 per bin, from an FP64 sin with full argument reduction).
// ---------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfWaves = 4;   // waves per workgroup (one 16-value Q block each)
constexpr int kMfImax = 5;    // count digits (counts below 2^40)
constexpr int kMfJ = 5;       // sinc digits
constexpr int kMfClasses = kMfImax + kMfJ - 1;

// count digits: cdig[i][row][kk] = byte i of 2 N_row,kz[kk] (0 past Kc or past P)
__global__ void p2_count_digits_kernel(const unsigned long long *__restrict__ hist, int P,
                                       int64_t n_bins, const int *__restrict__ kz, int Kc, int Kc_pad,
                                       int rows_pad, int I, __bf16 *__restrict__ cdig)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows_pad * Kc_pad) return;
    const int row = (int)(t / Kc_pad);
    const int kk = (int)(t - (int64_t)row * Kc_pad);
    unsigned long long c = 0;
    if (row < P && kk < Kc) c = 2ull * hist[(int64_t)row * n_bins + kz[kk]];
    for (int i = 0; i < I; i++)
        cdig[((int64_t)i * rows_pad + row) * Kc_pad + kk] = (__bf16)(float)((c >> (8 * i)) & 0xFFull);
}

// T rounded to a multiple of 2^-40, as 5 signed base-256 digits (exact in bf16)
__device__ __forceinline__ void fixed_digits(double T, float d[kMfJ])
{
    double r = rint(T * 1099511627776.0);  // 2^40
    const double t4 = rint(r * 0x1p-32);
    r = fma(-t4, 4294967296.0, r);
    const double t3 = rint(r * 0x1p-24);
    r = fma(-t3, 16777216.0, r);
    const double t2 = rint(r * 0x1p-16);
    r = fma(-t2, 65536.0, r);
    const double t1 = rint(r * 0x1p-8);
    const double t0 = fma(-t1, 256.0, r);
    d[0] = (float)t0;
    d[1] = (float)t1;
    d[2] = (float)t2;
    d[3] = (float)t3;
    d[4] = (float)t4;
}

// bin distances and their reciprocals: rr[kk] = kz[kk] * rstep, ir[kk] = 1 / rr[kk]
// (entries past Kc: 1, with zero counts)
__global__ void p2_rtab_kernel(const int *__restrict__ kz, int Kc, int Kc_pad, double rstep,
                               double *__restrict__ rr, double *__restrict__ ir)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Kc_pad) return;
    const double r = i < Kc ? (double)kz[i] * rstep : 1.0;
    rr[i] = r;
    ir[i] = 1.0 / r;
}

// One wave: 16 q values x RB blocks of 16 rows over this split's bins.  Operand layout of
// v_mfma_f32_16x16x32_bf16 (lane l): A[row l%16][k 8(l/16)..+7], B[k 8(l/16)..+7][col l%16],
// C[row 4(l/16)+v][col l%16].  Each lane generates the 8 sinc values of its B fragment.
template <int RB>
__global__ __launch_bounds__(kMfWaves * 64) void p2_debye_mfma_kernel(
    const double *__restrict__ rr, const double *__restrict__ ir, int Kc_pad, int k_per_split,
    const __bf16 *__restrict__ cdig, int I, int rows_pad, const double *__restrict__ q, int nq,
    int nrows, double *__restrict__ Spart)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int q0 = (blockIdx.x * kMfWaves + wave) * 16;
    if (q0 >= nq) return;
    const int col = lane & 15, kq = lane >> 4;
    const double Q = q[min(q0 + col, nq - 1)];
    const double iQ = Q != 0.0 ? 1.0 / Q : 0.0;
    const int kb = blockIdx.y * k_per_split;
    const int ke = min(Kc_pad, kb + k_per_split);
    f32x4 acc[RB][kMfClasses];
    double acc64[RB][4];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
#pragma unroll
        for (int c = 0; c < kMfClasses; c++) acc[rb][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int v = 0; v < 4; v++) acc64[rb][v] = 0.0;
    }
    for (int k0 = kb; k0 < ke; k0 += 32) {
        const int kk0 = k0 + kq * 8;
        bf16x8 B[kMfJ];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            // sinc(Q r) = sin(Q r) (1/Q)(1/r): FP64 reduction of the argument, no division
            const double T = Q != 0.0 ? sin_cw(Q * rr[kk0 + t]) * iQ * ir[kk0 + t] : 1.0;
            float d[kMfJ];
            fixed_digits(T, d);
#pragma unroll
            for (int j = 0; j < kMfJ; j++) B[j][t] = (__bf16)d[j];
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
#pragma unroll
            for (int i = 0; i < kMfImax; i++) {
                if (i < I) {
                    const bf16x8 A = *reinterpret_cast<const bf16x8 *>(
                        cdig + ((int64_t)i * rows_pad + rb * 16 + col) * Kc_pad + kk0);
#pragma unroll
                    for (int j = 0; j < kMfJ; j++)
                        acc[rb][i + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[j], acc[rb][i + j], 0, 0, 0);
                }
            }
        }
        // fold the exact FP32 class sums into FP64 (every 32 bins: class sums stay < 2^24)
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
#pragma unroll
            for (int c = 0; c < kMfClasses; c++) {
                if (c < I + kMfJ - 1) {
                    const double scale = __builtin_ldexp(1.0, 8 * c - 40);
#pragma unroll
                    for (int v = 0; v < 4; v++) acc64[rb][v] += scale * (double)acc[rb][c][v];
                    acc[rb][c] = f32x4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
    }
    const int qq = q0 + col;
    if (qq < nq) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb * 16 + kq * 4 + v;
                if (row < nrows) Spart[((int64_t)blockIdx.y * nrows + row) * nq + qq] = acc64[rb][v];
            }
    }
}

// S[p][q] += self[p] + sum_k low[p][k] sinc(q r_k)
__global__ void p2_add_low_kernel(const double *__restrict__ self, const double *__restrict__ low,
                                  int nlow, int P, double rstep, const double *__restrict__ q, int nq,
                                  double *__restrict__ S)
{
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    if (iq >= nq) return;
    for (int p = 0; p < P; p++) {
        double s = self[p];
        for (int k = 0; k < nlow; k++) {
            const double w = low[p * nlow + k];
            if (w != 0.0) {
                const double r = (double)k * rstep;
                const double x = q[iq] * r;
                s += r == 0.0 ? w : w * sin(x) / x;
            }
        }
        S[(int64_t)p * nq + iq] += s;
    }
}

__global__ void p2_combine_kernel(const double *__restrict__ ff, const double *__restrict__ S, int P,
                                  int nq, double *__restrict__ I)
{
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    if (iq >= nq) return;
    double acc = 0.0;
    for (int p = 0; p < P; p++) acc += S[(int64_t)p * nq + iq] * ff[(int64_t)p * nq + iq];
    I[iq] = acc;
}

// largest k with double(k)*rstep <= 0.0005 (the reference's test r[k] <= 0.0005)
static int self_bin_max(double rstep)
{
    int k = (int)std::floor(0.0005 / rstep);
    while ((double)(k + 1) * rstep <= 0.0005) k++;
    while (k >= 0 && (double)k * rstep > 0.0005) k--;
    return k;
}

void debye_structure(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep,
                     int first_n, const double *d_q, int nq, double *d_S)
{
    DBX_REQUIRE(P > 0 && nq > 0 && n_bins > 0, DBX_ERR_ARG, "empty pass-2 problem");
    const int ksm = self_bin_max(rstep);
    const int nlow = ksm + 1;
    const int64_t nblk = (n_bins + kScanBlock - 1) / kScanBlock;
    // scratch: block counts, total + max count, kz, self, low, weights / digits
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off = (off + b + 255) & ~(size_t)255; return o; };
    const size_t o_bc = take(sizeof(int) * (nblk + 1));
    const size_t o_tot = take(sizeof(unsigned long long) * 2);
    const size_t o_kz = take(sizeof(int) * n_bins);
    const size_t o_self = take(sizeof(double) * P);
    const size_t o_low = take(sizeof(double) * P * nlow);
    char *base = (char *)scratch(ctx, kSlotPass2, off);
    int *bc = (int *)(base + o_bc);
    unsigned long long *tot = (unsigned long long *)(base + o_tot);  // [0] Kc, [1] max count
    int *kz = (int *)(base + o_kz);
    double *self = (double *)(base + o_self);
    double *low = (double *)(base + o_low);
    const unsigned long long *h = (const unsigned long long *)d_hist;

    DBX_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    DBX_HIP(hipMemsetAsync(tot, 0, sizeof(unsigned long long) * 2, ctx->stream));
    hipLaunchKernelGGL(p2_count_kernel, dim3((unsigned)nblk), dim3(kScanBlock), 0, ctx->stream, h, P,
                       n_bins, ksm, bc, tot + 1);
    hipLaunchKernelGGL(p2_scan_kernel, dim3(1), dim3(64), 0, ctx->stream, bc, (int)nblk, (int *)tot);
    hipLaunchKernelGGL(p2_scatter_kernel, dim3((unsigned)nblk), dim3(kScanBlock), 0, ctx->stream, h,
                       P, n_bins, ksm, bc, kz);
    hipLaunchKernelGGL(p2_self_kernel, dim3((P + 63) / 64), dim3(64), 0, ctx->stream, h, P, n_bins,
                       ksm, first_n, self, low);
    unsigned long long hv[2];
    DBX_HIP(hipMemcpyAsync(hv, tot, sizeof(hv), hipMemcpyDeviceToHost, ctx->stream));
    DBX_HIP(hipStreamSynchronize(ctx->stream));
    const int Kc = (int)(hv[0] & 0xFFFFFFFFull);
    const unsigned long long max_count = hv[1];
    ctx->last_p2_bins = Kc;
    if (Kc > 0 && ctx->pass2_mode == DBX_PASS2_FP64) {
        double *A = (double *)scratch(ctx, kSlotDigits, sizeof(double) * (size_t)P * Kc);
        hipLaunchKernelGGL(p2_weights_kernel, dim3((Kc + 255) / 256), dim3(256), 0, ctx->stream, h, P,
                           n_bins, kz, Kc, A);
        const int qblocks = (nq + kDsThreads - 1) / kDsThreads;
        int nsplit = std::max(1, std::min((2048 + qblocks - 1) / qblocks, (Kc + 511) / 512));
        const int kper = (Kc + nsplit - 1) / nsplit;
        nsplit = (Kc + kper - 1) / kper;
        double *Spart = (double *)scratch(ctx, kSlotPass2b, sizeof(double) * (size_t)nsplit *
                                                                std::min(P, kDsPmax) * nq);
        for (int p0 = 0; p0 < P; p0 += kDsPmax) {
            const int np = std::min(kDsPmax, P - p0);
            hipLaunchKernelGGL(p2_debye_fp64_kernel, dim3(qblocks, nsplit), dim3(kDsThreads), 0,
                               ctx->stream, kz, Kc, kper, A, p0, np, rstep, d_q, nq, Spart);
            hipLaunchKernelGGL(p2_reduce_splits_kernel, dim3((np * nq + 255) / 256), dim3(256), 0,
                               ctx->stream, Spart, nsplit, p0, np, nq, d_S);
        }
    } else if (Kc > 0) {
        // count digits: 2 N < 2^(8 I)
        int I = 1;
        while (I < kMfImax && (2.0 * (double)max_count) >= std::ldexp(1.0, 8 * I)) I++;
        DBX_REQUIRE(2.0 * (double)max_count < std::ldexp(1.0, 8 * kMfImax), DBX_ERR_RANGE,
                    "bin count too large for the split-BF16 Debye sum");
        const int Kc_pad = (Kc + 31) & ~31;
        const int qgroups = (nq + 16 * kMfWaves - 1) / (16 * kMfWaves);
        const int ksteps = Kc_pad / 32;
        int nsplit = std::max(1, std::min((4096 + qgroups - 1) / qgroups, ksteps));
        const int kper = ((ksteps + nsplit - 1) / nsplit) * 32;
        nsplit = (Kc_pad + kper - 1) / kper;
        double *rtab = (double *)scratch(ctx, kSlotRtab, sizeof(double) * 2 * (size_t)Kc_pad);
        hipLaunchKernelGGL(p2_rtab_kernel, dim3((Kc_pad + 255) / 256), dim3(256), 0, ctx->stream, kz, Kc,
                           Kc_pad, rstep, rtab, rtab + Kc_pad);
        for (int p0 = 0; p0 < P; p0 += 32) {
            const int np = std::min(32, P - p0);
            const int RB = (np + 15) / 16;
            const int rows_pad = RB * 16;
            __bf16 *cdig = (__bf16 *)scratch(ctx, kSlotDigits, sizeof(__bf16) * (size_t)I * rows_pad * Kc_pad);
            double *Spart = (double *)scratch(ctx, kSlotPass2b, sizeof(double) * (size_t)nsplit * np * nq);
            const int64_t nd = (int64_t)rows_pad * Kc_pad;
            hipLaunchKernelGGL(p2_count_digits_kernel, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0,
                               ctx->stream, h + (int64_t)p0 * n_bins, np, n_bins, kz, Kc, Kc_pad, rows_pad, I,
                               cdig);
            const dim3 grid((unsigned)qgroups, (unsigned)nsplit);
            if (RB == 1)
                hipLaunchKernelGGL(p2_debye_mfma_kernel<1>, grid, dim3(kMfWaves * 64), 0, ctx->stream, rtab,
                                   rtab + Kc_pad, Kc_pad, kper, cdig, I, rows_pad, d_q, nq, np, Spart);
            else
                hipLaunchKernelGGL(p2_debye_mfma_kernel<2>, grid, dim3(kMfWaves * 64), 0, ctx->stream, rtab,
                                   rtab + Kc_pad, Kc_pad, kper, cdig, I, rows_pad, d_q, nq, np, Spart);
            hipLaunchKernelGGL(p2_reduce_splits_kernel, dim3((np * nq + 255) / 256), dim3(256), 0,
                               ctx->stream, Spart, nsplit, p0, np, nq, d_S);
        }
    } else {
        DBX_HIP(hipMemsetAsync(d_S, 0, sizeof(double) * (size_t)P * nq, ctx->stream));
    }
    hipLaunchKernelGGL(p2_add_low_kernel, dim3((nq + 255) / 256), dim3(256), 0, ctx->stream, self, low,
                       nlow, P, rstep, d_q, nq, d_S);
    DBX_HIP(hipGetLastError());
    DBX_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    ctx->p2_timed = true;
}

void combine_intensity(dbx_ctx *ctx, const double *d_ff, const double *d_S, int P, int nq,
                       double *d_I)
{
    hipLaunchKernelGGL(p2_combine_kernel, dim3((nq + 255) / 256), dim3(256), 0, ctx->stream, d_ff,
                       d_S, P, nq, d_I);
    DBX_HIP(hipGetLastError());
}

}  // namespace dbx

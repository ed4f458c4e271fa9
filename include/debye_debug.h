/*
 * debye_debug.h - diagnostic entry points of libdebye.so (no counterpart in the reference).
 * They expose the device layout of a model and the pass-1 work plan so that tests can check
 * the plan's geometry on the host.
 */
#ifndef DEBYE_DEBUG_H
#define DEBYE_DEBUG_H

#include "debye.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tiles of a model: 8 x 32-bit words each (cx, cy, cz, rad as float; start, count, group,
 * parent as int32).  Returns the tile count; fills `out` when cap >= count. */
int dbx_debug_model_tiles(const dbx_model *m, void *out, int cap);
/* Atom positions in device (tile) order as float4 x, y, z, 0.  Returns the atom count;
 * copies when cap >= count. */
int64_t dbx_debug_model_positions(dbx_ctx *ctx, const dbx_model *m, float *out4, int64_t cap);
/* Pass-1 plan for (m, rstep): work items of 12 int32 each (pair, a_off, a_cnt, b0, b1, same,
 * shell, k_lo, e2_off, degraded, win, pad) in launch order, and the shell-edge table (float,
 * squared A^2).  win: 1 no window below, 2 none above, 4 a window of 16-bit counters.
 * Returns DBX_OK and the counts; copies when the capacities suffice. */
int dbx_debug_pass1_plan(const dbx_model *m, double rstep, void *jobs, int cap_jobs, float *edge2,
                         int cap_edge2, int *n_jobs, int *n_edge2);

/* Pair evaluations of the context's last pass 1 (kernel counter): for every (A tile, B tile)
 * candidate, its B atoms inside the job's shell times the A tile's atoms, whether the pair's
 * bin falls in the job's LDS window or not.  Against the pairs counted (N(N+1)/2 per
 * subgroup, N_a N_b per pair of subgroups) this is the work spent per counted pair: about one
 * where one window holds a shell, more where fine bins split shells over several windows. */
int dbx_debug_pass1_evaluations(const dbx_ctx *ctx, uint64_t *evaluated);

/* The same tiling and plan computed on the host from arrays (no device call): tiles (all
 * 64-atom blocks of every subgroup), the tile-ordered positions
 * (float4), work items, shell edges and the A-tile index list the work items point into,
 * each copied when its capacity suffices; the counts are always returned. */
int dbx_debug_plan_host(const float *x, const float *y, const float *z, const int64_t *group_n,
                        int n_groups, double rstep, void *tiles, int cap_tiles, float *pos4,
                        int64_t cap_pos, void *jobs, int cap_jobs, float *edge2, int cap_edge2,
                        int32_t *alist, int cap_alist, int *n_tiles, int *n_jobs, int *n_edge2,
                        int *n_alist);

/* Floats v in [lo, hi) (bit patterns, 0 <= v < 2^31) where the device's one-instruction
 * round-half-up conversion (v_cvt_rpi_i32_f32) differs from roundf's half-away-from-zero
 * rounding of the bin rule (cuDebye-v15.cu:104); n_bad[0] receives the count. */
int dbx_selftest_round(uint32_t lo, uint32_t hi, uint64_t *n_bad);

/* Measured LDS atomic throughput of the current device (ds_add_u32 on conflict-free addresses,
 * all CUs), the peak against which pass 1's one LDS increment per pair is reported. */
int dbx_selftest_lds_atomic_rate(double *atomics_per_s);

/* The pass-2 FP64 sine (dbx_selftest_sin's device function) evaluated on the host, and the
 * sine-cosine pair the sinc generator starts each run of consecutive bins from. */
int dbx_debug_sin_host(const double *x, int n, double *out);
int dbx_debug_sincos_host(const double *x, int n, double *s_out, double *c_out);

/* The six signed sinc digits d_0..d_5 (digits[6 i + j]) that the matrix-core Debye sum forms
 * for sinc(q[i] r[i]): sinc = 3/8 + sum_j d_j 256^j / (3 2^46) within 2^-48 plus the error of
 * the FP64 sine (the device code, built for the host; q = 0 is taken as 1e-30).  n must be a
 * multiple of 4. */
int dbx_debug_sinc_digits_host(const double *q, const double *r, int n, signed char *digits);

/* Profiling of the matrix-core Debye sum on ctx (off by default).  When on, each pass 2
 * records HIP events around the contraction kernel and reads back its per-row-block bin
 * extents to count the matrix instructions it issues.  dbx_debug_pass2_profile returns,
 * for the last pass 2: the contraction kernel's time (ms), the 8-bit integer matrix operations
 * it issued (16 x 16 x 64 x 2 per v_mfma_i32_16x16x64_i8), the algorithmic flops (2 nQ per nonzero bin
 * of every row: the contraction counted once, whatever the digit count), the compacted bin
 * count, the largest count-digit number and the mean over the row tiles (each tile takes the
 * digits its own rows need). */
int dbx_debug_set_profiling(dbx_ctx *ctx, int on);
int dbx_debug_pass2_profile(dbx_ctx *ctx, double *kernel_ms, double *mfma_flops, double *useful_flops,
                            int *n_bins_compacted, int *count_digits, double *mean_digits);

/* Models whose partials the last pass 2 on ctx recomputed with FP64 arithmetic because their
 * matrix-core error bound exceeded the context's target (dbx_set_rel_err_target). */
int dbx_debug_pass2_refined(const dbx_ctx *ctx, int *n_models);

#ifdef __cplusplus
}
#endif
#endif /* DEBYE_DEBUG_H */

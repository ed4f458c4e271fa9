/*
 * debye.h - C ABI of libdebye.so, an MI355X (gfx950) Debye scattering engine with the
 * method and file surface of cuDebye v1.5 (Martin-Rudolph/cuDebye, cuDebye-v15.cu).
 *
 * The reference is one CUDA executable whose in-process seams are four host functions
 * (splitCalcDist, reduceCheck, atomicProp, intensity) plus main()'s parser and writers.
 * Each entry point below names the reference code it replaces.
 *
 * Conventions
 *   - Every function returns an int status: DBX_OK (0) or a negative DBX_ERR_* code;
 *     dbx_last_error() gives the message of the context's last call (empty after a success).
 *     Nothing here calls exit() (the reference's gpuErrchk does, cuDebye-v15.cu:58-66).
 *   - Host buffers belong to the caller.  Device memory belongs to the context and is
 *     released by dbx_model_destroy / dbx_close (which also releases the context's models
 *     still alive: their handles are invalid afterwards).  dbx_model_destroy of a handle
 *     that is not a live model of the context returns DBX_ERR_ARG.
 *   - A context serialises its synchronous calls; distinct contexts are independent.
 *   - Subgroup pairs are numbered in the reference's order (cuDebye-v15.cu:941-947):
 *     (0,0),(0,1),...,(0,S-1),(1,1),...,(S-1,S-1); P = S(S+1)/2.
 *   - Histograms are uint64 per bin (the reference uses uint32, cuDebye-v15.cu:872).
 */
#ifndef DEBYE_H
#define DEBYE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DBX_OK = 0,
    DBX_ERR_ARG = -1,      /* invalid argument */
    DBX_ERR_HIP = -2,      /* HIP runtime failure */
    DBX_ERR_IO = -3,       /* file missing / unwritable */
    DBX_ERR_PARSE = -4,    /* malformed .deb, or atom count mismatch (cuDebye-v15.cu:830-833) */
    DBX_ERR_RANGE = -5,    /* a pair distance fell at or beyond the last bin (rmax too small) */
    DBX_ERR_NOMEM = -6,    /* device or host allocation failed */
    DBX_ERR_BUDGET = -7,   /* request larger than the context's HBM budget */
    DBX_ERR_STATE = -8     /* unknown ticket / wrong state */
};

typedef struct dbx_ctx dbx_ctx;
typedef struct dbx_model dbx_model;

/* Scattering parameters of one subgroup: the subgroup line of a .deb
 * "n type occ Beq a1..a5 b1..b5 c" (cuDebye-v15.cu:786, README.md "Fifth Line"). */
typedef struct {
    double occ;
    double beq;
    double a[5];
    double b[5];
    double c;
} dbx_group_params;

/* ---- context ------------------------------------------------------------------- */

/* Opens device `device` (the .deb header's first number, cuDebye-v15.cu:644-646).
 * hbm_budget: bytes of device memory the context may hold (0 = 90% of free memory). */
int dbx_open(int device, size_t hbm_budget, dbx_ctx **out);
int dbx_close(dbx_ctx *ctx);
/* The message of the context's last call, empty after a success.  The string is the calling
 * thread's copy, valid until that thread's next dbx_last_error call. */
const char *dbx_last_error(const dbx_ctx *ctx);
/* HIP stream the context launches on (hipStream_t as void*). */
void *dbx_stream(dbx_ctx *ctx);

/* ---- grids (host) ----------------------------------------------------------------- */

/* 2theta and K grids, cuDebye-v15.cu:690-698: nTT = int((max-min)/step)+1,
 * 2theta_i = min + i*step, K_i = 4*pi*sin(2theta_i/2*pi/180)/lambda.
 * Returns nTT (>0) and fills tt/q when both are non-NULL and cap >= nTT;
 * returns DBX_ERR_ARG for a non-positive step or nTT < 1. */
int dbx_q_from_2theta(double lambda, double tt_min, double tt_max, double tt_step, double *tt,
                      double *q, int cap);
/* channels = int(rmax/rstep)+2 (cuDebye-v15.cu:734); <0 on bad input. */
int64_t dbx_channels(double rmax, double rstep);

/* ---- models ----------------------------------------------------------------------- */

/* Uploads a model: FP32 positions in subgroup order (SoA, as cuDebye-v15.cu:809-827),
 * group_n[n_groups] atoms per subgroup.  Coordinates are copied bit for bit (only their
 * order changes inside the device layout).  Non-finite coordinates are rejected. */
int dbx_model_create(dbx_ctx *ctx, const float *x, const float *y, const float *z,
                     const int64_t *group_n, int n_groups, dbx_model **out);
int dbx_model_destroy(dbx_ctx *ctx, dbx_model *m);
int dbx_model_info(const dbx_model *m, int64_t *n_atoms, int *n_groups, int *n_pairs);

/* ---- pass 1: pair-distance histograms ----------------------------------------------
 * Replaces splitCalcDist + the dist/dist2/care_dist2 kernels (cuDebye-v15.cu:88-272)
 * for every subgroup pair.  Bin rule (bit-exact with the reference):
 *   dx,dy,dz FP32 differences; s = (dx*dx + dy*dy) + dz*dz, each operation rounded;
 *   k = (int)roundf(sqrtf(s) * (float)(1/rstep)), sqrtf correctly rounded.
 * Same subgroup: pairs col <= row including self pairs (bin 0 += n); cross: n_a*n_b.
 * out: uint64[P][n_bins] (host memory, or device memory when out_is_device != 0);
 * 0 < n_bins < 2^31 - 1, else DBX_ERR_ARG.
 * Returns DBX_ERR_RANGE (histogram still written, offending pairs not counted) if any
 * pair falls at k >= n_bins; *n_out_of_range (optional) receives their number.
 * The model keeps its pass-1 work plan for the last rstep (host planning runs once per
 * model and bin width). */
int dbx_histogram(dbx_ctx *ctx, dbx_model *m, double rstep, int64_t n_bins, uint64_t *out,
                  int out_is_device, uint64_t *n_out_of_range);

/* ---- prefactor ----------------------------------------------------------------------
 * ffoobb[q] = occ_a occ_b exp(nh (B_a+B_b)) f_a f_b with nh = -Q^2/(16 pi^2) and
 * f = c + sum_5 a_i exp(b_i nh), FP64 (replaces atomicProp/atomicScatter,
 * cuDebye-v15.cu:447-479). */
int dbx_prefactor(dbx_ctx *ctx, const dbx_group_params *ga, const dbx_group_params *gb,
                  const double *q, int nq, double *ffoobb);

/* ---- pass 2: Debye sum ----------------------------------------------------------------
 * Replaces reduceCheck's compaction + intensity/debyeRowSum (cuDebye-v15.cu:283-346,
 * 493-577) for all P partials at once:
 *   S_p(q) = sum_k w_k N_pk sinc(q r_k), r_k = k*rstep, with the reference weight rule:
 *   the first `first_n` nonzero bins (first_n = the .deb block size) with r_k <= 0.0005
 *   count N_pk once with sinc 1, every other nonzero bin counts 2 N_pk sin(q r_k)/(q r_k);
 *   I(q) = sum_p ffoobb_p(q) S_p(q).
 * hist: uint64[P][n_bins] (host or device memory per hist_is_device).  ffoobb: [P][nq]
 * host FP64.  I_out[nq] host FP64 (overwritten).  S_out ([P][nq], optional) receives S_p.
 * rel_err_out ([nq], optional) receives for every I(q) a bound of its relative error against
 * the exact (FP64-oracle) sum over the same histograms and prefactors, derived from the
 * arithmetic used (see dbx_set_rel_err_target); every I returned here is within that bound.
 * An I of exactly 0 from terms of exactly 0 has bound 0.  The matrix-core sines hold for
 * q r < 2^20: past that their I and bound are NaN, which an error target recomputes in FP64. */
int dbx_debye(dbx_ctx *ctx, const uint64_t *hist, int hist_is_device, int P, int64_t n_bins,
              double rstep, int first_n, const double *q, int nq, const double *ffoobb,
              double *I_out, double *S_out, double *rel_err_out);

/* dbx_debye_noalloc: dbx_debye on device memory that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned memory during the call
 * (tight serving loops, harnesses that forbid hipMalloc).  d_hist [P][n_bins], d_q [nq],
 * d_ffoobb [P][nq] and d_I [nq] (and d_S [P][nq], d_rel_err [nq], optional) are device
 * pointers; the results equal dbx_debye's on the same inputs.  Every buffer pass 2 uses comes from one fixed 64 MiB
 * device region per device, set up by dbx_open; a problem whose buffers exceed it returns
 * DBX_ERR_BUDGET (dbx_debye has no such limit).  The region is shared by every call on a
 * device and keeps the last call's contents: calls that overlap on one device (two contexts
 * or threads at once) race.  The call waits for its own work before returning. */
int dbx_debye_noalloc(dbx_ctx *ctx, const uint64_t *d_hist, int P, int64_t n_bins, double rstep, int first_n,
                      const double *d_q, int nq, const double *d_ffoobb, double *d_I, double *d_S,
                      double *d_rel_err);

/* Pass-2 arithmetic of a context (default DBX_PASS2_MFMA):
 *   DBX_PASS2_MFMA: matrix-core contraction on signed 8-bit digits with exact 32-bit integer
 *     accumulation; sinc is evaluated in FP64, shifted by 3/8 and rounded to a multiple of
 *     1 / (3 2^46) (|error| < 2^-48 per bin, below the FP64 sine's own error);
 *   DBX_PASS2_FP64: FP64 vector arithmetic throughout (the reference's debyeRowSum). */
enum { DBX_PASS2_MFMA = 0, DBX_PASS2_FP64 = 1 };
int dbx_set_pass2_mode(dbx_ctx *ctx, int mode);
/* Error target of a context's pass 2 (default 1e-6).  Every I(q) comes with a bound of its
 * relative error: (sum_p |ffoobb_p(q)| B_p + rounding of the final sum) / |I(q)|, where B_p is
 * the sum of the partial's weights 2 N_pk times the largest error per unit weight of the
 * arithmetic (matrix cores: the sinc digits' rounding to 1 / (3 2^47) plus the FP64 sine and
 * sums; FP64: the rounding of the sinc values and sums).  A model whose bound exceeds the target at any q
 * has its partials recomputed with FP64 arithmetic (target <= 0: never). */
int dbx_set_rel_err_target(dbx_ctx *ctx, double target);

/* ---- whole model ----------------------------------------------------------------------
 * One model end to end on the device: pass 1, compaction, prefactors, pass 2.
 * params[n_groups]; I_out[nq]; hist_out ([P][n_bins], optional, host) receives the
 * histograms.  Equivalent to the reference's subgroup-pair loop (cuDebye-v15.cu:941-975). */
int dbx_intensity(dbx_ctx *ctx, dbx_model *m, const dbx_group_params *params,
                  double rstep, int64_t n_bins, int first_n, const double *q, int nq,
                  double *I_out, uint64_t *hist_out, double *rel_err_out);

/* ---- batch of models ------------------------------------------------------------------
 * Many models end to end in one pass-1 launch (a merged work list over all models) and one
 * pass-2 contraction over the rows of all models, sharing the sinc table.  No reference
 * counterpart in-process: it replaces running `cuDebye.exe` once per model
 * (Matlab/multisim.m:42-58).  Every item's intensity equals dbx_intensity on that item alone
 * (histograms bit-exact; I within the pass-2 mode's error bound).
 * items[n_items]: model, its params[n_groups], n_bins and first_n (all items share rstep and
 * the q grid).  I_out[n_items][nq].  n_out_of_range[n_items] (optional) receives each
 * item's count of pairs at or beyond its last bin; DBX_ERR_RANGE when any is nonzero.
 * hist_out (optional): hist_out[i], when not NULL, receives item i's [P_i][n_bins_i]. */
typedef struct dbx_batch_item {
    dbx_model *model;
    const dbx_group_params *params;
    int64_t n_bins;
    int first_n;
} dbx_batch_item;
int dbx_batch_intensity(dbx_ctx *ctx, const dbx_batch_item *items, int n_items, double rstep,
                        const double *q, int nq, double *I_out, uint64_t *n_out_of_range,
                        uint64_t *const *hist_out, double *rel_err_out);

/* ---- serving ---------------------------------------------------------------------------
 * A stream of structure models of uneven size (refinement / multi-simulation style), served
 * by continuous batching.  No reference counterpart: it replaces the one-process-per-model
 * loop of Matlab/multisim.m:42-58 (cuDebye.exe per .deb).
 *   - dbx_submit copies a request (host arrays) and returns at once.  Host threads tile and
 *     plan it; it is admitted when its device memory fits the server's HBM budget (a request
 *     larger than the whole budget is refused with DBX_ERR_BUDGET).
 *   - Two worker lanes (each with its own stream) take the admitted requests in steps: one
 *     pass-1 launch over all requests of the step, then one pass-2 contraction per q grid.
 *     Requests that arrive while a step is on the GPU join the next step; the two lanes'
 *     steps overlap on the device.
 *   - Results equal dbx_intensity on each request alone (histograms bit-exact, I within the
 *     pass-2 bound).  A request whose distances exceed its bins fails alone (DBX_ERR_RANGE).
 *   - With DBX_REQ_KEEP the histograms and the per-pair structure sums S_p(q) stay resident
 *     after completion: dbx_reweight recomputes I for new form-factor parameters by replaying
 *     a captured hipGraph (prefactor, combine and their copies); dbx_release frees them.
 *     Without it the request's device memory is freed at completion.
 * Calls are thread-safe. */
typedef struct dbx_server dbx_server;
typedef struct dbx_server_config {
    size_t hbm_budget;      /* device bytes for resident requests (0: 3/4 of free memory) */
    int max_batch;          /* requests per step (0: 256) */
    double max_step_pairs;  /* pair distances per step (0: 2e10) */
    int prep_threads;       /* host threads for tiling and planning (0: the host's hardware
                               threads less 4, between 2 and 12) */
    int pass2_mode;         /* DBX_PASS2_MFMA (default) or DBX_PASS2_FP64 */
    double batch_window_ms; /* a lane that finds admitted requests waits up to this long for
                               requests still being prepared or admitted, so that a burst of
                               submissions forms one step (0: start at once) */
    double rel_err_target;  /* as dbx_set_rel_err_target (0: 1e-6; negative: never recompute) */
} dbx_server_config;
typedef struct dbx_request {
    const float *x, *y, *z;           /* positions in subgroup order (copied) */
    const int64_t *group_n;           /* [n_groups] */
    int n_groups;
    const dbx_group_params *params;   /* [n_groups] (copied) */
    double rstep;
    int64_t n_bins;
    int first_n;
    const double *q;                  /* [nq] (copied) */
    int nq;
    int flags;                        /* DBX_REQ_KEEP */
} dbx_request;
enum { DBX_REQ_KEEP = 1 };
enum { DBX_REQ_QUEUED = 0, DBX_REQ_RUNNING = 1, DBX_REQ_DONE = 2, DBX_REQ_FAILED = 3 };
typedef struct dbx_server_stats {
    uint64_t submitted, completed, failed, steps;
    double pairs;                 /* pair distances histogrammed */
    uint64_t resident_bytes, budget_bytes;
    double max_step_requests;     /* most requests in one step */
    double max_step_qgrids;       /* most distinct q grids in one step */
    double step_enqueue_ms;       /* host time the lanes spent issuing their steps' work (sum) */
    double step_wait_ms;          /* host time the lanes waited for their steps to end (sum) */
} dbx_server_stats;
int dbx_server_open(int device, const dbx_server_config *cfg, dbx_server **out);
int dbx_server_close(dbx_server *s);
/* The message of the server's last failed call or request (the calling thread's copy, valid
 * until that thread's next dbx_server_last_error call). */
const char *dbx_server_last_error(dbx_server *s);
int dbx_submit(dbx_server *s, const dbx_request *req, uint64_t *ticket);
/* state: DBX_REQ_QUEUED (waiting for preparation or memory), _RUNNING, _DONE, _FAILED */
int dbx_poll(dbx_server *s, uint64_t ticket, int *state);
/* Blocks until the request ends; I_out[nq] (optional) receives I(q), latency_ms (optional)
 * the time from submission to completion, rel_err_out[nq] (optional) the relative error bound
 * of each I(q) (as dbx_debye's).  Returns the request's status. */
int dbx_wait(dbx_server *s, uint64_t ticket, double *I_out, double *latency_ms, double *rel_err_out);
/* I_out[nq] = intensity of a finished DBX_REQ_KEEP request for new params[n_groups];
 * rel_err_out[nq] (optional): its relative error bound. */
int dbx_reweight(dbx_server *s, uint64_t ticket, const dbx_group_params *params, double *I_out,
                 double *rel_err_out);
/* hist_out[P][n_bins] of a finished DBX_REQ_KEEP request (host). */
int dbx_request_histogram(dbx_server *s, uint64_t ticket, uint64_t *hist_out);
/* Frees a finished request (and its resident memory). */
int dbx_release(dbx_server *s, uint64_t ticket);
int dbx_server_stats_get(dbx_server *s, dbx_server_stats *out);

/* ---- .deb driver ------------------------------------------------------------------------
 * The reference program: `cuDebye.exe file.deb` (cuDebye-v15.cu:599-1003).  Parses the
 * .deb, runs every subgroup pair, writes <path>I (hist 0 or 2) and <path>x<r>x<c>
 * (hist 1 or 2) with the reference's formats.  flags: DBX_RUN_CRLF writes "\r\n" line
 * ends like the reference's Windows build; DBX_RUN_QUIET suppresses the log on stdout;
 * DBX_RUN_MFMA evaluates pass 2 on the matrix cores (default: FP64, as the reference). */
enum { DBX_RUN_CRLF = 1, DBX_RUN_QUIET = 2, DBX_RUN_MFMA = 4, DBX_RUN_FROM_DEBX = 8 };
int dbx_run_deb(const char *path, int flags);
/* DBX_RUN_FROM_DEBX: pass 2 only, from the <path>x<r>x<c> histograms a previous run wrote
 * (read with dbx_read_debx), for new header parameters; the .debx files are left as they are.
 *
 * dbx_read_debx: one .debx file (lines "%.7g\t%llu" of r_k and N_k, cuDebye-v15.cu:417-423)
 * into hist[n_bins] (zeroed first), k = rint(r / rstep); DBX_ERR_PARSE for a line whose r is
 * not a bin centre to seven digits, DBX_ERR_RANGE for a bin >= n_bins.  n_nonzero (optional)
 * receives the number of nonzero bins. */
int dbx_read_debx(const char *path, double rstep, int64_t n_bins, uint64_t *hist, int64_t *n_nonzero);
/* Folder run (Matlab/multisim.m:17-58): every .xyz in `dir` without its .deb is converted
 * first (dbx_xyz_to_deb with the default header and table), then every .deb without a .debI,
 * in name order, runs through one server (dbx_server_open); each file's outputs are those of
 * dbx_run_deb on it alone.  flags as for dbx_run_deb (DBX_RUN_FROM_DEBX excepted). */
int dbx_run_folder(const char *dir, int flags);

/* ---- host tools around the driver -------------------------------------------------------
 * .deb header values of the converter (the defaults of Matlab/xyz2deb.m:13-25: device 0,
 * block 16, grid 2048, wavelength 1.5418, 2theta 10 to 140 step 0.02, rstep 0.0001, hist 0). */
typedef struct dbx_deb_header {
    int device, block, grid;
    double wavelength, tt_min, tt_max, tt_step, rstep;
    int hist;
} dbx_deb_header;
void dbx_deb_header_defaults(dbx_deb_header *h);
/* xyz -> deb (Matlab/xyz2deb.m:169-266): lines "element x y z [occupancy [Beq]]" (the first line
 * fixes the column count; occupancy 1 and Beq 0 where absent); subgroups are the runs of equal
 * (element, occupancy, Beq) after a stable sort by element name, occupancy and Beq; coordinates
 * are centred on the middle of each axis' range; Rmax = 2 ceil(1.025 max r); scattering-factor
 * coefficients from table_path (rows "name a1..a5 c b1..b5" after '#' comments and one column
 * header; NULL: data/atomic_scattering_factors.txt beside libdebye.so, the Waasmaier-Kirfel
 * table the reference ships).  deb_path NULL: xyz_path with its last three characters replaced
 * by "deb".  hdr NULL: dbx_deb_header_defaults.  DBX_ERR_PARSE for an element not in the
 * table (the reference writes a .deb without its coefficients). */
int dbx_xyz_to_deb(const char *xyz_path, const char *deb_path, const char *table_path, const dbx_deb_header *hdr);
/* Average of .debI patterns (Matlab/plotdebi.m:13-63): the common 2theta range [max of the
 * files' first values (at least 0), min of their last values (at most 180)] at `step`
 * (0.02 in the reference); each raw pattern through its not-a-knot cubic spline, summed, then
 * divided by the file count (DBX_AVG_NONE), by its trapezoidal integral (DBX_AVG_INTEGRAL)
 * or by its maximum (DBX_AVG_MAX).  n_out receives the point count; tt_out / I_out NULL: size
 * query only, else capacity must hold n_out points. */
enum { DBX_AVG_NONE = 0, DBX_AVG_INTEGRAL = 1, DBX_AVG_MAX = 2 };
int dbx_average_debi(const char *const *paths, int n_files, int norm, double step, double *tt_out, double *I_out,
                     int capacity, int *n_out);

/* Structure view of a .deb file (Matlab/viewdeb.m:1-194): the atoms with lo < coordinate < hi on
 * every axis (limits[6] = xlo xhi ylo yhi zlo zhi; NULL: -10..10 on every axis, as the
 * reference) are drawn into svg_path as the reference's 3-D scatter (Matlab's default view,
 * discs of diameter 2 r 5 points in the element's colour, r and colour from radii_path: rows
 * "symbol radius red green blue"; NULL: data/atomic_radii.txt beside libdebye.so) and/or
 * listed in xyz_path ("symbol x y z" lines after a count and a title line).  Elements are the
 * subgroups' type fields without charge and label suffixes ("Al3+" -> "Al").  n_shown
 * (optional) receives the number of atoms kept.  Either output path may be NULL, not both. */
int dbx_view_deb(const char *deb_path, const double *limits, const char *svg_path, const char *xyz_path,
                 const char *radii_path, int64_t *n_shown);

/* ---- diagnostics --------------------------------------------------------------------
 * Self tests of the device arithmetic behind the pass-1 bin rule (no reference
 * counterpart: the reference relies on CUDA's sqrtf and roundf).
 * dbx_selftest_sqrt: for every float bit pattern in [lo, hi) with 2^-96 <= s <= FLT_MAX,
 *   n_bad[0] counts device square roots that are not correctly rounded, n_bad[1] those
 *   that differ from the compiler's correctly rounded sqrtf; n_bad[2] and n_bad[3] count the
 *   inputs where the bare hardware square root (before the correction) lies one float below /
 *   above the correctly rounded result.  n_bad has four entries.
 * dbx_selftest_bins: out[i] = device bin of the pair (a[3i..], b[3i..]) with rp. */
int dbx_selftest_sqrt(uint32_t lo, uint32_t hi, uint64_t *n_bad);
int dbx_selftest_bins(const float *a, const float *b, int n, float rp, int32_t *out);
/* dbx_selftest_sin: out[i] = the device FP64 sine used by the pass-2 sinc table, |x| < 2^20. */
int dbx_selftest_sin(const double *x, int n, double *out);
/* Statistics of the last pass-1 run on ctx: work items launched, pairs whose bin fell
 * outside their workgroup's LDS window (added to HBM directly) and exact-window work items
 * that met an out-of-window bin and were re-run by the checked kernel (performance figures). */
int dbx_last_pass1_stats(const dbx_ctx *ctx, int64_t *n_jobs, uint64_t *window_misses, uint64_t *retried);
/* Device times (HIP events on the context's stream) of the last pass-1 histogram kernel
 * and of the last pass-2 structure computation (compaction through the Debye sum), ms.  When
 * some models were recomputed with FP64 arithmetic to meet the error target, the pass-2 time
 * runs from the start of the first pass 2 to the end of the recomputation (read-back included). */
int dbx_last_kernel_times(dbx_ctx *ctx, double *pass1_ms, double *pass2_ms);

#ifdef __cplusplus
}
#endif
#endif /* DEBYE_H */

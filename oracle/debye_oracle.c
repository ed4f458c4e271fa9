/*
 * CPU oracle for the Debye scattering hot path (pass 1: pair-distance histogram,
 * pass 2: Debye sum) of cuDebye v1.5.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and nothing in the product (cudebye_amd/, include/) links, loads or calls it.
 *
 * Every function restates the reference's algorithm from its source text
 * (cuDebye-v15.cu, read as text; the CUDA program is not built or run here):
 *
 *   orc_q_from_2theta   cuDebye-v15.cu:690-698   2theta/K grid
 *   orc_channels, orc_rp cuDebye-v15.cu:732-740  bin grid, rp = float(1/rstep)
 *   orc_pair_bin        cuDebye-v15.cu:101-115   FP32 bin rule of dist2/care_dist2/dist
 *   orc_histogram       cuDebye-v15.cu:88-146,152-272  pair set (col <= row incl. self for
 *                                                 one subgroup, full rectangle otherwise)
 *   orc_prefactor       cuDebye-v15.cu:447-467   ffoobb[Q] in FP64
 *   orc_debye_sum       cuDebye-v15.cu:493-536   weight rule and Debye row sum, FP64
 *   orc_parse_deb_*     cuDebye-v15.cu:624-833   .deb header, subgroup lines, positions
 *   orc_write_debx      cuDebye-v15.cu:413-430   "%.7g\t%u" per nonzero bin
 *   orc_write_debi      cuDebye-v15.cu:989-998   "%g\t%g" per 2theta
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *   - orc_histogram / orc_pair_bin / parser: replay of the six
 *     Data/GammaAl2O3_10nm.debx* histograms (exact counts per bin, all 6.95e9
 *     pairs), closed-form totals n(n+1)/2 and n_a*n_b, bin 0 == n, and exact
 *     rational brute force on tiny models away from rounding ties.
 *   - orc_q_from_2theta, orc_prefactor, orc_debye_sum: the .debx fixtures plus
 *     the Gamma header reproduce Data/GammaAl2O3_10nm.debI at print precision;
 *     single-pair and two-atom closed forms; I(Q->0) -> (sum occ*f)^2;
 *     mpmath brute-force Debye double sums on small models.
 *   - writers: byte comparison with the fixture files (line ends aside).
 *
 * The FP32 sum of squares of the reference kernel (cuDebye-v15.cu:104,
 * "dx*dx + dy*dy + dz*dz", contracted by nvcc by default) is selectable:
 *   form 0:  (dx*dx + dy*dy) + dz*dz, every operation rounded
 *   form 1:  fmaf(dz, dz, fmaf(dx, dx, dy*dy))
 *   form 2:  fmaf(dz, dz, fmaf(dy, dy, dx*dx))
 * The Gamma fixture replay decides which one the reference used.
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (see oracle/Makefile).  Contraction
 * is off so every fused multiply-add is the explicit fmaf written here.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <omp.h>

#define ORC_PI 3.14159265358979323846 /* cuDebye-v15.cu:46 */

/* ------------------------------------------------------------------------ */
/* Grids                                                                      */
/* ------------------------------------------------------------------------ */

/* cuDebye-v15.cu:690-698.  Returns nTT; fills tt/K when non-NULL. */
int orc_q_from_2theta(double lambda, double tt_min, double tt_max, double tt_step,
                      double *tt, double *K)
{
    int ntt = (int)((tt_max - tt_min) / tt_step) + 1;
    if (tt && K) {
        for (int i = 0; i < ntt; i++) {
            tt[i] = tt_min + (double)i * tt_step;
            K[i] = 4 * ORC_PI * sin(tt[i] / 2 * ORC_PI / 180) / lambda;
        }
    }
    return ntt;
}

/* cuDebye-v15.cu:734 */
int64_t orc_channels(double rmax, double rstep) { return (int64_t)(int)(rmax / rstep) + 2; }

/* cuDebye-v15.cu:739 */
float orc_rp(double rstep) { return (float)(1 / rstep); }

/* ------------------------------------------------------------------------ */
/* Pass 1                                                                     */
/* ------------------------------------------------------------------------ */

static inline float orc_sumsq(float dx, float dy, float dz, int form)
{
    if (form == 1) return fmaf(dz, dz, fmaf(dx, dx, dy * dy));
    if (form == 2) return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float a = dx * dx;
    float b = dy * dy;
    float c = dz * dz;
    float ab = a + b;
    return ab + c;
}

/* round-half-away-from-zero of a non-negative float below 2^31, written as
 * trunc + (fraction >= 0.5); identical to roundf() on that domain because the
 * fraction v - trunc(v) is exact there. */
static inline int64_t orc_round_pos(float v)
{
    int64_t k = (int64_t)v;
    float f = v - (float)k;
    return k + (f >= 0.5f ? 1 : 0);
}

/* cuDebye-v15.cu:101-115: dx = x[col]-x[row] (sign irrelevant), r = sqrt(sumsq),
 * bin = int(round(r * rp)).  Returns -1 for a non-finite or negative value. */
int64_t orc_pair_bin(float xa, float ya, float za, float xb, float yb, float zb,
                     float rp, int form)
{
    float dx = xb - xa, dy = yb - ya, dz = zb - za;
    float s = orc_sumsq(dx, dy, dz, form);
    float v = sqrtf(s) * rp;
    if (!(v >= 0.0f) || !(v < 2147483520.0f)) return -1;
    return orc_round_pos(v);
}

/*
 * Histogram of one subgroup pair, cuDebye-v15.cu:88-146 with the tiling of
 * :152-272 removed (tiling does not change the pair set).
 *   rows:  atoms [row0, row0+nrow)   (CUDA Y, subgroup "row", :944-947)
 *   cols:  atoms [col0, col0+ncol)   (CUDA X, subgroup "col")
 *   same != 0: one subgroup (row0 == col0, nrow == ncol); pairs with
 *              col <= row including col == row (:137), i.e. n(n+1)/2 pairs.
 * hist[n_bins] is ADDED to (uint64).  Returns the number of pairs whose bin
 * was >= n_bins or invalid (they are not counted); the reference writes such
 * pairs out of bounds (README.md:102), the oracle reports them.
 */
int64_t orc_histogram(const float *x, const float *y, const float *z,
                      int64_t row0, int64_t nrow, int64_t col0, int64_t ncol,
                      int same, float rp, int form, uint64_t *hist, int64_t n_bins,
                      int nthreads)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t overflow = 0;
    const int64_t CH = 256;
#pragma omp parallel num_threads(nthreads) reduction(+ : overflow)
    {
        uint64_t *h = (uint64_t *)calloc((size_t)n_bins, sizeof(uint64_t));
        int64_t kb[256];
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < nrow; i++) {
            const int64_t ri = row0 + i;
            const float xa = x[ri], ya = y[ri], za = z[ri];
            const int64_t jn = same ? (i + 1) : ncol; /* col <= row incl. self */
            for (int64_t j0 = 0; j0 < jn; j0 += CH) {
                int64_t m = jn - j0 < CH ? jn - j0 : CH;
                const float *xc = x + col0 + j0, *yc = y + col0 + j0, *zc = z + col0 + j0;
                if (form == 1) {
                    for (int64_t t = 0; t < m; t++) {
                        float dx = xc[t] - xa, dy = yc[t] - ya, dz = zc[t] - za;
                        float s = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
                        float v = fminf(sqrtf(s) * rp, 2.0e9f);
                        int32_t k = (int32_t)v;
                        float f = v - (float)k;
                        kb[t] = (int64_t)k + (f >= 0.5f);
                    }
                } else if (form == 2) {
                    for (int64_t t = 0; t < m; t++) {
                        float dx = xc[t] - xa, dy = yc[t] - ya, dz = zc[t] - za;
                        float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        float v = fminf(sqrtf(s) * rp, 2.0e9f);
                        int32_t k = (int32_t)v;
                        float f = v - (float)k;
                        kb[t] = (int64_t)k + (f >= 0.5f);
                    }
                } else {
                    for (int64_t t = 0; t < m; t++) {
                        float dx = xc[t] - xa, dy = yc[t] - ya, dz = zc[t] - za;
                        float s = (dx * dx + dy * dy) + dz * dz;
                        float v = fminf(sqrtf(s) * rp, 2.0e9f);
                        int32_t k = (int32_t)v;
                        float f = v - (float)k;
                        kb[t] = (int64_t)k + (f >= 0.5f);
                    }
                }
                for (int64_t t = 0; t < m; t++) {
                    int64_t k = kb[t];
                    if (k >= 0 && k < n_bins)
                        h[k]++;
                    else
                        overflow++;
                }
            }
        }
#pragma omp critical
        {
            for (int64_t k = 0; k < n_bins; k++) hist[k] += h[k];
        }
        free(h);
    }
    return overflow;
}

/* ------------------------------------------------------------------------ */
/* Pass 2                                                                     */
/* ------------------------------------------------------------------------ */

/* cuDebye-v15.cu:447-467, one subgroup pair (ta, tb), FP64.
 * a_* and b_* point at the 5 coefficients of each subgroup. */
void orc_prefactor(const double *K, int nq, double occ_a, double occ_b, double beq_a,
                   double beq_b, const double *aa, const double *ba, double ca,
                   const double *ab, const double *bb, double cb, double *ff)
{
    const double rp16pi2 = -0.006332573977646; /* :452 */
    for (int q = 0; q < nq; q++) {
        double nh = K[q] * K[q] * rp16pi2;
        double v = occ_a * occ_b;
        v = v * exp(nh * (beq_a + beq_b));
        double f1 = ca, f2 = cb;
        for (int i = 0; i < 5; i++) {
            f1 += aa[i] * exp(ba[i] * nh);
            f2 += ab[i] * exp(bb[i] * nh);
        }
        ff[q] = v * f1 * f2;
    }
}

/*
 * Debye row sum of one compacted partial histogram, cuDebye-v15.cu:493-536.
 *   bins[i], cnt[i]: nonzero bins in ascending order (:315-333), r_i = double(bin)*rstep (:736)
 *   first_n: blockDim.x of the reference launch (the header's block size); entries with
 *            i < first_n and r_i <= 0.0005 contribute cnt_i (w=1, sinc=1), all others
 *            2*cnt_i*sin(K r_i)/(K r_i) (:508-522).
 *   I[q] += ff[q] * S(q)   (:534-536)
 * S(q) is summed with Neumaier compensation (the FP64 truth the GPU path is measured
 * against).  If absmag != NULL, absmag[q] += |ff[q]| * sum_i |term_i| (conditioning).
 * If s_out != NULL, s_out[q] = S(q) (structure function before weighting).
 */
void orc_debye_sum(const int64_t *bins, const double *cnt, int64_t nnz, double rstep,
                   const double *K, int nq, const double *ff, double *I, double *absmag,
                   double *s_out, int first_n, int nthreads)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int q = 0; q < nq; q++) {
        const double Kq = K[q];
        double sum = 0.0, comp = 0.0, mag = 0.0;
        for (int64_t i = 0; i < nnz; i++) {
            double r = (double)bins[i] * rstep;
            double term;
            if (i < first_n && !(r > 0.0005))
                term = cnt[i];
            else if (r == 0.0)
                term = 2 * cnt[i]; /* sinc(0) = 1; the reference would divide 0/0 here */
            else
                term = 2 * cnt[i] * sin(Kq * r) / (Kq * r);
            double t = sum + term;
            if (fabs(sum) >= fabs(term))
                comp += (sum - t) + term;
            else
                comp += (term - t) + sum;
            sum = t;
            mag += fabs(term);
        }
        double S = sum + comp;
        if (s_out) s_out[q] = S;
        I[q] += S * ff[q];
        if (absmag) absmag[q] += fabs(ff[q]) * mag;
    }
}

/* ------------------------------------------------------------------------ */
/* .deb parser, cuDebye-v15.cu:624-833 (header lines 1-4, S subgroup lines,    */
/* then one "x y z" line per atom; blank lines between positions skipped)     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int device, block, grid, hist, n_groups;
    int64_t n_atoms;
    double lambda, tt_min, tt_max, tt_step, rmax, rstep;
} orc_deb_header;

static char *orc_slurp(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) { fclose(f); return NULL; }
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = 0;
    *len = got;
    return buf;
}

/* next line (strips "\r\n" / "\n"); returns NULL at end */
static char *orc_next_line(char **cur)
{
    char *s = *cur;
    if (!s || !*s) return NULL;
    char *e = strchr(s, '\n');
    if (e) { *e = 0; *cur = e + 1; } else { *cur = s + strlen(s); }
    size_t n = strlen(s);
    if (n && s[n - 1] == '\r') s[n - 1] = 0;
    return s;
}

static int orc_blank(const char *s)
{
    for (; *s; s++)
        if (*s != ' ' && *s != '\t' && *s != '\r' && *s != '\v' && *s != '\f') return 0;
    return 1;
}

/* Parses header, subgroup table and positions.  Arrays may be NULL to query
 * the header only.  Returns 0, or a negative code:
 *  -1 missing file, -2 malformed header, -3 malformed subgroup line,
 *  -4 atom count mismatch / malformed position line. */
int orc_parse_deb(const char *path, orc_deb_header *h, int64_t *group_n, double *occ,
                  double *beq, double *a, double *b, double *c, float *x, float *y,
                  float *z)
{
    size_t len;
    char *buf = orc_slurp(path, &len);
    if (!buf) return -1;
    char *cur = buf, *ln;
    int rc = 0;
    memset(h, 0, sizeof(*h));
    if (!(ln = orc_next_line(&cur)) || sscanf(ln, "%d %d %d", &h->device, &h->block, &h->grid) != 3) { rc = -2; goto out; }
    if (!(ln = orc_next_line(&cur)) || sscanf(ln, "%lf %lf %lf %lf", &h->lambda, &h->tt_min, &h->tt_max, &h->tt_step) != 4) { rc = -2; goto out; }
    if (!(ln = orc_next_line(&cur)) || sscanf(ln, "%lf %lf %d", &h->rmax, &h->rstep, &h->hist) != 3) { rc = -2; goto out; }
    if (h->hist != 1 && h->hist != 2) h->hist = 0; /* :751-753 */
    if (!(ln = orc_next_line(&cur)) || sscanf(ln, "%d", &h->n_groups) != 1 || h->n_groups <= 0) { rc = -2; goto out; }
    h->n_atoms = 0;
    for (int g = 0; g < h->n_groups; g++) {
        long long n;
        char type[64];
        double v[13];
        if (!(ln = orc_next_line(&cur))) { rc = -3; goto out; }
        int got = sscanf(ln, "%lld %63s %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &n, type,
                         &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10],
                         &v[11], &v[12]);
        if (got != 15 || n < 0) { rc = -3; goto out; }
        if (group_n) group_n[g] = n;
        if (occ) occ[g] = v[0];
        if (beq) beq[g] = v[1];
        if (a) for (int i = 0; i < 5; i++) a[g * 5 + i] = v[2 + i];
        if (b) for (int i = 0; i < 5; i++) b[g * 5 + i] = v[7 + i];
        if (c) c[g] = v[12];
        h->n_atoms += n;
    }
    if (x) {
        int64_t i = 0;
        while ((ln = orc_next_line(&cur))) {
            if (orc_blank(ln)) continue;
            if (i >= h->n_atoms) { rc = -4; goto out; }
            char *p = ln, *e;
            float v3[3];
            for (int t = 0; t < 3; t++) {
                errno = 0;
                v3[t] = strtof(p, &e);
                if (e == p) { rc = -4; goto out; }
                p = e;
            }
            x[i] = v3[0]; y[i] = v3[1]; z[i] = v3[2];
            i++;
        }
        if (i != h->n_atoms) rc = -4;
    }
out:
    free(buf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Writers                                                                    */
/* ------------------------------------------------------------------------ */

/* cuDebye-v15.cu:417-423: "<r with 7 significant digits>\t<count>" per nonzero bin */
int orc_write_debx(const char *path, const int64_t *bins, const uint64_t *cnt, int64_t nnz,
                   double rstep, int crlf)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const char *eol = crlf ? "\r\n" : "\n";
    for (int64_t i = 0; i < nnz; i++)
        fprintf(f, "%.7g\t%llu%s", (double)bins[i] * rstep, (unsigned long long)cnt[i], eol);
    fclose(f);
    return 0;
}

/* cuDebye-v15.cu:993-997: "<2theta>\t<I>" with 6 significant digits each */
int orc_write_debi(const char *path, const double *tt, const double *I, int n, int crlf)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const char *eol = crlf ? "\r\n" : "\n";
    for (int i = 0; i < n; i++) fprintf(f, "%g\t%g%s", tt[i], I[i], eol);
    fclose(f);
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

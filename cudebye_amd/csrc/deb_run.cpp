// The .deb driver: the reference program `cuDebye.exe file.deb` (cuDebye-v15.cu:599-1003)
// on top of the C ABI.  Parses the .deb (cuDebye-v15.cu:624-833), runs every subgroup pair
// on the device, prints the reference's sanity checks (cuDebye-v15.cu:360-410) and writes
// <path>x<r>x<c> (cuDebye-v15.cu:413-430) and <path>I (cuDebye-v15.cu:973-998).  Also the
// reader of those .debx files (pass 2 again from written histograms) and the folder run of
// Matlab/multisim.m on the server.
#include <hip/hip_runtime.h>

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "debfile.h"
#include "internal.h"

namespace dbx {

namespace {

bool blank(const char *s)
{
    for (; *s; s++)
        if (!std::strchr(" \t\r\v\f", *s)) return false;
    return true;
}

// line reader over a whole file; strips "\n" and a trailing "\r"
struct Lines {
    std::string buf;
    size_t pos = 0;
    bool next(std::string &out)
    {
        if (pos >= buf.size()) return false;
        size_t e = buf.find('\n', pos);
        if (e == std::string::npos) e = buf.size();
        out.assign(buf, pos, e - pos);
        if (!out.empty() && out.back() == '\r') out.pop_back();
        pos = e + 1;
        return true;
    }
};

}  // namespace

int parse_deb(const char *path, DebFile &d, std::string &err)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        err = std::string(">>") + path + "<< does not exist";  // cuDebye-v15.cu:620
        return DBX_ERR_IO;
    }
    Lines L;
    char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) L.buf.append(tmp, got);
    std::fclose(f);
    std::string ln;
    // header lines 1-4 (cuDebye-v15.cu:644, 689, 730, 761)
    if (!L.next(ln) || std::sscanf(ln.c_str(), "%d %d %d", &d.device, &d.block, &d.grid) != 3) {
        err = "header line 1 must be 'device blockDimXY gridDimXY'";
        return DBX_ERR_PARSE;
    }
    if (!L.next(ln) ||
        std::sscanf(ln.c_str(), "%lf %lf %lf %lf", &d.lambda, &d.tt_min, &d.tt_max, &d.tt_step) != 4) {
        err = "header line 2 must be 'lambda 2theta_min 2theta_max 2theta_step'";
        return DBX_ERR_PARSE;
    }
    if (!L.next(ln) || std::sscanf(ln.c_str(), "%lf %lf %d", &d.rmax, &d.rstep, &d.hist) != 3) {
        err = "header line 3 must be 'rmax rstep hist'";
        return DBX_ERR_PARSE;
    }
    if (d.hist != 1 && d.hist != 2) d.hist = 0;  // cuDebye-v15.cu:751-753
    if (!L.next(ln) || std::sscanf(ln.c_str(), "%d", &d.S) != 1 || d.S < 1 || d.S > 64) {
        err = "header line 4 must hold the number of subgroups (1..64)";
        return DBX_ERR_PARSE;
    }
    int64_t N = 0;
    for (int g = 0; g < d.S; g++) {
        if (!L.next(ln)) {
            err = "missing subgroup line " + std::to_string(g);
            return DBX_ERR_PARSE;
        }
        long long n;
        char type[64];
        dbx_group_params p;
        int k = std::sscanf(ln.c_str(),
                            "%lld %63s %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &n, type,
                            &p.occ, &p.beq, &p.a[0], &p.a[1], &p.a[2], &p.a[3], &p.a[4], &p.b[0],
                            &p.b[1], &p.b[2], &p.b[3], &p.b[4], &p.c);
        if (k != 15 || n < 0) {
            err = "malformed subgroup line " + std::to_string(g) +
                  " (want 'n type occ Beq a1..a5 b1..b5 c')";
            return DBX_ERR_PARSE;
        }
        d.n.push_back(n);
        d.type.push_back(type);
        d.gp.push_back(p);
        N += n;
    }
    d.x.reserve(N);
    d.y.reserve(N);
    d.z.reserve(N);
    int64_t i = 0;
    while (L.next(ln)) {
        if (blank(ln.c_str())) continue;  // cuDebye-v15.cu:821
        if (i >= N) {
            err = "more atom positions than the subgroups declare";  // cuDebye-v15.cu:830-833
            return DBX_ERR_PARSE;
        }
        const char *p = ln.c_str();
        char *e;
        float v[3];
        for (int t = 0; t < 3; t++) {
            v[t] = std::strtof(p, &e);
            if (e == p) {
                err = "malformed position line for atom " + std::to_string(i);
                return DBX_ERR_PARSE;
            }
            p = e;
        }
        d.x.push_back(v[0]);
        d.y.push_back(v[1]);
        d.z.push_back(v[2]);
        i++;
    }
    if (i != N) {
        err = "Number of extracted Atoms (" + std::to_string(i) +
              ") doesn't match the number specified in the preamble (" + std::to_string(N) + ")";
        return DBX_ERR_PARSE;
    }
    return DBX_OK;
}

namespace {

__attribute__((format(printf, 2, 3))) void logf(bool quiet, const char *fmt, ...)
{
    if (quiet) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// per-pair checks (cuDebye-v15.cu:360-410), .debx files (cuDebye-v15.cu:413-430, unless the
// histograms were read from them) and the .debI file (cuDebye-v15.cu:973-998)
int write_outputs(const char *path, const DebFile &d, const std::vector<double> &tt, const uint64_t *H,
                  int64_t channels, const std::vector<double> &I, bool write_debx, const char *eol, bool quiet,
                  double *t_debx)
{
    int rc = DBX_OK;
    double t_w = 0;
    int p = 0;
    for (int row = 0; row < d.S; row++) {
        for (int col = row; col < d.S; col++, p++) {
            const uint64_t *h = &H[(size_t)p * channels];
            uint64_t sum = 0, zero = 0, near = 0;
            int64_t nnz = 0;
            for (int64_t k = 0; k < channels; k++) {
                if (!h[k]) continue;
                nnz++;
                sum += h[k];
                const double r = (double)k * d.rstep;
                if (r <= 0.0005) zero += h[k];
                if (r < 0.25) near += h[k];
            }
            const uint64_t na = d.n[row], nb = d.n[col];
            const uint64_t expect = row == col ? (na * na - na) / 2 + na : na * nb;
            logf(quiet, ">>>> Subgroups %d and %d: %lld unique distances, %llu observed, %llu expected\n", row,
                 col, (long long)nnz, (unsigned long long)sum, (unsigned long long)expect);
            if (sum != expect)
                logf(quiet, "WARNING: %llu Distances Expected and %llu Observed!!!\n", (unsigned long long)expect,
                     (unsigned long long)sum);
            if (row == col) {
                if (h[0] != na)
                    logf(quiet,
                         "WARNING: Number of Expected Zero Distances Missing or not Correct!!!\n"
                         "Number of Atoms within a Tolerance of 0.0005 A is %llu\n",
                         (unsigned long long)zero);
                if (near > na)
                    logf(quiet,
                         "WARNING: Possible Atom Collision!!! Found around %llu interatomic Distances "
                         "lower than the Radius of an Hydrogen Atom.\n",
                         (unsigned long long)(near - na));
            } else if (near > 0) {
                logf(quiet,
                     "WARNING: Possible Atom Collision!!! Found around %llu interatomic Distances lower "
                     "than the Radius of an Hydrogen Atom.\n",
                     (unsigned long long)near);
            }
            if (d.hist > 0 && write_debx) {
                const double tw0 = now_s();
                const std::string fn = std::string(path) + "x" + std::to_string(row) + "x" + std::to_string(col);
                FILE *f = std::fopen(fn.c_str(), "wb");
                if (!f) {
                    std::fprintf(stderr, "ERROR: cannot write %s\n", fn.c_str());
                    rc = DBX_ERR_IO;
                    continue;
                }
                for (int64_t k = 0; k < channels; k++)
                    if (h[k]) std::fprintf(f, "%.7g\t%llu%s", (double)k * d.rstep, (unsigned long long)h[k], eol);
                std::fclose(f);
                t_w += now_s() - tw0;
            }
        }
    }
    if (d.hist != 1) {
        const std::string fn = std::string(path) + "I";
        FILE *f = std::fopen(fn.c_str(), "wb");
        if (!f) {
            std::fprintf(stderr, "ERROR: cannot write %s\n", fn.c_str());
            rc = DBX_ERR_IO;
        } else {
            for (size_t i = 0; i < tt.size(); i++) std::fprintf(f, "%g\t%g%s", tt[i], I[i], eol);
            std::fclose(f);
        }
    }
    if (t_debx) *t_debx = t_w;
    return rc;
}

// the .deb's grids; false (message printed) when they are invalid
bool grids(const DebFile &d, std::vector<double> &tt, std::vector<double> &K, int64_t &channels)
{
    const int nTT = dbx_q_from_2theta(d.lambda, d.tt_min, d.tt_max, d.tt_step, nullptr, nullptr, 0);
    channels = dbx_channels(d.rmax, d.rstep);
    if (nTT < 1 || channels < 2) {
        std::fprintf(stderr, "ERROR: invalid 2theta or distance grid in the header\n");
        return false;
    }
    tt.resize(nTT);
    K.resize(nTT);
    dbx_q_from_2theta(d.lambda, d.tt_min, d.tt_max, d.tt_step, tt.data(), K.data(), nTT);
    return true;
}

}  // namespace

// Reads one .debx file (lines "%.7g\t%llu": r_k and N_k of each nonzero bin, the format of
// cuDebye-v15.cu:417-423) into hist[n_bins] (zeroed first).  k = rint(r / rstep), and r must
// round-trip: |r - k rstep| <= 1e-6 |r| + 1e-12 (seven significant digits).
extern "C" int dbx_read_debx(const char *path, double rstep, int64_t n_bins, uint64_t *hist, int64_t *n_nonzero)
{
    if (!path || !hist || n_bins <= 0 || !(rstep > 0)) return DBX_ERR_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return DBX_ERR_IO;
    std::memset(hist, 0, sizeof(uint64_t) * (size_t)n_bins);
    char line[256];
    int64_t nnz = 0;
    int rc = DBX_OK;
    while (std::fgets(line, sizeof(line), f)) {
        if (blank(line) || line[0] == '\n') continue;
        char *e = nullptr;
        const double r = std::strtod(line, &e);
        if (e == line) {
            rc = DBX_ERR_PARSE;
            break;
        }
        char *e2 = nullptr;
        const unsigned long long n = std::strtoull(e, &e2, 10);
        if (e2 == e || !std::isfinite(r) || r < 0) {
            rc = DBX_ERR_PARSE;
            break;
        }
        const double kd = std::rint(r / rstep);
        if (std::fabs(r - kd * rstep) > 1e-6 * std::fabs(r) + 1e-12) {
            rc = DBX_ERR_PARSE;
            break;
        }
        if (kd >= (double)n_bins) {
            rc = DBX_ERR_RANGE;
            break;
        }
        const int64_t k = (int64_t)kd;
        if (!hist[k]) nnz++;
        hist[k] += n;
    }
    std::fclose(f);
    if (n_nonzero) *n_nonzero = nnz;
    return rc;
}

static int run_deb_body(const char *path, int flags)
{
    if (!path) return DBX_ERR_ARG;
    const bool quiet = flags & DBX_RUN_QUIET;
    const bool from_debx = flags & DBX_RUN_FROM_DEBX;
    const char *eol = (flags & DBX_RUN_CRLF) ? "\r\n" : "\n";
    double t0 = now_s();
    DebFile d;
    std::string err;
    int rc = parse_deb(path, d, err);
    if (rc != DBX_OK) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return rc;
    }
    std::vector<double> tt, K;
    int64_t channels = 0;
    if (!grids(d, tt, K, channels)) return DBX_ERR_ARG;
    const int nTT = (int)tt.size();
    logf(quiet, "Loaded %s: %lld atoms, %d subgroups, %d 2theta values (K %g..%g), %lld channels, "
                "hist=%d (%.3f s)\n",
         path, (long long)d.x.size(), d.S, nTT, K[0], K[nTT - 1], (long long)channels, d.hist, now_s() - t0);

    // the context and the model are released on every way out, exceptions included
    struct Owned {
        dbx_ctx *ctx = nullptr;
        dbx_model *m = nullptr;
        ~Owned()
        {
            if (m) dbx_model_destroy(ctx, m);
            if (ctx) dbx_close(ctx);
        }
    } own;
    rc = dbx_open(d.device, 0, &own.ctx);
    dbx_ctx *const ctx = own.ctx;
    if (rc != DBX_OK) {
        std::fprintf(stderr, "ERROR: cannot open device %d\n", d.device);
        return rc;
    }
    // the reference evaluates the Debye sum in FP64; the driver does too unless asked for the
    // matrix-core path (integer digits of sinc, as accurate as FP64 but summed in another order)
    dbx_set_pass2_mode(ctx, (flags & DBX_RUN_MFMA) ? DBX_PASS2_MFMA : DBX_PASS2_FP64);
    const int P = d.S * (d.S + 1) / 2;
    std::vector<uint64_t> H((size_t)P * channels);
    std::vector<double> I(nTT, 0.0), rel(nTT, 0.0);
    const int first_n = d.block > 0 ? d.block : 1;
    double t1 = now_s();
    if (from_debx) {
        // pass 2 only, from the histograms a previous run wrote (<path>x<r>x<c>)
        int p = 0;
        for (int row = 0; row < d.S && rc == DBX_OK; row++)
            for (int col = row; col < d.S && rc == DBX_OK; col++, p++) {
                const std::string fn = std::string(path) + "x" + std::to_string(row) + "x" + std::to_string(col);
                rc = dbx_read_debx(fn.c_str(), d.rstep, channels, &H[(size_t)p * channels], nullptr);
                if (rc != DBX_OK) std::fprintf(stderr, "ERROR: cannot read histogram file %s\n", fn.c_str());
            }
        if (rc == DBX_OK && d.hist != 1) {
            std::vector<double> ff((size_t)P * nTT);
            p = 0;
            for (int a = 0; a < d.S && rc == DBX_OK; a++)
                for (int b = a; b < d.S && rc == DBX_OK; b++, p++)
                    rc = dbx_prefactor(ctx, &d.gp[a], &d.gp[b], K.data(), nTT, &ff[(size_t)p * nTT]);
            if (rc == DBX_OK)
                rc = dbx_debye(ctx, H.data(), 0, P, channels, d.rstep, first_n, K.data(), nTT, ff.data(), I.data(),
                               nullptr, rel.data());
            if (rc != DBX_OK) std::fprintf(stderr, "ERROR: %s\n", dbx_last_error(ctx));
        }
        if (rc != DBX_OK) return rc;
    } else {
        rc = dbx_model_create(ctx, d.x.data(), d.y.data(), d.z.data(), d.n.data(), d.S, &own.m);
        if (rc != DBX_OK) {
            std::fprintf(stderr, "ERROR: %s\n", dbx_last_error(ctx));
            return rc;
        }
        dbx_model *const m = own.m;
        if (d.hist != 1)
            rc = dbx_intensity(ctx, m, d.gp.data(), d.rstep, channels, first_n, K.data(), nTT, I.data(), H.data(),
                               rel.data());
        else
            rc = dbx_histogram(ctx, m, d.rstep, channels, H.data(), 0, nullptr);
        if (rc != DBX_OK) {
            std::fprintf(stderr, "ERROR: %s\n", dbx_last_error(ctx));
            if (rc != DBX_ERR_RANGE) return rc;
        }
    }
    const double t_calc = now_s() - t1;
    double t_debx = 0;
    const int wrc = write_outputs(path, d, tt, H.data(), channels, I, !from_debx, eol, quiet, &t_debx);
    if (rc == DBX_OK) rc = wrc;
    if (d.hist != 1) {
        double worst = 0;
        for (double v : rel) worst = std::max(worst, v);
        logf(quiet, "Largest relative error bound of I(2theta): %.3g\n", worst);
    }
    logf(quiet, "Time for Calculation of Distances and Evaluation of the Debye Double Sum: %.3f s\n", t_calc);
    // the reference times its subgroup-pair loop, which writes the .debx files inside it
    // (cuDebye-v15.cu:928-971)
    logf(quiet, "Time for writing the .debx files: %.3f s\n", t_debx);
    logf(quiet, "Time for Distances, Debye Double Sum and .debx files: %.3f s\n", t_calc + t_debx);
    logf(quiet, "Total time: %.3f s\n", now_s() - t0);
    return rc;
}

// Every .deb in `dir` that has no .debI yet (the skip-if-output-exists rule of
// Matlab/multisim.m:38-58), in name order, through one server: all files are submitted at
// once, so small models share steps and large ones overlap.  Each file's outputs are those of
// dbx_run_deb on it alone.  Returns DBX_OK, or the first failing file's status (the other
// files still run).
static int run_folder_body(const char *dir, int flags)
{
    if (!dir) return DBX_ERR_ARG;
    const bool quiet = flags & DBX_RUN_QUIET;
    const char *eol = (flags & DBX_RUN_CRLF) ? "\r\n" : "\n";
    DIR *dp = opendir(dir);
    if (!dp) {
        std::fprintf(stderr, "ERROR: >>%s<< is not a folder\n", dir);
        return DBX_ERR_IO;
    }
    // first every .xyz without its .deb, converted with the default header (multisim.m:17-36)
    std::vector<std::string> xyz;
    while (dirent *e = readdir(dp)) {
        const std::string name = e->d_name;
        if (name.size() < 5 || name.compare(name.size() - 4, 4, ".xyz") != 0) continue;
        const std::string full = std::string(dir) + "/" + name;
        struct stat sb;
        if (stat((full.substr(0, full.size() - 3) + "deb").c_str(), &sb) == 0) continue;
        xyz.push_back(full);
    }
    std::sort(xyz.begin(), xyz.end());
    for (size_t i = 0; i < xyz.size(); i++) {
        logf(quiet, "Convert File %zu of %zu -- %s\n", i + 1, xyz.size(), xyz[i].c_str());
        const int rc = dbx_xyz_to_deb(xyz[i].c_str(), nullptr, nullptr, nullptr);
        if (rc != DBX_OK) {
            closedir(dp);
            return rc;
        }
    }
    rewinddir(dp);
    std::vector<std::string> files;
    while (dirent *e = readdir(dp)) {
        const std::string name = e->d_name;
        if (name.size() < 5 || name.compare(name.size() - 4, 4, ".deb") != 0) continue;
        const std::string full = std::string(dir) + "/" + name;
        struct stat sb;
        if (stat((full + "I").c_str(), &sb) == 0) continue;  // already calculated
        files.push_back(full);
    }
    closedir(dp);
    std::sort(files.begin(), files.end());
    logf(quiet, "%zu file(s) to calculate in %s\n", files.size(), dir);
    if (files.empty()) return DBX_OK;

    struct Job {
        std::string path;
        DebFile d;
        std::vector<double> tt, K;
        int64_t channels = 0;
        uint64_t ticket = 0;
        int rc = DBX_OK;
    };
    std::vector<Job> jobs(files.size());
    int device = 0;
    for (size_t i = 0; i < files.size(); i++) {
        Job &j = jobs[i];
        j.path = files[i];
        std::string err;
        j.rc = parse_deb(j.path.c_str(), j.d, err);
        if (j.rc != DBX_OK)
            std::fprintf(stderr, "ERROR: %s: %s\n", j.path.c_str(), err.c_str());
        else if (!grids(j.d, j.tt, j.K, j.channels))
            j.rc = DBX_ERR_ARG;
        if (i == 0 || (jobs[0].rc != DBX_OK && j.rc == DBX_OK)) device = j.d.device;
    }
    dbx_server_config cfg{};
    cfg.pass2_mode = (flags & DBX_RUN_MFMA) ? DBX_PASS2_MFMA : DBX_PASS2_FP64;
    // the server (and its threads) is closed on every way out, exceptions included
    struct Owned {
        dbx_server *srv = nullptr;
        ~Owned()
        {
            if (srv) dbx_server_close(srv);
        }
    } own;
    int rc = dbx_server_open(device, &cfg, &own.srv);
    dbx_server *const srv = own.srv;
    if (rc != DBX_OK) {
        std::fprintf(stderr, "ERROR: cannot open device %d\n", device);
        return rc;
    }
    for (Job &j : jobs) {
        if (j.rc != DBX_OK) continue;
        dbx_request q{};
        q.x = j.d.x.data();
        q.y = j.d.y.data();
        q.z = j.d.z.data();
        q.group_n = j.d.n.data();
        q.n_groups = j.d.S;
        q.params = j.d.gp.data();
        q.rstep = j.d.rstep;
        q.n_bins = j.channels;
        q.first_n = j.d.block > 0 ? j.d.block : 1;
        q.q = j.K.data();
        q.nq = (int)j.K.size();
        q.flags = DBX_REQ_KEEP;
        j.rc = dbx_submit(srv, &q, &j.ticket);
        if (j.rc != DBX_OK) std::fprintf(stderr, "ERROR: %s: %s\n", j.path.c_str(), dbx_server_last_error(srv));
    }
    int first_bad = DBX_OK;
    for (Job &j : jobs) {
        if (j.rc == DBX_OK) {
            std::vector<double> I(j.K.size());
            double ms = 0;
            j.rc = dbx_wait(srv, j.ticket, I.data(), &ms, nullptr);
            if (j.rc == DBX_OK) {
                const int P = j.d.S * (j.d.S + 1) / 2;
                std::vector<uint64_t> H((size_t)P * j.channels);
                j.rc = dbx_request_histogram(srv, j.ticket, H.data());
                if (j.rc == DBX_OK) {
                    logf(quiet, "Calculated %s (%.3f s)\n", j.path.c_str(), ms * 1e-3);
                    j.rc = write_outputs(j.path.c_str(), j.d, j.tt, H.data(), j.channels, I, true, eol, quiet,
                                         nullptr);
                }
            } else {
                std::fprintf(stderr, "ERROR: %s: %s\n", j.path.c_str(), dbx_server_last_error(srv));
            }
            dbx_release(srv, j.ticket);
        }
        if (j.rc != DBX_OK && first_bad == DBX_OK) first_bad = j.rc;
    }
    return first_bad;
}

// The C boundary: no exception leaves it (a failed host allocation, say for the histograms of
// a header whose rmax is huge, returns DBX_ERR_NOMEM instead of ending the caller's process)
template <class F>
static int no_throw(F &&f)
{
    try {
        return f();
    } catch (const dbx::Error &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        std::fprintf(stderr, "ERROR: host allocation failed\n");
        return DBX_ERR_NOMEM;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return DBX_ERR_ARG;
    }
}

extern "C" int dbx_run_deb(const char *path, int flags)
{
    return no_throw([&] { return run_deb_body(path, flags); });
}

extern "C" int dbx_run_folder(const char *dir, int flags)
{
    return no_throw([&] { return run_folder_body(dir, flags); });
}

}  // namespace dbx

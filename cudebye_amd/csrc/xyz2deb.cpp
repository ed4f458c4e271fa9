// Host tools around the .deb driver: the xyz -> deb converter (Matlab/xyz2deb.m:169-266 with
// its atomicscatter table lookup, :268-311) and the averaging of several .debI patterns
// (Matlab/plotdebi.m:13-63).  Plain C++, no device code.
#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/debye.h"

namespace {

std::vector<std::string> split_ws(const std::string &line)
{
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

bool to_double(const std::string &s, double &v)
{
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && end != s.c_str();
}

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct Coeffs {
    double a[5], c, b[5];
};

// the coefficient table: '#' comment lines, one column-header line, then rows
// "name a1 a2 a3 a4 a5 c b1 b2 b3 b4 b5"
bool read_table(const std::string &path, std::vector<std::pair<std::string, Coeffs>> &rows, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot open the scattering-factor table " + path;
        return false;
    }
    std::string line;
    bool header = false;
    while (std::getline(f, line)) {
        const auto t = split_ws(line);
        if (t.empty()) continue;
        if (!header) {
            if (t[0][0] == '#') continue;
            header = true;  // the column-header line
            continue;
        }
        if (t.size() < 12) {
            err = "short table row: " + line;
            return false;
        }
        Coeffs c;
        double v[11];
        for (int i = 0; i < 11; i++)
            if (!to_double(t[1 + i], v[i])) {
                err = "bad number in table row: " + line;
                return false;
            }
        for (int i = 0; i < 5; i++) c.a[i] = v[i];
        c.c = v[5];
        for (int i = 0; i < 5; i++) c.b[i] = v[6 + i];
        rows.emplace_back(t[0], c);
    }
    return true;
}

// directory of libdebye.so (its data/ folder holds the default table)
std::string lib_dir()
{
    Dl_info info;
    if (dladdr((void *)&dbx_xyz_to_deb, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t s = p.rfind('/');
        return s == std::string::npos ? std::string(".") : p.substr(0, s);
    }
    return ".";
}

std::string fmt(const char *f, double v)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), f, v);
    return buf;
}

// "%.<P>g" as the reference's converter prints coordinates (Matlab's dlmwrite, precision 7, on
// Windows): the value is first converted to 17 significant digits, correctly rounded, and that
// decimal string is then rounded half away from zero to P digits.  This double rounding differs
// from a correctly rounded %.7g where the 17-digit string ends in 5000... (e.g. -12.256335, stored
// as -12.2563349999...98, prints as -12.25634); Data/UH3_7nm.deb has 2,052 such coordinates.
// Exponents are printed with at least two digits.
std::string fmt_g17(double v, int P)
{
    if (v == 0.0 || !std::isfinite(v)) return fmt("%g", v);
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.16e", std::fabs(v));  // d.dddddddddddddddde+XX
    std::string dig(1, buf[0]);
    dig.append(buf + 2, 16);
    int e10 = std::atoi(std::strchr(buf, 'e') + 1);
    std::string d = dig.substr(0, P);
    if (dig[P] >= '5') {
        int i = P - 1;
        while (i >= 0 && d[i] == '9') d[i--] = '0';
        if (i < 0) {
            d.insert(d.begin(), '1');
            d.pop_back();
            e10++;
        } else {
            d[i]++;
        }
    }
    std::string out = v < 0 ? "-" : "";
    auto strip = [](std::string m) {
        if (m.find('.') != std::string::npos) {
            while (!m.empty() && m.back() == '0') m.pop_back();
            if (!m.empty() && m.back() == '.') m.pop_back();
        }
        return m;
    };
    if (e10 < -4 || e10 >= P) {
        std::string m = d.substr(0, 1) + "." + d.substr(1);
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
        return out + strip(m) + eb;
    }
    std::string m;
    if (e10 >= 0)
        m = d.substr(0, e10 + 1) + "." + d.substr(e10 + 1);
    else
        m = "0." + std::string(-e10 - 1, '0') + d;
    return out + strip(m);
}

// cubic spline through (x, y), x strictly increasing, with not-a-knot ends (the third
// derivative continuous at x_1 and x_{n-2}); two points: the line, three: the parabola
struct Spline {
    std::vector<double> x, y, M;  // M: second derivatives at the knots

    bool build(const std::vector<double> &xs, const std::vector<double> &ys)
    {
        x = xs;
        y = ys;
        const size_t n = x.size();
        if (n < 2) return false;
        for (size_t i = 1; i < n; i++)
            if (!(x[i] > x[i - 1])) return false;
        M.assign(n, 0.0);
        if (n == 2) return true;
        if (n == 3) {
            // one parabola: constant second derivative 2 c2
            const double d0 = (y[1] - y[0]) / (x[1] - x[0]), d1 = (y[2] - y[1]) / (x[2] - x[1]);
            const double c2 = (d1 - d0) / (x[2] - x[0]);
            M.assign(3, 2.0 * c2);
            return true;
        }
        std::vector<double> h(n - 1), d(n - 1);
        for (size_t i = 0; i + 1 < n; i++) {
            h[i] = x[i + 1] - x[i];
            d[i] = (y[i + 1] - y[i]) / h[i];
        }
        // unknowns M_1 .. M_{n-2}; rows i = 1 .. n-2:
        //   h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = 6 (d_i - d_{i-1})
        // not-a-knot: M_0 = ((h0 + h1) M_1 - h0 M_2) / h1, and the mirror image at the end,
        // substituted into the first and last rows
        const size_t m = n - 2;
        std::vector<double> lo(m, 0.0), di(m, 0.0), up(m, 0.0), r(m, 0.0);
        for (size_t k = 0; k < m; k++) {
            const size_t i = k + 1;
            lo[k] = h[i - 1];
            di[k] = 2.0 * (h[i - 1] + h[i]);
            up[k] = h[i];
            r[k] = 6.0 * (d[i] - d[i - 1]);
        }
        const double h0 = h[0], h1 = h[1];
        di[0] += h0 * (h0 + h1) / h1;
        up[0] -= h0 * h0 / h1;
        const double ha = h[n - 3], hb = h[n - 2];  // h_{n-3}, h_{n-2}
        di[m - 1] += hb * (ha + hb) / ha;
        lo[m - 1] -= hb * hb / ha;
        // Thomas algorithm (m >= 2 here)
        for (size_t k = 1; k < m; k++) {
            const double w = lo[k] / di[k - 1];
            di[k] -= w * up[k - 1];
            r[k] -= w * r[k - 1];
        }
        M[m] = r[m - 1] / di[m - 1];
        for (size_t k = m - 1; k-- > 0;) M[k + 1] = (r[k] - up[k] * M[k + 2]) / di[k];
        M[0] = ((h0 + h1) * M[1] - h0 * M[2]) / h1;
        M[n - 1] = ((ha + hb) * M[n - 2] - hb * M[n - 3]) / ha;
        return true;
    }

    double operator()(double t) const
    {
        const size_t n = x.size();
        size_t i = (size_t)(std::upper_bound(x.begin(), x.end(), t) - x.begin());
        i = i == 0 ? 0 : i - 1;
        if (i > n - 2) i = n - 2;
        const double hi = x[i + 1] - x[i];
        const double a = x[i + 1] - t, b = t - x[i];
        return M[i] * a * a * a / (6.0 * hi) + M[i + 1] * b * b * b / (6.0 * hi) +
               (y[i] - M[i] * hi * hi / 6.0) * a / hi + (y[i + 1] - M[i + 1] * hi * hi / 6.0) * b / hi;
    }
};

bool read_debi(const char *path, std::vector<double> &tt, std::vector<double> &I)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        const auto t = split_ws(line);
        if (t.empty()) continue;
        double a, b;
        if (t.size() < 2 || !to_double(t[0], a) || !to_double(t[1], b)) return false;
        tt.push_back(a);
        I.push_back(b);
    }
    return !tt.empty();
}

// The C boundary: no exception leaves it (a failed host allocation returns DBX_ERR_NOMEM
// instead of ending the caller's process)
template <class F>
int no_throw(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        std::fprintf(stderr, "ERROR: host allocation failed\n");
        return DBX_ERR_NOMEM;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return DBX_ERR_ARG;
    }
}

}  // namespace

extern "C" {

void dbx_deb_header_defaults(dbx_deb_header *h)
{
    if (!h) return;
    h->device = 0;
    h->block = 16;
    h->grid = 2048;
    h->wavelength = 1.5418;
    h->tt_min = 10.0;
    h->tt_max = 140.0;
    h->tt_step = 0.02;
    h->rstep = 0.0001;
    h->hist = 0;
}

static int xyz_to_deb_body(const char *xyz_path, const char *deb_path, const char *table_path,
                           const dbx_deb_header *hdr)
{
    if (!xyz_path) return DBX_ERR_ARG;
    dbx_deb_header H;
    dbx_deb_header_defaults(&H);
    if (hdr) H = *hdr;
    std::string out = deb_path ? deb_path : "";
    if (out.empty()) {
        // the xyz name with its last three characters replaced (xyz2deb.m:228)
        out = xyz_path;
        if (out.size() < 3) return DBX_ERR_ARG;
        out = out.substr(0, out.size() - 3) + "deb";
    }
    const std::string table = table_path ? table_path : lib_dir() + "/data/atomic_scattering_factors.txt";
    std::string err;
    std::vector<std::pair<std::string, Coeffs>> trows;
    if (!read_table(table, trows, err)) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return DBX_ERR_IO;
    }

    // atoms: element, x, y, z [, occupancy [, Beq]] per line; the first line fixes the
    // column count (4, 5 or 6; xyz2deb.m:172-197): occupancy 1 and Beq 0 where absent
    std::ifstream f(xyz_path);
    if (!f) {
        std::fprintf(stderr, "ERROR: cannot open %s\n", xyz_path);
        return DBX_ERR_IO;
    }
    struct Atom {
        int type;
        double x, y, z, occ, beq;
    };
    std::vector<std::string> names;
    std::vector<Atom> atoms;
    std::vector<std::string> elem;
    size_t cols = 0;
    std::string line;
    long lineno = 0;
    while (std::getline(f, line)) {
        lineno++;
        const auto t = split_ws(line);
        if (t.empty()) continue;
        if (cols == 0) {
            cols = t.size();
            if (cols < 4 || cols > 6) {
                std::fprintf(stderr, "ERROR: %s: %zu columns (4 to 6 expected)\n", xyz_path, cols);
                return DBX_ERR_PARSE;
            }
        }
        if (t.size() != cols) {
            std::fprintf(stderr, "ERROR: %s:%ld: %zu columns, the first line has %zu\n", xyz_path, lineno, t.size(),
                         cols);
            return DBX_ERR_PARSE;
        }
        double v[5] = {0, 0, 0, 1.0, 0.0};
        for (size_t i = 1; i < cols; i++)
            if (!to_double(t[i], v[i - 1])) {
                std::fprintf(stderr, "ERROR: %s:%ld: bad number %s\n", xyz_path, lineno, t[i].c_str());
                return DBX_ERR_PARSE;
            }
        elem.push_back(t[0]);
        atoms.push_back(Atom{0, v[0], v[1], v[2], v[3], v[4]});
    }
    if (atoms.empty()) {
        std::fprintf(stderr, "ERROR: %s holds no atoms\n", xyz_path);
        return DBX_ERR_PARSE;
    }
    // element types: rank in the sorted list of distinct names (xyz2deb.m:184-189)
    names = elem;
    std::sort(names.begin(), names.end());
    names.erase(std::unique(names.begin(), names.end()), names.end());
    for (size_t i = 0; i < atoms.size(); i++)
        atoms[i].type = (int)(std::lower_bound(names.begin(), names.end(), elem[i]) - names.begin());
    // stable order by (type, occupancy, Beq) (xyz2deb.m:199)
    std::vector<size_t> order(atoms.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Atom &A = atoms[a], &B = atoms[b];
        if (A.type != B.type) return A.type < B.type;
        if (A.occ != B.occ) return A.occ < B.occ;
        return A.beq < B.beq;
    });
    // subgroups: runs of equal (type, occupancy, Beq)
    std::vector<size_t> start{0};
    for (size_t k = 1; k < order.size(); k++) {
        const Atom &A = atoms[order[k - 1]], &B = atoms[order[k]];
        if (A.type != B.type || A.occ != B.occ || A.beq != B.beq) start.push_back(k);
    }
    start.push_back(order.size());
    // centre each axis on the middle of its range (xyz2deb.m:207-211); Rmax from the
    // largest centred radius (xyz2deb.m:212-213)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const Atom &a : atoms) {
        const double v[3] = {a.x, a.y, a.z};
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], v[c]);
            mx[c] = std::max(mx[c], v[c]);
        }
    }
    double rmax = 0;
    for (Atom &a : atoms) {
        a.x = a.x - mn[0] - (mx[0] - mn[0]) / 2;
        a.y = a.y - mn[1] - (mx[1] - mn[1]) / 2;
        a.z = a.z - mn[2] - (mx[2] - mn[2]) / 2;
        rmax = std::max(rmax, std::hypot(std::hypot(a.x, a.y), a.z));
    }
    const double Rmax = 2.0 * std::ceil(rmax * 1.025);

    // header and subgroup lines (xyz2deb.m:229-255), positions with seven significant
    // digits, tab separated (xyz2deb.m:259)
    std::string s;
    s += fmt("%.0f", H.device) + " " + fmt("%.0f", H.block) + " " + fmt("%.0f", H.grid);
    s += "\n" + fmt("%.6f", H.wavelength) + " " + fmt("%.2f", H.tt_min) + " " + fmt("%.2f", H.tt_max) + " " +
         fmt("%.2f", H.tt_step);
    s += "\n" + fmt("%.0f", Rmax) + " " + fmt("%.4f", H.rstep) + " " + fmt("%.0f", H.hist);
    s += "\n" + fmt("%.0f", (double)(start.size() - 1));
    for (size_t g = 0; g + 1 < start.size(); g++) {
        const Atom &A = atoms[order[start[g]]];
        const std::string &name = names[A.type];
        const Coeffs *cf = nullptr;
        for (const auto &r : trows)
            if (lower(r.first) == lower(name)) {
                cf = &r.second;
                break;
            }
        if (!cf) {
            std::fprintf(stderr, "ERROR: element %s is not in %s\n", name.c_str(), table.c_str());
            return DBX_ERR_PARSE;
        }
        s += "\n" + fmt("%.0f", (double)(start[g + 1] - start[g])) + " " + name + " " + fmt("%.4f", A.occ) + " " +
             fmt("%.4f", A.beq);
        for (int j = 0; j < 5; j++) s += " " + fmt("%.6f", cf->a[j]);
        for (int j = 0; j < 5; j++) s += " " + fmt("%.6f", cf->b[j]);
        s += " " + fmt("%.6f", cf->c);
    }
    s += "\n";
    FILE *o = std::fopen(out.c_str(), "wb");
    if (!o) {
        std::fprintf(stderr, "ERROR: cannot write %s\n", out.c_str());
        return DBX_ERR_IO;
    }
    bool ok = std::fwrite(s.data(), 1, s.size(), o) == s.size();
    for (size_t k = 0; k < order.size() && ok; k++) {
        const Atom &a = atoms[order[k]];
        const std::string l = fmt_g17(a.x, 7) + "\t" + fmt_g17(a.y, 7) + "\t" + fmt_g17(a.z, 7) + "\n";
        ok = std::fwrite(l.data(), 1, l.size(), o) == l.size();
    }
    ok = (std::fclose(o) == 0) && ok;
    return ok ? DBX_OK : DBX_ERR_IO;
}

static int average_debi_body(const char *const *paths, int n_files, int norm, double step, double *tt_out,
                             double *I_out, int capacity, int *n_out)
{
    if (!paths || n_files < 1 || !n_out || !(step > 0) || norm < DBX_AVG_NONE || norm > DBX_AVG_MAX)
        return DBX_ERR_ARG;
    std::vector<std::vector<double>> X(n_files), Y(n_files);
    double xmin = 0.0, xmax = 180.0;  // plotdebi.m:13-27
    for (int i = 0; i < n_files; i++) {
        if (!paths[i] || !read_debi(paths[i], X[i], Y[i])) {
            std::fprintf(stderr, "ERROR: cannot read %s\n", paths[i] ? paths[i] : "(null)");
            return DBX_ERR_IO;
        }
        xmin = std::max(xmin, *std::min_element(X[i].begin(), X[i].end()));
        xmax = std::min(xmax, *std::max_element(X[i].begin(), X[i].end()));
    }
    if (!(xmax >= xmin)) return DBX_ERR_RANGE;
    const int n = (int)std::floor((xmax - xmin) / step + 1e-9) + 1;
    *n_out = n;
    if (!tt_out || !I_out) return DBX_OK;  // size query
    if (capacity < n) return DBX_ERR_ARG;
    std::vector<double> t(n), acc(n, 0.0);
    for (int k = 0; k < n; k++) t[k] = xmin + k * step;
    // the sum of the raw patterns, each through its not-a-knot cubic spline (plotdebi.m:52-56)
    for (int i = 0; i < n_files; i++) {
        Spline sp;
        if (!sp.build(X[i], Y[i])) {
            std::fprintf(stderr, "ERROR: %s: 2theta values must increase\n", paths[i]);
            return DBX_ERR_PARSE;
        }
        for (int k = 0; k < n; k++) acc[k] += sp(t[k]);
    }
    // normalisation (plotdebi.m:57-64): by the trapezoidal integral, by the maximum, or the mean
    double div = (double)n_files;
    if (norm == DBX_AVG_INTEGRAL) {
        div = 0;
        for (int k = 0; k + 1 < n; k++) div += 0.5 * (t[k + 1] - t[k]) * (acc[k] + acc[k + 1]);
    } else if (norm == DBX_AVG_MAX) {
        div = *std::max_element(acc.begin(), acc.end());
    }
    for (int k = 0; k < n; k++) {
        tt_out[k] = t[k];
        I_out[k] = acc[k] / div;
    }
    return DBX_OK;
}

int dbx_xyz_to_deb(const char *xyz_path, const char *deb_path, const char *table_path, const dbx_deb_header *hdr)
{
    return no_throw([&] { return xyz_to_deb_body(xyz_path, deb_path, table_path, hdr); });
}

int dbx_average_debi(const char *const *paths, int n_files, int norm, double step, double *tt_out, double *I_out,
                     int capacity, int *n_out)
{
    return no_throw([&] { return average_debi_body(paths, n_files, norm, step, tt_out, I_out, capacity, n_out); });
}

}  // extern "C"

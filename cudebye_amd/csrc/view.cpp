// Structure view of a .deb model (Matlab/viewdeb.m:1-194): the atoms inside per-axis limits,
// rendered as a 3-D scatter seen from Matlab's default 3-D view into an SVG file, and/or
// listed as an .xyz file that molecular viewers open.  Plain C++, no device code.
//
// The rules that decide what is drawn follow the reference script: the element of a subgroup is
// its type field without charge, label and count suffixes (viewdeb.m:121-127); an atom is kept
// when lo < coordinate < hi on every axis (viewdeb.m:47-54), the limits -10..10 on every axis
// unless given; each atom is a filled disc of diameter 2 r MarkerSize points (MarkerSize 5,
// r the element's radius from the radius/colour table) in the element's colour with a black
// edge (viewdeb.m:64-83); the axes span the kept atoms with equal scales, and the elements are
// listed in their colours beside the plot (viewdeb.m:84-91).
#include <dlfcn.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "debfile.h"

namespace dbx {

namespace {

bool word_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

// The element symbol of a subgroup's type field: every '-', '+', '(' or ':' is removed with the
// word characters after it, then every digit with the word characters after it ("Al3+" -> "Al",
// "O2-" -> "O", "Fe:oct" -> "Fe"), the substitutions of viewdeb.m:122-126 in their order.
std::string element_of_type(const std::string &type)
{
    std::string s = type;
    auto strip = [&](auto starts) {
        std::string out;
        for (size_t i = 0; i < s.size();) {
            if (starts(s[i])) {
                i++;
                while (i < s.size() && word_char(s[i])) i++;
            } else {
                out += s[i++];
            }
        }
        s = out;
    };
    for (char c : {'-', '+', '(', ':'}) strip([c](char x) { return x == c; });
    strip([](char x) { return std::isdigit((unsigned char)x) != 0; });
    return s;
}

struct Style {
    double radius = 0;        // Angstrom; 0 for an element missing from the table (not drawn)
    double rgb[3] = {0, 0, 0};
};

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// directory of libdebye.so (its data/ folder holds the default table)
std::string lib_dir()
{
    Dl_info info;
    if (dladdr((void *)&element_of_type, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t k = p.rfind('/');
        return k == std::string::npos ? "." : p.substr(0, k);
    }
    return ".";
}

// rows "symbol radius red green blue" after '#' comments, keyed by the lower-case symbol
bool load_styles(const std::string &path, std::map<std::string, Style> &out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string sym;
        Style st;
        if (is >> sym >> st.radius >> st.rgb[0] >> st.rgb[1] >> st.rgb[2]) out[lower(sym)] = st;
    }
    return !out.empty();
}

struct Atom {
    double x, y, z;
    int elem;  // index into the element list
};

// Matlab's default 3-D view (azimuth -37.5 deg, elevation 30 deg): screen x, screen y (up) and
// the distance towards the viewer, the rows of its orthographic view matrix
void project(double x, double y, double z, double &sx, double &sy, double &depth)
{
    const double az = -37.5 * M_PI / 180, el = 30.0 * M_PI / 180;
    sx = std::cos(az) * x + std::sin(az) * y;
    sy = -std::sin(el) * std::sin(az) * x + std::sin(el) * std::cos(az) * y + std::cos(el) * z;
    depth = std::cos(el) * std::sin(az) * x - std::cos(el) * std::cos(az) * y + std::sin(el) * z;
}

std::string xml_escape(const std::string &t)
{
    std::string out;
    for (char c : t) {
        if (c == '&') out += "&amp;";
        else if (c == '<') out += "&lt;";
        else if (c == '>') out += "&gt;";
        else out += c;
    }
    return out;
}

std::string hex_rgb(const double rgb[3])
{
    char b[8];
    auto c = [](double v) { return (int)std::lround(std::min(1.0, std::max(0.0, v)) * 255); };
    std::snprintf(b, sizeof(b), "#%02x%02x%02x", c(rgb[0]), c(rgb[1]), c(rgb[2]));
    return b;
}

int write_svg(const std::string &path, const std::string &title, const std::vector<Atom> &atoms,
              const std::vector<std::string> &elems, const std::vector<Style> &styles)
{
    constexpr double kMarkerSize = 5.0;            // viewdeb.m:69
    constexpr double kW = 640, kH = 560, kPlot = 440, kLeft = 60, kTop = 60;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return DBX_ERR_IO;
    std::fprintf(f,
                 "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
                 "<svg xmlns=\"http://www.w3.org/2000/svg\" width=\"%gpt\" height=\"%gpt\" "
                 "viewBox=\"0 0 %g %g\">\n",
                 kW, kH, kW, kH);
    std::fprintf(f, "<title>%s</title>\n<rect width=\"100%%\" height=\"100%%\" fill=\"white\"/>\n",
                 xml_escape(title).c_str());
    // the axes span the kept atoms (viewdeb.m:93-94); every axis on the same scale
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (size_t i = 0; i < atoms.size(); i++) {
        const double v[3] = {atoms[i].x, atoms[i].y, atoms[i].z};
        for (int a = 0; a < 3; a++) {
            lo[a] = i ? std::min(lo[a], v[a]) : v[a];
            hi[a] = i ? std::max(hi[a], v[a]) : v[a];
        }
    }
    double smin[2] = {1e300, 1e300}, smax[2] = {-1e300, -1e300};
    std::vector<std::array<double, 3>> corner;
    for (int c = 0; c < 8; c++) {
        double sx, sy, d;
        project((c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2], sx, sy, d);
        corner.push_back({sx, sy, d});
        smin[0] = std::min(smin[0], sx), smax[0] = std::max(smax[0], sx);
        smin[1] = std::min(smin[1], sy), smax[1] = std::max(smax[1], sy);
    }
    const double span = std::max({smax[0] - smin[0], smax[1] - smin[1], 1e-9});
    const double scale = kPlot / span;
    // the projected box centred in the plot square
    const double ox = 0.5 * (kPlot - (smax[0] - smin[0]) * scale), oy = 0.5 * (kPlot - (smax[1] - smin[1]) * scale);
    auto px = [&](double sx) { return kLeft + ox + (sx - smin[0]) * scale; };
    auto py = [&](double sy) { return kTop + kPlot - oy - (sy - smin[1]) * scale; };
    // the box of the axis limits
    std::fprintf(f, "<g stroke=\"#b0b0b0\" stroke-width=\"0.5\" fill=\"none\">\n");
    for (int c = 0; c < 8; c++)
        for (int b = 0; b < 3; b++) {
            const int d = c | (1 << b);
            if (d == c) continue;
            std::fprintf(f, "<line x1=\"%.2f\" y1=\"%.2f\" x2=\"%.2f\" y2=\"%.2f\"/>\n", px(corner[c][0]),
                         py(corner[c][1]), px(corner[d][0]), py(corner[d][1]));
        }
    std::fprintf(f, "</g>\n");
    // the axis labels at the middle of one edge per axis, as scatter3's xlabel/ylabel/zlabel
    const char *label[3] = {"x [&#197;]", "y [&#197;]", "z [&#197;]"};
    const int edge[3][2] = {{0, 1}, {1, 3}, {0, 4}};
    for (int a = 0; a < 3; a++) {
        const double mx = 0.5 * (corner[edge[a][0]][0] + corner[edge[a][1]][0]);
        const double my = 0.5 * (corner[edge[a][0]][1] + corner[edge[a][1]][1]);
        std::fprintf(f,
                     "<text x=\"%.2f\" y=\"%.2f\" font-family=\"sans-serif\" font-size=\"11\" "
                     "fill=\"#404040\" text-anchor=\"middle\">%s</text>\n",
                     px(mx) + (a == 2 ? -18 : 0), py(my) + (a == 2 ? 0 : 16), label[a]);
    }
    // the atoms, farthest first, each a disc of diameter 2 r MarkerSize with a black edge
    std::vector<std::pair<double, size_t>> order(atoms.size());
    for (size_t i = 0; i < atoms.size(); i++) {
        double sx, sy, d;
        project(atoms[i].x, atoms[i].y, atoms[i].z, sx, sy, d);
        order[i] = {d, i};
    }
    std::stable_sort(order.begin(), order.end());
    std::fprintf(f, "<g stroke=\"black\" stroke-width=\"0.5\">\n");
    for (const auto &o : order) {
        const Atom &at = atoms[o.second];
        const Style &st = styles[at.elem];
        if (st.radius <= 0) continue;
        double sx, sy, d;
        project(at.x, at.y, at.z, sx, sy, d);
        std::fprintf(f, "<circle cx=\"%.2f\" cy=\"%.2f\" r=\"%.3f\" fill=\"%s\"/>\n", px(sx), py(sy),
                     st.radius * kMarkerSize, hex_rgb(st.rgb).c_str());
    }
    std::fprintf(f, "</g>\n");
    // the elements in their colours beside the plot (viewdeb.m:88-90)
    for (size_t e = 0; e < elems.size(); e++)
        std::fprintf(f,
                     "<text x=\"%g\" y=\"%g\" font-family=\"sans-serif\" font-size=\"24\" "
                     "font-weight=\"bold\" fill=\"%s\">%s</text>\n",
                     kLeft + kPlot + 40, kTop + 24 + 44.0 * (double)e, hex_rgb(styles[e].rgb).c_str(),
                     xml_escape(elems[e]).c_str());
    std::fprintf(f, "</svg>\n");
    return std::fclose(f) == 0 ? DBX_OK : DBX_ERR_IO;
}

int write_xyz(const std::string &path, const std::string &title, const std::vector<Atom> &atoms,
              const std::vector<std::string> &elems)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return DBX_ERR_IO;
    std::fprintf(f, "%zu\n%s\n", atoms.size(), title.c_str());  // count, title line
    for (const Atom &a : atoms)
        std::fprintf(f, "%s %.7g %.7g %.7g\n", elems[a.elem].c_str(), a.x, a.y, a.z);
    return std::fclose(f) == 0 ? DBX_OK : DBX_ERR_IO;
}

int view_body(const char *deb_path, const double *limits, const char *svg_path, const char *xyz_path,
              const char *radii_path, int64_t *n_shown)
{
    if (!deb_path || (!svg_path && !xyz_path)) return DBX_ERR_ARG;
    DebFile d;
    std::string err;
    int rc = parse_deb(deb_path, d, err);
    if (rc != DBX_OK) {
        std::fprintf(stderr, "ERROR: %s\n", err.c_str());
        return rc;
    }
    std::map<std::string, Style> table;
    const std::string tpath = radii_path ? radii_path : lib_dir() + "/data/atomic_radii.txt";
    if (svg_path && !load_styles(tpath, table)) {
        std::fprintf(stderr, "ERROR: cannot read the radius table %s\n", tpath.c_str());
        return DBX_ERR_IO;
    }
    // the elements of the subgroups, first appearance order, and their styles
    std::vector<std::string> elems;
    std::vector<Style> styles;
    std::vector<int> group_elem(d.S);
    for (int g = 0; g < d.S; g++) {
        const std::string e = element_of_type(d.type[g]);
        auto it = std::find(elems.begin(), elems.end(), e);
        if (it == elems.end()) {
            elems.push_back(e);
            auto t = table.find(lower(e));
            styles.push_back(t == table.end() ? Style{} : t->second);
            it = elems.end() - 1;
        }
        group_elem[g] = (int)(it - elems.begin());
    }
    double lim[6] = {-10, 10, -10, 10, -10, 10};  // viewdeb.m:17,25
    if (limits) std::copy(limits, limits + 6, lim);
    std::vector<Atom> atoms;
    int64_t i = 0;
    for (int g = 0; g < d.S; g++)
        for (int64_t k = 0; k < d.n[g]; k++, i++) {
            const double v[3] = {d.x[i], d.y[i], d.z[i]};
            bool in = true;
            for (int a = 0; a < 3; a++) in = in && v[a] > lim[2 * a] && v[a] < lim[2 * a + 1];
            if (in) atoms.push_back({v[0], v[1], v[2], group_elem[g]});
        }
    std::string title = deb_path;
    const size_t slash = title.rfind('/');
    if (slash != std::string::npos) title = title.substr(slash + 1);
    if (svg_path && (rc = write_svg(svg_path, title, atoms, elems, styles)) != DBX_OK) return rc;
    if (xyz_path && (rc = write_xyz(xyz_path, title, atoms, elems)) != DBX_OK) return rc;
    if (n_shown) *n_shown = (int64_t)atoms.size();
    return DBX_OK;
}

}  // namespace

}  // namespace dbx

extern "C" int dbx_view_deb(const char *deb_path, const double *limits, const char *svg_path, const char *xyz_path,
                            const char *radii_path, int64_t *n_shown)
{
    try {
        return dbx::view_body(deb_path, limits, svg_path, xyz_path, radii_path, n_shown);
    } catch (const std::bad_alloc &) {
        return DBX_ERR_NOMEM;
    } catch (const std::exception &) {
        return DBX_ERR_ARG;
    }
}

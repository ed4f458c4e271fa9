// `debye file.deb [--crlf] [--quiet] [--mfma] [--from-debx]`: the command-line program of
// cuDebye (cuDebye.exe <file.deb>, cuDebye-v15.cu:599-614) on top of libdebye.
// `debye --folder DIR [...]`: every .xyz in DIR without a .deb converted, then every .deb without
// a .debI calculated (Matlab/multisim.m), served together.
// `debye --xyz2deb FILE.xyz [--table PATH] [--rstep R]`: the converter (Matlab/xyz2deb.m).
// `debye --average [--norm none|int|max] [--step S] -o OUT A.debI B.debI ...`: the averaged
// pattern of several .debI files (Matlab/plotdebi.m), written as "%.10g\t%.10g" lines.
// `debye --view FILE.deb [--limits LO HI | --limits XLO XHI YLO YHI ZLO ZHI] [--svg OUT]
// [--xyz OUT] [--radii PATH]`: the structure view (Matlab/viewdeb.m); without --svg or --xyz it
// draws FILE.deb.svg.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/debye.h"

static int usage(const char *prog)
{
    std::fprintf(stderr,
                 "usage: %s file.deb [--crlf] [--quiet] [--mfma] [--from-debx]\n"
                 "       %s --folder DIR [--crlf] [--quiet] [--mfma]\n"
                 "       %s --xyz2deb FILE.xyz [--table PATH] [--rstep R]\n"
                 "       %s --average [--norm none|int|max] [--step S] -o OUT A.debI B.debI ...\n"
                 "       %s --view FILE.deb [--limits LO HI | --limits XLO XHI YLO YHI ZLO ZHI] [--svg OUT] "
                 "[--xyz OUT] [--radii PATH]\n",
                 prog, prog, prog, prog, prog);
    return 2;
}

static int average(int argc, char **argv)
{
    int norm = DBX_AVG_NONE;
    double step = 0.02;
    const char *out = nullptr;
    std::vector<const char *> files;
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--norm") && i + 1 < argc) {
            const char *v = argv[++i];
            if (!std::strcmp(v, "int") || !std::strcmp(v, "integral"))
                norm = DBX_AVG_INTEGRAL;
            else if (!std::strcmp(v, "max") || !std::strcmp(v, "maximum"))
                norm = DBX_AVG_MAX;
            else if (std::strcmp(v, "none") != 0)
                return usage(argv[0]);
        } else if (!std::strcmp(argv[i], "--step") && i + 1 < argc) {
            step = std::atof(argv[++i]);
        } else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) {
            out = argv[++i];
        } else {
            files.push_back(argv[i]);
        }
    }
    if (!out || files.empty()) return usage(argv[0]);
    int n = 0;
    int rc = dbx_average_debi(files.data(), (int)files.size(), norm, step, nullptr, nullptr, 0, &n);
    if (rc != DBX_OK) return 1;
    std::vector<double> tt(n), I(n);
    rc = dbx_average_debi(files.data(), (int)files.size(), norm, step, tt.data(), I.data(), n, &n);
    if (rc != DBX_OK) return 1;
    FILE *f = std::fopen(out, "w");
    if (!f) {
        std::fprintf(stderr, "ERROR: cannot write %s\n", out);
        return 1;
    }
    for (int k = 0; k < n; k++) std::fprintf(f, "%.10g\t%.10g\n", tt[k], I[k]);
    return std::fclose(f) == 0 ? 0 : 1;
}

static int view(int argc, char **argv)
{
    const char *deb = nullptr, *svg = nullptr, *xyz = nullptr, *radii = nullptr;
    std::vector<double> lim;
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--limits")) {
            // two numbers (every axis) or six (per axis)
            while (i + 1 < argc && lim.size() < 6) {
                char *e = nullptr;
                const double v = std::strtod(argv[i + 1], &e);
                if (e == argv[i + 1] || *e) break;
                lim.push_back(v);
                i++;
            }
            if (lim.size() != 2 && lim.size() != 6) return usage(argv[0]);
        } else if (!std::strcmp(argv[i], "--svg") && i + 1 < argc) {
            svg = argv[++i];
        } else if (!std::strcmp(argv[i], "--xyz") && i + 1 < argc) {
            xyz = argv[++i];
        } else if (!std::strcmp(argv[i], "--radii") && i + 1 < argc) {
            radii = argv[++i];
        } else {
            deb = argv[i];
        }
    }
    if (!deb) return usage(argv[0]);
    std::string svg_default;
    if (!svg && !xyz) {
        svg_default = std::string(deb) + ".svg";
        svg = svg_default.c_str();
    }
    if (lim.size() == 2) lim = {lim[0], lim[1], lim[0], lim[1], lim[0], lim[1]};
    int64_t n = 0;
    const int rc = dbx_view_deb(deb, lim.empty() ? nullptr : lim.data(), svg, xyz, radii, &n);
    if (rc != DBX_OK) return 1;
    std::printf("%lld atoms inside the limits\n", (long long)n);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && !std::strcmp(argv[1], "--average")) return average(argc, argv);
    if (argc > 1 && !std::strcmp(argv[1], "--view")) return view(argc, argv);
    const char *path = nullptr;
    const char *table = nullptr;
    double rstep = 0;
    int flags = 0;
    bool folder = false, convert = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--crlf"))
            flags |= DBX_RUN_CRLF;
        else if (!std::strcmp(argv[i], "--quiet"))
            flags |= DBX_RUN_QUIET;
        else if (!std::strcmp(argv[i], "--mfma"))
            flags |= DBX_RUN_MFMA;
        else if (!std::strcmp(argv[i], "--from-debx"))
            flags |= DBX_RUN_FROM_DEBX;
        else if (!std::strcmp(argv[i], "--folder"))
            folder = true;
        else if (!std::strcmp(argv[i], "--xyz2deb"))
            convert = true;
        else if (!std::strcmp(argv[i], "--table") && i + 1 < argc)
            table = argv[++i];
        else if (!std::strcmp(argv[i], "--rstep") && i + 1 < argc)
            rstep = std::atof(argv[++i]);
        else
            path = argv[i];
    }
    if (!path) return usage(argv[0]);
    if (convert) {
        dbx_deb_header h;
        dbx_deb_header_defaults(&h);
        if (rstep > 0) h.rstep = rstep;
        return dbx_xyz_to_deb(path, nullptr, table, &h) == DBX_OK ? 0 : 1;
    }
    const int rc = folder ? dbx_run_folder(path, flags & ~DBX_RUN_FROM_DEBX) : dbx_run_deb(path, flags);
    return rc == DBX_OK ? 0 : 1;
}

// A parsed .deb model file (README.md "Input File"; cuDebye-v15.cu:624-833), shared by the .deb
// driver and the structure view.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/debye.h"

namespace dbx {

struct DebFile {
    int device = 0, block = 0, grid = 0, hist = 0, S = 0;
    double lambda = 0, tt_min = 0, tt_max = 0, tt_step = 0, rmax = 0, rstep = 0;
    std::vector<int64_t> n;           // atoms per subgroup
    std::vector<std::string> type;    // the subgroup lines' type field
    std::vector<dbx_group_params> gp;
    std::vector<float> x, y, z;       // positions in subgroup order
};

// header lines 1-4, S subgroup lines, then x y z lines (blank lines skipped); DBX_OK, DBX_ERR_IO
// (missing file) or DBX_ERR_PARSE (err says why)
int parse_deb(const char *path, DebFile &d, std::string &err);

}  // namespace dbx

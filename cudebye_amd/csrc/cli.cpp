// `debye file.deb [--crlf] [--quiet] [--mfma] [--from-debx]`: the command-line program of
// cuDebye (cuDebye.exe <file.deb>, cuDebye-v15.cu:599-614) on top of libdebye.
// `debye --folder DIR [...]`: every .deb in DIR without a .debI (Matlab/multisim.m), served
// together.
#include <cstdio>
#include <cstring>

#include "../../include/debye.h"

int main(int argc, char **argv)
{
    const char *path = nullptr;
    int flags = 0;
    bool folder = false;
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
        else
            path = argv[i];
    }
    if (!path) {
        std::fprintf(stderr, "usage: %s file.deb [--crlf] [--quiet] [--mfma] [--from-debx]\n"
                             "       %s --folder DIR [--crlf] [--quiet] [--mfma]\n", argv[0], argv[0]);
        return 2;
    }
    const int rc = folder ? dbx_run_folder(path, flags & ~DBX_RUN_FROM_DEBX) : dbx_run_deb(path, flags);
    return rc == DBX_OK ? 0 : 1;
}

"""Writes the golden fixtures under tests/golden/ from the reference's own data files.

The fixtures are data only (inputs and expected outputs of the reference's example
runs, Data/ of Martin-Rudolph/cuDebye v1.5):
  gamma/GammaAl2O3_10nm.deb.gz   the input model (117,936 atoms, 3 subgroups), gzip of the text
  gamma/debx.npz                 the six partial histograms of Data/GammaAl2O3_10nm.debx<r>x<c>
                                 as (bin index int32, count uint32) per pair
  gamma/debx_sha256.json         sha256 and byte length of each full .debx file (CRLF text)
  gamma/GammaAl2O3_10nm.debI.gz  the reference intensity file (gzip of the CRLF text)
  gamma/log_values.json          numbers printed in Data/cmd-logfile.txt for this run
  uh3/UH3_7nm.deb.gz, uh3/UH3_7nm.debI.gz   the second example (36,903 atoms, 2 subgroups)
  uh3/UH3_7nm.xyz.gz, uh3/UH3_7nm_reduced.xyz.gz   the converter's inputs for that .deb (six
                                 and four columns; Data/UH3_7nm_reduced.deb is identical to
                                 Data/UH3_7nm.deb)

Run in the build container only (it reads /root/reference/Data):
    python tests/golden/make_fixtures.py
"""
import gzip
import hashlib
import json
import os

import numpy as np

SRC = "/root/reference/Data"
HERE = os.path.dirname(os.path.abspath(__file__))


def gz_copy(src, dst):
    with open(src, "rb") as f:
        data = f.read()
    with gzip.GzipFile(dst, "wb", mtime=0) as g:
        g.write(data)


def main():
    os.makedirs(os.path.join(HERE, "gamma"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "uh3"), exist_ok=True)
    gz_copy(f"{SRC}/GammaAl2O3_10nm.deb", f"{HERE}/gamma/GammaAl2O3_10nm.deb.gz")
    gz_copy(f"{SRC}/GammaAl2O3_10nm.debI", f"{HERE}/gamma/GammaAl2O3_10nm.debI.gz")
    gz_copy(f"{SRC}/UH3_7nm.deb", f"{HERE}/uh3/UH3_7nm.deb.gz")
    gz_copy(f"{SRC}/UH3_7nm.debI", f"{HERE}/uh3/UH3_7nm.debI.gz")
    gz_copy(f"{SRC}/UH3_7nm.xyz", f"{HERE}/uh3/UH3_7nm.xyz.gz")
    gz_copy(f"{SRC}/UH3_7nm_reduced.xyz", f"{HERE}/uh3/UH3_7nm_reduced.xyz.gz")
    with open(f"{SRC}/UH3_7nm_reduced.deb", "rb") as a, open(f"{SRC}/UH3_7nm.deb", "rb") as b:
        assert a.read() == b.read()
    rstep = 1e-4
    arrays, shas = {}, {}
    for r in range(3):
        for c in range(r, 3):
            path = f"{SRC}/GammaAl2O3_10nm.debx{r}x{c}"
            with open(path, "rb") as f:
                raw = f.read()
            shas[f"{r}x{c}"] = {"sha256": hashlib.sha256(raw).hexdigest(), "bytes": len(raw),
                                "lines": raw.count(b"\n")}
            d = np.loadtxt(path, dtype=np.float64)
            k = np.rint(d[:, 0] / rstep).astype(np.int64)
            assert np.all(np.abs(d[:, 0] - k * rstep) < 1e-9)
            assert np.all(np.diff(k) > 0)
            arrays[f"bins_{r}x{c}"] = k.astype(np.int32)
            arrays[f"counts_{r}x{c}"] = d[:, 1].astype(np.uint32)
    np.savez_compressed(f"{HERE}/gamma/debx.npz", **arrays)
    with open(f"{HERE}/gamma/debx_sha256.json", "w") as f:
        json.dump(shas, f, indent=1, sort_keys=True)
    # values printed by the reference run (Data/cmd-logfile.txt)
    log = {
        "source": "Data/cmd-logfile.txt",
        "nTT": 6501,                      # log:51
        "K_first": "0.710359",           # log:52 (printed with 6 significant digits)
        "K_last": "7.65892",              # log:52
        "channels": 3560002,              # log:58
        "unique_distances": {"0x0": 71767, "0x1": 71762, "0x2": 120216, "1x1": 61523,
                             "1x2": 62020, "2x2": 23371},
        "distances_total": 6954509016,
        # "Maximum Frequency of Distances: <count> at <bin>", log:198,275,333,394,440,480
        "max_frequency": {"0x0": [721964, 358475], "0x1": [716524, 360072],
                          "0x2": [238025, 254797], "1x1": [239762, 322533],
                          "1x2": [210185, 267967], "2x2": [122959, 340980]},
    }
    with open(f"{HERE}/gamma/log_values.json", "w") as f:
        json.dump(log, f, indent=1)


if __name__ == "__main__":
    main()

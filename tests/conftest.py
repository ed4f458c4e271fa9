"""Shared pytest configuration: the `gpu` marker and access to the golden fixtures."""
import gzip
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than ~10 s on the CPU")


def _gunzip_to(src, dst):
    if not os.path.exists(dst):
        with gzip.open(src, "rb") as f, open(dst + ".tmp", "wb") as g:
            g.write(f.read())
        os.replace(dst + ".tmp", dst)
    return dst


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    """A temp dir holding the decompressed reference example inputs/outputs."""
    d = tmp_path_factory.mktemp("golden")
    _gunzip_to(os.path.join(GOLDEN, "gamma", "GammaAl2O3_10nm.deb.gz"), str(d / "GammaAl2O3_10nm.deb"))
    _gunzip_to(os.path.join(GOLDEN, "gamma", "GammaAl2O3_10nm.debI.gz"), str(d / "GammaAl2O3_10nm.debI.ref"))
    _gunzip_to(os.path.join(GOLDEN, "uh3", "UH3_7nm.deb.gz"), str(d / "UH3_7nm.deb"))
    _gunzip_to(os.path.join(GOLDEN, "uh3", "UH3_7nm.debI.gz"), str(d / "UH3_7nm.debI.ref"))
    _gunzip_to(os.path.join(GOLDEN, "uh3", "UH3_7nm.xyz.gz"), str(d / "UH3_7nm.xyz"))
    _gunzip_to(os.path.join(GOLDEN, "uh3", "UH3_7nm_reduced.xyz.gz"), str(d / "UH3_7nm_reduced.xyz"))
    return d


@pytest.fixture(scope="session")
def gamma_debx():
    """dict 'rxc' -> (bins int64, counts uint64) of Data/GammaAl2O3_10nm.debx<r>x<c>."""
    z = np.load(os.path.join(GOLDEN, "gamma", "debx.npz"))
    out = {}
    for r in range(3):
        for c in range(r, 3):
            out[f"{r}x{c}"] = (z[f"bins_{r}x{c}"].astype(np.int64), z[f"counts_{r}x{c}"].astype(np.uint64))
    return out


@pytest.fixture(scope="session")
def gamma_log():
    with open(os.path.join(GOLDEN, "gamma", "log_values.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gamma_debx_sha():
    with open(os.path.join(GOLDEN, "gamma", "debx_sha256.json")) as f:
        return json.load(f)


def read_debi(path):
    """(2theta strings, intensity strings) of a .debI file, line ends ignored."""
    with open(path, "rb") as f:
        lines = f.read().decode().replace("\r", "").splitlines()
    return lines

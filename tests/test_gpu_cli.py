"""The .deb driver end to end on the GPU (`debye file.deb`, the reference's cuDebye.exe):
the reference's example inputs give its example outputs."""
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

import cudebye_amd as cd
from tests.conftest import read_debi

pytestmark = pytest.mark.gpu


def test_gamma_cli_writes_reference_debx_and_debI(golden_dir, tmp_path, gamma_debx_sha):
    deb = tmp_path / "GammaAl2O3_10nm.deb"
    shutil.copy(golden_dir / "GammaAl2O3_10nm.deb", deb)
    r = subprocess.run([cd.CLI_PATH, str(deb), "--crlf"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert "WARNING" not in r.stdout
    for key, ref in gamma_debx_sha.items():
        with open(f"{deb}x{key}", "rb") as f:
            raw = f.read()
        assert hashlib.sha256(raw).hexdigest() == ref["sha256"], key
    got = read_debi(f"{deb}I")
    assert got == read_debi(str(golden_dir / "GammaAl2O3_10nm.debI.ref"))


def test_uh3_run_deb_matches_reference_debI(golden_dir, tmp_path):
    deb = tmp_path / "UH3_7nm.deb"
    shutil.copy(golden_dir / "UH3_7nm.deb", deb)
    assert cd.run_deb(str(deb)) == 0
    assert not os.path.exists(f"{deb}x0x0")  # hist = 0: intensity only
    got = read_debi(f"{deb}I")
    ref = read_debi(str(golden_dir / "UH3_7nm.debI.ref"))
    assert len(got) == len(ref)
    bad = [(a, b) for a, b in zip(ref, got) if a != b]
    for a, b in bad:  # a last printed digit may flip with summation order
        va, vb = float(a.split("\t")[1]), float(b.split("\t")[1])
        assert abs(va - vb) <= 1.0001e-5 * abs(va), (a, b)
    assert len(bad) <= 3


def test_cli_errors(tmp_path):
    r = subprocess.run([cd.CLI_PATH, str(tmp_path / "missing.deb")], capture_output=True, text=True)
    assert r.returncode != 0 and "does not exist" in r.stderr
    bad = tmp_path / "bad.deb"
    bad.write_text("0 8 1024\n1.5418 10 140 0.02\n50 0.001 0\n1\n3 X 1 0 1 1 1 1 1 1 1 1 1 1 0\n0 0 0\n1 1 1\n")
    r = subprocess.run([cd.CLI_PATH, str(bad)], capture_output=True, text=True)
    assert r.returncode != 0 and "match" in r.stderr


def test_gamma_cli_mfma_pass2_within_bound(golden_dir, tmp_path):
    """--mfma: pass 2 on the matrix cores; every printed I(Q) within one unit of the 6th
    significant digit of the reference's printed value (both files round to 6 digits)."""
    deb = tmp_path / "GammaAl2O3_10nm.deb"
    shutil.copy(golden_dir / "GammaAl2O3_10nm.deb", deb)
    r = subprocess.run([cd.CLI_PATH, str(deb), "--mfma", "--quiet"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    got = np.array([float(l.split("\t")[1]) for l in read_debi(f"{deb}I")])
    ref = np.array([float(l.split("\t")[1]) for l in read_debi(str(golden_dir / "GammaAl2O3_10nm.debI.ref"))])
    assert len(got) == len(ref) == 6501
    rel = np.abs(got - ref) / np.abs(ref)
    worst = np.argsort(-rel)[:5]
    assert rel.max() <= 1.0001e-5, [(int(i), got[i], ref[i]) for i in worst]

"""The .deb driver end to end on the GPU (`debye file.deb`, the reference's cuDebye.exe):
the reference's example inputs give its example outputs."""
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth
from oracle import oracle as O
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


def test_gamma_pass2_from_written_debx(golden_dir, tmp_path, gamma_debx):
    """--from-debx: pass 2 only, from the reference's own .debx histograms (written in its
    exact format next to the .deb) -> the reference's .debI."""
    deb = tmp_path / "GammaAl2O3_10nm.deb"
    shutil.copy(golden_dir / "GammaAl2O3_10nm.deb", deb)
    g = O.parse_deb(str(deb))
    for key, (bins, cnt) in gamma_debx.items():
        O.write_debx(f"{deb}x{key}", bins, cnt, g.rstep, crlf=True)
    before = {k: open(f"{deb}x{k}", "rb").read() for k in gamma_debx}
    r = subprocess.run([cd.CLI_PATH, str(deb), "--from-debx", "--crlf"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    assert read_debi(f"{deb}I") == read_debi(str(golden_dir / "GammaAl2O3_10nm.debI.ref"))
    assert {k: open(f"{deb}x{k}", "rb").read() for k in gamma_debx} == before  # inputs untouched


def _write_deb(path, m, hist=2, block=16):
    tt0, tt1, n = 10.0, 150.0, 300
    step = (tt1 - tt0) / (n - 1) * (1 + 1e-12)
    rmax = (m.n_bins - 2) * m.rstep
    lines = ["0 %d 1" % block, f"1.5406 {tt0} {tt1} {step!r}", f"{rmax!r} {m.rstep!r} {hist}", str(len(m.group_n))]
    for g, p in zip(m.group_n, m.params):
        a = list(p["a"]) + [0.0] * (5 - len(p["a"]))
        b = list(p["b"]) + [0.0] * (5 - len(p["b"]))
        lines.append(" ".join([str(int(g)), "X", repr(float(p["occ"])), repr(float(p["beq"]))]
                              + [repr(float(v)) for v in a] + [repr(float(v)) for v in b] + [repr(float(p["c"]))]))
    for x, y, z in zip(m.x, m.y, m.z):
        lines.append("%.9g %.9g %.9g" % (x, y, z))
    path.write_text("\n".join(lines) + "\n")


def test_folder_run_equals_each_file_alone(tmp_path):
    """--folder (Matlab/multisim.m): every .xyz without a .deb converted, then every .deb without a
    .debI, through one server; each file's .debx equal byte for byte and its .debI within one
    unit of the printed sixth digit of a run of that file alone (the FP64 sums of a batch split
    their bins differently)."""
    models = [synth.small_random(n, S, 900 + i, spread=sp) for i, (n, S, sp) in
              enumerate([(400, 1, 6.0), (1500, 3, 12.0), (900, 2, 9.0), (2500, 4, 14.0)])]
    solo, batch = tmp_path / "solo", tmp_path / "batch"
    solo.mkdir()
    batch.mkdir()
    for i, m in enumerate(models):
        for d in (solo, batch):
            _write_deb(d / f"m{i}.deb", m)
    (batch / "m3.debI").write_text("already calculated\n")  # skipped, as multisim.m does
    # an .xyz without its .deb: converted first with the default header, then calculated
    rng = np.random.default_rng(77)
    xyz = "".join(f"{e}  {x:.6f}  {y:.6f}  {z:.6f}\n" for e, (x, y, z) in
                  zip(rng.choice(["O", "Al"], 300), rng.uniform(-8.0, 8.0, (300, 3))))
    for d in (solo, batch):
        (d / "x5.xyz").write_text(xyz)
    cd.xyz_to_deb(solo / "x5.xyz")
    for i in range(len(models)):
        assert cd.run_deb(solo / f"m{i}.deb") == 0
    assert cd.run_deb(solo / "x5.deb") == 0
    r = subprocess.run([cd.CLI_PATH, "--folder", str(batch)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert (batch / "m3.debI").read_text() == "already calculated\n"
    assert not (batch / "m3.debx0x0").exists()
    for i in range(3):
        S = len(models[i].group_n)
        for a in range(S):
            for b in range(a, S):
                assert (batch / f"m{i}.debx{a}x{b}").read_bytes() == (solo / f"m{i}.debx{a}x{b}").read_bytes()
        got = np.array([[float(v) for v in l.split("\t")] for l in read_debi(str(batch / f"m{i}.debI"))])
        ref = np.array([[float(v) for v in l.split("\t")] for l in read_debi(str(solo / f"m{i}.debI"))])
        assert got.shape == ref.shape and np.array_equal(got[:, 0], ref[:, 0])
        assert np.max(np.abs(got[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) <= 1.0001e-5
    assert (batch / "x5.deb").read_bytes() == (solo / "x5.deb").read_bytes()
    got = np.array([[float(v) for v in l.split("\t")] for l in read_debi(str(batch / "x5.debI"))])
    ref = np.array([[float(v) for v in l.split("\t")] for l in read_debi(str(solo / "x5.debI"))])
    assert got.shape == ref.shape == (6501, 2)
    assert np.max(np.abs(got[:, 1] - ref[:, 1]) / np.abs(ref[:, 1])) <= 1.0001e-5

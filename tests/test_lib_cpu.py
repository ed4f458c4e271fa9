"""CPU-side checks of the product library: it loads, exports every entry point that
include/debye.h declares, and its host-only grid functions follow the reference's rules
(cuDebye-v15.cu:690-698, 734).  No device call is made here."""
import os
import re

import numpy as np
import pytest

import cudebye_amd as cd
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dbx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = cd.lib()
    names = _declared("debye.h")
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(cd.EXPORTS)
    for n in _declared("debye_debug.h"):
        assert hasattr(L, n), n


def test_cli_built():
    assert os.access(cd.CLI_PATH, os.X_OK)


@pytest.mark.parametrize("hdr", [(1.5418, 10.0, 140.0, 0.02), (1.5406, 10.0, 150.0, 140.0 / 4095),
                                 (0.7093, 2.0, 120.0, 0.005), (1.0, 0.0, 180.0, 1.0)])
def test_q_grid_matches_oracle(hdr):
    tt, q = cd.q_from_2theta(*hdr)
    ott, oq = O.q_from_2theta(*hdr)
    assert np.array_equal(tt, ott) and np.array_equal(q, oq)


@pytest.mark.parametrize("rmax,rstep", [(356.0, 1e-4), (122.0, 1e-4), (104.0, 1e-3), (208.0, 5e-4),
                                        (0.0, 1e-3)])
def test_channels_matches_oracle(rmax, rstep):
    assert cd.channels(rmax, rstep) == O.channels(rmax, rstep)


def test_bad_grids_rejected():
    with pytest.raises(cd.DebyeError):
        cd.q_from_2theta(1.5, 10.0, 140.0, 0.0)
    with pytest.raises(cd.DebyeError):
        cd.channels(10.0, -1.0)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(cd, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(cd, "_lib", None)
    with pytest.raises(ImportError):
        cd.lib()


def test_read_debx_round_trips_reference_files(tmp_path, golden_dir, gamma_debx):
    """The reader of .debx files (host only): the reference's Gamma histograms, written in its
    exact format (the oracle writer reproduces the six files byte for byte), read back."""
    g = O.parse_deb(str(golden_dir / "GammaAl2O3_10nm.deb"))
    nb = O.channels(g.rmax, g.rstep)
    for key, (bins, cnt) in gamma_debx.items():
        path = tmp_path / f"g.debx{key}"
        O.write_debx(str(path), bins, cnt, g.rstep, crlf=True)
        h = cd.read_debx(path, g.rstep, nb)
        ref = np.zeros(nb, np.uint64)
        ref[bins] = cnt
        assert np.array_equal(h, ref), key


def test_read_debx_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.debx"
    p.write_text("1.2345\t7\n")          # not a bin centre at 0.001 A... 1.2345 / 0.001 = 1234.5
    with pytest.raises(cd.DebyeError) as e:
        cd.read_debx(p, 1e-3, 5000)
    assert e.value.code == -4
    p.write_text("4.5\t3\n")
    with pytest.raises(cd.DebyeError) as e:
        cd.read_debx(p, 1e-3, 1000)         # bin 4500 >= 1000
    assert e.value.code == cd.DBX_ERR_RANGE
    p.write_text("0.001\t3\n0.002\t5\n\n")
    h = cd.read_debx(p, 1e-3, 10)
    assert h[1] == 3 and h[2] == 5 and h.sum() == 8

"""Pins the CPU oracle against the reference's own example outputs.

Gate 1: the oracle's pass 1 (FP32 bin rule, cuDebye-v15.cu:101-115) reproduces the six
        Data/GammaAl2O3_10nm.debx* histograms bin for bin (6,954,509,016 pairs).
Gate 2: the oracle's pass 2 (cuDebye-v15.cu:447-536) fed the .debx fixtures and the
        Gamma header reproduces Data/GammaAl2O3_10nm.debI line for line.
End to end: Data/UH3_7nm.deb -> Data/UH3_7nm.debI.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import read_debi


@pytest.fixture(scope="module")
def gamma(golden_dir):
    return O.parse_deb(str(golden_dir / "GammaAl2O3_10nm.deb"))


def _dense(bins, counts, n):
    h = np.zeros(n, np.uint64)
    h[bins] = counts
    return h


def test_gamma_header_and_grids(gamma, gamma_log):
    m = gamma
    assert (m.device, m.block, m.grid, m.hist) == (0, 8, 1024, 2)
    assert list(m.group_n) == [67392, 33696, 16848]
    assert m.types == ["O2-", "Al3+", "Al3+"]
    tt, K = O.q_from_2theta(m.lam, m.tt_min, m.tt_max, m.tt_step)
    assert len(tt) == gamma_log["nTT"]
    assert f"{K[0]:g}" == gamma_log["K_first"] and f"{K[-1]:g}" == gamma_log["K_last"]
    assert O.channels(m.rmax, m.rstep) == gamma_log["channels"]
    assert O.rp_of(m.rstep) == np.float32(10000.0)


def test_gamma_fixture_invariants(gamma, gamma_debx, gamma_log):
    """The fixtures themselves satisfy the reference's closed forms (cuDebye-v15.cu:372,396,382)."""
    n = gamma.group_n
    total = 0
    for key, (bins, cnt) in gamma_debx.items():
        r, c = map(int, key.split("x"))
        s = int(cnt.sum())
        total += s
        assert s == (n[r] * (n[r] + 1) // 2 if r == c else n[r] * n[c])
        assert len(bins) == gamma_log["unique_distances"][key]
        mx = int(np.argmax(cnt))
        assert [int(cnt[mx]), int(bins[mx])] == gamma_log["max_frequency"][key]
        if r == c:
            assert bins[0] == 0 and cnt[0] == n[r]
    assert total == gamma_log["distances_total"]


@pytest.mark.parametrize("key", ["0x0", "0x1", "0x2", "1x1", "1x2", "2x2"])
def test_gate1_gamma_histograms_bit_exact(gamma, gamma_debx, key):
    m = gamma
    r, c = map(int, key.split("x"))
    nb = O.channels(m.rmax, m.rstep)
    st = m.starts
    h, ov = O.histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c,
                        O.rp_of(m.rstep), nb)
    assert ov == 0
    ref = _dense(*gamma_debx[key], nb)
    assert np.array_equal(h, ref)


def test_gate1_contracted_forms_do_not_match(gamma, gamma_debx):
    """Planted mistake: an FMA-contracted sum of squares (what a compiler may do to
    cuDebye-v15.cu:104) breaks the fixture replay, so the replay does decide the form."""
    m = gamma
    nb = O.channels(m.rmax, m.rstep)
    st = m.starts
    ref = _dense(*gamma_debx["2x2"], nb)
    for form in (1, 2):
        h, _ = O.histogram(m.x, m.y, m.z, st[2], m.group_n[2], st[2], m.group_n[2], True,
                           O.rp_of(m.rstep), nb, form=form)
        assert int(h.sum()) == int(ref.sum())
        assert np.count_nonzero(h != ref) > 1000


def test_debx_writer_bytes(tmp_path, gamma_debx, gamma_debx_sha):
    """.debx text (cuDebye-v15.cu:417-423), CRLF as in the fixtures: sha256 of the full file."""
    for key, (bins, cnt) in gamma_debx.items():
        p = str(tmp_path / f"x{key}")
        O.write_debx(p, bins, cnt, 1e-4, crlf=True)
        with open(p, "rb") as f:
            raw = f.read()
        assert len(raw) == gamma_debx_sha[key]["bytes"]
        assert hashlib.sha256(raw).hexdigest() == gamma_debx_sha[key]["sha256"]


def test_gate2_gamma_debI_from_fixture_histograms(gamma, gamma_debx, golden_dir, tmp_path):
    m = gamma
    tt, K = O.q_from_2theta(m.lam, m.tt_min, m.tt_max, m.tt_step)
    nb = O.channels(m.rmax, m.rstep)
    H = np.stack([_dense(*gamma_debx[f"{r}x{c}"], nb) for r, c in O.pair_list(m.group_n)])
    I = O.intensity_from_histograms(H, m, K, m.rstep, first_n=m.block)
    out = str(tmp_path / "g.debI")
    O.write_debi(out, tt, I)
    ref = read_debi(str(golden_dir / "GammaAl2O3_10nm.debI.ref"))
    got = read_debi(out)
    assert len(ref) == len(got) == 6501
    assert got == ref


def test_uh3_end_to_end(golden_dir, tmp_path):
    m = O.parse_deb(str(golden_dir / "UH3_7nm.deb"))
    assert list(m.group_n) == [27972, 8931] and m.hist == 0
    tt, K = O.q_from_2theta(m.lam, m.tt_min, m.tt_max, m.tt_step)
    H, ovs = O.all_histograms(m)
    assert ovs == [0, 0, 0]
    n = m.group_n
    assert [int(h.sum()) for h in H] == [n[0] * (n[0] + 1) // 2, n[0] * n[1], n[1] * (n[1] + 1) // 2]
    I = O.intensity_from_histograms(H, m, K, m.rstep, first_n=m.block)
    out = str(tmp_path / "u.debI")
    O.write_debi(out, tt, I)
    ref = read_debi(str(golden_dir / "UH3_7nm.debI.ref"))
    got = read_debi(out)
    assert len(got) == len(ref)
    # 6 significant digits printed; allow a last-digit flip from summation order only
    bad = [(a, b) for a, b in zip(ref, got) if a != b]
    for a, b in bad:
        va, vb = float(a.split("\t")[1]), float(b.split("\t")[1])
        assert abs(va - vb) <= 1.0001e-5 * abs(va), (a, b)
    assert len(bad) <= 3

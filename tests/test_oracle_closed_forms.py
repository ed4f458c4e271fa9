"""Pins the oracle where the fixtures leave it open: exact rational brute force of the
bin rule on small models (away from rounding ties), closed forms of the Debye sum, and
mpmath brute-force double sums over all atom pairs."""
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from oracle import oracle as O

mpmath.mp.dps = 40


def _rand_model(n, seed, spread=12.0):
    rng = np.random.default_rng(seed)
    return (rng.uniform(-spread, spread, size=(n, 3))).astype(np.float32)


def test_bin_rule_matches_exact_distance_away_from_ties():
    """bin = round(|a-b| / rstep) computed exactly from the float32 coordinates; the FP32
    chain of cuDebye-v15.cu:101-115 agrees wherever the exact value is farther from a
    rounding boundary than the chain's FP32 error bound (about 8 ulp relative, taken as
    1e-6 of the bin value)."""
    pts = _rand_model(300, 1)
    for rstep in (1e-3, 5e-4, 1e-4):
        rp = O.rp_of(rstep)
        checked = 0
        for i in range(0, 300, 7):
            for j in range(300):
                d2 = sum((Fraction(float(pts[i][t])) - Fraction(float(pts[j][t]))) ** 2 for t in range(3))
                d = mpmath.sqrt(mpmath.mpf(d2.numerator) / d2.denominator)
                v = d * mpmath.mpf(float(np.float32(rp)))
                frac = v - mpmath.floor(v)
                if abs(frac - mpmath.mpf("0.5")) < 1e-6 * v + 1e-9:
                    continue
                exact = int(mpmath.floor(v + mpmath.mpf("0.5")))
                assert O.pair_bin(pts[i], pts[j], rp) == exact
                checked += 1
        assert checked > 5000


def test_pair_set_closed_forms():
    """same subgroup: n(n+1)/2 pairs incl. n self pairs in bin 0 (cuDebye-v15.cu:137,372,382);
    cross: n_a*n_b (cuDebye-v15.cu:396)."""
    pts = _rand_model(500, 2)
    x, y, z = pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()
    nb = O.channels(50.0, 1e-3)
    rp = O.rp_of(1e-3)
    h, ov = O.histogram(x, y, z, 0, 300, 0, 300, True, rp, nb)
    assert ov == 0 and int(h.sum()) == 300 * 301 // 2 and int(h[0]) == 300
    h, ov = O.histogram(x, y, z, 0, 300, 300, 200, False, rp, nb)
    assert ov == 0 and int(h.sum()) == 300 * 200 and int(h[0]) == 0
    # brute force the same histograms pair by pair through orc_pair_bin
    ref = np.zeros(nb, np.uint64)
    for i in range(300):
        for j in range(300, 500):
            ref[O.pair_bin(pts[i], pts[j], rp)] += 1
    assert np.array_equal(h, ref)


def test_histogram_overflow_is_reported():
    pts = _rand_model(50, 3, spread=30.0)
    x, y, z = pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()
    nb = O.channels(10.0, 1e-3)  # far below the real maximum distance (README.md:102)
    h, ov = O.histogram(x, y, z, 0, 50, 0, 50, True, O.rp_of(1e-3), nb)
    assert ov > 0 and int(h.sum()) + ov == 50 * 51 // 2


def test_q_grid_rule():
    """cuDebye-v15.cu:690-698: nTT = int((max-min)/step)+1 (truncation), K = 4 pi sin(theta)/lambda."""
    tt, K = O.q_from_2theta(1.5406, 10.0, 150.0, 140.0 / 4095)
    assert len(tt) == 4095  # truncation drops the last point of a 4096-point grid
    tt, K = O.q_from_2theta(1.5418, 10.0, 140.0, 0.02)
    assert len(tt) == 6501
    th = np.deg2rad(tt / 2)
    assert np.allclose(K, 4 * np.pi * np.sin(th) / 1.5418, rtol=1e-15, atol=0)


def _ff_mp(Q, occ, beq, a, b, c):
    nh = mpmath.mpf(Q) ** 2 * mpmath.mpf(-0.006332573977646)
    f = mpmath.mpf(c) + sum(mpmath.mpf(a[i]) * mpmath.exp(mpmath.mpf(b[i]) * nh) for i in range(5))
    return mpmath.mpf(occ) * mpmath.exp(nh * mpmath.mpf(beq)) * f


def test_prefactor_against_mpmath():
    rng = np.random.default_rng(4)
    K = np.linspace(0.3, 20.0, 50)
    for _ in range(5):
        pa = dict(occ=rng.uniform(0.5, 1), beq=rng.uniform(0, 2), a=rng.uniform(1, 20, 5),
                  b=rng.uniform(0.5, 60, 5), c=rng.uniform(0, 2))
        pb = dict(occ=rng.uniform(0.5, 1), beq=rng.uniform(0, 2), a=rng.uniform(1, 20, 5),
                  b=rng.uniform(0.5, 60, 5), c=rng.uniform(0, 2))
        ff = O.prefactor(K, pa["occ"], pb["occ"], pa["beq"], pb["beq"], pa["a"], pa["b"], pa["c"],
                         pb["a"], pb["b"], pb["c"])
        for q in range(len(K)):
            ref = _ff_mp(K[q], pa["occ"], pa["beq"], pa["a"], pa["b"], pa["c"]) * \
                _ff_mp(K[q], pb["occ"], pb["beq"], pb["a"], pb["b"], pb["c"])
            assert abs(ff[q] - float(ref)) <= 1e-13 * abs(float(ref))


def test_two_atom_closed_form():
    """Two atoms a, b of one subgroup (f=1): I(Q) = 2 + 2 sinc(Q r_k), r_k the bin centre."""
    K = np.linspace(0.5, 10, 64)
    bins = np.array([0, 12345], np.int64)
    cnt = np.array([2.0, 1.0])
    I = O.debye_sum(bins, cnt, 1e-3, K, np.ones_like(K), first_n=8)
    r = 12.345
    assert np.allclose(I, 2 + 2 * np.sin(K * r) / (K * r), rtol=1e-14, atol=0)


def test_weight_rule_near_zero_bins():
    """cuDebye-v15.cu:508-514: compacted entries below blockDim.x with r <= 0.0005 count once
    without sinc; the first entry above 0.0005 is doubled with sinc."""
    K = np.array([1.0, 5.0])
    bins = np.array([0, 5, 6], np.int64)  # r = 0, 0.0005, 0.0006 at rstep 1e-4
    cnt = np.array([10.0, 3.0, 1.0])
    I = O.debye_sum(bins, cnt, 1e-4, K, np.ones(2), first_n=8)
    r = 6 * 1e-4
    assert np.allclose(I, 13 + 2 * np.sin(K * r) / (K * r), rtol=1e-15, atol=0)


def test_low_q_limit_is_square_of_total_scattering():
    """I(Q->0) -> (sum_i occ_i f_i(0))^2 when every pair is counted (cross pairs doubled)."""
    pts = _rand_model(120, 5, spread=5.0)
    x, y, z = pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()

    class P:
        group_n = np.array([70, 50])
        occ = np.array([1.0, 0.5])
        beq = np.array([0.0, 0.0])
        a = np.array([[1.0, 2.0, 0, 0, 0], [3.0, 0, 0, 0, 0]])
        b = np.array([[1.0, 2.0, 0, 0, 0], [3.0, 0, 0, 0, 0]])
        c = np.array([0.5, 0.25])

    nb = O.channels(30.0, 1e-3)
    rp = O.rp_of(1e-3)
    H = np.zeros((3, nb), np.uint64)
    for p, (r, c) in enumerate(O.pair_list(P.group_n)):
        st = [0, 70]
        O.histogram(x, y, z, st[r], P.group_n[r], st[c], P.group_n[c], r == c, rp, nb, out=H[p])
    K = np.array([1e-7])
    I = O.intensity_from_histograms(H, P, K, 1e-3)
    f0 = np.array([P.c[g] + P.a[g].sum() for g in range(2)])
    tot = (P.group_n * P.occ * f0).sum()
    assert abs(I[0] - tot ** 2) <= 1e-9 * tot ** 2


def test_debye_sum_against_mpmath_all_pairs():
    """Histogram route (bins rounded to rstep) vs an mpmath double sum over all atom pairs
    at the bin-centre distances: identical up to FP64 rounding."""
    pts = _rand_model(60, 6, spread=6.0)
    x, y, z = pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy()
    rstep = 1e-3
    rp = O.rp_of(rstep)
    nb = O.channels(25.0, rstep)
    h, _ = O.histogram(x, y, z, 0, 60, 0, 60, True, rp, nb)
    nz = np.nonzero(h)[0]
    K = np.array([0.7, 3.1, 9.9, 17.5])
    S = np.zeros(len(K))
    O.debye_sum(nz, h[nz].astype(np.float64), rstep, K, np.ones(len(K)), s_out=S, first_n=16)
    for q, Kq in enumerate(K):
        acc = mpmath.mpf(60)
        for i in range(60):
            for j in range(i):
                k = O.pair_bin(pts[i], pts[j], rp)
                r = mpmath.mpf(k) * mpmath.mpf(rstep)
                acc += 2 * mpmath.sin(mpmath.mpf(Kq) * r) / (mpmath.mpf(Kq) * r)
        assert abs(S[q] - float(acc)) <= 1e-12 * float(abs(acc)) + 1e-12

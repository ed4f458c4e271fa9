"""The FP64 sine of the pass-2 sinc table (Cody-Waite reduction by pi + a Taylor polynomial,
cudebye_amd/csrc/fastmath.h), host build, against a 50-digit mpmath reference: absolute error
at the level of a few ulp over the argument range of the Debye sum (q r up to ~1e4) and next
to multiples of pi/2 (where the reduction cancels, and where the polynomial's argument is
largest)."""
import ctypes as C

import mpmath
import numpy as np

import cudebye_amd as cd

mpmath.mp.dps = 50


def sin_host(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    fn = cd.lib().dbx_debug_sin_host
    fn.restype = C.c_int
    assert fn(x.ctypes.data_as(C.c_void_p), len(x), out.ctypes.data_as(C.c_void_p)) == 0
    return out


def _ref(x):
    return np.array([float(mpmath.sin(mpmath.mpf(float(v)))) for v in x])


def test_sin_random_arguments():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 4, 3000), rng.uniform(0, 200, 3000), rng.uniform(0, 2e4, 3000),
                        -rng.uniform(0, 100, 500)])
    err = np.abs(sin_host(x) - _ref(x))
    assert err.max() < 4e-16, err.max()


def test_sin_near_quadrant_boundaries():
    n = np.arange(1, 20000, 7)
    base = n * (np.pi / 2)
    x = np.concatenate([base, np.nextafter(base, 0), np.nextafter(base, np.inf), base + 1e-9])
    got = sin_host(x)
    ref = _ref(x)
    assert np.abs(got - ref).max() < 4e-16
    # relative accuracy is kept where sin is tiny (no cancellation loss in the reduction)
    small = np.abs(ref) < 1e-6
    assert np.all(np.abs(got[small] - ref[small]) <= 1e-10 * np.abs(ref[small]) + 1e-22)


def test_sincos_pair_accuracy():
    """sincos_cw (start of each rotated run of the sinc generator): both outputs within a few ulp
    of 50-digit references, over the Debye sum's argument range and near multiples of pi/2."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0, 4, 2000), rng.uniform(0, 2e4, 3000),
                        np.arange(1, 5000, 3) * (np.pi / 2)])
    x = np.ascontiguousarray(x)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    fn = cd.lib().dbx_debug_sincos_host
    fn.restype = C.c_int
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert fn(p(x), len(x), p(s), p(c)) == 0
    sref = _ref(x)
    cref = np.array([float(mpmath.cos(mpmath.mpf(float(v)))) for v in x])
    assert np.abs(s - sref).max() < 4e-16 and np.abs(c - cref).max() < 4e-16
    assert np.array_equal(s, sin_host(x))  # the same sine as the direct path


def test_sinc_digits_encoding():
    """The matrix-core Debye sum's sinc digits (mfdigits.h, host build of the device code):
    sinc(q r) = 3/8 + sum_j d_j 256^j / (3 2^46) to within half a grid step (below 2^-48) plus
    the FP64 sine's error, for six signed bytes d_j, against a 50-digit sinc.  Covers q = 0
    (sinc exactly 1), tiny q r, the sinc minimum (q r = 4.4934, sinc = -0.21723), and q r up to
    2e4: the ends of the range the six digits must hold.  A dropped or misplaced top digit
    would be off by 256^4 / c or more."""
    rng = np.random.default_rng(5)
    xmin = 4.493409457909064
    qr = np.concatenate([rng.uniform(0, 3, 400), rng.uniform(0, 2e4, 800), xmin + rng.uniform(-1e-6, 1e-6, 40),
                         10.0 ** rng.uniform(-9, -2, 100), np.full(4, 1e-30)])
    r = rng.uniform(0.5, 300.0, len(qr))
    q = qr / r
    q[:8] = 0.0
    n = len(q) // 4 * 4
    q, r = np.ascontiguousarray(q[:n]), np.ascontiguousarray(r[:n])
    d = np.zeros((n, 6), np.int8)
    fn = cd.lib().dbx_debug_sinc_digits_host
    fn.restype = C.c_int
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert fn(p(q), p(r), n, p(d)) == 0
    c = 3.0 * 2.0 ** 46
    V = sum(d[:, j].astype(np.int64) * 256 ** j for j in range(6))
    got = 0.375 + V / c
    ref = []
    for qi, ri in zip(q, r):
        x = mpmath.mpf(float(qi)) * mpmath.mpf(float(ri))
        ref.append(1.0 if x == 0 else float(mpmath.sin(x) / x))
    ref = np.array(ref)
    assert np.abs(got - ref).max() <= 0.5 / c + 4e-15
    assert np.all(got[:8] == 1.0)
    assert got.min() < -0.2172 and V.max() > 0.46 * 2.0 ** 48

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

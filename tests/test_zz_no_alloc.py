"""dbx_debye_noalloc, the allocation-free pass 2, against dbx_debye on the same inputs.

Inputs and outputs live in device memory that the test allocates through the HIP runtime the
library itself loads (libamdhip64.so.7); the entry point takes every buffer of its kernels from
its fixed region.  Two calls in a row, the second on a smaller problem (fewer partials, bins and
q points), each equal to dbx_debye bit for bit (same kernels, same order of every FP64 sum), in
both pass-2 modes; a problem beyond the region is refused with DBX_ERR_BUDGET."""
import ctypes as C

import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """Device buffers through the library's own HIP runtime."""

    def __init__(self):
        cd.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def _gp(p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


def _problem(ctx, m):
    """Histograms of model m (pass 1 on the device) and its prefactors."""
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    mod.close()
    pr = [_gp(p) for p in m.params]
    ff = np.array([ctx.prefactor(pr[a], pr[b], m.q) for a in range(len(pr)) for b in range(a, len(pr))])
    return H, ff


def _noalloc(ctx, dev, H, m, ff, with_S=True):
    P, nb = H.shape
    nq = len(m.q)
    dH, dq, dff = dev.put(H), dev.put(m.q), dev.put(ff)
    dI = dev.alloc(8 * nq)
    dS = dev.alloc(8 * P * nq) if with_S else None
    drel = dev.alloc(8 * nq)
    rc = cd.lib().dbx_debye_noalloc(ctx.handle, dH, P, nb, m.rstep, m.first_n, dq, nq, dff, dI, dS, drel)
    assert rc == cd.DBX_OK, cd.lib().dbx_last_error(ctx.handle).decode()
    I = dev.get(dI, nq, np.float64)
    S = dev.get(dS, (P, nq), np.float64) if with_S else None
    return I, S, dev.get(drel, nq, np.float64)


@pytest.mark.parametrize("mode", [cd.DBX_PASS2_MFMA, cd.DBX_PASS2_FP64], ids=["mfma", "fp64"])
def test_noalloc_equals_debye_two_calls_second_smaller(mode):
    big = synth.small_random(6000, 3, 71, spread=14.0, nq=1500)
    small = synth.small_random(1500, 2, 72, spread=7.0, nq=300)
    dev = Dev()
    with cd.Context(0) as ctx:
        ctx.set_pass2_mode(mode)
        probs = [(m,) + _problem(ctx, m) for m in (big, small)]
        assert probs[1][1].shape[0] < probs[0][1].shape[0] and probs[1][1].shape[1] < probs[0][1].shape[1]
        try:
            for i, (m, H, ff) in enumerate(probs):
                Iref, Sref, rref = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_S=True, with_err=True)
                I, S, rel = _noalloc(ctx, dev, H, m, ff, with_S=(i == 0))
                assert np.array_equal(I, Iref), (i, np.max(np.abs(I - Iref) / np.abs(Iref)))
                assert np.array_equal(rel, rref), i
                if S is not None:
                    assert np.array_equal(S, Sref), i
        finally:
            dev.free()


def test_noalloc_refuses_problems_beyond_its_region():
    """A problem whose buffers exceed the fixed region: DBX_ERR_BUDGET before any of its
    kernels run; the context stays usable."""
    dev = Dev()
    with cd.Context(0) as ctx:
        nb, nq = 16 * 1024 * 1024, 64
        try:
            dH = dev.alloc(8 * nb)
            assert dev.hip.hipMemset(dH, 0, 8 * nb) == 0
            q = np.linspace(0.5, 5.0, nq)
            dq, dff = dev.put(q), dev.put(np.ones(nq))
            dI = dev.alloc(8 * nq)
            rc = cd.lib().dbx_debye_noalloc(ctx.handle, dH, 1, nb, 1e-3, 8, dq, nq, dff, dI, None, None)
            assert rc == cd.DBX_ERR_BUDGET
            assert "fixed scratch region" in cd.lib().dbx_last_error(ctx.handle).decode()
            m = synth.small_random(500, 1, 73, nq=64)
            H, ff = _problem(ctx, m)
            Iref = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n)
            I, _, _ = _noalloc(ctx, dev, H, m, ff, with_S=False)
            assert np.array_equal(I, Iref)
        finally:
            dev.free()


def test_noalloc_equals_debye_on_rows_of_mixed_digit_classes():
    """40 rows whose counts need one to five digits in shuffled order: the rows are ordered by
    class on the device and both calls size their launches alike, so the call equals dbx_debye
    bit for bit."""
    from types import SimpleNamespace
    rng = np.random.default_rng(78)
    nb, nq = 12000, 200
    caps = np.repeat([2 ** 5, 2 ** 13, 2 ** 21, 2 ** 29, 2 ** 36], 8)
    rng.shuffle(caps)
    P = len(caps)
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), 1000, replace=False)
        H[p, idx] = rng.integers(1, caps[p], 1000).astype(np.uint64)
    m = SimpleNamespace(q=np.linspace(0.3, 20.0, nq), rstep=1e-3, first_n=8)
    ff = rng.uniform(0.5, 2.0, (P, nq))
    dev = Dev()
    with cd.Context(0) as ctx:
        ctx.set_rel_err_target(0)  # the matrix-core sums, never recomputed in FP64
        try:
            Iref, Sref, rref = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_S=True, with_err=True)
            I, S, rel = _noalloc(ctx, dev, H, m, ff)
            assert np.array_equal(I, Iref) and np.array_equal(S, Sref) and np.array_equal(rel, rref)
        finally:
            dev.free()


@pytest.mark.parametrize("mode", [cd.DBX_PASS2_MFMA, cd.DBX_PASS2_FP64], ids=["mfma", "fp64"])
def test_noalloc_sized_from_the_compacted_bins_equals_debye(mode):
    """Six sparse rows of 1.5M bins (20,000 occupied each): buffers sized from the bins' upper
    bound (five digit planes or FP64 weights over 1.5M bins) exceed the fixed region, so the
    call reads the compacted-bin count back and sizes them exactly, while dbx_debye sizes them
    from the bound with no read-back; the splits are chosen from the bound in both, so the sums
    are the same bit for bit."""
    from types import SimpleNamespace
    rng = np.random.default_rng(79)
    P, nb, nq = 6, 1_500_000, 500
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), 20000, replace=False)
        H[p, idx] = rng.integers(1, [2 ** 6, 2 ** 14, 2 ** 22][p % 3], 20000).astype(np.uint64)
        H[p, 0] = 1000
    m = SimpleNamespace(q=np.linspace(0.3, 20.0, nq), rstep=1e-4, first_n=8)
    ff = rng.uniform(0.5, 2.0, (P, nq))
    dev = Dev()
    with cd.Context(0) as ctx:
        ctx.set_pass2_mode(mode)
        ctx.set_rel_err_target(0)  # the mode's own sums, never recomputed
        try:
            Iref, Sref, rref = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_S=True, with_err=True)
            I, S, rel = _noalloc(ctx, dev, H, m, ff)
            assert np.array_equal(S, Sref), np.max(np.abs(S - Sref))
            assert np.array_equal(I, Iref) and np.array_equal(rel, rref)
        finally:
            dev.free()

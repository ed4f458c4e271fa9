"""Pass 2 on the GPU against the FP64 oracle: prefactors, the Debye sum with the reference
weight rule, and I(Q) (relative error bound written in each test)."""
import copy

import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth
from oracle import oracle as O
from tests.sweep_models import SWEEP, sweep_model
from tests.conftest import read_debi

pytestmark = pytest.mark.gpu

# FP64 mode: only summation order and the device sin differ from the oracle's compensated
# FP64 sum.  MFMA mode (signed 8-bit digits, exact 32-bit integer accumulation): the sinc table
# is rounded to a multiple of 1 / (3 2^46), held to 1e-6 relative on every I(Q).
TOL = {cd.DBX_PASS2_FP64: 1e-9, cd.DBX_PASS2_MFMA: 1e-6}
MODES = [cd.DBX_PASS2_MFMA, cd.DBX_PASS2_FP64]


@pytest.fixture(scope="module")
def ctx():
    c = cd.Context(0)
    yield c
    c.close()


@pytest.fixture(params=MODES, ids=["mfma", "fp64"])
def mode_ctx(request, ctx):
    ctx.set_pass2_mode(request.param)
    yield ctx, request.param
    ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)


def _gp(p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


# Matrix-core error per unit weight 2 N of a row's sum: the sinc digits' rounding, 1 / (2 c) with
# c = 3 2^46, the FP64 sine and its recurrence (4e-15), and 1e-14 for the FP64 partial sums over
# the splits and folds (a few hundred roundings of sums below the weight at most).
MF_UNIT_ERR = 0.5 / (3.0 * 2.0 ** 46) + 4e-15 + 1e-14


def _quant_bound(h, Sref):
    """Bound of |S - Sref| for a row with counts h on the matrix cores: the unit error above
    times the row's weights, plus 1e-14 of the largest |Sref| for the oracle's own rounding."""
    return 2.0 * float(h.astype(np.float64).sum()) * MF_UNIT_ERR + 1e-14 * np.abs(Sref).max()


def _oracle_I(m, H, first_n):
    class Pm:
        pass
    Pm.group_n = m.group_n
    Pm.occ = np.array([p["occ"] for p in m.params])
    Pm.beq = np.array([p["beq"] for p in m.params])
    Pm.a = np.array([p["a"] for p in m.params])
    Pm.b = np.array([p["b"] for p in m.params])
    Pm.c = np.array([p["c"] for p in m.params])
    return O.intensity_from_histograms(H, Pm, m.q, m.rstep, first_n=first_n, with_cond=True)


def test_prefactor_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    q = np.linspace(0.3, 20, 777)
    for _ in range(4):
        pa, pb = synth.cromer_mann(rng), synth.cromer_mann(rng)
        pa["occ"], pb["beq"] = 0.8333, 2.0
        got = ctx.prefactor(_gp(pa), _gp(pb), q)
        ref = O.prefactor(q, pa["occ"], pb["occ"], pa["beq"], pb["beq"], pa["a"], pa["b"], pa["c"],
                          pb["a"], pb["b"], pb["c"])
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-14


def test_gamma_debI_from_fixture_histograms(mode_ctx, golden_dir, gamma_debx):
    """Pass-2-only known answer: the reference's .debx histograms + header -> its .debI."""
    ctx, mode = mode_ctx
    g = O.parse_deb(str(golden_dir / "GammaAl2O3_10nm.deb"))
    tt, K = cd.q_from_2theta(g.lam, g.tt_min, g.tt_max, g.tt_step)
    nb = O.channels(g.rmax, g.rstep)
    pairs = O.pair_list(g.group_n)
    H = np.zeros((len(pairs), nb), np.uint64)
    ff = np.zeros((len(pairs), len(K)))
    for p, (r, c) in enumerate(pairs):
        bins, cnt = gamma_debx[f"{r}x{c}"]
        H[p][bins] = cnt
        ga = cd.GroupParams.make(g.occ[r], g.beq[r], g.a[r], g.b[r], g.c[r])
        gb = cd.GroupParams.make(g.occ[c], g.beq[c], g.a[c], g.b[c], g.c[c])
        ff[p] = ctx.prefactor(ga, gb, K)
    I = ctx.debye(H, g.rstep, K, ff, first_n=g.block)
    Iref, mag = O.intensity_from_histograms(H, g, K, g.rstep, first_n=g.block, with_cond=True)
    rel = np.abs(I - Iref) / np.abs(Iref)
    print(f"gamma pass 2 mode {mode}: max rel err {rel.max():.3e}, max cond {(mag / np.abs(Iref)).max():.3e}")
    assert rel.max() < TOL[mode]
    ref = read_debi(str(golden_dir / "GammaAl2O3_10nm.debI.ref"))
    got = [f"{a:g}\t{b:g}" for a, b in zip(tt, I)]
    if mode == cd.DBX_PASS2_FP64:
        assert got == ref
    else:  # printed with 6 significant digits: within print rounding plus the 1e-6 bound
        vref = np.array([float(l.split("\t")[1]) for l in ref])
        assert np.all(np.abs(I - vref) <= (5e-6 + 1e-6) * np.abs(vref))


@pytest.mark.parametrize("maker", ["c1", "rand2", "rand5"])
def test_intensity_matches_oracle(mode_ctx, maker):
    ctx, mode = mode_ctx
    m = {"c1": lambda: synth.fcc_sphere(nq=1024),
         "rand2": lambda: synth.small_random(3000, 2, 31, nq=512),
         "rand5": lambda: synth.small_random(4000, 5, 32, nq=300)}[maker]()
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    I, H = ctx.intensity(mod, [_gp(p) for p in m.params], m.rstep, m.n_bins, m.q,
                         first_n=m.first_n, with_hist=True)
    Iref, mag = _oracle_I(m, H, m.first_n)
    rel = np.abs(I - Iref) / np.abs(Iref)
    assert rel.max() < TOL[mode], (rel.max(), (mag / np.abs(Iref)).max())


def test_weight_rule_first_n(mode_ctx):
    """More low bins than first_n: the extra entries at r <= 0.0005 take weight 2 with sinc."""
    nb = 100
    H = np.zeros((1, nb), np.uint64)
    H[0, [0, 1, 2, 3, 4, 5, 6, 40]] = [10, 3, 4, 5, 6, 7, 8, 9]
    K = np.array([0.7, 3.0, 11.0])
    ff = np.ones((1, 3))
    ctx, mode = mode_ctx
    for first_n in (1, 3, 8):
        I = ctx.debye(H, 1e-4, K, ff, first_n=first_n)
        nz = np.nonzero(H[0])[0]
        Iref = O.debye_sum(nz, H[0][nz].astype(float), 1e-4, K, ff[0], first_n=first_n)
        assert np.allclose(I, Iref, rtol=1e-13 if mode == cd.DBX_PASS2_FP64 else 1e-9, atol=0)


@pytest.mark.parametrize("P,cmax,nq", [(5, 2 ** 30, 16), (5, 2 ** 30, 100), (5, 2 ** 30, 1000),
                                        (24, 2 ** 22, 100), (40, 2 ** 36, 100), (130, 2 ** 30, 100),
                                        (256, 2 ** 30, 100)])
def test_mfma_error_within_quantisation_bound(ctx, P, cmax, nq):
    """Exact accumulation: with the sinc table rounded to multiples of 1 / (3 2^46), the MFMA
    Debye sum differs from the oracle's FP64 sum by at most sum_k 2 N_k times the unit error
    (about 1.6e-14, the FP64 sine and sums included).  A lossy FP32 accumulation of large counts
    would exceed that bound by seven orders of magnitude, a lost top digit by more.
    Rows and count sizes cover every workgroup tile (16, 32 and 128 rows) and 3 to 5 count
    digits (counts below 2^22, 2^30, 2^36)."""
    rng = np.random.default_rng(nq + P)
    nb = 20000
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), 3000, replace=False)
        H[p, idx] = rng.integers(1, cmax, 3000).astype(np.uint64)
    K = np.linspace(0.3, 20.0, nq)
    ff = np.ones((P, nq))
    ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    I, S = ctx.debye(H, 1e-3, K, ff, first_n=8, with_S=True)
    for p in range(P):
        nz = np.nonzero(H[p])[0]
        Sref = np.zeros(nq)
        O.debye_sum(nz, H[p][nz].astype(float), 1e-3, K, np.ones(nq), s_out=Sref, first_n=8)
        bound = _quant_bound(H[p], Sref)
        err = np.abs(S[p] - Sref)
        assert err.max() <= bound, (p, err.max(), bound)


def test_device_sin_equals_host_build():
    """The pass-2 sine uses only IEEE fma/rint/mul: device and host builds agree bit for bit
    (the host build is pinned against mpmath in test_fastmath_cpu.py)."""
    import ctypes as C
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(0, 2e4, 200000), np.arange(1, 5000) * (np.pi / 2)])
    dev = cd.selftest_sin(x)
    host = np.zeros_like(x)
    fn = cd.lib().dbx_debug_sin_host
    fn.restype = C.c_int
    assert fn(x.ctypes.data_as(C.c_void_p), len(x), host.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(dev.view(np.int64), host.view(np.int64))


@pytest.mark.parametrize("maker", ["c1", "rand2", "rand5"])
def test_error_bound_covers_error_against_oracle(ctx, maker):
    """Every I(q) comes with a relative error bound; the error against the oracle's compensated
    FP64 sums stays within it, on the matrix cores (no FP64 recomputation) and in FP64."""
    m = {"c1": lambda: synth.fcc_sphere(nq=1024),
         "rand2": lambda: synth.small_random(3000, 2, 31, nq=512),
         "rand5": lambda: synth.small_random(4000, 5, 32, nq=300)}[maker]()
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    params = [_gp(p) for p in m.params]
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    Iref, mag = _oracle_I(m, H, m.first_n)
    try:
        ctx.set_rel_err_target(0)
        for mode in MODES:
            ctx.set_pass2_mode(mode)
            I, rel = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_err=True)
            err = np.abs(I - Iref) / np.abs(Iref)
            print(f"{maker} mode {mode}: error {err.max():.2e}, bound {rel.max():.2e}, worst ratio "
                  f"{(err / rel).max():.3f}")
            # 1e-14: the oracle's prefactors (its own exp) differ from the device's by an ulp or two
            assert np.all(err <= rel + 1e-14), (mode, np.argmax(err / rel))
            assert ctx.pass2_refined() == 0
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
    mod.close()


@pytest.mark.parametrize("P,cmax,nq", [(5, 2 ** 30, 100), (40, 2 ** 36, 100), (130, 2 ** 30, 100)])
def test_error_bound_covers_quantisation_error(ctx, P, cmax, nq):
    """Rows of 3000 bins with counts up to 2^36 (up to 5 count digits, every workgroup tile): the
    matrix-core sum of each I(q) = sum_p S_p(q) is within its bound of the oracle's FP64 sum."""
    rng = np.random.default_rng(nq + P)
    nb = 20000
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), 3000, replace=False)
        H[p, idx] = rng.integers(1, cmax, 3000).astype(np.uint64)
    K = np.linspace(0.3, 20.0, nq)
    ff = np.ones((P, nq))
    Iref = np.zeros(nq)
    for p in range(P):
        nz = np.nonzero(H[p])[0]
        Iref += O.debye_sum(nz, H[p][nz].astype(float), 1e-3, K, np.ones(nq), first_n=8)
    ctx.set_rel_err_target(0)
    try:
        I, rel = ctx.debye(H, 1e-3, K, ff, first_n=8, with_err=True)
    finally:
        ctx.set_rel_err_target(1e-6)
    err = np.abs(I - Iref) / np.abs(Iref)
    print(f"P={P}: error {err.max():.2e}, bound {rel.max():.2e}, worst ratio {(err / rel).max():.3f}")
    assert np.all(err <= rel)


def test_error_target_recomputes_in_fp64(ctx):
    """A target below the matrix-core bound: the model's partials are recomputed with FP64
    arithmetic, giving the FP64 mode's I bit for bit and its bound; a loose target leaves the
    matrix-core result.  With six sinc digits the matrix-core bound is the tighter of the two."""
    m = synth.small_random(3000, 3, 33, nq=400)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    params = [_gp(p) for p in m.params]
    args = (mod, params, m.rstep, m.n_bins, m.q)
    try:
        ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
        I64, rel64 = ctx.intensity(*args, first_n=m.first_n, with_err=True)
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1.0)
        Imf, relmf = ctx.intensity(*args, first_n=m.first_n, with_err=True)
        assert ctx.pass2_refined() == 0 and relmf.max() < rel64.max()
        ctx.set_rel_err_target(relmf.min() / 2)
        I, rel = ctx.intensity(*args, first_n=m.first_n, with_err=True)
        assert ctx.pass2_refined() == 1
        assert np.array_equal(I, I64) and np.array_equal(rel, rel64)
    finally:
        ctx.set_rel_err_target(1e-6)
    mod.close()


def test_error_target_in_a_batch_recomputes_only_the_models_that_miss_it(ctx):
    """In a batch, only the models whose bound misses the target are recomputed in FP64: each
    result within 1e-12 of the FP64 mode on that model alone, the others equal to the batch
    without a target."""
    models = [synth.small_random(n, s, 40 + i, nq=200) for i, (n, s) in enumerate([(800, 2), (3000, 3), (1500, 1),
                                                                                     (2500, 2)])]
    items = [(ctx.model(m.x, m.y, m.z, m.group_n), [_gp(p) for p in m.params], m.n_bins, m.first_n)
             for m in models]
    q = models[0].q
    try:
        ctx.set_rel_err_target(0)
        I0, rel0 = ctx.batch_intensity(items, 1e-3, q, with_err=True)
        worst = rel0.max(axis=1)
        target = float(np.median(worst))
        ctx.set_rel_err_target(target)
        I1, rel1 = ctx.batch_intensity(items, 1e-3, q, with_err=True)
        redo = worst > target
        assert ctx.pass2_refined() == int(redo.sum()) > 0
        ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
        for i, (m, it) in enumerate(zip(models, items)):
            if redo[i]:
                I64, rel64 = ctx.intensity(it[0], it[1], 1e-3, it[2], q, first_n=it[3], with_err=True)
                assert np.max(np.abs(I1[i] - I64) / np.abs(I64)) < 1e-12, i
                # the FP64 kind of bound (its summation term grows with the batch's bins)
                assert rel1[i].max() <= 8 * rel64.max(), i
            else:
                assert np.array_equal(I1[i], I0[i]) and np.array_equal(rel1[i], rel0[i]), i
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
        for it in items:
            it[0].close()


@pytest.mark.parametrize("top", [3, 4, 5])
def test_rows_of_mixed_digit_classes(ctx, top):
    """200 rows whose largest counts need 1 to `top` count digits, in shuffled order: the rows
    are ordered by digit class on the device and each row tile's matrix instructions cover only
    the digits of its own rows.  200 rows take the 128-row tile for classes up to 3 and the
    32-row tile for classes 4 and 5, so each of the contraction's two launches runs for some
    value of `top` and exits at once for the others.  Every row's sum lands in the caller's row
    within the quantisation bound of the oracle's FP64 sum, and the bound of I covers its
    error."""
    rng = np.random.default_rng(77 + top)
    nb, nq = 20000, 150
    caps = np.repeat([2 ** 13, 2 ** 5, 2 ** 21, 2 ** 29, 2 ** 36][:top], [100, 40, 30, 20, 10][:top])
    caps = np.concatenate([caps, np.full(200 - len(caps), 2 ** 5)])
    rng.shuffle(caps)
    P = len(caps)
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), 1500, replace=False)
        H[p, idx] = rng.integers(1, caps[p], 1500).astype(np.uint64)
    K = np.linspace(0.3, 20.0, nq)
    ff = np.ones((P, nq))
    ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    ctx.set_rel_err_target(0)
    try:
        I, S, rel = ctx.debye(H, 1e-3, K, ff, first_n=8, with_S=True, with_err=True)
    finally:
        ctx.set_rel_err_target(1e-6)
    Iref = np.zeros(nq)
    for p in range(P):
        nz = np.nonzero(H[p])[0]
        Sref = np.zeros(nq)
        O.debye_sum(nz, H[p][nz].astype(float), 1e-3, K, np.ones(nq), s_out=Sref, first_n=8)
        Iref += Sref
        bound = _quant_bound(H[p], Sref)
        err = np.abs(S[p] - Sref)
        assert err.max() <= bound, (p, caps[p], err.max(), bound)
    assert np.all(np.abs(I - Iref) / np.abs(Iref) <= rel)



@pytest.mark.parametrize("top", [3, 5])
def test_more_rows_than_one_ordering_pass(ctx, top):
    """2,500 rows, more than the 1,024 rows that the device's ordering by digit class handles
    per pass, with largest counts needing 1 to `top` count digits in shuffled order (top 3: the
    128-row tile; top 5: the 32-row tile).  Every row's sum lands in the caller's row within
    the quantisation bound of the oracle's FP64 sum, and the bound of I covers its error."""
    rng = np.random.default_rng(700 + top)
    P, nb, nz, nq = 2500, 4000, 300, 40
    caps = np.array([2 ** 5, 2 ** 13, 2 ** 21, 2 ** 29, 2 ** 36][:top])[rng.integers(0, top, P)]
    H = np.zeros((P, nb), np.uint64)
    for p in range(P):
        idx = rng.choice(np.arange(1, nb), nz, replace=False)
        H[p, idx] = rng.integers(1, caps[p], nz).astype(np.uint64)
    K = np.linspace(0.3, 20.0, nq)
    ff = np.ones((P, nq))
    ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    ctx.set_rel_err_target(0)
    try:
        I, S, rel = ctx.debye(H, 1e-3, K, ff, first_n=8, with_S=True, with_err=True)
    finally:
        ctx.set_rel_err_target(1e-6)
    Iref = np.zeros(nq)
    for p in range(P):
        idx = np.nonzero(H[p])[0]
        Sref = np.zeros(nq)
        O.debye_sum(idx, H[p][idx].astype(float), 1e-3, K, np.ones(nq), s_out=Sref, first_n=8)
        Iref += Sref
        err = np.abs(S[p] - Sref)
        assert err.max() <= _quant_bound(H[p], Sref), (p, caps[p], err.max())
    assert np.all(np.abs(I - Iref) / np.abs(Iref) <= rel)

# the sweep's q grids: nq and range per model (fewer points where the oracle's sums are long)
_SWEEP_Q = {"hollow_shell_fine": (17, 0.5, 12.0), "fine_bins_wide": (5, 0.3, 9.0), "very_coarse_bins": (1, 1.7, 1.7),
            "plane": (257, 0.2, 30.0)}


@pytest.mark.parametrize("kind", SWEEP)
def test_geometry_and_bin_width_sweep_within_bound(ctx, kind):
    """The pass-1 sweep's models through pass 2 against the oracle's FP64 sums: on the matrix
    cores (no FP64 recomputation) and in FP64, every I(q) within its reported bound, and the
    default call (recomputation at 1e-6) within 1e-6."""
    m = sweep_model(kind, 900 + SWEEP.index(kind))
    nq, q0, q1 = _SWEEP_Q.get(kind, (97, 0.3, 25.0))
    m.q = np.linspace(q0, q1, nq)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    params = [_gp(p) for p in m.params]
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    Iref, _ = _oracle_I(m, H, m.first_n)
    try:
        ctx.set_rel_err_target(0)
        for mode in MODES:
            ctx.set_pass2_mode(mode)
            I, rel = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_err=True)
            err = np.abs(I - Iref) / np.abs(Iref)
            assert np.all(err <= rel + 1e-14), (mode, err.max(), rel.max())
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
        I = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
        assert np.max(np.abs(I - Iref) / np.abs(Iref)) < 1e-6
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
    mod.close()


def test_q_zero_and_tiny_q_give_the_square_of_the_total_scattering(ctx):
    """At Q = 0 every sinc is 1 (the reference's sin(x)/x would be 0/0 there): I(0) is
    (sum_i occ_i f_i(0))^2 when every pair is counted, and I(1e-7) agrees with it; the oracle
    pins the other points.  Both modes, within the reported bounds."""
    m = synth.small_random(900, 3, 61, spread=9.0)
    q = np.array([0.0, 1e-7, 1e-3, 0.3, 5.0])
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    params = [_gp(p) for p in m.params]
    f0 = np.array([p["c"] + sum(p["a"]) for p in m.params])
    occ = np.array([p["occ"] for p in m.params])
    tot = float((m.group_n * occ * f0).sum())
    m2 = copy.copy(m)
    m2.q = q[1:]
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    Iref, _ = _oracle_I(m2, H, m.first_n)
    try:
        for mode in MODES:
            ctx.set_pass2_mode(mode)
            I, rel = ctx.intensity(mod, params, m.rstep, m.n_bins, q, first_n=m.first_n, with_err=True)
            assert np.all(np.isfinite(I)) and np.all(np.isfinite(rel))
            assert abs(I[0] - tot ** 2) <= max(rel[0], 1e-12) * tot ** 2, (mode, I[0], tot ** 2)
            assert abs(I[1] - tot ** 2) <= 1e-9 * tot ** 2
            err = np.abs(I[1:] - Iref) / np.abs(Iref)
            assert np.all(err <= rel[1:] + 1e-14), (mode, err, rel)
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    mod.close()


def test_q_r_beyond_the_sine_range_recomputed_in_fp64(ctx):
    """Two clusters 4000 A apart and Q up to 400 A^-1: Q r passes 2^20, past which the
    matrix-core path's sines do not hold.  Its tiles there return NaN, which the error target
    turns into an FP64 recomputation (libm sines), equal to the FP64 mode and to the oracle;
    with the target off the NaN reaches the caller instead of a wrong value."""
    rng = np.random.default_rng(4000)
    a = rng.uniform(-3, 3, (10, 3))
    b = rng.uniform(-3, 3, (10, 3)) + [4000.0, 0, 0]
    p = np.concatenate([a, b]).astype(np.float32)
    gn = np.array([12, 8], np.int64)
    rstep = 0.01
    nb = O.channels(4020.0, rstep)
    q = np.array([0.5, 50.0, 400.0])
    gps = [synth.cromer_mann(rng) for _ in range(2)]
    for g in gps:  # form factors that do not underflow at 400 A^-1
        g["b"] = g["b"] * 1e-3
        g["beq"] = 0.0
    params = [_gp(g) for g in gps]
    mod = ctx.model(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), gn)
    try:
        ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
        I64, rel64 = ctx.intensity(mod, params, rstep, nb, q, first_n=8, with_err=True)
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        I, rel = ctx.intensity(mod, params, rstep, nb, q, first_n=8, with_err=True)
        assert ctx.pass2_refined() == 1
        assert np.array_equal(I, I64) and np.array_equal(rel, rel64)
        ctx.set_rel_err_target(0)
        Inan = ctx.intensity(mod, params, rstep, nb, q, first_n=8)
        assert np.isnan(Inan).all()
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
    H = ctx.histogram(mod, rstep, nb)
    mod.close()

    class M:
        pass
    M.group_n, M.q, M.rstep = gn, q, rstep
    M.params = gps
    Iref, _ = _oracle_I(M, H, 8)
    assert np.all(np.abs(I - Iref) / np.abs(Iref) <= rel + 1e-14)


def test_exact_zero_intensity_has_zero_bound(ctx):
    """Form factors of exactly 0 at one q (they underflow at large Q): I is exactly 0 there and
    its relative error bound 0, not 0 / 0; the other q keep their bounds.  Both modes."""
    rng = np.random.default_rng(12)
    H = np.zeros((3, 5000), np.uint64)
    for p in range(3):
        idx = rng.choice(np.arange(1, 5000), 400, replace=False)
        H[p, idx] = rng.integers(1, 1000, 400).astype(np.uint64)
    K = np.array([0.7, 3.0, 11.0])
    ff = rng.uniform(0.5, 2.0, (3, 3))
    ff[:, 1] = 0.0
    try:
        for mode in MODES:
            ctx.set_pass2_mode(mode)
            I, rel = ctx.debye(H, 1e-3, K, ff, first_n=8, with_err=True)
            assert I[1] == 0.0 and rel[1] == 0.0, (mode, I, rel)
            assert np.all(np.isfinite(rel)) and rel[0] > 0 and rel[2] > 0
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)

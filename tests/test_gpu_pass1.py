"""Pass 1 on the GPU against the CPU oracle: bit-exact uint64 histograms.

Fixture-anchored: the reference's own Data/GammaAl2O3_10nm.debx* (all six partials)."""
import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth
from oracle import oracle as O
from tests.sweep_models import SWEEP, sweep_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cd.Context(0)
    yield c
    c.close()


def _oracle_hist(m, n_bins=None):
    st = np.concatenate([[0], np.cumsum(m.group_n)[:-1]])
    S = len(m.group_n)
    nb = n_bins or m.n_bins
    rp = O.rp_of(m.rstep)
    H = np.zeros((S * (S + 1) // 2, nb), np.uint64)
    ov = 0
    p = 0
    for r in range(S):
        for c in range(r, S):
            _, o = O.histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c, rp, nb,
                               out=H[p])
            ov += o
            p += 1
    return H, ov


def test_sqrt_correctly_rounded_exhaustive():
    """Every float from 2^-96 up to FLT_MAX (bit patterns 0x0F800000..0x7F7FFFFF): the device
    square root of pass 1 is correctly rounded and equals the compiler's sqrtf."""
    bad, mism = cd.selftest_sqrt(0x0F800000, 0x7F800000)
    assert bad == 0 and mism == 0
    # s == 0 (self pairs, coincident atoms) gives +0
    assert cd.selftest_sqrt(0, 1)[0] == 0


def test_round_half_up_exhaustive():
    """Every float v in [0, 2^31): the one-instruction round-half-up conversion the bin rule
    uses (v_cvt_rpi_i32_f32) equals roundf's half-away-from-zero (trunc + (fract >= 0.5))."""
    assert cd.selftest_round(0, 0x4F000000) == 0


def test_bin_rule_matches_oracle_random_and_near_ties():
    rng = np.random.default_rng(11)
    n = 400000
    a = rng.uniform(-150, 150, (n, 3)).astype(np.float32)
    b = rng.uniform(-150, 150, (n, 3)).astype(np.float32)
    # near-tie pairs: b along x from a at distances (k + 0.5) * rstep +- a few ulp
    for rstep in (1e-3, 5e-4, 1e-4):
        rp = O.rp_of(rstep)
        k = rng.integers(1, 150000, 50000)
        d = ((k + 0.5) * rstep).astype(np.float32)
        jit = rng.integers(-3, 4, 50000).astype(np.int32)
        d = (d.view(np.int32) + jit).view(np.float32)
        a2 = rng.uniform(-10, 10, (50000, 3)).astype(np.float32)
        b2 = a2.copy()
        b2[:, 0] = (a2[:, 0] + d).astype(np.float32)
        A = np.concatenate([a, a2])
        B = np.concatenate([b, b2])
        got = cd.selftest_bins(A, B, rp)
        want = np.array([O.pair_bin(A[i], B[i], rp) for i in range(0, len(A), 7)])
        assert np.array_equal(got[::7], want)


def test_c1_fcc_sphere_bit_exact(ctx):
    m = synth.fcc_sphere()
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)
    n = m.n_atoms
    assert int(H.sum()) == n * (n + 1) // 2 and int(H[0, 0]) == n


@pytest.mark.parametrize("n,S,seed", [(1, 1, 0), (2, 1, 1), (63, 1, 2), (64, 2, 3), (65, 3, 4),
                                      (300, 2, 5), (1000, 4, 6), (2500, 5, 7)])
def test_small_random_models_bit_exact(ctx, n, S, seed):
    m = synth.small_random(n, S, seed)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)


def test_empty_subgroup_and_duplicates(ctx):
    """A subgroup with no atoms, and coincident atoms (all in bin 0)."""
    x = np.array([0, 0, 0, 1.5, 1.5], np.float32)
    y = np.array([0, 0, 0, 0.0, 0.0], np.float32)
    z = np.zeros(5, np.float32)
    gn = np.array([3, 0, 2], np.int64)
    mod = ctx.model(x, y, z, gn)
    H = ctx.histogram(mod, 1e-3, 2000)
    # pairs: (0,0): 3*4/2 = 6 all in bin 0; (0,2): 3*2 = 6 at 1.5 A; (2,2): 3 in bin 0
    assert H.shape == (6, 2000)
    assert H[0, 0] == 6 and H[0].sum() == 6
    assert H[1].sum() == 0 and H[3].sum() == 0 and H[4].sum() == 0
    assert H[2, 1500] == 6 and H[2].sum() == 6
    assert H[5, 0] == 3 and H[5].sum() == 3


def test_single_atom_and_far_from_origin(ctx):
    """One atom (its self pair only); a cluster about 7000 A from the origin, where FP32
    coordinate differences carry ~5e-4 A of rounding and the bin of each pair depends on the
    exact FP32 operation order."""
    one = np.array([3.25], np.float32)
    mod = ctx.model(one, one, one, np.array([1], np.int64))
    H = ctx.histogram(mod, 1e-3, 10)
    assert H.shape == (1, 10) and H[0, 0] == 1 and H.sum() == 1
    m = synth.small_random(2000, 2, 77, spread=15.0)
    x = (m.x.astype(np.float64) + 5000.0).astype(np.float32)
    y = (m.y.astype(np.float64) - 3000.0).astype(np.float32)
    z = (m.z.astype(np.float64) + 4000.0).astype(np.float32)
    m.x, m.y, m.z = x, y, z
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)


def test_rmax_too_small_reports_range_error(ctx):
    m = synth.small_random(500, 2, 9, spread=20.0)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    nb = O.channels(15.0, m.rstep)
    H, oor = ctx.histogram(mod, m.rstep, nb, return_oor=True)
    ref, ov = _oracle_hist(m, nb)
    assert oor == ov and ov > 0
    assert np.array_equal(H, ref)
    with pytest.raises(cd.DebyeError) as e:
        ctx.histogram(mod, m.rstep, nb)
    assert e.value.code == cd.DBX_ERR_RANGE


def test_nonfinite_coordinates_rejected(ctx):
    x = np.array([0, np.nan], np.float32)
    with pytest.raises(cd.DebyeError):
        ctx.model(x, x, x, np.array([2], np.int64))


def test_gamma_fixture_histograms_bit_exact(ctx, golden_dir, gamma_debx):
    """The reference's own outputs: Data/GammaAl2O3_10nm.debx0x0 ... debx2x2 (rstep 1e-4)."""
    g = O.parse_deb(str(golden_dir / "GammaAl2O3_10nm.deb"))
    nb = O.channels(g.rmax, g.rstep)
    mod = ctx.model(g.x, g.y, g.z, g.group_n)
    H = ctx.histogram(mod, g.rstep, nb)
    for p, (r, c) in enumerate(O.pair_list(g.group_n)):
        bins, cnt = gamma_debx[f"{r}x{c}"]
        ref = np.zeros(nb, np.uint64)
        ref[bins] = cnt
        assert np.array_equal(H[p], ref), (r, c)


def test_c2_spinel_bit_exact_and_totals(ctx):
    m = synth.spinel_sphere()
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    n = m.group_n
    assert [int(h.sum()) for h in H] == [n[0] * (n[0] + 1) // 2, n[0] * n[1], n[1] * (n[1] + 1) // 2]
    jobs, misses = ctx.pass1_stats()
    assert misses == 0
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)


def test_repeatable_and_permutation_invariant(ctx):
    """Integer histograms are identical across runs and under any atom order within a subgroup."""
    m = synth.small_random(3000, 2, 21, spread=15.0)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H1 = ctx.histogram(mod, m.rstep, m.n_bins)
    H2 = ctx.histogram(mod, m.rstep, m.n_bins)
    rng = np.random.default_rng(3)
    n0 = int(m.group_n[0])
    perm = np.concatenate([rng.permutation(n0), n0 + rng.permutation(int(m.group_n[1]))])
    mod2 = ctx.model(m.x[perm], m.y[perm], m.z[perm], m.group_n)
    H3 = ctx.histogram(mod2, m.rstep, m.n_bins)
    assert np.array_equal(H1, H2) and np.array_equal(H1, H3)


def test_bins_beyond_2_24_bit_exact(ctx):
    """Fine bins over a wide cloud: bin indices above 2^24 (1e-5 A bins, distances to ~200 A),
    where a single window's clamp bounds are no longer exact in FP32 and the planner sends those
    jobs to the checked variant."""
    m = synth.small_random(300, 2, 61, spread=60.0, rstep=1e-5)
    assert m.n_bins > 2 ** 24
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)
    assert H[:, 2 ** 24:].sum() > 0


@pytest.mark.parametrize("maker", ["c3_small", "c4_model5", "c5_req7", "sparse"])
def test_wide_tiles_and_window_groups_bit_exact(ctx, maker):
    """Workloads whose sparse subgroups and fine bins need 16-bit windows and several windows per
    shell: a 10 nm sphere of the rocksalt-like three-species lattice at 0.0005 A bins, two
    request models of the batched/serving streams, a sparse cloud."""
    m = {"c3_small": lambda: synth.rocksalt_sphere(diameter=100.0, nq=64),
         "c4_model5": lambda: synth.batch_model(5),
         "c5_req7": lambda: synth.serve_request(7),
         "sparse": lambda: synth.small_random(1500, 2, 43, spread=45.0)}[maker]()
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    n = m.group_n
    S = len(n)
    want = [n[a] * (n[a] + 1) // 2 if a == b else n[a] * n[b] for a in range(S) for b in range(a, S)]
    assert [int(h.sum()) for h in H] == want
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)


def test_model_lifetime_destroy_twice_and_close_with_live_models():
    """A second dbx_model_destroy of a handle is an argument error (no double free), and
    dbx_close releases the models still alive on the context."""
    m = synth.small_random(500, 2, 5)
    with cd.Context(0) as c:
        mod = c.model(m.x, m.y, m.z, m.group_n)
        h = mod.handle
        assert cd.lib().dbx_model_destroy(c.handle, h) == cd.DBX_OK
        mod.handle = None
        assert cd.lib().dbx_model_destroy(c.handle, h) == cd.DBX_ERR_ARG
        keep = [c.model(m.x, m.y, m.z, m.group_n) for _ in range(3)]
        H = c.histogram(keep[0], m.rstep, m.n_bins)
        a, b = (int(v) for v in m.group_n)
        assert [int(r.sum()) for r in H] == [a * (a + 1) // 2, a * b, b * (b + 1) // 2]
    # closed with three live models; a later close() of them is a no-op
    for k in keep:
        k.close()


def _device_plan_windows(m, mod):
    """win field of every pass-1 job the model's plan runs (dbx_debug_pass1_plan)."""
    import ctypes as C
    fn = cd.lib().dbx_debug_pass1_plan
    fn.restype = C.c_int
    nj, ne = C.c_int(), C.c_int()
    assert fn(mod.handle, C.c_double(m.rstep), None, 0, None, 0, C.byref(nj), C.byref(ne)) == 0
    jobs = np.zeros((nj.value, 12), np.int32)
    e2 = np.zeros(ne.value, np.float32)
    assert fn(mod.handle, C.c_double(m.rstep), jobs.ctypes.data_as(C.c_void_p), nj.value,
              e2.ctypes.data_as(C.c_void_p), ne.value, C.byref(nj), C.byref(ne)) == 0
    return jobs


@pytest.mark.parametrize("maker", ["stacked_lattice", "shell_around_stack"])
def test_16bit_windows_under_heavy_counts_bit_exact(ctx, maker):
    """16-bit LDS counters whose fields overflow again and again within one workgroup: a perfect
    lattice with eight coincident atoms per site (single 16-bit windows, over 65536 counts in a
    bin per job), and atoms on a sphere around 60,000 coincident atoms (multi-window 16-bit
    jobs; every cross pair of a workgroup lands in the one field of bin rho, over 10^6 per job,
    incremented by all 1024 threads at once).  Bit-exact against the oracle, closed forms."""
    m = getattr(synth, maker)()
    print(f"{m.name}: {m.n_atoms} atoms, {m.pair_count:.3g} pairs, {m.n_bins} bins", flush=True)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    jobs = _device_plan_windows(m, mod)
    w = jobs[:, 10]
    if maker == "stacked_lattice":
        assert np.sum((w & 4 != 0) & (w & 3 == 3)) > 100
    else:
        cross = w[jobs[:, 0] == 1]
        assert np.all(cross & 4) and np.all(cross & 3 != 3)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    n = [int(v) for v in m.group_n]
    S = len(n)
    assert [int(h.sum()) for h in H] == [n[a] * (n[a] + 1) // 2 if a == b else n[a] * n[b]
                                         for a in range(S) for b in range(a, S)]
    if maker == "shell_around_stack":
        k_rho = int(np.float32(m.meta["rho"]) * np.float32(1.0 / m.rstep) + np.float32(0.5))
        assert int(H[1, k_rho]) == n[0] * n[1]           # all cross pairs in one bin
        assert int(H[2, 0]) == n[1] * (n[1] + 1) // 2    # the stack: every pair at distance 0
    else:
        assert int(H[0, 0]) == n[0] * (8 + 1) // 2       # each site: its 8 atoms' 36 pairs
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)
    mod.close()


@pytest.mark.parametrize("kind", SWEEP)
def test_geometry_and_bin_width_sweep_bit_exact(ctx, kind):
    m = sweep_model(kind, 900 + SWEEP.index(kind))
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    mod.close()
    ref, ov = _oracle_hist(m)
    assert ov == 0
    assert np.array_equal(H, ref)
    n = int(m.group_n.sum())
    assert int(H.sum()) == n * (n + 1) // 2


def test_geometry_sweep_in_one_batch_bit_exact(ctx):
    """The sweep's models of 1e-3 A bins in one batch call (one merged pass-1 launch list)
    give each model's histograms, bit-exact."""
    kinds = ["rod", "two_clusters", "cubic_lattice_exact", "duplicates", "ragged_species"]
    models = [sweep_model(k, 900 + SWEEP.index(k)) for k in kinds]
    assert all(m.rstep == 1e-3 for m in models)
    mods = [ctx.model(m.x, m.y, m.z, m.group_n) for m in models]
    gp = lambda p: cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])
    items = [(mod, [gp(p) for p in m.params], m.n_bins, m.first_n) for mod, m in zip(mods, models)]
    _, hists = ctx.batch_intensity(items, 1e-3, models[0].q, with_hist=True)
    for mod in mods:
        mod.close()
    for m, H in zip(models, hists):
        ref, ov = _oracle_hist(m)
        assert ov == 0
        assert np.array_equal(H, ref), m.name

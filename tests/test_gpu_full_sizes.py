"""The benchmark configurations at their full sizes, checked through properties that do not
need the CPU oracle to finish (it would take minutes here): closed-form pair totals,
invariance of every histogram under a relabelling of the atoms (which changes the tiling and
the whole pass-1 plan), the matrix-core Debye sum against the FP64 one on the same
histograms, and each model of the full 256-model batch against the same model alone."""
import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cd.Context(0)
    yield c
    c.close()


def _totals(group_n):
    n = [int(v) for v in group_n]
    return [n[a] * (n[a] + 1) // 2 if a == b else n[a] * n[b] for a in range(len(n)) for b in range(a, len(n))]


def _params(m):
    return [cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"]) for p in m.params]


@pytest.mark.parametrize("diameter", [200.0, 256.0])
def test_rocksalt_sphere_full_size(ctx, diameter):
    """The three-species rocksalt-like sphere, 6 partials at 0.0005 A, 8192 Q: 20 nm
    (452,719 atoms, 1.02e11 pair distances) and 25.6 nm (948,557 atoms, 4.5e11)."""
    m = synth.rocksalt_sphere(diameter=diameter)
    print(f"rocksalt sphere D={diameter}: {m.n_atoms} atoms, {m.n_bins} bins, {len(m.q)} Q", flush=True)
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    H = ctx.histogram(mod, m.rstep, m.n_bins)
    assert [int(h.sum()) for h in H] == _totals(m.group_n)
    assert np.array_equal(H[[0, 3, 5], 0], np.asarray(m.group_n, np.uint64))  # self pairs in bin 0
    # the same atoms in another order within each subgroup: other tiles, same histograms
    rng = np.random.default_rng(2024)
    starts = np.concatenate([[0], np.cumsum(m.group_n)[:-1]])
    perm = np.concatenate([s + rng.permutation(int(n)) for s, n in zip(starts, m.group_n)])
    mod2 = ctx.model(m.x[perm], m.y[perm], m.z[perm], m.group_n)
    assert np.array_equal(ctx.histogram(mod2, m.rstep, m.n_bins), H)
    # pass 2 on the matrix cores against FP64 arithmetic on the same histograms
    pr = _params(m)
    P = len(H)
    ff = np.zeros((P, len(m.q)))
    p = 0
    for a in range(len(pr)):
        for b in range(a, len(pr)):
            ff[p] = ctx.prefactor(pr[a], pr[b], m.q)
            p += 1
    # the matrix-core sums with their error bound (no FP64 recomputation), against FP64 sums
    ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
    try:
        I64, b64 = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_err=True)
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(0)
        I, bmf = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_err=True)
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
        ctx.set_rel_err_target(1e-6)
    err = np.abs(I - I64) / np.abs(I64)
    print(f"max rel I(Q) matrix cores vs FP64: {err.max():.3e}, matrix-core bound {bmf.max():.3e}, "
          f"FP64 bound {b64.max():.3e}", flush=True)
    assert err.max() < 1e-6
    assert np.all(err <= bmf + b64)
    # the default target: every I within 1e-6 by its bound, recomputed in FP64 where needed
    Id, bd = ctx.debye(H, m.rstep, m.q, ff, first_n=m.first_n, with_err=True)
    assert bd.max() <= 1e-6 and ctx.pass2_refined() == int(bmf.max() > 1e-6)
    assert np.all(np.abs(Id - I64) <= (bd + b64) * np.abs(I64))


def test_batch_of_256_models_full_size(ctx):
    """The 256-model batch (1k-60k atoms each, 5.1e10 pair distances) in one call: every
    model's totals, and 32 of them (every 8th) bit for bit and within 1e-9 against the
    same model alone."""
    models = synth.batch(256)
    items = [(ctx.model(m.x, m.y, m.z, m.group_n), _params(m), m.n_bins, m.first_n) for m in models]
    I, H = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_hist=True)
    for m, h in zip(models, H):
        assert [int(r.sum()) for r in h] == _totals(m.group_n)
    for i in range(0, 256, 8):
        m, it = models[i], items[i]
        Ia, Ha = ctx.intensity(it[0], it[1], m.rstep, m.n_bins, m.q, first_n=m.first_n, with_hist=True)
        assert np.array_equal(H[i], Ha), f"model {i}"
        assert np.max(np.abs(I[i] - Ia) / np.abs(Ia)) < 1e-9, f"model {i}"

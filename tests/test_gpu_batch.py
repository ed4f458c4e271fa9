"""Batches of models (dbx_batch_intensity) on the GPU: one pass-1 launch over the merged work
list of every model and one pass-2 contraction over the rows of every model.  Each item must
match the oracle (histograms bit-exact, I within the pass-2 mode's bound) and must match the
same model run alone (histograms bit-exact, I within 1e-9 relative: only the FP64 order in
which split partial sums are added differs)."""
import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth
from oracle import oracle as O
from tests.test_gpu_pass1 import _oracle_hist

pytestmark = pytest.mark.gpu

TOL = {cd.DBX_PASS2_FP64: 1e-9, cd.DBX_PASS2_MFMA: 1e-6}
ALONE_TOL = 1e-9


@pytest.fixture(scope="module")
def ctx():
    c = cd.Context(0)
    yield c
    c.close()


def _gp(p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


def _oracle_I(m, H):
    class Pm:
        pass
    Pm.group_n = m.group_n
    Pm.occ = np.array([p["occ"] for p in m.params])
    Pm.beq = np.array([p["beq"] for p in m.params])
    Pm.a = np.array([p["a"] for p in m.params])
    Pm.b = np.array([p["b"] for p in m.params])
    Pm.c = np.array([p["c"] for p in m.params])
    return O.intensity_from_histograms(H, Pm, m.q, m.rstep, first_n=m.first_n)


def _items(ctx, models, n_bins=None):
    out = []
    for i, m in enumerate(models):
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        nb = m.n_bins if n_bins is None else n_bins[i]
        out.append((mod, [_gp(p) for p in m.params], nb, m.first_n))
    return out


def _mixed_models(n_models, seed, nq=256):
    rng = np.random.default_rng(seed)
    models = []
    for i in range(n_models):
        n = int(np.exp(rng.uniform(np.log(150), np.log(2500))))
        S = int(rng.integers(1, 5))
        spread = float(rng.uniform(5.0, 16.0))
        models.append(synth.small_random(n, S, seed * 100 + i, spread=spread, nq=nq))
    return models


def _rel(a, b):
    return np.max(np.abs(a - b) / np.abs(b))


@pytest.mark.parametrize("mode", [cd.DBX_PASS2_MFMA, cd.DBX_PASS2_FP64], ids=["mfma", "fp64"])
@pytest.mark.parametrize("n_models,seed", [(3, 11), (12, 12)], ids=["narrow", "wide"])
def test_batch_matches_oracle(ctx, mode, n_models, seed):
    models = _mixed_models(n_models, seed)
    rows = sum(m.n_pairs for m in models)
    print(f"batch {n_models} models, {rows} rows, mode {mode}")
    ctx.set_pass2_mode(mode)
    try:
        items = _items(ctx, models)
        I, H = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_hist=True)
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    worst = 0.0
    for i, m in enumerate(models):
        ref, ov = _oracle_hist(m)
        assert ov == 0
        assert np.array_equal(H[i], ref), f"item {i} histogram"
        Iref = _oracle_I(m, ref)
        rel = _rel(I[i], Iref)
        worst = max(worst, rel)
        assert rel < TOL[mode], f"item {i}: rel {rel:.3e}"
    print(f"worst rel err vs oracle {worst:.3e}")


def test_batch_equals_each_model_alone(ctx):
    models = _mixed_models(14, 21)
    items = _items(ctx, models)
    I, H = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_hist=True)
    for i, (m, it) in enumerate(zip(models, items)):
        Ia, Ha = ctx.intensity(it[0], it[1], m.rstep, m.n_bins, m.q, first_n=m.first_n, with_hist=True)
        assert np.array_equal(H[i], Ha)
        assert _rel(I[i], Ia) < ALONE_TOL, f"item {i}"


def test_batch_permutation_invariant(ctx):
    models = _mixed_models(9, 31)
    items = _items(ctx, models)
    I1, H1 = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_hist=True)
    perm = np.random.default_rng(3).permutation(len(items))
    I2, H2 = ctx.batch_intensity([items[j] for j in perm], models[0].rstep, models[0].q, with_hist=True)
    for k, j in enumerate(perm):
        assert np.array_equal(H1[j], H2[k])
        assert _rel(I2[k], I1[j]) < ALONE_TOL


def test_batch_duplicate_model_and_single_item(ctx):
    models = _mixed_models(2, 41)
    items = _items(ctx, models)
    I, H = ctx.batch_intensity([items[0], items[1], items[0]], models[0].rstep, models[0].q, with_hist=True)
    assert np.array_equal(H[0], H[2]) and _rel(I[0], I[2]) < ALONE_TOL
    I1 = ctx.batch_intensity([items[1]], models[0].rstep, models[0].q)
    assert _rel(I1[0], I[1]) < ALONE_TOL


def test_batch_range_error_is_per_item(ctx):
    models = _mixed_models(4, 51)
    short = O.channels(6.0, models[2].rstep)
    nb = [m.n_bins for m in models]
    nb[2] = short
    items = _items(ctx, models, nb)
    I, H, oor = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_hist=True, return_oor=True)
    ref2, ov2 = _oracle_hist(models[2], short)
    assert ov2 > 0 and oor[2] == ov2
    assert np.array_equal(H[2], ref2)
    for i in (0, 1, 3):
        assert oor[i] == 0
        ref, _ = _oracle_hist(models[i])
        assert np.array_equal(H[i], ref)
    with pytest.raises(cd.DebyeError) as e:
        ctx.batch_intensity(items, models[0].rstep, models[0].q)
    assert e.value.code == cd.DBX_ERR_RANGE


def test_batch_of_fault_and_ellipsoid_models(ctx):
    """Models of the 256-model batched workload (FCC/HCP spheres and ellipsoids with stacking
    faults), the smaller ones: histograms equal each model alone, I within 1e-6 of the FP64
    oracle evaluated on those histograms."""
    sizes = [(i, synth.batch_model(i)) for i in range(40)]
    picked = [m for _, m in sorted(sizes, key=lambda t: t[1].n_atoms)[:8]]
    print("atoms", [m.n_atoms for m in picked])
    items = _items(ctx, picked)
    I, H = ctx.batch_intensity(items, picked[0].rstep, picked[0].q, with_hist=True)
    worst = 0.0
    for i, (m, it) in enumerate(zip(picked, items)):
        Ha = ctx.histogram(it[0], m.rstep, m.n_bins)
        assert np.array_equal(H[i], Ha)
        assert int(H[i].sum()) == m.pair_count
        Iref = _oracle_I(m, H[i])
        rel = _rel(I[i], Iref)
        worst = max(worst, rel)
        assert rel < 1e-6, f"item {i}: {rel:.3e}"
    print(f"worst rel err vs oracle {worst:.3e}")

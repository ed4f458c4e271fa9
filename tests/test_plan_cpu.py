"""Pass-1 work plan on the host (no device): the planner's jobs cover every pair of every
subgroup pair exactly once, and every job the fast kernel runs holds its pairs' bins: a
single-window job all of them inside its LDS window, the consecutive windows of a wide shell
all of them inside their union (each pair counted by the one window that contains it).  The
kernel's shell test and bin rule are emulated in NumPy float32 (IEEE, unfused), so these
checks see the same per-pair values as the device."""
import ctypes as C

import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth

W = 35328  # LDS window bins (kP1Bins)


def host_plan(m):
    L = cd.lib()
    fn = L.dbx_debug_plan_host
    fn.restype = C.c_int
    x, y, z = (np.ascontiguousarray(v, np.float32) for v in (m.x, m.y, m.z))
    gn = np.ascontiguousarray(m.group_n, np.int64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    nt, nj, ne, na = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = fn(p(x), p(y), p(z), p(gn), len(gn), C.c_double(m.rstep), None, 0, None, C.c_int64(0),
            None, 0, None, 0, None, 0, C.byref(nt), C.byref(nj), C.byref(ne), C.byref(na))
    assert rc == 0
    tiles = np.zeros((nt.value, 8), np.int32)
    pos = np.zeros((m.n_atoms, 4), np.float32)
    jobs = np.zeros((nj.value, 12), np.int32)
    e2 = np.zeros(ne.value, np.float32)
    al = np.zeros(max(1, na.value), np.int32)
    rc = fn(p(x), p(y), p(z), p(gn), len(gn), C.c_double(m.rstep), p(tiles), nt.value, p(pos),
            C.c_int64(m.n_atoms), p(jobs), nj.value, p(e2), ne.value, p(al), na.value,
            C.byref(nt), C.byref(nj), C.byref(ne), C.byref(na))
    assert rc == 0
    return tiles, pos, jobs, e2, al[:na.value]


def _bins(A, B, rp):
    d = (B[None, :, :] - A[:, None, :]).astype(np.float32)
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    v = np.sqrt(s) * np.float32(rp)
    return v.astype(np.int64) + ((v - np.floor(v)) >= 0.5)


def check_plan(m):
    tiles, pos, jobs, e2, alist = host_plan(m)
    tf = tiles.view(np.float32)
    cen, rad = tf[:, :3], tf[:, 3]
    start, cnt, grp, parent = tiles[:, 4], tiles[:, 5], tiles[:, 6], tiles[:, 7]
    rp = np.float32(1.0 / m.rstep)
    S = len(m.group_n)
    assert np.all(cnt >= 1) and np.all(cnt <= 64)
    pairs = [(a, b) for a in range(S) for b in range(a, S)]
    seen = {}
    counted = np.zeros(len(pairs), np.int64)
    for J in jobs:
        pair, a_off, a_cnt, b0, b1, same, shell, klo, eoff, degraded, win = J[:11]
        ga, gb = pairs[pair]
        assert (win & ~7) == 0
        assert not degraded or win == 3
        width = 2 * W if win & 4 else W  # 16-bit counters: twice the bins
        assert not (win & 4) or klo % 2 == 0  # 16-bit windows address bin k's field from k alone
        lo, hi = e2[eoff + shell], e2[eoff + shell + 1]
        for t in alist[a_off:a_off + a_cnt]:
            assert grp[t] == ga and cnt[t] <= 64
            A = pos[start[t]:start[t] + cnt[t], :3]
            ia = np.arange(start[t], start[t] + cnt[t])
            for j in range(max(b0, parent[t]) if same else b0, b1):
                assert grp[j] == gb and cnt[j] <= 64
                key = (pair, t, j, int(shell), int(klo))
                assert key not in seen, key  # each (A tile, B tile, shell, window) once
                seen[key] = True
                B = pos[start[j]:start[j] + cnt[j], :3]
                e = (B - cen[t]).astype(np.float32)
                d2c = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
                sel = (d2c >= lo) & (d2c < hi)
                if not np.any(sel):
                    continue
                k = _bins(A, B[sel], rp)
                if same and j == parent[t]:
                    ib = np.arange(start[j], start[j] + cnt[j])[sel]
                    valid = ia[:, None] <= ib[None, :]
                else:
                    valid = np.ones(k.shape, bool)
                if degraded:
                    counted[pair] += int(valid.sum())  # the checked kernel counts every pair
                    continue
                inside = (k >= klo) & (k < klo + width)
                counted[pair] += int((valid & inside).sum())
                if win & 1:
                    assert np.all((k >= klo) | ~valid), (pair, t, j, shell)
                if win & 2:
                    assert np.all((k < klo + width) | ~valid), (pair, t, j, shell)
    n = m.group_n
    want = [n[a] * (n[a] + 1) // 2 if a == b else n[a] * n[b] for a, b in pairs]
    assert list(counted) == want
    return jobs


@pytest.mark.parametrize("maker", ["c1", "rand3", "batch3", "batch0", "sparse", "fine", "huge"])
def test_plan_covers_every_pair_once_and_windows_hold(maker):
    m = {"c1": synth.fcc_sphere,
         "rand3": lambda: synth.small_random(3000, 3, 41, spread=20.0),
         "batch3": lambda: synth.batch_model(3),
         "batch0": lambda: synth.batch_model(0),
         # sparse species: a few atoms spread over a large volume (16-bit and multiple windows)
         "sparse": lambda: synth.small_random(1500, 2, 43, spread=45.0),
         # fine bins: narrow windows (several per shell)
         "fine": lambda: synth.small_random(1200, 1, 44, spread=12.0, rstep=1e-4),
         # bins above 2^24: windows there, any single one degraded (its FP32 clamp is inexact)
         "huge": lambda: synth.small_random(300, 2, 61, spread=60.0, rstep=1e-5)}[maker]()
    jobs = check_plan(m)
    assert len(jobs) > 0
    if maker == "huge":
        single_high = (jobs[:, 10] & 3 == 3) & (jobs[:, 7].astype(np.int64) + np.where(jobs[:, 10] & 4, 2 * W, W) >= 2 ** 24)
        assert (jobs[:, 7] >= 2 ** 24).any() and np.all(jobs[single_high, 9] == 1)


def test_plan_uses_several_windows_where_needed():
    """Sparse species (wide 64-atom tiles) and fine bins: shells counted in several windows;
    the bench workload: one window for nearly every job, no checked jobs."""
    for m in (synth.small_random(1500, 2, 43, spread=45.0),
              synth.small_random(1200, 1, 44, spread=12.0, rstep=1e-4)):
        _, _, jobs, _, _ = host_plan(m)
        assert np.any(jobs[:, 10] != 3) and np.any(jobs[:, 10] & 4) and jobs[:, 9].sum() == 0
    _, _, jobs, _, _ = host_plan(synth.spinel_sphere())
    assert jobs[:, 9].sum() == 0 and np.mean(jobs[:, 10] == 3) > 0.97


def _leaf_sets(x, y, z, ids, unit=64):
    """The tiling rule restated: median bisection along the longest float32 bounding-box axis
    (the first of equal extents), the left part the nt // 2 * unit atoms first in (coordinate,
    atom index) order, depth first, left before right."""
    xyz = np.stack([x, y, z], 1).astype(np.float32)
    out = []
    stack = [ids]
    while stack:
        r = stack.pop()
        if len(r) <= unit:
            out.append(r)
            continue
        ext = xyz[r].max(0) - xyz[r].min(0)
        ax = 0
        for a in (1, 2):
            if ext[a] > ext[ax]:
                ax = a
        nl = (-(-len(r) // unit) // 2) * unit
        o = r[np.lexsort((r, xyz[r, ax]))]
        stack.append(o[nl:])
        stack.append(o[:nl])
    return out


def test_tiles_are_the_median_bisection_leaves():
    """Every 64-atom tile holds exactly the atoms of one leaf of the median bisection, for a
    subgroup large enough to be split on several threads (a 46k-atom lattice with exact
    coordinate ties) and a small one."""
    rng = np.random.default_rng(12)
    a = 4.08
    g = np.arange(-14, 14)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    p = ((cells[:, None, :] + base[None]) * a).reshape(-1, 3)
    p = p[(p ** 2).sum(1) <= (14 * a) ** 2]
    small = rng.uniform(-10, 10, (700, 3))
    xyz = np.concatenate([p, small]).astype(np.float32)
    gn = np.array([len(p), len(small)], np.int64)

    class M:
        pass
    m = M()
    m.x, m.y, m.z = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
    m.group_n, m.rstep, m.n_atoms = gn, 1e-3, len(xyz)
    assert gn[0] > 32768
    tiles, pos, _, _, _ = host_plan(m)
    got = sorted(tuple(sorted(map(tuple, pos[t[4]:t[4] + t[5], :3].tolist()))) for t in tiles)
    want = []
    start = 0
    for n in gn:
        ids = np.arange(start, start + n)
        for leaf in _leaf_sets(m.x, m.y, m.z, ids):
            want.append(tuple(sorted(map(tuple, xyz[leaf].tolist()))))
        start += n
    assert got == sorted(want)


def _job_bin_counts(m, tiles, pos, J, e2, alist):
    """Counts per window bin that one job adds (its shell's pairs inside its window), emulated
    like check_plan: the kernel's shell test and bin rule in float32."""
    tf = tiles.view(np.float32)
    cen, start, cnt, parent = tf[:, :3], tiles[:, 4], tiles[:, 5], tiles[:, 7]
    rp = np.float32(1.0 / m.rstep)
    pair, a_off, a_cnt, b0, b1, same, shell, klo, eoff, degraded, win = J[:11]
    lo, hi = e2[eoff + shell], e2[eoff + shell + 1]
    width = 2 * W if win & 4 else W
    acc = np.zeros(width, np.int64)
    for t in alist[a_off:a_off + a_cnt]:
        j0 = max(b0, parent[t]) if same else b0
        if j0 >= b1:
            continue
        A = pos[start[t]:start[t] + cnt[t], :3]
        ia = np.arange(start[t], start[t] + cnt[t])
        ib = np.arange(start[j0], start[b1 - 1] + cnt[b1 - 1])  # tiles j0..b1-1 are contiguous
        B = pos[ib, :3]
        e = (B - cen[t]).astype(np.float32)
        d2c = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
        sel = (d2c >= lo) & (d2c < hi)
        if not sel.any():
            continue
        k = _bins(A, B[sel], rp)
        valid = ia[:, None] <= ib[sel][None, :] if same else np.ones(k.shape, bool)
        kk = k[valid] - klo
        kk = kk[(kk >= 0) & (kk < width)]
        acc += np.bincount(kk, minlength=width)[:width]
    return acc


def test_plan_16bit_windows_take_heavy_counts():
    """The workloads of the 16-bit counter stress tests on the GPU: a perfect lattice with eight
    coincident atoms per site runs single 16-bit windows in which one job adds more than 65536
    counts to a bin (every such field drains at least twice), and atoms on a sphere around a
    stack of coincident atoms run multi-window 16-bit jobs whose every cross pair falls into
    the same bin, more than 10^6 of them per job."""
    m = synth.stacked_lattice()
    tiles, pos, jobs, e2, al = host_plan(m)
    w = jobs[:, 10]
    single16 = np.nonzero((w & 4 != 0) & (w & 3 == 3))[0]
    assert len(single16) > 100 and jobs[:, 9].sum() == 0
    heavy = max(int(_job_bin_counts(m, tiles, pos, jobs[j], e2, al).max()) for j in single16[:48])
    assert heavy > 65536, heavy

    m = synth.shell_around_stack()
    tiles, pos, jobs, e2, al = host_plan(m)
    cross = jobs[jobs[:, 0] == 1]
    assert np.all(cross[:, 10] & 4) and np.all(cross[:, 10] & 3 != 3) and cross[:, 9].sum() == 0
    cnt = tiles[:, 5]
    # every cross pair is at distance rho: the window of a shell that holds rho's bin counts
    # all the job's cross pairs in one field, A atoms x B atoms of its range
    k_rho = int(np.float32(m.meta["rho"]) * np.float32(1.0 / m.rstep) + np.float32(0.5))
    mine = cross[(cross[:, 7] <= k_rho) & (k_rho < cross[:, 7] + 2 * W)]
    per_job = [int(cnt[al[J[1]:J[1] + J[2]]].sum()) * int(cnt[J[3]:J[4]].sum()) for J in mine]
    assert len(mine) > 0 and max(per_job) > 10 ** 6, max(per_job)
    for J, n in list(zip(mine, per_job))[:6]:
        h = _job_bin_counts(m, tiles, pos, J, e2, al)
        assert h.max() == n == h.sum() and h[k_rho - J[7]] == n

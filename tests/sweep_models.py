"""Models of unusual geometry, subgroup layout and bin width for the parity sweeps of pass 1
and pass 2 (synthetic, seeded; none comes from the reference's files)."""
import numpy as np

from cudebye_amd import synth


def sweep_model(kind, seed):
    """One model of the sweep: each exercises a different part of the planner (shell widths,
    window groups, 16-bit windows, empty and tiny subgroups, exact distance ties, heavy bins)
    and of the pass-2 digit classes (counts from 1 to ~870k per bin)."""
    rng = np.random.default_rng(seed)
    S, rstep, sp = 2, 1e-3, None
    if kind == "rod":  # 400 A long, 0.3 A thick: one long row of tiles
        p = np.zeros((3000, 3))
        p[:, 0] = rng.uniform(0, 400, 3000)
        p[:, 1:] = rng.normal(0, 0.3, (3000, 2))
    elif kind == "plane":  # every z exactly 0
        p = np.zeros((4000, 3))
        p[:, :2] = rng.uniform(-50, 50, (4000, 2))
        S, rstep = 3, 5e-4
    elif kind == "two_clusters":  # 300 A apart: a long empty stretch of bins
        p = np.concatenate([rng.uniform(-10, 10, (2000, 3)), rng.uniform(-10, 10, (2000, 3)) + [300.0, 0, 0]])
    elif kind == "hollow_shell_fine":  # a 40 A sphere surface at the reference's 1e-4 A bins
        v = rng.normal(size=(5000, 3))
        p = 40.0 * v / np.linalg.norm(v, axis=1, keepdims=True)
        S, rstep = 1, 1e-4
    elif kind == "cubic_lattice_exact":  # sigma 0: many exactly equal distances, heavy bins
        g = np.arange(20) * 2.0
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        sp = ((X + Y + Z).ravel() / 2.0).astype(np.int64) % 2
    elif kind == "duplicates":  # every point eight times: heavy bin 0 and repeated distances
        p = np.repeat(rng.uniform(-6, 6, (500, 3)), 8, axis=0)
        S = 3
    elif kind == "coarse_bins":
        p = rng.uniform(-15, 15, (5000, 3))
        rstep = 0.05
    elif kind == "very_coarse_bins":  # ~60 bins, up to ~1e6 counts each
        p = rng.uniform(-15, 15, (6000, 3))
        S, rstep = 1, 1.0
    elif kind == "large_sparse_six_species":
        v = rng.normal(size=(20000, 3))
        r = 75.0 * rng.uniform(0, 1, 20000) ** (1 / 3)
        p = v / np.linalg.norm(v, axis=1, keepdims=True) * r[:, None]
        S, rstep = 6, 2e-3
    elif kind == "ragged_species":  # subgroups of 1, 2, 5000, 3, 0 and 64 atoms
        sizes = [1, 2, 5000, 3, 0, 64]
        p = rng.uniform(-12, 12, (sum(sizes), 3))
        sp = np.repeat(np.arange(len(sizes)), sizes)
        S = len(sizes)
    elif kind == "fine_bins_wide":  # 3.6M bins of 1e-4 A
        p = rng.uniform(-100, 100, (3000, 3))
        rstep = 1e-4
    else:
        raise ValueError(kind)
    if sp is None:
        sp = rng.integers(0, S, len(p))
    groups = [p[sp == s] for s in range(S)]
    return synth._finish(kind, groups, [synth.cromer_mann(rng) for _ in range(S)], rstep, synth.q_grid_2theta(64),
                         dict(seed=seed))


SWEEP = ["rod", "plane", "two_clusters", "hollow_shell_fine", "cubic_lattice_exact", "duplicates", "coarse_bins",
         "very_coarse_bins", "large_sparse_six_species", "ragged_species", "fine_bins_wide"]

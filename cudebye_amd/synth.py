"""Seeded synthetic structure models for the benchmark workloads (seed = 1000*config + index).

    C1  FCC sphere, a = 4.08 A, diameter 40 A, one species, bins 0.001 A, 4096 Q
    C2  spinel-like sphere, diameter 100 A: O on the FCC sublattice of a = 7.91 A, cations on a
        seeded random 2/3 of the tetrahedral + octahedral sites, Gaussian displacements
        sigma = 0.03 A; two species, bins 0.001 A, 4096 Q
    C3  rocksalt-like sphere a = 4.2 A, diameter 200 A, species fractions 0.5/0.3/0.2,
        bins 0.0005 A, 8192 Q uniform in 0.3..20 A^-1
    C4  batch of FCC/HCP spheres or ellipsoids with stacking faults (probability 0.1 per
        close-packed layer) and displacements sigma in [0.01, 0.08] A; two species;
        atom counts log-uniform in [1k, 60k]; bins 0.001 A, 4096 Q
    C5  request stream: atom counts log-uniform in [500, 200k], 1-4 species, 2048/4096/8192 Q

Coordinates are generated in FP64 and rounded once to FP32; they are never recentred
afterwards.  Form factors are 4-Gaussian + constant Cromer-Mann models with seeded
coefficients a_i in [1, 20], b_i in [0.5, 60] A^2, c in [0, 2] (a5 = b5 = 0).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

LAMBDA_CU = 1.5406
PI = 3.14159265358979323846


@dataclass
class SynthModel:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    group_n: np.ndarray          # int64[S]
    params: list                 # per subgroup: dict(occ, beq, a[5], b[5], c)
    rstep: float
    n_bins: int
    q: np.ndarray                # FP64 Q grid (A^-1)
    first_n: int = 16
    meta: dict = field(default_factory=dict)

    @property
    def n_atoms(self):
        return int(self.group_n.sum())

    @property
    def n_pairs(self):
        S = len(self.group_n)
        return S * (S + 1) // 2

    @property
    def pair_count(self):
        """pair distances evaluated, self pairs included: N(N+1)/2"""
        n = self.n_atoms
        return n * (n + 1) // 2


def q_grid_2theta(n, tt_min=10.0, tt_max=150.0, lam=LAMBDA_CU):
    """n points of K = 4 pi sin(theta)/lambda over 2theta in [tt_min, tt_max] (the formula of
    cuDebye-v15.cu:696 with an explicit point count)."""
    tt = tt_min + np.arange(n, dtype=np.float64) * ((tt_max - tt_min) / (n - 1))
    return 4 * PI * np.sin(tt / 2 * PI / 180) / lam


def q_grid_uniform(n, qmin, qmax):
    return qmin + np.arange(n, dtype=np.float64) * ((qmax - qmin) / (n - 1))


def cromer_mann(rng):
    a = np.zeros(5)
    b = np.zeros(5)
    a[:4] = rng.uniform(1, 20, 4)
    b[:4] = rng.uniform(0.5, 60, 4)
    return dict(occ=1.0, beq=float(rng.uniform(0.0, 1.0)), a=a, b=b, c=float(rng.uniform(0, 2)))


def _rmax_bins(pts, rstep):
    """rmax = 2 ceil(1.025 max|p - c|) about the bounding-box centre (the rule of
    Matlab/xyz2deb.m), channels = int(rmax/rstep) + 2 (cuDebye-v15.cu:734)."""
    p = pts.astype(np.float64)
    c = 0.5 * (p.min(axis=0) + p.max(axis=0))
    maxr = float(np.sqrt(((p - c) ** 2).sum(1).max())) if len(p) else 0.0
    rmax = 2.0 * math.ceil(1.025 * maxr) + 2.0
    return rmax, int(rmax / rstep) + 2


def _finish(name, species_pts, params, rstep, q, meta, first_n=16):
    pts = [np.asarray(p, np.float64) for p in species_pts]
    allp = np.concatenate([p for p in pts if len(p)], axis=0).astype(np.float32)
    rmax, nb = _rmax_bins(allp, rstep)
    gn = np.array([len(p) for p in pts], np.int64)
    meta = dict(meta, rmax=rmax)
    return SynthModel(name, allp[:, 0].copy(), allp[:, 1].copy(), allp[:, 2].copy(), gn, params,
                      rstep, nb, q, first_n, meta)


def fcc_sphere(a=4.08, diameter=40.0, seed=1001, rstep=1e-3, nq=4096):
    """C1: single-species FCC sphere."""
    rng = np.random.default_rng(seed)
    R = diameter / 2
    n = int(math.ceil(R / a)) + 1
    g = np.arange(-n, n + 1, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    p = (cells[:, None, :] + basis[None]).reshape(-1, 3) * a
    p = p[(p ** 2).sum(1) <= R * R]
    return _finish("C1-fcc-sphere", [p], [cromer_mann(rng)], rstep, q_grid_2theta(nq),
                   dict(config=1, a=a, diameter=diameter, seed=seed))


def spinel_sphere(a=7.91, diameter=100.0, sigma=0.03, frac=2.0 / 3.0, seed=2001, rstep=1e-3, nq=4096):
    """C2: O on the FCC sublattice (a/2) of the cubic cell a, cations on a seeded random
    `frac` of the tetrahedral + octahedral interstices of that sublattice."""
    rng = np.random.default_rng(seed)
    ao = a / 2
    R = diameter / 2
    n = int(math.ceil(R / ao)) + 1
    g = np.arange(-n, n + 1, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    o_basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    tet = np.array([[x, y, z] for x in (0.25, 0.75) for y in (0.25, 0.75) for z in (0.25, 0.75)])
    octa = np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5], [0.5, 0.5, 0.5]])
    O = (cells[:, None, :] + o_basis[None]).reshape(-1, 3) * ao
    sites = (cells[:, None, :] + np.concatenate([tet, octa])[None]).reshape(-1, 3) * ao
    O = O[(O ** 2).sum(1) <= R * R]
    sites = sites[(sites ** 2).sum(1) <= R * R]
    M = sites[rng.random(len(sites)) < frac]
    O = O + rng.normal(0, sigma, O.shape)
    M = M + rng.normal(0, sigma, M.shape)
    return _finish("C2-spinel-sphere", [O, M], [cromer_mann(rng), cromer_mann(rng)], rstep,
                   q_grid_2theta(nq), dict(config=2, a=a, diameter=diameter, sigma=sigma, seed=seed))


def rocksalt_sphere(a=4.2, diameter=200.0, fractions=(0.5, 0.3, 0.2), seed=3001, rstep=5e-4,
                    nq=8192, qmin=0.3, qmax=20.0, sigma=0.0):
    """C3: rocksalt-like lattice, species drawn per site with the given fractions."""
    rng = np.random.default_rng(seed)
    R = diameter / 2
    h = a / 2
    n = int(math.ceil(R / h)) + 1
    g = np.arange(-n, n + 1, dtype=np.float64) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    p = p[(p ** 2).sum(1) <= R * R]
    if sigma > 0:
        p = p + rng.normal(0, sigma, p.shape)
    sp = rng.choice(len(fractions), size=len(p), p=np.asarray(fractions) / sum(fractions))
    groups = [p[sp == s] for s in range(len(fractions))]
    return _finish("C3-rocksalt-sphere", groups, [cromer_mann(rng) for _ in fractions], rstep,
                   q_grid_uniform(nq, qmin, qmax), dict(config=3, a=a, diameter=diameter, seed=seed))


def _close_packed(rng, n_target, kind, d_nn, fault_p, shape_axes):
    """Close-packed layers stacked along z with stacking faults; cut to an ellipsoid."""
    # in-plane triangular lattice, layer spacing sqrt(2/3) d
    dz = math.sqrt(2.0 / 3.0) * d_nn
    vol_per_atom = d_nn ** 3 / math.sqrt(2.0)
    ax = np.asarray(shape_axes, np.float64)
    R = (3 * n_target * vol_per_atom / (4 * PI * ax.prod())) ** (1 / 3)
    semi = R * ax
    nl = int(math.ceil(semi[2] / dz)) + 1
    nxy = int(math.ceil(1.5 * max(semi[0], semi[1]) / d_nn)) + 2  # rhombic patch covers the ellipse
    i, j = np.meshgrid(np.arange(-nxy, nxy + 1), np.arange(-nxy, nxy + 1), indexing="ij")
    base = np.stack([(i + 0.5 * j).ravel() * d_nn, (j * math.sqrt(3) / 2).ravel() * d_nn], 1)
    shifts = np.array([[0, 0], [0.5 * d_nn, d_nn / (2 * math.sqrt(3))],
                       [d_nn, d_nn / math.sqrt(3)]])
    seq = []
    prev2, prev = 0, 1
    seq = [prev2, prev]
    direction = 1
    for _ in range(2 * nl + 1 - 2):
        if kind == "fcc":
            if rng.random() < fault_p:
                direction = -direction
            nxt = (prev + direction) % 3
        else:  # hcp: ABAB..., a fault puts the layer on the third position
            nxt = prev2 if rng.random() >= fault_p else 3 - prev - prev2
        seq.append(nxt)
        prev2, prev = prev, nxt
    layers = []
    for li, s in enumerate(seq):
        zc = (li - nl) * dz
        xy = base + shifts[s]
        layers.append(np.column_stack([xy, np.full(len(xy), zc)]))
    p = np.concatenate(layers)
    p = p[((p / semi) ** 2).sum(1) <= 1.0]
    return p


def batch_model(index, seed_base=4000, n_lo=1000, n_hi=60000, rstep=1e-3, nq=4096, q=None):
    """C4 model `index`: two species on an FCC or HCP sphere/ellipsoid with stacking faults."""
    rng = np.random.default_rng(seed_base + index)
    n_target = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
    kind = "fcc" if rng.random() < 0.5 else "hcp"
    axes = (1.0, 1.0, 1.0) if rng.random() < 0.5 else tuple(rng.uniform(0.75, 1.3, 3))
    sigma = float(rng.uniform(0.01, 0.08))
    p = _close_packed(rng, n_target, kind, 2.885, 0.1, axes)
    p = p + rng.normal(0, sigma, p.shape)
    f = rng.uniform(0.2, 0.8)
    sp = rng.random(len(p)) < f
    params = [cromer_mann(rng), cromer_mann(rng)]
    return _finish(f"C4-{index:03d}-{kind}", [p[sp], p[~sp]], params, rstep,
                   q if q is not None else q_grid_2theta(nq),
                   dict(config=4, index=index, kind=kind, sigma=sigma, n_target=n_target,
                        seed=seed_base + index))


def batch(n_models=256, **kw):
    q = q_grid_2theta(kw.pop("nq", 4096))
    return [batch_model(i, q=q, **kw) for i in range(n_models)]


def serve_request(index, seed_base=5000, n_lo=500, n_hi=200000, rstep=1e-3):
    """C5 request `index`: 1-4 species, Q grid of 2048/4096/8192 points."""
    rng = np.random.default_rng(seed_base + index)
    n_target = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
    S = int(rng.integers(1, 5))
    nq = int(rng.choice([2048, 4096, 8192]))
    kind = "fcc" if rng.random() < 0.5 else "hcp"
    sigma = float(rng.uniform(0.01, 0.08))
    p = _close_packed(rng, n_target, kind, 2.885, 0.1, (1.0, 1.0, 1.0))
    p = p + rng.normal(0, sigma, p.shape)
    sp = rng.integers(0, S, len(p))
    params = [cromer_mann(rng) for _ in range(S)]
    return _finish(f"C5-{index:04d}", [p[sp == s] for s in range(S)], params, rstep,
                   q_grid_2theta(nq), dict(config=5, index=index, nq=nq, seed=seed_base + index))


def stacked_lattice(spacing=9.0, radius=100.0, stack=8, rstep=1e-3, seed=6001, nq=64):
    """Perfect simple cubic lattice (sigma = 0) cut to a sphere, `stack` coincident atoms on
    every site, one species: many pairs per bin (shell multiplicity x stack^2 per site pair),
    and 64-atom tiles wide enough for 16-bit pass-1 windows."""
    rng = np.random.default_rng(seed)
    n = int(math.ceil(radius / spacing)) + 1
    g = np.arange(-n, n + 1, dtype=np.float64) * spacing
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    p = np.repeat(p[(p ** 2).sum(1) <= radius * radius], stack, axis=0)
    return _finish(f"stack-{spacing:g}-{radius:g}-{stack}", [p], [cromer_mann(rng)], rstep, q_grid_2theta(nq),
                   dict(spacing=spacing, radius=radius, stack=stack, seed=seed))


def shell_around_stack(n_shell=2000, n_stack=60000, rho=20.00003, rstep=1e-4, seed=6002, nq=64):
    """Subgroup 0: n_shell atoms at distance rho (FP64, then rounded to FP32) from the origin in
    seeded random directions; subgroup 1: n_stack atoms all at the origin.  Every cross pair
    falls in the one bin of rho, so the (0, 1) partial puts all its counts into a single bin,
    and its A tiles (patches of the sphere) are wide enough for 16-bit windows."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_shell, 3))
    d *= rho / np.sqrt((d ** 2).sum(1))[:, None]
    return _finish(f"shell-{n_shell}-stack-{n_stack}", [d, np.zeros((n_stack, 3))],
                   [cromer_mann(rng), cromer_mann(rng)], rstep, q_grid_2theta(nq),
                   dict(n_shell=n_shell, n_stack=n_stack, rho=rho, seed=seed))


def small_random(n, n_groups=2, seed=0, spread=8.0, rstep=1e-3, nq=64):
    """Small random cluster for tests."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-spread, spread, (n, 3))
    sp = rng.integers(0, n_groups, n)
    return _finish(f"rand-{n}-{seed}", [p[sp == s] for s in range(n_groups)],
                   [cromer_mann(rng) for _ in range(n_groups)], rstep, q_grid_2theta(nq),
                   dict(seed=seed))

"""ctypes front end of the CPU oracle (oracle/debye_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; the product never imports it.

The functions restate cuDebye-v15.cu (file:line given on each) and are pinned
by the reference's own fixtures (Data/GammaAl2O3_10nm.debx*, .debI,
Data/UH3_7nm.debI) and by closed forms; see the header of debye_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# FP32 sum-of-squares form that reproduces the reference's .debx fixtures
# (cuDebye-v15.cu:104 as compiled by nvcc 8.0); fixed by the Gamma replay test.
REFERENCE_FORM = 0


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        i64, f32, f64, i32 = C.c_int64, C.c_float, C.c_double, C.c_int
        vp = C.c_void_p
        L.orc_q_from_2theta.restype = i32
        L.orc_q_from_2theta.argtypes = [f64, f64, f64, f64, vp, vp]
        L.orc_channels.restype = i64
        L.orc_channels.argtypes = [f64, f64]
        L.orc_rp.restype = f32
        L.orc_rp.argtypes = [f64]
        L.orc_pair_bin.restype = i64
        L.orc_pair_bin.argtypes = [f32] * 6 + [f32, i32]
        L.orc_histogram.restype = i64
        L.orc_histogram.argtypes = [vp, vp, vp, i64, i64, i64, i64, i32, f32, i32, vp, i64, i32]
        L.orc_prefactor.restype = None
        L.orc_prefactor.argtypes = [vp, i32, f64, f64, f64, f64, vp, vp, f64, vp, vp, f64, vp]
        L.orc_debye_sum.restype = None
        L.orc_debye_sum.argtypes = [vp, vp, i64, f64, vp, i32, vp, vp, vp, vp, i32, i32]
        L.orc_parse_deb.restype = i32
        L.orc_parse_deb.argtypes = [C.c_char_p, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_write_debx.restype = i32
        L.orc_write_debx.argtypes = [C.c_char_p, vp, vp, i64, f64, i32]
        L.orc_write_debi.restype = i32
        L.orc_write_debi.argtypes = [C.c_char_p, vp, vp, i32, i32]
        L.orc_max_threads.restype = i32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class _Header(C.Structure):
    _fields_ = [
        ("device", C.c_int), ("block", C.c_int), ("grid", C.c_int), ("hist", C.c_int),
        ("n_groups", C.c_int), ("n_atoms", C.c_int64),
        ("lam", C.c_double), ("tt_min", C.c_double), ("tt_max", C.c_double),
        ("tt_step", C.c_double), ("rmax", C.c_double), ("rstep", C.c_double),
    ]


@dataclass
class DebModel:
    """A parsed .deb model (cuDebye-v15.cu:624-833)."""
    device: int
    block: int
    grid: int
    lam: float
    tt_min: float
    tt_max: float
    tt_step: float
    rmax: float
    rstep: float
    hist: int
    group_n: np.ndarray           # int64[S]
    occ: np.ndarray               # f64[S]
    beq: np.ndarray               # f64[S]
    a: np.ndarray                 # f64[S,5]
    b: np.ndarray                 # f64[S,5]
    c: np.ndarray                 # f64[S]
    x: np.ndarray                 # f32[N]
    y: np.ndarray
    z: np.ndarray
    types: list = field(default_factory=list)

    @property
    def starts(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.group_n)[:-1]]).astype(np.int64)


def parse_deb(path: str) -> DebModel:
    L = lib()
    h = _Header()
    rc = L.orc_parse_deb(path.encode(), C.byref(h), None, None, None, None, None, None, None, None, None)
    if rc != 0:
        raise ValueError(f"oracle parse of {path} failed: {rc}")
    S, N = h.n_groups, h.n_atoms
    gn = np.zeros(S, np.int64)
    occ, beq, c = np.zeros(S), np.zeros(S), np.zeros(S)
    a, b = np.zeros((S, 5)), np.zeros((S, 5))
    x, y, z = np.zeros(N, np.float32), np.zeros(N, np.float32), np.zeros(N, np.float32)
    rc = L.orc_parse_deb(path.encode(), C.byref(h), _p(gn), _p(occ), _p(beq), _p(a), _p(b), _p(c),
                         _p(x), _p(y), _p(z))
    if rc != 0:
        raise ValueError(f"oracle parse of {path} failed: {rc}")
    types = []
    with open(path, "rb") as f:
        lines = f.read().split(b"\n")
    for g in range(S):
        types.append(lines[4 + g].split()[1].decode())
    return DebModel(h.device, h.block, h.grid, h.lam, h.tt_min, h.tt_max, h.tt_step, h.rmax,
                    h.rstep, h.hist, gn, occ, beq, a, b, c, x, y, z, types)


def q_from_2theta(lam, tt_min, tt_max, tt_step):
    """cuDebye-v15.cu:690-698 -> (2theta[nTT], K[nTT])"""
    L = lib()
    n = L.orc_q_from_2theta(lam, tt_min, tt_max, tt_step, None, None)
    tt, K = np.zeros(n), np.zeros(n)
    L.orc_q_from_2theta(lam, tt_min, tt_max, tt_step, _p(tt), _p(K))
    return tt, K


def channels(rmax, rstep) -> int:
    return int(lib().orc_channels(rmax, rstep))


def rp_of(rstep) -> float:
    return float(lib().orc_rp(rstep))


def pair_bin(pa, pb, rp, form=REFERENCE_FORM) -> int:
    return int(lib().orc_pair_bin(*[float(v) for v in pa], *[float(v) for v in pb], rp, form))


def histogram(x, y, z, row0, nrow, col0, ncol, same, rp, n_bins, form=REFERENCE_FORM,
              nthreads=0, out=None):
    """Pair-distance histogram of one subgroup pair (cuDebye-v15.cu:88-146).
    Returns (hist uint64[n_bins], overflow_count)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    z = np.ascontiguousarray(z, np.float32)
    h = np.zeros(n_bins, np.uint64) if out is None else out
    ov = lib().orc_histogram(_p(x), _p(y), _p(z), row0, nrow, col0, ncol, int(bool(same)),
                             float(rp), int(form), _p(h), n_bins, nthreads)
    return h, int(ov)


def pair_list(group_n):
    """Subgroup pairs in the reference's order (cuDebye-v15.cu:941-947): (row, col), row <= col."""
    S = len(group_n)
    return [(r, c) for r in range(S) for c in range(r, S)]


def all_histograms(m: DebModel, n_bins=None, form=REFERENCE_FORM, nthreads=0):
    """All S(S+1)/2 partial histograms, uint64[P, n_bins], in reference pair order."""
    if n_bins is None:
        n_bins = channels(m.rmax, m.rstep)
    rp = rp_of(m.rstep)
    st = m.starts
    pairs = pair_list(m.group_n)
    H = np.zeros((len(pairs), n_bins), np.uint64)
    ovs = []
    for p, (r, c) in enumerate(pairs):
        _, ov = histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c, rp,
                          n_bins, form, nthreads, out=H[p])
        ovs.append(ov)
    return H, ovs


def prefactor(K, occ_a, occ_b, beq_a, beq_b, aa, ba, ca, ab, bb, cb):
    """ffoobb[Q] (cuDebye-v15.cu:447-467), FP64."""
    K = np.ascontiguousarray(K, np.float64)
    aa, ba, ab, bb = (np.ascontiguousarray(v, np.float64) for v in (aa, ba, ab, bb))
    ff = np.zeros(len(K))
    lib().orc_prefactor(_p(K), len(K), occ_a, occ_b, beq_a, beq_b, _p(aa), _p(ba), ca, _p(ab),
                        _p(bb), cb, _p(ff))
    return ff


def debye_sum(bins, counts, rstep, K, ff, I=None, first_n=8, absmag=None, s_out=None,
              nthreads=0):
    """I[q] += ff[q] * S(q) for one compacted histogram (cuDebye-v15.cu:493-536), FP64
    with compensated summation."""
    bins = np.ascontiguousarray(bins, np.int64)
    cnt = np.ascontiguousarray(counts, np.float64)
    K = np.ascontiguousarray(K, np.float64)
    ff = np.ascontiguousarray(ff, np.float64)
    if I is None:
        I = np.zeros(len(K))
    lib().orc_debye_sum(_p(bins), _p(cnt), len(bins), rstep, _p(K), len(K), _p(ff), _p(I),
                        None if absmag is None else _p(absmag),
                        None if s_out is None else _p(s_out), first_n, nthreads)
    return I


def intensity_from_histograms(H, m_params, K, rstep, first_n=8, nthreads=0, with_cond=False):
    """I(Q) = sum over subgroup pairs of ffoobb_p(Q) * S_p(Q) (cuDebye-v15.cu:941-967).
    H: uint64[P, n_bins] in reference pair order; m_params: object with occ, beq, a, b, c,
    group_n."""
    I = np.zeros(len(K))
    mag = np.zeros(len(K))
    pairs = pair_list(m_params.group_n)
    for p, (r, c) in enumerate(pairs):
        nz = np.nonzero(H[p])[0]
        ff = prefactor(K, m_params.occ[r], m_params.occ[c], m_params.beq[r], m_params.beq[c],
                       m_params.a[r], m_params.b[r], m_params.c[r], m_params.a[c],
                       m_params.b[c], m_params.c[c])
        debye_sum(nz, H[p][nz].astype(np.float64), rstep, K, ff, I, first_n, mag,
                  nthreads=nthreads)
    return (I, mag) if with_cond else I


def write_debx(path, bins, counts, rstep, crlf=False):
    bins = np.ascontiguousarray(bins, np.int64)
    cnt = np.ascontiguousarray(counts, np.uint64)
    if lib().orc_write_debx(path.encode(), _p(bins), _p(cnt), len(bins), rstep, int(crlf)) != 0:
        raise OSError(path)


def write_debi(path, tt, I, crlf=False):
    tt = np.ascontiguousarray(tt, np.float64)
    I = np.ascontiguousarray(I, np.float64)
    if lib().orc_write_debi(path.encode(), _p(tt), _p(I), len(tt), int(crlf)) != 0:
        raise OSError(path)

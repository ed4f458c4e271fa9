"""cudebye_amd: MI355X (gfx950) Debye scattering engine with the method and file surface of
cuDebye v1.5 (Martin-Rudolph/cuDebye).

This module is a thin ctypes front end of libdebye.so (C ABI in include/debye.h).  Every
computation runs in the HIP kernels of that library; if the library is missing or has
not been built, importing the library raises - there is no CPU fallback.

    import cudebye_amd as cd
    with cd.Context(0) as ctx:
        m = ctx.model(x, y, z, group_n)
        H = ctx.histogram(m, rstep=1e-3, n_bins=cd.channels(rmax, 1e-3))
        I = ctx.intensity(m, params, rstep, n_bins, q, first_n=8)
    cd.run_deb("model.deb")          # the reference program, writes .debI / .debx files
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdebye.so")
CLI_PATH = os.path.join(_HERE, "bin", "debye")

DBX_OK = 0
ERRORS = {-1: "DBX_ERR_ARG", -2: "DBX_ERR_HIP", -3: "DBX_ERR_IO", -4: "DBX_ERR_PARSE",
          -5: "DBX_ERR_RANGE", -6: "DBX_ERR_NOMEM", -7: "DBX_ERR_BUDGET", -8: "DBX_ERR_STATE"}
DBX_ERR_ARG, DBX_ERR_HIP, DBX_ERR_IO, DBX_ERR_PARSE, DBX_ERR_RANGE, DBX_ERR_NOMEM = -1, -2, -3, -4, -5, -6
DBX_RUN_CRLF = 1
DBX_RUN_QUIET = 2
DBX_RUN_MFMA = 4
DBX_RUN_FROM_DEBX = 8
DBX_PASS2_MFMA = 0
DBX_PASS2_FP64 = 1

_lib = None


class DebyeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class BatchItem(C.Structure):
    """dbx_batch_item: one model of a batch."""
    _fields_ = [("model", C.c_void_p), ("params", C.c_void_p), ("n_bins", C.c_int64),
                ("first_n", C.c_int)]


# dbx_batch_item as a NumPy record, so that a batch's items are packed column by column
_BATCH_ITEM = np.dtype([("model", "<u8"), ("params", "<u8"), ("n_bins", "<i8"), ("first_n", "<i4"),
                        ("pad", "<i4")])


class ServerConfig(C.Structure):
    """dbx_server_config."""
    _fields_ = [("hbm_budget", C.c_size_t), ("max_batch", C.c_int), ("max_step_pairs", C.c_double),
                ("prep_threads", C.c_int), ("pass2_mode", C.c_int), ("batch_window_ms", C.c_double),
                ("rel_err_target", C.c_double)]


class RequestDesc(C.Structure):
    """dbx_request."""
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("group_n", C.c_void_p),
                ("n_groups", C.c_int), ("params", C.c_void_p), ("rstep", C.c_double),
                ("n_bins", C.c_int64), ("first_n", C.c_int), ("q", C.c_void_p), ("nq", C.c_int),
                ("flags", C.c_int)]


class ServerStats(C.Structure):
    """dbx_server_stats."""
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("steps", C.c_uint64), ("pairs", C.c_double), ("resident_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("max_step_requests", C.c_double),
                ("max_step_qgrids", C.c_double), ("step_enqueue_ms", C.c_double),
                ("step_wait_ms", C.c_double)]


DBX_REQ_KEEP = 1
DBX_REQ_QUEUED, DBX_REQ_RUNNING, DBX_REQ_DONE, DBX_REQ_FAILED = 0, 1, 2, 3
DBX_ERR_BUDGET = -7
DBX_ERR_STATE = -8


class GroupParams(C.Structure):
    """dbx_group_params: occ, Beq, a1..a5, b1..b5, c of one subgroup."""
    _fields_ = [("occ", C.c_double), ("beq", C.c_double), ("a", C.c_double * 5),
                ("b", C.c_double * 5), ("c", C.c_double)]

    @classmethod
    def make(cls, occ, beq, a, b, c):
        g = cls()
        g.occ, g.beq, g.c = float(occ), float(beq), float(c)
        for i in range(5):
            g.a[i] = float(a[i]) if i < len(a) else 0.0
            g.b[i] = float(b[i]) if i < len(b) else 0.0
        return g


def lib():
    """The loaded libdebye.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cudebye_amd._build` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, f64, f32, sz = (C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double,
                                           C.c_float, C.c_size_t)
        P = C.POINTER
        sig = {
            "dbx_open": (i32, [i32, sz, P(vp)]),
            "dbx_close": (i32, [vp]),
            "dbx_last_error": (C.c_char_p, [vp]),
            "dbx_stream": (vp, [vp]),
            "dbx_q_from_2theta": (i32, [f64, f64, f64, f64, vp, vp, i32]),
            "dbx_channels": (i64, [f64, f64]),
            "dbx_model_create": (i32, [vp, vp, vp, vp, vp, i32, P(vp)]),
            "dbx_model_destroy": (i32, [vp, vp]),
            "dbx_model_info": (i32, [vp, P(i64), P(i32), P(i32)]),
            "dbx_histogram": (i32, [vp, vp, f64, i64, vp, i32, P(u64)]),
            "dbx_prefactor": (i32, [vp, vp, vp, vp, i32, vp]),
            "dbx_debye": (i32, [vp, vp, i32, i32, i64, f64, i32, vp, i32, vp, vp, vp, vp]),
            "dbx_debye_noalloc": (i32, [vp, vp, i32, i64, f64, i32, vp, i32, vp, vp, vp, vp]),
            "dbx_intensity": (i32, [vp, vp, vp, f64, i64, i32, vp, i32, vp, vp, vp]),
            "dbx_set_rel_err_target": (i32, [vp, f64]),
            "dbx_view_deb": (i32, [C.c_char_p, vp, C.c_char_p, C.c_char_p, C.c_char_p, P(i64)]),
            "dbx_debug_pass2_refined": (i32, [vp, P(i32)]),
            "dbx_debug_pass1_evaluations": (i32, [vp, P(u64)]),
            "dbx_run_deb": (i32, [C.c_char_p, i32]),
            "dbx_run_folder": (i32, [C.c_char_p, i32]),
            "dbx_read_debx": (i32, [C.c_char_p, f64, i64, vp, P(i64)]),
            "dbx_selftest_sqrt": (i32, [C.c_uint32, C.c_uint32, P(u64)]),
            "dbx_selftest_bins": (i32, [vp, vp, i32, f32, vp]),
            "dbx_last_pass1_stats": (i32, [vp, P(i64), P(u64), P(u64)]),
            "dbx_last_kernel_times": (i32, [vp, P(f64), P(f64)]),
            "dbx_set_pass2_mode": (i32, [vp, i32]),
            "dbx_selftest_sin": (i32, [vp, i32, vp]),
            "dbx_batch_intensity": (i32, [vp, vp, i32, f64, vp, i32, vp, vp, vp, vp]),
            "dbx_server_open": (i32, [i32, vp, P(vp)]),
            "dbx_server_close": (i32, [vp]),
            "dbx_server_last_error": (C.c_char_p, [vp]),
            "dbx_submit": (i32, [vp, vp, P(u64)]),
            "dbx_poll": (i32, [vp, u64, P(i32)]),
            "dbx_wait": (i32, [vp, u64, vp, P(f64), vp]),
            "dbx_reweight": (i32, [vp, u64, vp, vp, vp]),
            "dbx_request_histogram": (i32, [vp, u64, vp]),
            "dbx_release": (i32, [vp, u64]),
            "dbx_server_stats_get": (i32, [vp, vp]),
            "dbx_debug_set_profiling": (i32, [vp, i32]),
            "dbx_selftest_round": (i32, [C.c_uint32, C.c_uint32, P(u64)]),
            "dbx_selftest_lds_atomic_rate": (i32, [P(f64)]),
            "dbx_debug_pass2_profile": (i32, [vp, P(f64), P(f64), P(f64), P(i32), P(i32), P(f64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


EXPORTS = ("dbx_open", "dbx_close", "dbx_last_error", "dbx_stream", "dbx_q_from_2theta",
           "dbx_channels", "dbx_model_create", "dbx_model_destroy", "dbx_model_info",
           "dbx_histogram", "dbx_prefactor", "dbx_debye", "dbx_intensity", "dbx_run_deb",
           "dbx_selftest_sqrt", "dbx_selftest_bins", "dbx_last_pass1_stats",
           "dbx_last_kernel_times", "dbx_set_pass2_mode", "dbx_selftest_sin",
           "dbx_batch_intensity", "dbx_server_open", "dbx_server_close", "dbx_server_last_error",
           "dbx_submit", "dbx_poll", "dbx_wait", "dbx_reweight", "dbx_request_histogram",
           "dbx_release", "dbx_server_stats_get", "dbx_read_debx", "dbx_run_folder",
           "dbx_deb_header_defaults", "dbx_xyz_to_deb", "dbx_average_debi", "dbx_debye_noalloc",
           "dbx_set_rel_err_target", "dbx_view_deb")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _results(first, *optional):
    """first alone, or a tuple of first and the (value, wanted) items that are wanted"""
    out = (first,) + tuple(v for v, wanted in optional if wanted)
    return out[0] if len(out) == 1 else out


def q_from_2theta(lam, tt_min, tt_max, tt_step):
    """(2theta, K) grids of cuDebye-v15.cu:690-698."""
    L = lib()
    n = L.dbx_q_from_2theta(lam, tt_min, tt_max, tt_step, None, None, 0)
    if n < 1:
        raise DebyeError(n, "bad 2theta grid")
    tt, q = np.zeros(n), np.zeros(n)
    L.dbx_q_from_2theta(lam, tt_min, tt_max, tt_step, _p(tt), _p(q), n)
    return tt, q


def channels(rmax, rstep) -> int:
    n = int(lib().dbx_channels(rmax, rstep))
    if n < 0:
        raise DebyeError(n, "bad distance grid")
    return n


def run_deb(path, crlf=False, quiet=True, mfma=False, from_debx=False) -> int:
    """Runs a .deb file like `cuDebye.exe path` (writes path+'I' and/or path+'x<r>x<c>').
    from_debx: pass 2 only, from the path+'x<r>x<c>' histograms a previous run wrote."""
    flags = ((DBX_RUN_CRLF if crlf else 0) | (DBX_RUN_QUIET if quiet else 0) | (DBX_RUN_MFMA if mfma else 0)
             | (DBX_RUN_FROM_DEBX if from_debx else 0))
    return int(lib().dbx_run_deb(os.fspath(path).encode(), flags))


def run_folder(folder, crlf=False, quiet=True, mfma=False) -> int:
    """Every .deb in `folder` without a .debI, through one server (Matlab/multisim.m)."""
    flags = (DBX_RUN_CRLF if crlf else 0) | (DBX_RUN_QUIET if quiet else 0) | (DBX_RUN_MFMA if mfma else 0)
    return int(lib().dbx_run_folder(os.fspath(folder).encode(), flags))


def read_debx(path, rstep, n_bins):
    """uint64[n_bins] histogram of one .debx file."""
    h = np.zeros(int(n_bins), np.uint64)
    nnz = C.c_int64(0)
    rc = lib().dbx_read_debx(os.fspath(path).encode(), rstep, int(n_bins), _p(h), C.byref(nnz))
    if rc != DBX_OK:
        raise DebyeError(rc, f"cannot read {path}")
    return h


class DebHeader(C.Structure):
    """dbx_deb_header: the .deb header values the converter writes."""
    _fields_ = [("device", C.c_int), ("block", C.c_int), ("grid", C.c_int), ("wavelength", C.c_double),
                ("tt_min", C.c_double), ("tt_max", C.c_double), ("tt_step", C.c_double), ("rstep", C.c_double),
                ("hist", C.c_int)]


DBX_AVG_NONE, DBX_AVG_INTEGRAL, DBX_AVG_MAX = 0, 1, 2


def xyz_to_deb(xyz, deb=None, table=None, **header):
    """Converts an .xyz file to .deb (Matlab/xyz2deb.m); header fields of DebHeader override the
    converter's defaults.  Returns the .deb path."""
    h = DebHeader()
    lib().dbx_deb_header_defaults(C.byref(h))
    for k, v in header.items():
        setattr(h, k, v)
    xyz = os.fspath(xyz)
    enc = lambda v: None if v is None else os.fspath(v).encode()
    rc = lib().dbx_xyz_to_deb(xyz.encode(), enc(deb), enc(table), C.byref(h))
    if rc != DBX_OK:
        raise DebyeError(rc, f"cannot convert {xyz}")
    return os.fspath(deb) if deb is not None else xyz[:-3] + "deb"


def average_debi(paths, norm="none", step=0.02):
    """(2theta, I) average of several .debI patterns (Matlab/plotdebi.m); norm: "none" (mean),
    "int" (divided by its trapezoidal integral) or "max" (by its maximum)."""
    code = {"none": DBX_AVG_NONE, "int": DBX_AVG_INTEGRAL, "integral": DBX_AVG_INTEGRAL,
            "max": DBX_AVG_MAX, "maximum": DBX_AVG_MAX}[norm]
    arr = (C.c_char_p * len(paths))(*[os.fspath(p).encode() for p in paths])
    n = C.c_int(0)
    fn = lib().dbx_average_debi
    rc = fn(arr, len(paths), code, C.c_double(step), None, None, 0, C.byref(n))
    if rc != DBX_OK:
        raise DebyeError(rc, "average_debi")
    tt, I = np.zeros(n.value), np.zeros(n.value)
    rc = fn(arr, len(paths), code, C.c_double(step), _p(tt), _p(I), n.value, C.byref(n))
    if rc != DBX_OK:
        raise DebyeError(rc, "average_debi")
    return tt, I


def view_deb(deb, limits=None, svg=None, xyz=None, radii=None):
    """Structure view of a .deb file (Matlab/viewdeb.m): the atoms inside `limits` (lo, hi) for
    every axis or ((xlo, xhi), (ylo, yhi), (zlo, zhi)); default -10..10) drawn into `svg` and/or
    listed in `xyz`.  Returns the number of atoms kept."""
    lim = None
    if limits is not None:
        a = np.asarray(limits, np.float64)
        a = np.tile(a.reshape(1, 2), (3, 1)) if a.size == 2 else a.reshape(3, 2)
        lim = np.ascontiguousarray(a.ravel())
    enc = lambda v: None if v is None else os.fspath(v).encode()
    n = C.c_int64(0)
    rc = lib().dbx_view_deb(os.fspath(deb).encode(), None if lim is None else _p(lim), enc(svg), enc(xyz),
                            enc(radii), C.byref(n))
    if rc != DBX_OK:
        raise DebyeError(rc, f"cannot view {deb}")
    return int(n.value)


def lds_atomic_rate():
    """Measured LDS atomic throughput of device 0 (ds_add_u32 per second, conflict-free)."""
    v = C.c_double(0)
    rc = lib().dbx_selftest_lds_atomic_rate(C.byref(v))
    if rc != DBX_OK:
        raise DebyeError(rc, "lds_atomic_rate")
    return v.value


def selftest_round(lo=0, hi=0x4F000000):
    """Count of floats in [lo, hi) where the hardware round-half-up convert differs from roundf."""
    n = C.c_uint64(0)
    rc = lib().dbx_selftest_round(lo, hi, C.byref(n))
    if rc != DBX_OK:
        raise DebyeError(rc, "selftest_round")
    return int(n.value)


def selftest_sqrt(lo, hi, with_raw=False):
    """(not correctly rounded, differs from sqrtf) counts over float bit patterns [lo, hi);
    with_raw adds the counts where the bare hardware square root lies below / above."""
    n = (C.c_uint64 * 4)()
    rc = lib().dbx_selftest_sqrt(lo, hi, n)
    if rc != DBX_OK:
        raise DebyeError(rc, "selftest_sqrt")
    if with_raw:
        return int(n[0]), int(n[1]), int(n[2]), int(n[3])
    return int(n[0]), int(n[1])


def selftest_bins(a, b, rp):
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(b, np.float32).reshape(-1, 3)
    out = np.zeros(len(a), np.int32)
    rc = lib().dbx_selftest_bins(_p(a), _p(b), len(a), float(np.float32(rp)), _p(out))
    if rc != DBX_OK:
        raise DebyeError(rc, "selftest_bins")
    return out


def selftest_sin(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    rc = lib().dbx_selftest_sin(_p(x), len(x), _p(out))
    if rc != DBX_OK:
        raise DebyeError(rc, "selftest_sin")
    return out


class Model:
    def __init__(self, ctx, handle, n_atoms, n_groups):
        self.ctx, self.handle, self.n_atoms, self.n_groups = ctx, handle, n_atoms, n_groups

    @property
    def n_pairs(self):
        return self.n_groups * (self.n_groups + 1) // 2

    def close(self):
        if self.handle:
            lib().dbx_model_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.close()
        except Exception:
            pass


class Context:
    """A device context (dbx_open).  Use as a context manager or call close()."""

    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        rc = lib().dbx_open(device, hbm_budget, C.byref(h))
        if rc != DBX_OK:
            raise DebyeError(rc, f"cannot open device {device}")
        self.handle = h

    def _check(self, rc, what, allow=()):
        if rc != DBX_OK and rc not in allow:
            raise DebyeError(rc, f"{what}: {lib().dbx_last_error(self.handle).decode()}")
        return rc

    def close(self):
        if self.handle:
            lib().dbx_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def model(self, x, y, z, group_n) -> Model:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        z = np.ascontiguousarray(z, np.float32)
        gn = np.ascontiguousarray(group_n, np.int64)
        assert len(x) == len(y) == len(z) == int(gn.sum())
        h = C.c_void_p()
        self._check(lib().dbx_model_create(self.handle, _p(x), _p(y), _p(z), _p(gn), len(gn),
                                           C.byref(h)), "model_create")
        return Model(self, h, int(gn.sum()), len(gn))

    def histogram(self, m: Model, rstep, n_bins, return_oor=False):
        """uint64[P, n_bins] partial histograms (reference pair order)."""
        out = np.zeros((m.n_pairs, n_bins), np.uint64)
        oor = C.c_uint64(0)
        self._check(lib().dbx_histogram(self.handle, m.handle, rstep, n_bins, _p(out), 0,
                                        C.byref(oor)), "histogram",
                    allow=(DBX_ERR_RANGE,) if return_oor else ())
        return (out, int(oor.value)) if return_oor else out

    def pass1_stats(self, with_retried=False):
        """(work items launched, pairs outside their LDS window[, exact-window items re-run by the
        checked kernel]) of the last pass 1."""
        nj, miss, rt = C.c_int64(0), C.c_uint64(0), C.c_uint64(0)
        lib().dbx_last_pass1_stats(self.handle, C.byref(nj), C.byref(miss), C.byref(rt))
        if with_retried:
            return int(nj.value), int(miss.value), int(rt.value)
        return int(nj.value), int(miss.value)

    def pass1_evaluations(self):
        """Pair evaluations of the last pass 1 (staged B atoms times A-tile atoms, in or out of
        their job's window): about the pairs counted where one window holds each shell."""
        v = C.c_uint64(0)
        self._check(lib().dbx_debug_pass1_evaluations(self.handle, C.byref(v)), "pass1_evaluations")
        return int(v.value)

    def set_pass2_mode(self, mode):
        """DBX_PASS2_MFMA (8-bit integer digits on the matrix cores, default) or DBX_PASS2_FP64."""
        self._check(lib().dbx_set_pass2_mode(self.handle, mode), "set_pass2_mode")

    def set_rel_err_target(self, target):
        """Models whose I(q) error bound exceeds `target` (default 1e-6) are recomputed with FP64
        arithmetic; target <= 0: never."""
        self._check(lib().dbx_set_rel_err_target(self.handle, float(target)), "set_rel_err_target")

    def pass2_refined(self):
        """Models the last pass 2 recomputed with FP64 arithmetic to meet the error target."""
        n = C.c_int(0)
        self._check(lib().dbx_debug_pass2_refined(self.handle, C.byref(n)), "pass2_refined")
        return n.value

    def kernel_times(self):
        """(pass-1 histogram kernel ms, pass-2 structure ms) of the last calls (HIP events)."""
        a, b = C.c_double(0), C.c_double(0)
        self._check(lib().dbx_last_kernel_times(self.handle, C.byref(a), C.byref(b)), "kernel_times")
        return a.value, b.value

    def set_profiling(self, on=True):
        """Events around the matrix-core contraction kernel and a count of its instructions."""
        self._check(lib().dbx_debug_set_profiling(self.handle, int(on)), "set_profiling")

    def pass2_profile(self):
        """(kernel ms, issued int8 matrix ops, useful flops, compacted bins, largest count-digit
        number, mean count digits over the row tiles) of the last profiled matrix-core pass 2."""
        ms, fl, us, md = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        kc, dg = C.c_int(), C.c_int()
        self._check(lib().dbx_debug_pass2_profile(self.handle, C.byref(ms), C.byref(fl), C.byref(us), C.byref(kc),
                                                  C.byref(dg), C.byref(md)), "pass2_profile")
        return ms.value, fl.value, us.value, kc.value, dg.value, md.value

    def prefactor(self, ga: GroupParams, gb: GroupParams, q):
        q = np.ascontiguousarray(q, np.float64)
        ff = np.zeros(len(q))
        self._check(lib().dbx_prefactor(self.handle, C.byref(ga), C.byref(gb), _p(q), len(q),
                                        _p(ff)), "prefactor")
        return ff

    def debye(self, hist, rstep, q, ffoobb, first_n=8, with_S=False, with_err=False):
        """I(q) [, S_p(q)] [, relative error bound of each I(q)]."""
        hist = np.ascontiguousarray(hist, np.uint64)
        P, nb = hist.shape
        q = np.ascontiguousarray(q, np.float64)
        ff = np.ascontiguousarray(ffoobb, np.float64).reshape(P, len(q))
        I = np.zeros(len(q))
        S = np.zeros((P, len(q))) if with_S else None
        rel = np.zeros(len(q)) if with_err else None
        self._check(lib().dbx_debye(self.handle, _p(hist), 0, P, nb, rstep, first_n, _p(q), len(q),
                                    _p(ff), _p(I), None if S is None else _p(S),
                                    None if rel is None else _p(rel)), "debye")
        return _results(I, (S, with_S), (rel, with_err))

    def intensity(self, m: Model, params, rstep, n_bins, q, first_n=8, with_hist=False, with_err=False):
        """I(q) [, histograms] [, relative error bound of each I(q)]."""
        arr = (GroupParams * m.n_groups)(*params)
        q = np.ascontiguousarray(q, np.float64)
        I = np.zeros(len(q))
        H = np.zeros((m.n_pairs, n_bins), np.uint64) if with_hist else None
        rel = np.zeros(len(q)) if with_err else None
        self._check(lib().dbx_intensity(self.handle, m.handle, arr, rstep, n_bins, first_n, _p(q),
                                        len(q), _p(I), None if H is None else _p(H),
                                        None if rel is None else _p(rel)), "intensity")
        return _results(I, (H, with_hist), (rel, with_err))

    def batch_intensity(self, items, rstep, q, with_hist=False, return_oor=False, with_err=False):
        """Many models in one pass-1 launch and one pass-2 contraction (dbx_batch_intensity).
        items: sequence of (Model, params, n_bins, first_n).  Returns I[n_items, nq]
        (+ list of histograms, + out-of-range counts per item)."""
        n = len(items)
        q = np.ascontiguousarray(q, np.float64)
        assert _BATCH_ITEM.itemsize == C.sizeof(BatchItem)
        for m, params, _, _ in items:
            if len(params) != m.n_groups:
                raise ValueError("one GroupParams per subgroup of each model")
        # every item's parameters in one buffer, the items as one record array
        pbuf = np.frombuffer(b"".join(bytes(g) for _, params, _, _ in items for g in params), np.uint8).copy()
        ng = np.fromiter((len(params) for _, params, _, _ in items), np.int64, n)
        arr = np.zeros(n, _BATCH_ITEM)
        arr["model"] = np.fromiter((m.handle.value for m, _, _, _ in items), np.uint64, n)
        arr["params"] = pbuf.ctypes.data + (np.cumsum(ng) - ng) * C.sizeof(GroupParams)
        arr["n_bins"] = np.fromiter((int(nb) for _, _, nb, _ in items), np.int64, n)
        arr["first_n"] = np.fromiter((int(f) for _, _, _, f in items), np.int32, n)
        hists = [np.zeros((m.n_pairs, int(nb)), np.uint64) for m, _, nb, _ in items] if with_hist else []
        I = np.zeros((n, len(q)))
        oor = np.zeros(n, np.uint64)
        hp = (C.c_void_p * n)(*[h.ctypes.data for h in hists]) if with_hist else None
        rel = np.zeros((n, len(q))) if with_err else None
        self._check(lib().dbx_batch_intensity(self.handle, _p(arr), n, rstep, _p(q), len(q), _p(I), _p(oor),
                                              hp, None if rel is None else _p(rel)), "batch_intensity",
                    allow=(DBX_ERR_RANGE,) if return_oor else ())
        return _results(I, (hists, with_hist), ([int(v) for v in oor], return_oor), (rel, with_err))


class Server:
    """Continuous-batching server (dbx_server_open): submit requests from any thread, wait for
    their intensities, re-weight resident ones with new form factors."""

    def __init__(self, device=0, hbm_budget=0, max_batch=0, max_step_pairs=0.0, prep_threads=0,
                 pass2_mode=DBX_PASS2_MFMA, batch_window_ms=0.0, rel_err_target=0.0):
        cfg = ServerConfig(hbm_budget, max_batch, max_step_pairs, prep_threads, pass2_mode, batch_window_ms,
                           rel_err_target)
        h = C.c_void_p()
        rc = lib().dbx_server_open(device, C.byref(cfg), C.byref(h))
        if rc != DBX_OK:
            raise DebyeError(rc, "cannot open server")
        self.handle = h
        self._nq = {}

    def _err(self, rc, what):
        raise DebyeError(rc, f"{what}: {lib().dbx_server_last_error(self.handle).decode()}")

    def close(self):
        if self.handle:
            lib().dbx_server_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, x, y, z, group_n, params, rstep, n_bins, q, first_n=16, keep=False):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        z = np.ascontiguousarray(z, np.float32)
        gn = np.ascontiguousarray(group_n, np.int64)
        q = np.ascontiguousarray(q, np.float64)
        pa = (GroupParams * len(gn))(*params)
        req = RequestDesc(x.ctypes.data, y.ctypes.data, z.ctypes.data, gn.ctypes.data, len(gn),
                          C.cast(pa, C.c_void_p).value, rstep, int(n_bins), first_n, q.ctypes.data, len(q),
                          DBX_REQ_KEEP if keep else 0)
        t = C.c_uint64()
        rc = lib().dbx_submit(self.handle, C.byref(req), C.byref(t))
        if rc != DBX_OK:
            self._err(rc, "submit")
        self._nq[t.value] = (len(q), len(gn), int(n_bins))
        return t.value

    def poll(self, ticket):
        st = C.c_int()
        rc = lib().dbx_poll(self.handle, ticket, C.byref(st))
        if rc != DBX_OK:
            self._err(rc, "poll")
        return st.value

    def _info(self, ticket):
        """(nq, n_groups, n_bins) of a ticket submitted here and not yet released"""
        if ticket not in self._nq:
            raise DebyeError(DBX_ERR_ARG, f"unknown ticket {ticket}")
        return self._nq[ticket]

    def wait(self, ticket, allow=(), with_err=False):
        """(I, latency_ms, status[, relative error bound of each I(q)]).  DBX_ERR_BUDGET: the
        request cannot be admitted until finished DBX_REQ_KEEP requests are released (it stays
        queued)."""
        known = ticket in self._nq
        I = np.zeros(self._nq[ticket][0]) if known else None
        rel = np.zeros(self._nq[ticket][0]) if known and with_err else None
        lat = C.c_double()
        rc = lib().dbx_wait(self.handle, ticket, None if I is None else _p(I), C.byref(lat),
                            None if rel is None else _p(rel))
        if rc != DBX_OK and rc not in allow:
            self._err(rc, f"request {ticket}")
        return (I, lat.value, rc, rel) if with_err else (I, lat.value, rc)

    def reweight(self, ticket, params, with_err=False):
        """I(q) for new form factors [, its relative error bound]."""
        nq, S, _ = self._info(ticket)
        pa = (GroupParams * S)(*params)
        I = np.zeros(nq)
        rel = np.zeros(nq) if with_err else None
        rc = lib().dbx_reweight(self.handle, ticket, pa, _p(I), None if rel is None else _p(rel))
        if rc != DBX_OK:
            self._err(rc, "reweight")
        return (I, rel) if with_err else I

    def histogram(self, ticket):
        nq, S, nb = self._info(ticket)
        H = np.zeros((S * (S + 1) // 2, nb), np.uint64)
        rc = lib().dbx_request_histogram(self.handle, ticket, _p(H))
        if rc != DBX_OK:
            self._err(rc, "histogram")
        return H

    def release(self, ticket):
        rc = lib().dbx_release(self.handle, ticket)
        if rc != DBX_OK:
            self._err(rc, "release")
        self._nq.pop(ticket, None)

    def stats(self):
        st = ServerStats()
        lib().dbx_server_stats_get(self.handle, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

"""Named workloads on the device, for profiling runs (rocprofv3) and quick checks.

    python -m cudebye_amd.workloads NAME [--reps N] [--fp64]

    c1        single-species FCC sphere, 4 nm (~2k atoms), 4096 Q
    c2        two-species spinel-like sphere, 10 nm (~102k atoms), 4096 Q
    c2-fine   the c2 sphere at the reference's own bin width, 1e-4 A (1.06M bins), 4096 Q
    c3        three-species rocksalt-like sphere, 20 nm (~453k atoms), 8192 Q (0.3-20 A^-1)
    c4-batch  256 two-species models (1k-60k atoms) in one dbx_batch_intensity call
    c4-plain  the same 256 models, one dbx_intensity call each
    c5-serve  150 mixed requests (500-200k atoms, 1-4 species) through the server, all at once
    copy      a 1 GiB device-to-device copy (the HIP runtime's copy kernel): the reference for
              scaling the HBM traffic counters FETCH_SIZE and WRITE_SIZE to bytes

Each repetition prints one line with the wall time and the library's device times.
"""
import argparse
import ctypes as C
import time

import cudebye_amd as cd
from cudebye_amd import synth


def _gp(p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


def _single(m, reps, mode):
    with cd.Context(0) as ctx:
        ctx.set_pass2_mode(mode)
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        params = [_gp(p) for p in m.params]
        for r in range(reps):
            t = time.perf_counter()
            ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
            p1, p2 = ctx.kernel_times()
            print(f"{m.name} rep {r}: {(time.perf_counter() - t) * 1e3:.2f} ms  pass1 {p1:.3f} ms  pass2 {p2:.3f} ms",
                  flush=True)
        mod.close()


def _c4(reps, mode, batched):
    models = synth.batch(256)
    with cd.Context(0) as ctx:
        ctx.set_pass2_mode(mode)
        items = [(ctx.model(m.x, m.y, m.z, m.group_n), [_gp(p) for p in m.params], m.n_bins, m.first_n)
                 for m in models]
        for r in range(reps):
            t = time.perf_counter()
            if batched:
                ctx.batch_intensity(items, models[0].rstep, models[0].q)
            else:
                for m, it in zip(models, items):
                    ctx.intensity(it[0], it[1], m.rstep, it[2], m.q, first_n=it[3])
            p1, p2 = ctx.kernel_times()
            print(f"c4-{'batch' if batched else 'plain'} rep {r}: {(time.perf_counter() - t) * 1e3:.2f} ms  "
                  f"last pass1 {p1:.3f} ms  last pass2 {p2:.3f} ms", flush=True)
        for it in items:
            it[0].close()


def _c5(reps, mode):
    models = [synth.serve_request(i) for i in range(150)]
    with cd.Server(0, pass2_mode=mode) as srv:
        for r in range(reps):
            t = time.perf_counter()
            ts = [srv.submit(m.x, m.y, m.z, m.group_n, [_gp(p) for p in m.params], m.rstep, m.n_bins, m.q,
                             first_n=m.first_n) for m in models]
            for tk in ts:
                srv.wait(tk)
                srv.release(tk)
            print(f"c5-serve rep {r}: {(time.perf_counter() - t) * 1e3:.2f} ms  {srv.stats()}", flush=True)


def _copy(reps):
    cd.lib()  # loads the HIP runtime the library links
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    n = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0 and hip.hipDeviceSynchronize() == 0
    for r in range(reps):
        t = time.perf_counter()
        assert hip.hipMemcpy(b, a, n, 3) == 0  # device to device
        assert hip.hipDeviceSynchronize() == 0
        dt = time.perf_counter() - t
        print(f"copy rep {r}: {n} bytes in {dt * 1e3:.3f} ms ({2 * n / dt / 1e9:.0f} GB/s read + write)", flush=True)
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=["c1", "c2", "c2-fine", "c3", "c4-batch", "c4-plain", "c5-serve", "copy"])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--fp64", action="store_true", help="FP64 pass 2 instead of the matrix cores")
    a = ap.parse_args()
    mode = cd.DBX_PASS2_FP64 if a.fp64 else cd.DBX_PASS2_MFMA
    if a.name == "c1":
        _single(synth.fcc_sphere(), a.reps, mode)
    elif a.name == "c2":
        _single(synth.spinel_sphere(), a.reps, mode)
    elif a.name == "c2-fine":
        _single(synth.spinel_sphere(rstep=1e-4), a.reps, mode)
    elif a.name == "c3":
        _single(synth.rocksalt_sphere(), a.reps, mode)
    elif a.name == "c4-batch":
        _c4(a.reps, mode, True)
    elif a.name == "c4-plain":
        _c4(a.reps, mode, False)
    elif a.name == "copy":
        _copy(a.reps)
    else:
        _c5(a.reps, mode)


if __name__ == "__main__":
    main()

"""Benchmark of the Debye hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-streams] [--no-workloads]

Primary line (`value`): one step = one structure model through the whole hot path on the
device: pass 1 (all subgroup-pair distance histograms, uint64), compaction, prefactors and
pass 2 (I(Q)), with the model's atoms already resident in HBM.  Workload: the two-species
spinel-like sphere (diameter 10 nm, bins 0.001 A, 4096 Q points), checked in the same run
against the CPU oracle: bit-exact histograms and the max relative I(Q) error against its FP64
sums.  `value` is G pair-distances/s, a pair-distance counted as the reference counts it
(N(N+1)/2 per model, self pairs included).

`roofline` is for the pass-1 histogram kernel (9 FP32 flops per pair against the FP32 vector
peak), timed with HIP events on the library's stream, beside its LDS-atomic rate (one per pair)
against the device's measured LDS-atomic peak.  `pass2_roofline` is for the matrix-core
contraction: algorithmic flops (2 nQ per nonzero bin of every row, the contraction counted once
whatever the digit count) against the dense int8 matrix peak (the kernel computes on 8-bit
digits), with the issued matrix operations (MFMA utilisation) beside it.  `cpu_baseline` is the OpenMP oracle (a port of the reference path) on
the host cores of the same machine.

`batched`: the 256-model stream (FCC/HCP spheres and ellipsoids with stacking faults, 1k-60k
atoms, 2 species) through the continuous-batching server, all requests submitted at once,
against the plain path (one model per call: model creation + dbx_intensity, sequential).
`serving`: the mixed stream (500-200k atoms, 1-4 species, 2048/4096/8192 Q) with Poisson
arrivals at a fixed rate (80% of the server's capacity measured on the same requests), server
against the plain path fed the same arrival times: models/s and p50/p99 latency (submission to
result).
`workloads`: the other configurations, each timed on its own: the full rocksalt-like sphere
(6 partials, 0.0005 A bins, 8192 Q) with its matrix-core I(Q) against FP64 sums, the primary
sphere at the reference's 1e-4 A bins, the primary model cold (tiling, upload and planning
inside the timing), and the reference's own Gamma example through the `debye` program against
the reference's logged times.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak
I8_PEAK_TOPS = 5000.0      # MI355X dense int8 matrix peak
FLOPS_PER_PAIR = 9         # 3 sub, 3 mul, 2 add, 1 scale-mul
VALU_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9  # unpacked VALU lane operations/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
SERVE_LOAD = 0.8           # Poisson arrival rate of the mixed stream, as a fraction of the
                           # server's measured capacity on the same requests (all at once)
SERVE_N = 600              # requests in the mixed stream


def _gp(cd, p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


def primary(args, cd, synth, np):
    m = synth.spinel_sphere()
    params = [_gp(cd, p) for p in m.params]
    pairs = m.pair_count
    with cd.Context(0) as ctx:
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        for _ in range(args.warmup):
            ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
        p1_ms, p2_ms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
            a, b = ctx.kernel_times()
            p1_ms.append(a)
            p2_ms.append(b)
        elapsed = time.perf_counter() - t0  # intensity() synchronises the stream on return
        I, H, bound = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_hist=True,
                                    with_err=True)
        refined = ctx.pass2_refined()
        n_jobs, misses = ctx.pass1_stats()
        evals = ctx.pass1_evaluations()
        ctx.set_profiling(True)
        ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
        mf_ms, mf_flops, mf_useful, kc, digits, mean_digits = ctx.pass2_profile()
        ctx.set_profiling(False)
        mod.close()
    lds_peak = cd.lds_atomic_rate()

    ms_per_step = elapsed / args.steps * 1e3
    value = pairs * args.steps / elapsed / 1e9
    k_ms = float(np.mean(p1_ms))
    achieved = FLOPS_PER_PAIR * pairs / (k_ms * 1e-3) / 1e12
    mf_ach = mf_useful / (mf_ms * 1e-3) / 1e12
    mf_issued = mf_flops / (mf_ms * 1e-3) / 1e12
    p2_mean = float(np.mean(p2_ms))
    out = {
        "metric": "G pair-distances/s (end to end: pass 1 histograms + pass 2 I(Q)); "
                  "batched models/s + p50/p99 latency",
        "value": round(value, 3),
        "unit": "G pair-distances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spinel-like sphere)",
        "config": {
            "workload": "spinel-like sphere D=10 nm (O on FCC sublattice of a=7.91 A, cations on "
                        "a seeded 2/3 of tet+oct sites, sigma=0.03 A), 2 species -> 3 partials, "
                        "bin 0.001 A, 4096 Q, uint64 counts",
            "atoms": m.n_atoms,
            "pairs_per_step": pairs,
            "bins": m.n_bins,
            "nq": int(len(m.q)),
            "parallelism": "single GPU",
        },
        "pass1_kernel_ms": round(k_ms, 4),
        "pass1_gpairs_per_s": round(pairs / (k_ms * 1e-3) / 1e9, 2),
        "pass1_lds_atomics_per_s": round(pairs / (k_ms * 1e-3), 0),
        "lds_atomic_peak_per_s": round(lds_peak, 0),
        "pass1_lds_atomic_frac": round(pairs / (k_ms * 1e-3) / lds_peak, 4),
        "pass2_ms": round(p2_mean, 4),
        "pass2_gterms_per_s": round(mf_useful / 2 / (p2_mean * 1e-3) / 1e9, 2),
        "sinc_gelements_per_s": round(kc * len(m.q) / (mf_ms * 1e-3) / 1e9, 2),
        "pass1_jobs": n_jobs,
        "pass1_window_misses": misses,
        "pass1_evaluations_per_pair": round(evals / pairs, 3),
        "roofline": {
            "bound": "valu",
            "kernel": "p1_histogram_kernel",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": _profile_traffic("p1_histogram_kernel"),
            "valu_issue": _valu_issue(pairs, k_ms),
            "unit_figure": f"{FLOPS_PER_PAIR} FP32 flops per pair x {pairs} pairs per launch",
        },
        "pass2_roofline": {
            "bound": "mfma",
            "kernel": "p2_debye_mfma_kernel",
            "kernel_ms": round(mf_ms, 4),
            "achieved": round(mf_ach, 3),
            "peak": I8_PEAK_TOPS,
            "unit": "TOP/s",
            "frac": round(mf_ach / I8_PEAK_TOPS, 5),
            "unit_figure": f"2 nQ flops per nonzero bin of every row ({mf_useful:.4g} per launch), "
                           f"{kc} compacted bins, {digits} count digits",
            "mfma_issued_tops": round(mf_issued, 2),
            "mfma_issued_frac": round(mf_issued / I8_PEAK_TOPS, 4),
            "note": "one model has 3 rows: the contraction is bound by generating the sinc table "
                    "(FP64 VALU), not by the matrix cores; see the batched stream for many rows",
        },
    }
    out["error_bound"] = {"max_rel_err_bound_I": float(bound.max()), "pass2_refined_models": refined,
                          "target": 1e-6}
    return out, m, (I, bound), H, pairs


def stream_c4(cd, synth, np):
    models = synth.batch(256)
    params = [[_gp(cd, p) for p in m.params] for m in models]
    # the batched contraction of all 256 models (768 rows) in one dbx_batch_intensity call
    with cd.Context(0) as ctx:
        items = [(ctx.model(m.x, m.y, m.z, m.group_n), pa, m.n_bins, m.first_n) for m, pa in zip(models, params)]
        ctx.batch_intensity(items, models[0].rstep, models[0].q)
        t0 = time.perf_counter()
        ctx.batch_intensity(items, models[0].rstep, models[0].q)
        t_batch = time.perf_counter() - t0
        p1_b, p2_b = ctx.kernel_times()
        ctx.set_profiling(True)
        _, batch_bound = ctx.batch_intensity(items, models[0].rstep, models[0].q, with_err=True)
        batch_refined = ctx.pass2_refined()
        mf_ms, mf_flops, mf_useful, kc, digits, mean_digits = ctx.pass2_profile()
        ctx.set_profiling(False)
        # the same resident models, one dbx_intensity call each (plans already built): the
        # batching gain without host tiling on either side
        for m, it in zip(models, items):
            ctx.intensity(it[0], it[1], m.rstep, it[2], m.q, first_n=it[3])
        t0 = time.perf_counter()
        for m, it in zip(models, items):
            ctx.intensity(it[0], it[1], m.rstep, it[2], m.q, first_n=it[3])
        t_seq = time.perf_counter() - t0
        for it in items:
            it[0].close()
    batch_call = {
        "models_per_s": round(len(models) / t_batch, 1),
        "resident_one_call_per_model_models_per_s": round(len(models) / t_seq, 1),
        "gpu_only_ratio": round(t_seq / t_batch, 2),
        "max_rel_err_bound_I": float(batch_bound.max()), "pass2_refined_models": batch_refined,
        "pass1_kernel_ms": round(p1_b, 3),
        "pass2_ms": round(p2_b, 3),
        "pass2_roofline": {"bound": "mfma", "kernel": "p2_debye_mfma_kernel", "kernel_ms": round(mf_ms, 3),
                           "achieved": round(mf_useful / (mf_ms * 1e-3) / 1e12, 2), "peak": I8_PEAK_TOPS,
                           "unit": "TOP/s", "frac": round(mf_useful / (mf_ms * 1e-3) / 1e12 / I8_PEAK_TOPS, 4),
                           "mfma_issued_tops": round(mf_flops / (mf_ms * 1e-3) / 1e12, 1),
                           "mfma_issued_frac": round(mf_flops / (mf_ms * 1e-3) / 1e12 / I8_PEAK_TOPS, 4),
                           "unit_figure": f"2 nQ flops per nonzero bin of every row ({mf_useful:.4g}), "
                                          f"768 rows, {kc} compacted bins, {digits} count digits"},
    }
    # both paths in steady state: one untimed pass of the whole stream first (it grows the
    # context's and the lanes' scratch to their working sizes), then the timed pass
    # plain: one model per call (model creation + dbx_intensity), sequential
    with cd.Context(0) as ctx:
        for rep in range(2):
            t0 = time.perf_counter()
            I_plain = []
            for m, pa in zip(models, params):
                mod = ctx.model(m.x, m.y, m.z, m.group_n)
                I_plain.append(ctx.intensity(mod, pa, m.rstep, m.n_bins, m.q, first_n=m.first_n))
                mod.close()
            t_plain = time.perf_counter() - t0
    # server: the median of three timed passes (its steps depend on when the host threads finish
    # preparing each request, which varies from pass to pass)
    passes = []
    with cd.Server(0) as srv:
        for rep in range(4):
            st0 = srv.stats()
            steps0 = st0["steps"]
            t0 = time.perf_counter()
            ts = [srv.submit(m.x, m.y, m.z, m.group_n, pa, m.rstep, m.n_bins, m.q, first_n=m.first_n)
                  for m, pa in zip(models, params)]
            res = [srv.wait(t) for t in ts]
            t_pass = time.perf_counter() - t0
            st1 = srv.stats()
            steps = st1["steps"] - steps0
            host = ((st1["step_enqueue_ms"] - st0["step_enqueue_ms"]) / max(1, steps),
                    (st1["step_wait_ms"] - st0["step_wait_ms"]) / max(1, steps))
            for t in ts:
                srv.release(t)
            if rep:
                passes.append((t_pass, res, steps, host))
    passes.sort(key=lambda p: p[0])
    t_srv, res, steps, host = passes[len(passes) // 2]
    lats = np.array([r[1] for r in res])
    rel = max(float(np.max(np.abs(r[0] - ip) / np.abs(ip))) for r, ip in zip(res, I_plain))
    n = len(models)
    return {
        "workload": "256 two-species models (seeded FCC/HCP spheres and ellipsoids, stacking faults p=0.1, "
                    "sigma 0.01-0.08 A, 1k-60k atoms log-uniform), bin 0.001 A, 4096 Q; all submitted at once; "
                    "after one untimed pass of each path (steady state): plain timed once, server the median "
                    "of three passes",
        "pairs": float(sum(m.pair_count for m in models)),
        "plain_models_per_s": round(n / t_plain, 1),
        "server_models_per_s": round(n / t_srv, 1),
        "ratio": round(t_plain / t_srv, 2),
        "server_p50_ms": round(float(np.percentile(lats, 50)), 2),
        "server_p99_ms": round(float(np.percentile(lats, 99)), 2),
        "server_steps": int(steps),
        "step_enqueue_ms_mean": round(host[0], 3),
        "step_wait_ms_mean": round(host[1], 3),
        "max_rel_I_server_vs_plain": rel,
        "plain": "per model: dbx_model_create (host tiling + upload) + dbx_intensity, one at a time",
        "batch_call": batch_call,
    }


def _server_burst(srv, models, params):
    """Models/s of the server with every request submitted at once (after one warm-up request)."""
    m0 = models[0]
    t = srv.submit(m0.x, m0.y, m0.z, m0.group_n, params[0], m0.rstep, m0.n_bins, m0.q, first_n=m0.first_n)
    srv.wait(t)
    srv.release(t)
    t0 = time.perf_counter()
    ts = [srv.submit(m.x, m.y, m.z, m.group_n, pa, m.rstep, m.n_bins, m.q, first_n=m.first_n)
          for m, pa in zip(models, params)]
    for t in ts:
        srv.wait(t)
        srv.release(t)
    return len(models) / (time.perf_counter() - t0)


def stream_c5(cd, synth, np):
    models = [synth.serve_request(i) for i in range(SERVE_N)]
    params = [[_gp(cd, p) for p in m.params] for m in models]
    m0 = models[0]
    # one server for the capacity burst and the timed stream: the stream meets lanes whose
    # scratch memory the burst has grown (steady-state serving)
    with cd.Server(0) as srv:
        capacity = _server_burst(srv, models, params)
        rate = SERVE_LOAD * capacity
        rng = np.random.default_rng(5150)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(models)))
        # plain: a FIFO worker serving one model per call at the same arrival times
        with cd.Context(0) as ctx:
            mod = ctx.model(m0.x, m0.y, m0.z, m0.group_n)
            ctx.intensity(mod, params[0], m0.rstep, m0.n_bins, m0.q, first_n=m0.first_n)
            mod.close()
            t0 = time.perf_counter()
            lat_plain, I_plain = [], []
            for m, pa, a in zip(models, params, arrivals):
                wait = t0 + a - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                mod = ctx.model(m.x, m.y, m.z, m.group_n)
                I_plain.append(ctx.intensity(mod, pa, m.rstep, m.n_bins, m.q, first_n=m.first_n))
                mod.close()
                lat_plain.append((time.perf_counter() - t0 - a) * 1e3)
            t_plain = time.perf_counter() - t0
        st0 = srv.stats()
        ts = [None] * len(models)
        t0 = time.perf_counter()

        def feed():
            for i, (m, pa, a) in enumerate(zip(models, params, arrivals)):
                wait = t0 + a - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                ts[i] = srv.submit(m.x, m.y, m.z, m.group_n, pa, m.rstep, m.n_bins, m.q, first_n=m.first_n)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = [srv.wait(t, with_err=True) for t in ts]
        t_srv = time.perf_counter() - t0
        lat_srv = [r[1] for r in res]
        st1 = srv.stats()
        st = {k: st1[k] - st0[k] for k in ("steps", "step_enqueue_ms", "step_wait_ms")}
        st["max_step_qgrids"] = st1["max_step_qgrids"]
        for t in ts:
            srv.release(t)
    n = len(models)
    return {
        "workload": f"{n} seeded requests, 500-200k atoms log-uniform, 1-4 species, FCC/HCP spheres, "
                    "2048/4096/8192 Q, bin 0.001 A",
        "server_capacity_per_s": round(capacity, 1),
        "arrival_rate_per_s": round(rate, 1),
        "pairs": float(sum(m.pair_count for m in models)),
        "server": {"models_per_s": round(n / t_srv, 1),
                   "p50_ms": round(float(np.percentile(lat_srv, 50)), 2),
                   "p99_ms": round(float(np.percentile(lat_srv, 99)), 2),
                   "steps": int(st["steps"]),
                   # host time per step: issuing its work (one wait per step, nothing read back
                   # before it) and waiting for it; the other lane's step runs meanwhile
                   "step_enqueue_ms_mean": round(st["step_enqueue_ms"] / max(1, st["steps"]), 3),
                   "step_wait_ms_mean": round(st["step_wait_ms"] / max(1, st["steps"]), 3)},
        "plain": {"models_per_s": round(n / t_plain, 1),
                  "p50_ms": round(float(np.percentile(lat_plain, 50)), 2),
                  "p99_ms": round(float(np.percentile(lat_plain, 99)), 2)},
        "ratio": round(t_plain / t_srv, 2),
        "max_rel_I_server_vs_plain": max(float(np.max(np.abs(r[0] - ip) / np.abs(ip)))
                                         for r, ip in zip(res, I_plain)),
        "max_rel_err_bound_I": max(float(r[3].max()) for r in res),
        "max_step_qgrids": int(st["max_step_qgrids"]),
    }


def _resident_steps(cd, ctx, m, params, reps=3):
    """(median wall ms, median pass-1 kernel ms, median pass-2 ms, last (I, bound)) of `reps` steps of a
    resident model after one warm-up step (which builds its pass-1 plan)."""
    import numpy as np
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
    wall, p1, p2 = [], [], []
    for _ in range(reps):
        t = time.perf_counter()
        I = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_err=True)
        wall.append((time.perf_counter() - t) * 1e3)
        a, b = ctx.kernel_times()
        p1.append(a)
        p2.append(b)
    mod.close()
    return float(np.median(wall)), float(np.median(p1)), float(np.median(p2)), I


def workload_legs(cd, synth, np):
    """The other configurations, each timed on its own (not part of `value`)."""
    out = {}
    # the three-species rocksalt-like sphere at full size, and its I(Q) on the matrix cores
    # against FP64 arithmetic on the same histograms
    m = synth.rocksalt_sphere()
    params = [_gp(cd, p) for p in m.params]
    with cd.Context(0) as ctx:
        wall, p1, p2, (I, bound) = _resident_steps(cd, ctx, m, params)
        refined = ctx.pass2_refined()
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        # the matrix-core sums alone (no FP64 recomputation) and FP64 sums, each with its bound
        ctx.set_rel_err_target(0)
        Imf, bmf = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_err=True)
        ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
        I64, b64 = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_err=True)
        mod.close()
    out["rocksalt_sphere_20nm"] = {
        "workload": "rocksalt-like sphere D=20 nm (a=4.2 A, sigma 0), species 0.5/0.3/0.2 -> 6 partials, "
                    "bin 0.0005 A, 8192 Q uniform in 0.3-20 A^-1",
        "atoms": m.n_atoms, "pairs_per_step": m.pair_count, "bins": m.n_bins,
        "ms_per_step": round(wall, 2), "gpairs_per_s": round(m.pair_count / wall / 1e6, 1),
        "pass1_kernel_ms": round(p1, 2), "pass1_gpairs_per_s": round(m.pair_count / p1 / 1e6, 1),
        "pass2_ms": round(p2, 3),
        "max_rel_err_bound_I": float(bound.max()), "pass2_refined_models": refined,
        "matrix_cores_alone": {"max_rel_I_vs_fp64": float(np.max(np.abs(Imf - I64) / np.abs(I64))),
                               "max_rel_err_bound_I": float(bmf.max())},
        "fp64_max_rel_err_bound_I": float(b64.max()),
    }
    # the spinel-like sphere at the reference's default bin width (1e-4 A: 1.06M bins)
    m = synth.spinel_sphere(rstep=1e-4)
    params = [_gp(cd, p) for p in m.params]
    with cd.Context(0) as ctx:
        wall, p1, p2, _ = _resident_steps(cd, ctx, m, params)
        evals = ctx.pass1_evaluations()
    out["spinel_sphere_1e-4"] = {
        "workload": "the primary spinel-like sphere at bin 1e-4 A (the reference's default), 4096 Q",
        "bins": m.n_bins, "ms_per_step": round(wall, 2), "gpairs_per_s": round(m.pair_count / wall / 1e6, 1),
        "pass1_kernel_ms": round(p1, 2), "pass1_gpairs_per_s": round(m.pair_count / p1 / 1e6, 1),
        "pass2_ms": round(p2, 3),
        # pair evaluations per counted pair: shells wider than one LDS window are counted window
        # by window, each job evaluating every pair its staged atoms can form
        "pass1_evaluations_per_pair": round(evals / m.pair_count, 3),
    }
    # the primary model cold: host tiling, upload, pass-1 planning and the step, per model
    m = synth.spinel_sphere()
    params = [_gp(cd, p) for p in m.params]
    with cd.Context(0) as ctx:
        walls, creates = [], []
        for rep in range(4):
            t = time.perf_counter()
            mod = ctx.model(m.x, m.y, m.z, m.group_n)
            t1 = time.perf_counter()
            ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
            t2 = time.perf_counter()
            mod.close()
            if rep:
                walls.append((t2 - t) * 1e3)
                creates.append((t1 - t) * 1e3)
    w = float(np.median(walls))
    out["spinel_sphere_cold"] = {
        "workload": "the primary model not resident: dbx_model_create (host tiling + upload) + dbx_intensity "
                    "(pass-1 plan built on first use) per step",
        "ms_per_step": round(w, 2), "model_create_ms": round(float(np.median(creates)), 2),
        "gpairs_per_s": round(m.pair_count / w / 1e6, 1),
    }
    out["gamma_example_cli"] = _gamma_cli(cd)
    return out


def _gamma_cli(cd):
    """The reference's own example (Data/GammaAl2O3_10nm.deb: 117,936 atoms, 3 subgroups,
    1e-4 A bins, 6501 2-theta points, hist=2) through the `debye` program, FP64 pass 2 as the
    reference, CRLF outputs; the program's own timers against the reference's log."""
    import gzip
    import re
    import subprocess
    import tempfile
    gold = os.path.join(ROOT, "tests", "golden", "gamma")
    with tempfile.TemporaryDirectory() as d:
        deb = os.path.join(d, "GammaAl2O3_10nm.deb")
        with gzip.open(os.path.join(gold, "GammaAl2O3_10nm.deb.gz"), "rb") as f, open(deb, "wb") as g:
            g.write(f.read())
        runs = []
        for _ in range(2):
            t = time.perf_counter()
            r = subprocess.run([cd.CLI_PATH, deb, "--crlf"], capture_output=True, text=True, check=True)
            wall = time.perf_counter() - t
            v = {k: float(x) for k, x in re.findall(r"Time for ([^:]+): ([0-9.]+) s", r.stdout)}
            tot = re.search(r"Total time: ([0-9.]+) s", r.stdout)
            runs.append((wall, v, float(tot.group(1)) if tot else None))
        with gzip.open(os.path.join(gold, "GammaAl2O3_10nm.debI.gz"), "rb") as f:
            ref = f.read().decode().replace("\r", "").splitlines()
        with open(deb + "I", "rb") as f:
            got = f.read().decode().replace("\r", "").splitlines()
    wall, v, tot = runs[-1]
    calc = v.get("Calculation of Distances and Evaluation of the Debye Double Sum")
    region = v.get("Distances, Debye Double Sum and .debx files")
    return {
        "workload": "Data/GammaAl2O3_10nm.deb (117,936 atoms, 6 partials, 3.56M bins of 1e-4 A, 6501 2-theta), "
                    "hist=2, `debye --crlf` (FP64 pass 2); second of two runs",
        "pairs": 6954509016,
        "calc_s": calc, "debx_write_s": v.get("writing the .debx files"), "timed_region_s": region,
        "program_s": tot, "process_wall_s": round(wall, 3),
        "reference_calc_excl_debx_s": 6.832, "reference_timed_region_s": 11.563,
        "speedup_timed_region": round(11.563 / region, 1) if region else None,
        "speedup_calc": round(6.832 / calc, 1) if calc else None,
        "debI_identical_to_reference": got == ref,
    }


def cpu_leg(out, m, I_bound, H, pairs, np):
    """The oracle on the primary model; I_bound = (I, its error bound) from the device."""
    from oracle import oracle as O
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    st = np.concatenate([[0], np.cumsum(m.group_n)[:-1]])
    rp = O.rp_of(m.rstep)
    times = []
    for rep in range(4):  # one warm-up, then the median of three
        t = time.perf_counter()
        Href = np.zeros_like(H)
        p = 0
        for r in range(len(m.group_n)):
            for c in range(r, len(m.group_n)):
                O.histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c, rp,
                            m.n_bins, nthreads=nthreads, out=Href[p])
                p += 1
        times.append(time.perf_counter() - t)
    t1 = float(np.median(times[1:]))

    class Pm:
        group_n = m.group_n
        occ = np.array([q["occ"] for q in m.params])
        beq = np.array([q["beq"] for q in m.params])
        a = np.array([q["a"] for q in m.params])
        b = np.array([q["b"] for q in m.params])
        c = np.array([q["c"] for q in m.params])

    # pass 2 on a fixed sample of 512 of the 4096 Q points, scaled to the full grid
    qs = np.arange(0, len(m.q), 8)
    t = time.perf_counter()
    Iref = O.intensity_from_histograms(Href, Pm, m.q[qs], m.rstep, first_n=m.first_n, nthreads=nthreads)
    t2 = (time.perf_counter() - t) * len(m.q) / len(qs)
    cpu_s = t1 + t2
    out["cpu_baseline"] = {
        "value": round(pairs / cpu_s / 1e9, 4),
        "unit": "G pair-distances/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"full pass 1 of the same model (median of 3 after a warm-up: {t1:.2f} s); pass 2 "
                  f"timed on every 8th of the 4096 Q points and scaled x8 ({t2:.2f} s est.)",
        "cpu": _cpu_model(),
    }
    I, bound = I_bound
    err = np.abs(I[qs] - Iref) / np.abs(Iref)
    out["parity"] = {
        "histograms_bit_exact": bool(np.array_equal(H, Href)),
        "max_rel_err_I_vs_fp64_oracle": float(err.max()),
        "max_rel_err_bound_I": float(bound.max()),
        "error_within_bound": bool(np.all(err <= bound[qs] + 1e-14)),
    }


def cpu_batched(out, synth, np):
    """The oracle on the batched configuration: every 16th of the 256 models, each end to end
    (all partial histograms, then I at all 4096 Q), for models/s beside the GPU batch."""
    from oracle import oracle as O
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    models = synth.batch(256)[::16]
    t0 = time.perf_counter()
    pairs = 0
    for m in models:
        st = np.concatenate([[0], np.cumsum(m.group_n)[:-1]])
        rp = O.rp_of(m.rstep)
        S = len(m.group_n)
        H = np.zeros((S * (S + 1) // 2, m.n_bins), np.uint64)
        p = 0
        for r in range(S):
            for c in range(r, S):
                O.histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c, rp, m.n_bins,
                            nthreads=nthreads, out=H[p])
                p += 1

        class Pm:
            group_n = m.group_n
            occ = np.array([q["occ"] for q in m.params])
            beq = np.array([q["beq"] for q in m.params])
            a = np.array([q["a"] for q in m.params])
            b = np.array([q["b"] for q in m.params])
            c = np.array([q["c"] for q in m.params])

        O.intensity_from_histograms(H, Pm, m.q, m.rstep, first_n=m.first_n, nthreads=nthreads)
        pairs += m.pair_count
    dt = time.perf_counter() - t0
    out["cpu_baseline"]["batched"] = {
        "value": round(len(models) / dt, 2), "unit": "models/s", "cores": nthreads, "kind": "port",
        "pair_distances_per_s": round(pairs / dt / 1e9, 3),
        "sample": f"every 16th of the 256 batched models ({len(models)} models, {pairs:.3g} pair "
                  f"distances), each end to end: histograms then I at 4096 Q ({dt:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-streams", action="store_true", help="skip the batched and serving streams")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the other configurations (full rocksalt sphere, fine bins, cold model, Gamma example)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cudebye_amd as cd
    from cudebye_amd import synth

    out, m, I_bound, H, pairs = primary(args, cd, synth, np)
    if not args.no_workloads:
        out["workloads"] = workload_legs(cd, synth, np)
    if not args.no_streams:
        out["batched"] = stream_c4(cd, synth, np)
        out["serving"] = stream_c5(cd, synth, np)
    if not args.no_cpu:
        cpu_leg(out, m, I_bound, H, pairs, np)
        if not args.no_streams:
            cpu_batched(out, synth, np)
    print(json.dumps(out))


def _valu_issue(pairs, k_ms):
    """What bounds the pass-1 kernel: VALU lane operations issued per second (lane operations per
    pair from the committed SQ_INSTS_VALU count, profiles/traffic.json, times this run's pair
    rate) against the unpacked issue peak."""
    per_pair = _profile_traffic("p1_histogram_kernel_valu_lane_ops_per_pair")
    if not per_pair:
        return None
    rate = per_pair * pairs / (k_ms * 1e-3)
    out = {"lane_ops_per_pair": per_pair, "achieved": round(rate / 1e12, 2), "peak": round(VALU_ISSUE_PEAK / 1e12, 2),
           "unit": "T lane-ops/s", "frac": round(rate / VALU_ISSUE_PEAK, 4)}
    # the engine clock at which this rate is one VALU instruction per SIMD every cycle; the chip
    # does not hold its 2.4 GHz peak under this load, so compare with the clock the profiled
    # run held (GRBM_GUI_ACTIVE over the kernel's time), which varies from box to box
    out["clock_for_full_issue_ghz"] = round(rate / (VALU_ISSUE_PEAK / 2.4), 3)
    clock = _profile_traffic("p1_histogram_kernel_clock_ghz")
    if clock:
        out["profiled_clock_ghz"] = clock
    return out


def _profile_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same
    counters' reading for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()

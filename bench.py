"""Benchmark of the Debye hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One step = one structure model through the whole hot path on the device: pass 1 (all
subgroup-pair distance histograms, uint64), compaction, prefactors and pass 2 (I(Q)),
with the model's atoms already resident in HBM.  Workload: the two-species spinel-like
sphere (diameter 10 nm, bins 0.001 A, 4096 Q points), checked in the same run against the
CPU oracle: bit-exact histograms and the max relative I(Q) error against its FP64 sums.

Prints ONE JSON line.  `value` is G pair-distances/s, a pair-distance counted as the
reference counts it (N(N+1)/2 per model, self pairs included).  `roofline` is for the
pass-1 histogram kernel (9 FP32 flops per pair against the FP32 vector peak), timed with
HIP events on the library's stream.  `cpu_baseline` is the OpenMP oracle (a port of the
reference path) on the host cores of the same machine.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector/matrix peak
FLOPS_PER_PAIR = 9         # 3 sub, 3 mul, 2 add, 1 scale-mul


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cudebye_amd as cd
    from cudebye_amd import synth

    m = synth.spinel_sphere()
    params = [cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"]) for p in m.params]
    pairs = m.pair_count
    with cd.Context(0) as ctx:
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        for _ in range(args.warmup):
            ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
        p1_ms, p2_ms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            I = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n)
            a, b = ctx.kernel_times()
            p1_ms.append(a)
            p2_ms.append(b)
        elapsed = time.perf_counter() - t0  # intensity() synchronises the stream on return
        I, H = ctx.intensity(mod, params, m.rstep, m.n_bins, m.q, first_n=m.first_n, with_hist=True)
        n_jobs, misses = ctx.pass1_stats()
        mod.close()

    ms_per_step = elapsed / args.steps * 1e3
    value = pairs * args.steps / elapsed / 1e9
    k_ms = float(np.mean(p1_ms))
    achieved = FLOPS_PER_PAIR * pairs / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "G pair-distances/s (end to end: pass 1 histograms + pass 2 I(Q))",
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
        "pass2_ms": round(float(np.mean(p2_ms)), 4),
        "pass1_jobs": n_jobs,
        "pass1_window_misses": misses,
        "roofline": {
            "bound": "valu",
            "kernel": "p1_histogram_kernel",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": None,
            "unit_figure": f"{FLOPS_PER_PAIR} FP32 flops per pair x {pairs} pairs per launch",
        },
    }

    if not args.no_cpu:
        from oracle import oracle as O
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        st = np.concatenate([[0], np.cumsum(m.group_n)[:-1]])
        rp = O.rp_of(m.rstep)
        t = time.perf_counter()
        Href = np.zeros_like(H)
        p = 0
        for r in range(len(m.group_n)):
            for c in range(r, len(m.group_n)):
                O.histogram(m.x, m.y, m.z, st[r], m.group_n[r], st[c], m.group_n[c], r == c, rp,
                            m.n_bins, nthreads=nthreads, out=Href[p])
                p += 1
        t1 = time.perf_counter() - t

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
        Iref = O.intensity_from_histograms(Href, Pm, m.q[qs], m.rstep, first_n=m.first_n,
                                           nthreads=nthreads)
        t2 = (time.perf_counter() - t) * len(m.q) / len(qs)
        cpu_s = t1 + t2
        out["cpu_baseline"] = {
            "value": round(pairs / cpu_s / 1e9, 4),
            "unit": "G pair-distances/s",
            "cores": nthreads,
            "kind": "port",
            "sample": f"full pass 1 of the same model ({t1:.2f} s); pass 2 timed on every 8th of "
                      f"the 4096 Q points and scaled x8 ({t2:.2f} s est.)",
            "cpu": _cpu_model(),
        }
        out["parity"] = {
            "histograms_bit_exact": bool(np.array_equal(H, Href)),
            "max_rel_err_I_vs_fp64_oracle": float(np.max(np.abs(I[qs] - Iref) / np.abs(Iref))),
        }
    print(json.dumps(out))


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

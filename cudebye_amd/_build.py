"""Builds libdebye.so and the `debye` command-line program for gfx950 with hipcc.

    python -m cudebye_amd._build

Every HIP source is compiled for gfx950 only, with FP contraction off (pass 1's bin rule
depends on unfused FP32 arithmetic) and packed FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdebye.so")
CLI = os.path.join(HERE, "bin", "debye")
SOURCES = ["pass1.hip", "pass2.hip", "noalloc.hip", "api.cpp", "server.cpp", "deb_run.cpp", "xyz2deb.cpp",
           "view.cpp", "selftest.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # matrix instructions read and write ordinary VGPRs: no per-step copies of the
         # accumulators into AGPRs in the pass-2 contraction
         "-mllvm", "-amdgpu-mfma-vgpr-form=1",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
# pass 1 aggregates its uniform-address atomics itself (one lane per wave); the compiler's
# wave-level atomic rewriting would only add scan code around them, inside the pair loops too
FILE_FLAGS = {"pass1.hip": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC] + FLAGS + FILE_FLAGS.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + ["-ldl"])
    os.replace(tmp, LIB)
    _run([HIPCC, "-O2", "-std=c++17", "-o", CLI + ".tmp", os.path.join(CSRC, "cli.cpp"),
          "-L" + HERE, "-ldebye", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)

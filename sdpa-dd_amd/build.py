"""Builds libsdpadd_hip.so (dd Newton-step kernels + C ABI), the
sdpa_dd_gpu command-line solver for gfx950 with hipcc, and
libmpblas_dd_hip.so (MPLAPACK's dd BLAS/LAPACK symbols forwarded to the C
ABI), in-tree.

Usage: python sdpa-dd_amd/build.py [--jobs N] [--debug]
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsdpadd_hip.so")
CLI = os.path.join(HERE, "sdpa_dd_gpu")
SHIM = os.path.join(HERE, "libmpblas_dd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "ozaki.hip", "smallblk.hip", "problem.hip", "sparse_schur.hip", "solver.hip", "adjoint.hip",
           "capi.hip", "steqr.cpp"]
# Every rounding step of the dd arithmetic is part of the result: no
# contraction into FMAs.  Packed FP32 instructions are disabled (the dd code
# is FP64; this keeps the compiler from emitting packed forms anywhere).
FLAGS = [
    "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise SystemExit("build failed: %s" % cmd[-1])
    return r


def build(jobs=8, debug=False):
    os.makedirs(OBJ, exist_ok=True)
    opt = ["-O0", "-g"] if debug else ["-O3"]
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + FLAGS + opt + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(run, cmds))
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs)
    run([HIPCC] + FLAGS + opt + ["-x", "hip", os.path.join(CSRC, "cli_main.cpp"), "-o", CLI,
                                  "-L" + HERE, "-lsdpadd_hip", "-Wl,-rpath,$ORIGIN"])
    # MPLAPACK-symbol forwarding library (host C++ over the C ABI)
    run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", os.path.join(CSRC, "mpblas_shim.cpp"),
         "-o", SHIM, "-L" + HERE, "-lsdpadd_hip", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.debug))

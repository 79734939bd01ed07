"""Benchmark of the dd interior-point Newton step on one MI355X.

Workload: the max-cut relaxation of a 40x20 toroidal grid with +-1 weights
(n = m = 800, one SDP block, A_i = e_i e_i^T), synthetic, seed 20261203+1.
A step is one full interior-point solve from the standard starting point
(X = Z = 1e4 I) until the stopping rule with epsilonStar = epsilonDash =
1e-25 (relative gap and feasibility 1e-25), all other parameters at the
reference defaults.  Inputs are resident on the GPU before timing starts.

value = solve time in seconds (lower is better).  Also reported: Schur
dd-GFLOP/s (assembly + Cholesky algorithmic dd-flops over their time),
the roofline of the dominant kernel (dd GEMM) from HIP-event timing in the
timed region, and the CPU oracle's time for the same workload measured on a
bounded sample (one interior-point iteration) on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sdpa-dd_amd"))

FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 (vector and matrix), dense
DD_FP64_SLOTS_PER_DDFLOP = 20  # a dd multiply-add is 20 FP64 instructions = 40 flop-slots of the FMA peak
METRIC = "solve time to rel. gap 1e-25 (Ozaki-MFMA vs plain dd HIP); Schur dd-GFLOP/s"


def workload_files(tmp):
    from sdpadd import gen
    data = os.path.join(tmp, "maxcut_torus_40x20.dat-s")
    if not os.path.exists(data):
        gen.maxcut_torus(data)
    param = os.path.join(tmp, "param_gap1e-25.sdpa")
    gen.param_file(param, epsilon_star=1e-25, epsilon_dash=1e-25)
    return data, param


def schur_ddflops(n, m, iterations):
    """Algorithmic dd-flops of the Schur work per predictor step for A_i = e_i e_i^T:
    F3 pairs (j <= i) with one stored entry each cost 3 dd ops per pair; the
    Cholesky costs m^3/3."""
    pairs = m * (m + 1) / 2
    return iterations * (3.0 * pairs + m ** 3 / 3.0)


def cpu_baseline(data, param, iterations):
    """Oracle (CPU restatement, OpenMP GEMM) on one interior-point iteration of
    the same workload, scaled by the GPU solve's iteration count."""
    exe = os.path.join(ROOT, "oracle", "_build", "sdpa_oracle")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"], check=True, capture_output=True)
    cores = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.time()
    import sdpadd
    r = subprocess.run([exe, data, param, "--max-iter", "1", "--dot-threads", str(sdpadd.DEFAULT_DOT_PARTITION)],
                       capture_output=True, text=True, env=env, timeout=900)
    wall = time.time() - t0
    out = json.loads(r.stdout)
    per_it = out["seconds"]
    return {"value": per_it * iterations, "unit": "s", "cores": cores, "kind": "port",
            "sample": "1 interior-point iteration of the same workload by the CPU oracle (OpenMP GEMM), "
                      "%.2f s, scaled by the GPU solve's %d iterations" % (per_it, iterations),
            "wall_s": round(wall, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        return 2
    import sdpadd
    tmp = os.path.join(tempfile.gettempdir(), "sdpadd_bench")
    os.makedirs(tmp, exist_ok=True)
    data, param = workload_files(tmp)
    dev = sdpadd.Device(0)
    prob = sdpadd.Problem(dev, data)
    n, m = prob.block_struct[0], prob.m
    for _ in range(a.warmup):
        prob.solve(param)
    dev.stats(True)
    times, res = [], None
    t_all = time.time()
    for _ in range(a.steps):
        t0 = time.time()
        res = prob.solve(param)
        times.append(time.time() - t0)
    t_all = time.time() - t_all
    g = dev.stats_get("gemm")
    fam = {k: dev.stats_get(k) for k in ("gemm", "potrf", "trsm", "trmm", "trsv", "mv", "dot")}
    dev.stats(False)
    solve_s = sum(times) / len(times)
    it = res["iterations"]
    sch_flops = schur_ddflops(n, m, it)
    sch_time = res["t_schur"] + res["t_chol"]
    gemm_avg_ms = g["ms"] / max(1, g["launches"])
    gemm_dd_tflops = g["ddflops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    achieved = gemm_dd_tflops * DD_FP64_SLOTS_PER_DDFLOP
    line = {
        "metric": METRIC,
        "value": round(solve_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(solve_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "dd (f64x2)",
        "data": "synthetic",
        "config": {"workload": "max-cut SDP relaxation, 40x20 toroidal grid, +-1 weights (n=m=800), "
                               "gap and feasibility 1e-25, plain dd HIP path",
                   "n": n, "m": m, "path": "plain"},
        "iterations": it,
        "phase": res["phase"],
        "relgap": res["relgap"][0],
        "schur_dd_gflops": round(sch_flops / sch_time / 1e9, 3) if sch_time > 0 else None,
        "kernel_ms_per_family": {k: round(v["ms"] / a.steps, 2) for k, v in fam.items()},
        "roofline": {
            "bound": "mfma", "kernel": "k_gemm (dd GEMM on FP64 VALU)",
            "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "avg_launch_ms": round(gemm_avg_ms, 4), "dd_tflops": round(gemm_dd_tflops, 4),
        },
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(data, param, it)
    print(json.dumps(line))
    prob.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

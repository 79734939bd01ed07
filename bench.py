"""Benchmark of the dd interior-point Newton step on one MI355X.

Default workload: the max-cut relaxation of a 40x20 toroidal grid with +-1
weights (n = m = 800, one SDP block, A_i = e_i e_i^T), synthetic, seed
20261203+1.  --workload theta: the Lovasz-theta SDP of G(300, 0.1) (seed
20261203+4, m = |E|+1), whose dense Schur Cholesky dominates.

A step is one full interior-point solve from the standard starting point
(X = Z = 1e4 I) until the stopping rule with epsilonStar = epsilonDash =
1e-25 (relative gap and feasibility 1e-25), other parameters at the
reference defaults.  Inputs are resident on the GPU before timing starts.
Both forms of the Newton step are timed on the same problem: the plain dd
path (the reference's operation order) and the Ozaki form (Cholesky
trailing updates on INT8 MFMA with 16 digit planes and dd refinement of the
Schur solve; on max-cut also the n x n GEMMs of the direction).

value = solve time of the Ozaki form in seconds (lower is better);
plain_solve_s = the plain form; ozaki_speedup = plain / Ozaki.  Also:
Schur dd-GFLOP/s of the plain form, the roofline of the dominant kernel of
each form from HIP-event timing in the timed region, and (max-cut only) the
CPU oracle's time on a bounded sample (one interior-point iteration).

The reference's own algorithm stalls short of gap 1e-25 on the theta
workload (the plain form, which reproduces it, ends pFEAS near 1e-23; the
CPU oracle does the same on G(150, 0.1)), so --gap sets the stopping
tolerance for both forms (epsilonStar = epsilonDash); 1e-20 is the
reference's published SDPLIB setting (SDPLIB/param.sdpa:2,10).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload maxcut|theta] [--gap G]
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

FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 (vector and matrix), dense
INT8_PEAK_TOPS = 5000.0        # MI355X INT8 MFMA, dense
DD_FP64_SLOTS_PER_DDFLOP = 20   # a dd multiply-add is 20 FP64 instructions = 40 flop-slots of the FMA peak
OZ_SLICES = 16
METRIC = "solve time to rel. gap 1e-25 (Ozaki-MFMA vs plain dd HIP); Schur dd-GFLOP/s"
FAMILIES = ("gemm", "ozaki", "potrf", "trsm", "trmm", "trsv", "mv", "dot")
# HBM bytes per launch of the dominant kernels on the max-cut workload, from
# two rocprofv3 --pmc passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0: the factors
# a 256 MiB copy needs to read its true size); profiles/gemm_hbm_traffic_maxcut800_theta300.md
TRAFFIC = {"maxcut": {"oz::k_gemm": 30.52e6 + 2.83e6, "sdk::k_gemm": 38.86e6 + 2.45e6},
           "theta": {"oz::k_gemm": 70.49e6 + 23.54e6}}
TRAFFIC_SOURCE = "profiles/gemm_hbm_traffic_maxcut800_theta300.md"


def workload_files(tmp, workload, gap):
    from sdpadd import gen
    if workload == "theta":
        data = os.path.join(tmp, "theta_G300_0.1.dat-s")
        if not os.path.exists(data):
            gen.theta(data, n=300, p=0.1)
    else:
        data = os.path.join(tmp, "maxcut_torus_40x20.dat-s")
        if not os.path.exists(data):
            gen.maxcut_torus(data)
    param = os.path.join(tmp, "param_gap%g.sdpa" % gap)
    gen.param_file(param, epsilon_star=gap, epsilon_dash=gap)
    return data, param


def cpu_baseline(data, param, iterations):
    """Oracle (CPU restatement, OpenMP GEMM) on one interior-point iteration of
    the same workload, scaled by the GPU solve's iteration count."""
    import sdpadd
    exe = os.path.join(ROOT, "oracle", "_build", "sdpa_oracle")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"], check=True, capture_output=True)
    cores = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.time()
    r = subprocess.run([exe, data, param, "--max-iter", "1", "--dot-threads", str(sdpadd.DEFAULT_DOT_PARTITION)],
                       capture_output=True, text=True, env=env, timeout=900)
    wall = time.time() - t0
    out = json.loads(r.stdout)
    per_it = out["seconds"]
    return {"value": round(per_it * iterations, 3), "unit": "s", "cores": cores, "kind": "port",
            "sample": "1 interior-point iteration of the same workload by the CPU oracle (OpenMP GEMM), "
                      "%.2f s, scaled by the GPU solve's %d iterations" % (per_it, iterations),
            "wall_s": round(wall, 2)}


def run_form(dev, prob, param, path, steps, warmup, family, breakdown=True):
    """Mean solve time over `steps` timed solves, with HIP events on the
    launches of the measured kernel's family only (events on every family
    cost ~15% of a solve); then (breakdown) one untimed solve with every
    family timed, for the per-family breakdown."""
    dev.set_path(path, OZ_SLICES, 2)
    for _ in range(warmup):
        prob.solve(param)
    dev.stats(True, families=[family])
    times, res = [], None
    for _ in range(steps):
        t0 = time.time()
        res = prob.solve(param)
        times.append(time.time() - t0)
    timed = dev.stats_get(family)
    fam = {}
    if breakdown:
        dev.stats(True)
        prob.solve(param)
        fam = {k: dev.stats_get(k) for k in FAMILIES}
    dev.stats(False)
    dev.set_path(0)
    return sum(times) / len(times), res, timed, fam


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("maxcut", "theta"), default="maxcut")
    ap.add_argument("--gap", type=float, default=1e-25,
                    help="stopping tolerance: relative gap and feasibility (default 1e-25)")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the extra untimed solve per form that times every kernel family")
    ap.add_argument("--ozaki-path", type=int, choices=(1, 2), default=None,
                    help="1: Ozaki Cholesky; 2: Ozaki Cholesky and GEMMs (default: 2 for max-cut, 1 for theta)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        return 2
    import sdpadd
    tmp = os.path.join(tempfile.gettempdir(), "sdpadd_bench")
    os.makedirs(tmp, exist_ok=True)
    data, param = workload_files(tmp, a.workload, a.gap)
    dev = sdpadd.Device(0)
    prob = sdpadd.Problem(dev, data)
    n, m = prob.block_struct[0], prob.m
    plain_s, plain, g, fam_p = run_form(dev, prob, param, 0, a.steps, a.warmup, "gemm", not a.no_breakdown)
    # Ozaki form: Cholesky trailing updates and (max-cut, dense iterates) the
    # n x n GEMMs on INT8 MFMA; theta keeps its n x n GEMMs in dd
    oz_path = a.ozaki_path if a.ozaki_path else (2 if a.workload == "maxcut" else 1)
    oz_s, ozk, o, fam_o = run_form(dev, prob, param, oz_path, a.steps, a.warmup, "ozaki", not a.no_breakdown)

    gemm_dd_tflops = g["ddflops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    # int8 multiply-adds of the Ozaki kernel: S(S+1)/2 plane products per dd multiply-add
    oz_int8_tops = (o["ddflops"] * OZ_SLICES * (OZ_SLICES + 1) / 2) / (o["ms"] * 1e-3) / 1e12 if o["ms"] > 0 else 0.0
    chol_flops = plain["iterations"] * m ** 3 / 3.0
    ozaki_ok = ozk["phase"] == "pdOPT" and ozk["relgap"][0] <= a.gap
    value = oz_s if ozaki_ok else plain_s
    line = {
        "metric": METRIC if a.gap == 1e-25 else METRIC.replace("1e-25", "%g" % a.gap),
        "value": round(value, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(value * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "dd (f64x2); Ozaki form: int8 digit planes recombined in dd",
        "data": "synthetic",
        "config": {"workload": ("Lovasz theta SDP of G(300, 0.1) (n=300, m=%d)" % m) if a.workload == "theta" else
                   "max-cut SDP relaxation, 40x20 toroidal grid, +-1 weights (n=m=800)",
                   "gap_and_feasibility": a.gap, "n": n, "m": m,
                   "value_form": "ozaki" if ozaki_ok else "plain (the Ozaki form did not reach the gap)",
                   "ozaki_slices": OZ_SLICES, "ozaki_path": oz_path},
        "plain_solve_s": round(plain_s, 4),
        "ozaki_solve_s": round(oz_s, 4),
        "ozaki_speedup": round(plain_s / oz_s, 3),
        "iterations": {"plain": plain["iterations"], "ozaki": ozk["iterations"]},
        "phase": {"plain": plain["phase"], "ozaki": ozk["phase"]},
        "relgap": {"plain": plain["relgap"][0], "ozaki": ozk["relgap"][0]},
        "schur_dd_gflops": round(chol_flops / (plain["t_chol"] + plain["t_schur"]) / 1e9, 3)
        if plain["t_chol"] + plain["t_schur"] > 0 else None,
        "kernel_ms_per_solve": {"plain": {k: round(v["ms"], 2) for k, v in fam_p.items() if v["launches"]},
                                "ozaki": {k: round(v["ms"], 2) for k, v in fam_o.items() if v["launches"]}},
        "roofline": {
            "bound": "mfma", "kernel": "oz::k_gemm (Ozaki dd product, INT8 MFMA v_mfma_i32_16x16x64_i8)",
            "achieved": round(oz_int8_tops, 2), "peak": INT8_PEAK_TOPS, "unit": "TOP/s (int8 multiply-adds x2)",
            "frac": round(oz_int8_tops / INT8_PEAK_TOPS, 4),
            "traffic": TRAFFIC[a.workload].get("oz::k_gemm"),
            "traffic_source": TRAFFIC_SOURCE,
            "avg_launch_ms": round(o["ms"] / max(1, o["launches"]), 4),
        },
        "roofline_plain": {
            "bound": "mfma", "kernel": "sdk::k_gemm (dd GEMM on the FP64 VALU)",
            "achieved": round(gemm_dd_tflops * DD_FP64_SLOTS_PER_DDFLOP, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(gemm_dd_tflops * DD_FP64_SLOTS_PER_DDFLOP / FP64_PEAK_TFLOPS, 4),
            "traffic": TRAFFIC[a.workload].get("sdk::k_gemm"),
            "traffic_source": TRAFFIC_SOURCE if TRAFFIC[a.workload].get("sdk::k_gemm") else None,
            "avg_launch_ms": round(g["ms"] / max(1, g["launches"]), 4),
            "dd_tflops": round(gemm_dd_tflops, 4),
        },
    }
    if not a.no_cpu and a.workload == "maxcut":
        line["cpu_baseline"] = cpu_baseline(data, param, plain["iterations"])
    print(json.dumps(line))
    prob.close()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Benchmark of the dd interior-point Newton step on one MI355X.

Workload (default, the headline): the Lovasz-theta SDP of G(300, 0.1)
(n = 300, m = |E|+1 = 4551, seed 20261203+4), whose dense 4551 x 4551 dd
Schur Cholesky dominates.  Secondary (in the same line, --no-secondary to
skip): the same workload asked for gap 1e-25, and the max-cut relaxation of
a 40x20 toroidal grid with +-1 weights (n = m = 800, seed 20261203+1) at gap
1e-25.

A step is one full interior-point solve from the standard starting point
(X = Z = 1e4 I) until the stopping rule; inputs are resident on the GPU
before timing starts.  Both forms of the Newton step are timed on the same
problem: the plain dd path (the reference's operation order) and the Ozaki
form (Cholesky trailing updates on INT8 MFMA with 16 digit planes and dd
refinement of the Schur solve; on max-cut also the n x n GEMMs of the
direction).  value = solve time of the Ozaki form (lower is better).

The stopping tolerance of the headline is epsilonStar = epsilonDash = 1e-20,
the reference's own SDPLIB setting (SDPLIB/param.sdpa:2,10): asked for
1e-25, the reference's algorithm (which the plain form reproduces) stalls
near 1e-23 on this workload; the 1e-25 outcome of both forms is in the line
("theta_gap_1e-25").

Also in the line: Schur dd-GFLOP/s ((assembly dd-flops + m^3/3) per
iteration over the measured assembly + Cholesky time), the roofline of the
dominant kernel family of each form (HIP events on that family's launches
inside the timed region), a per-family kernel breakdown (one extra untimed
solve per form with every family timed) and the CPU oracle on a bounded
sample (one interior-point iteration of the same workload on the host
cores).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload theta|maxcut] [--gap G]
                [--no-breakdown] [--no-secondary] [--forms plain,ozaki]
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
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
DD_SLOTS = 20                  # FP64 flop-slots of the FMA peak per dd-flop (a dd multiply-add: 20 FP64 instructions)
OZ_SLICES = 16
SAMPLE_ITERS = 3               # oracle iterations timed for the CPU baseline
METRIC = "solve time to rel. gap 1e-25 (Ozaki-MFMA vs plain dd HIP); Schur dd-GFLOP/s"
FAMILIES = ("gemm", "ozaki", "potrf", "panel", "trsm", "trmm", "trsv", "trsv_prep", "mv", "dot")

# Per family: (bound, kernel names, how the algorithmic work per launch is
# priced).  "fp64": dd-flops x 20 FP64 flop-slots against the FP64 peak;
# "int8": dd multiply-adds x S(S+1)/2 plane products x 2 against the INT8
# peak; "bytes8"/"bytes16": 8 or 16 bytes per dd-flop read from HBM (a
# triangular mat-vec or solve reads 16 n^2/2 bytes for n^2 dd-flops; a dot
# 32 n bytes for 2n).
MODEL = {
    "gemm": ("mfma", "sdk::k_gemm (dd GEMM on the FP64 VALU)", "fp64"),
    "ozaki": ("mfma", "oz::k_gemm / oz::k_gemm32 (Ozaki dd product, INT8 MFMA v_mfma_i32_16x16x64_i8)", "int8"),
    "potrf": ("mfma", "sdk::k_potrf_fused (a 64-column diagonal Cholesky block and the panel rows below it in one "
              "launch on the FP64 VALU; the diagonal block's chain of 64 pivots sets its time)", "fp64"),
    "panel": ("mfma", "sdk::k_panel_rltn (Cholesky panel solve, Rtrsm R,L,T,N on the FP64 VALU)", "fp64"),
    "trsm": ("mfma", "sdk::k_trsm_llnn_diag + GEMM (Rtrsm L,L,N,N)", "fp64"),
    "trmm": ("mfma", "sdk::k_trmm_lltn (Rtrmm L,L,T,N)", "fp64"),
    "trsv": ("hbm", "sdk::k_trsv_*_coop / k_trsv_binv (Schur triangular solves)", "bytes8"),
    "trsv_prep": ("mfma", "sdk::k_blk_inv / k_blk_ops (block operators of the Ozaki-form solves)", "fp64"),
    "mv": ("hbm", "sdk::k_trmv_* / k_gemv_n / k_mvf_* (Lanczos mat-vecs)", "bytes8"),
    "dot": ("hbm", "sdk::k_dot_* (dot products)", "bytes16"),
}
# HBM bytes per launch of a kernel family on the theta workload (read +
# written, averaged over the family's launches in the form it dominates),
# from two rocprofv3 --pmc passes (FETCH_SIZE and WRITE_SIZE, scaled by the
# factors a 256 MiB copy needs to read its true size).
TRAFFIC = {
    "theta": {"ozaki": 35.26e6 + 12.58e6, "gemm": 78.77e6 + 42.21e6, "potrf": 4.35e6 + 5.64e6,
              "trsv": 176.97e6 + 0.36e6},
}
TRAFFIC_SOURCE = "profiles/hbm_traffic_theta300.md"


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


def host_cores():
    """Threads the CPU leg uses: OMP_NUM_THREADS when set (the box sets
    it to its CPU share), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline(data, param, iterations, workload):
    """The CPU oracle (the C++ restatement of the reference, OpenMP over the
    columns of its GEMM, SYRK, TRSM and TRMM loops and the rows of its Schur
    assembly) on the first SAMPLE_ITERS interior-point iterations of the same
    workload; their mean, scaled by the GPU solve's iteration count.  (The
    oracle's iterations cost about the same from the first on: 39, 50, 36
    and 46 s at 4 threads for the theta workload.)"""
    import sdpadd
    exe = os.path.join(ROOT, "oracle", "_build", "sdpa_oracle")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"], check=True, capture_output=True)
    cores = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.time()
    r = subprocess.run([exe, data, param, "--max-iter", str(SAMPLE_ITERS), "--dot-threads",
                        str(sdpadd.DEFAULT_DOT_PARTITION)], capture_output=True, text=True, env=env, timeout=900)
    wall = time.time() - t0
    out = json.loads(r.stdout)
    done = max(1, len(out["trace"]))
    per_it = out["seconds"] / done
    return {"value": round(per_it * iterations, 3), "unit": "s", "cores": cores, "kind": "port",
            "sample": "first %d interior-point iterations of the %s workload by the CPU oracle (%d OpenMP threads; "
                      "host nproc %d, %s), %.2f s each, scaled by the GPU solve's %d iterations" %
                      (done, workload, cores, os.cpu_count() or 0, cpu_model(), per_it, iterations),
            "wall_s": round(wall, 2)}


def solve_timed(dev, prob, param, path, steps, warmup, family=None):
    """Mean wall time of `steps` solves after `warmup` untimed ones; HIP
    events only on the launches of `family` (None: none)."""
    dev.set_path(path, OZ_SLICES, 2)
    try:
        for _ in range(warmup):
            prob.solve(param)
        if family:
            dev.stats(True, families=[family])
        times, res = [], None
        for _ in range(steps):
            t0 = time.time()
            res = prob.solve(param)
            times.append(time.time() - t0)
        fam = dev.stats_get(family) if family else None
        dev.stats(False)
    finally:
        dev.set_path(0)
    return sum(times) / len(times), res, fam


def breakdown(dev, prob, param, path):
    """One untimed solve with every kernel family timed."""
    dev.set_path(path, OZ_SLICES, 2)
    try:
        dev.stats(True)
        prob.solve(param)
        out = {k: dev.stats_get(k) for k in FAMILIES}
        dev.stats(False)
    finally:
        dev.set_path(0)
    return {k: v for k, v in out.items() if v["launches"]}


def roofline(family, fam, workload):
    bound, kernel, price = MODEL[family]
    secs = fam["ms"] * 1e-3
    out = {"bound": bound, "kernel": kernel, "family": family, "launches": fam["launches"],
           "avg_launch_ms": round(fam["ms"] / max(1, fam["launches"]), 4)}
    if price == "fp64":
        a = fam["ddflops"] * DD_SLOTS / secs / 1e12
        out.update(achieved=round(a, 4), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=round(a / FP64_PEAK_TFLOPS, 5),
                   work="%.4g dd-flops per launch x 20 FP64 flop-slots" % (fam["ddflops"] / max(1, fam["launches"])))
    elif price == "int8":
        a = fam["ddflops"] * OZ_SLICES * (OZ_SLICES + 1) / 2 / secs / 1e12
        out.update(achieved=round(a, 2), peak=INT8_PEAK_TOPS, unit="TOP/s (int8 multiply-adds x2)",
                   frac=round(a / INT8_PEAK_TOPS, 4),
                   work="%.4g dd-flops per launch x S(S+1)/2 = 136 plane products" %
                        (fam["ddflops"] / max(1, fam["launches"])))
    else:
        per = 8.0 if price == "bytes8" else 16.0
        a = fam["ddflops"] * per / secs / 1e9
        out.update(achieved=round(a, 2), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(a / HBM_PEAK_GBS, 4),
                   work="%.4g algorithmic bytes per launch" % (fam["ddflops"] * per / max(1, fam["launches"])))
    t = TRAFFIC.get(workload, {}).get(family)
    out["traffic"] = t
    out["traffic_source"] = TRAFFIC_SOURCE if t else None
    return out


def dominant(fams):
    return max(fams.items(), key=lambda kv: kv[1]["ms"])[0] if fams else "gemm"


def run_workload(dev, workload, gap, steps, warmup, forms, do_breakdown):
    import sdpadd
    tmp = os.path.join(tempfile.gettempdir(), "sdpadd_bench")
    os.makedirs(tmp, exist_ok=True)
    data, param = workload_files(tmp, workload, gap)
    prob = sdpadd.Problem(dev, data)
    oz_path = 2 if workload == "maxcut" else 1
    out = {"data": data, "param": param, "m": prob.m, "n": prob.block_struct[0],
           "schur_ddflops": prob.schur_ddflops(), "oz_path": oz_path}
    for form, path in (("plain", 0), ("ozaki", oz_path)):
        if form not in forms:
            continue
        fams = {}
        if do_breakdown:
            dev.set_path(path, OZ_SLICES, 2)
            prob.solve(param)  # warm the workspaces before the breakdown solve
            dev.set_path(0)
            fams = breakdown(dev, prob, param, path)
        fam_name = dominant(fams) if fams else ("ozaki" if form == "ozaki" else "gemm")
        secs, res, fam = solve_timed(dev, prob, param, path, steps, warmup, fam_name)
        out[form] = {"s": secs, "res": res, "fams": fams, "dom": fam_name, "dom_stats": fam}
    prob.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("theta", "maxcut"), default="theta")
    ap.add_argument("--gap", type=float, default=None,
                    help="stopping tolerance: relative gap and feasibility (default 1e-20 theta, 1e-25 max-cut)")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="skip the untimed solve per form that times every kernel family")
    ap.add_argument("--no-secondary", action="store_true", help="skip the theta 1e-25 and max-cut runs")
    ap.add_argument("--forms", default="plain,ozaki", help="comma-separated subset of plain,ozaki")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (requested --gpus %d)" % a.gpus,
              file=sys.stderr)
        return 2
    import sdpadd
    forms = set(a.forms.split(","))
    gap = a.gap if a.gap is not None else (1e-20 if a.workload == "theta" else 1e-25)
    dev = sdpadd.Device(0)
    w = run_workload(dev, a.workload, gap, a.steps, a.warmup, forms, not a.no_breakdown)
    m, n = w["m"], w["n"]
    plain, ozk = w.get("plain"), w.get("ozaki")
    ozaki_ok = ozk is not None and ozk["res"]["phase"] == "pdOPT" and ozk["res"]["relgap"][0] <= gap
    value_form = "ozaki" if ozaki_ok else "plain"
    head = w[value_form] if value_form in w else (ozk or plain)
    value = head["s"]
    desc = ("Lovasz theta SDP of G(300, 0.1) (n=300, m=%d)" % m) if a.workload == "theta" else \
        "max-cut SDP relaxation, 40x20 toroidal grid, +-1 weights (n=m=800)"
    line = {
        "metric": METRIC,
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
        "config": {"workload": desc, "gap_and_feasibility": gap, "n": n, "m": m, "value_form": value_form,
                   "ozaki_slices": OZ_SLICES, "ozaki_path": w["oz_path"],
                   "gap_note": "epsilonStar = epsilonDash = %g (the reference's SDPLIB setting); the reference's "
                               "algorithm stalls near 1e-23 here when asked for 1e-25 (see theta_gap_1e-25)" % gap
                   if a.workload == "theta" else "epsilonStar = epsilonDash = %g" % gap},
    }
    for form, f in (("plain", plain), ("ozaki", ozk)):
        if f:
            line["%s_solve_s" % form] = round(f["s"], 4)
    if plain and ozk:
        line["ozaki_speedup"] = round(plain["s"] / ozk["s"], 3)
    # the solver's own wall-clock split of the last timed solve (Schur
    # assembly and Cholesky from events; Lanczos step lengths and the X/Z
    # inverse factors as host wall time around their launches)
    line["phase_s"] = {k: {q: round(w[k]["res"][q], 4) for q in ("t_schur", "t_chol", "t_lanczos", "t_inverse")}
                       for k in ("plain", "ozaki") if k in w}
    line["iterations"] = {k: w[k]["res"]["iterations"] for k in ("plain", "ozaki") if k in w}
    line["phase"] = {k: w[k]["res"]["phase"] for k in ("plain", "ozaki") if k in w}
    line["relgap"] = {k: w[k]["res"]["relgap"][0] for k in ("plain", "ozaki") if k in w}
    for k in ("plain", "ozaki"):
        if k in w:
            r = w[k]["res"]
            t = r["t_schur"] + r["t_chol"]
            if t > 0:
                line["schur_dd_gflops" if k == "plain" else "schur_dd_gflops_ozaki"] = round(
                    r["iterations"] * (w["schur_ddflops"] + m ** 3 / 3.0) / t / 1e9, 3)
    line["schur_ddflops_per_assembly"] = w["schur_ddflops"]
    if not a.no_breakdown:
        line["kernel_ms_per_solve"] = {k: {f: round(v["ms"], 2) for f, v in w[k]["fams"].items()}
                                       for k in ("plain", "ozaki") if k in w}
    if head.get("dom_stats") and head["dom_stats"]["launches"]:
        line["roofline"] = roofline(head["dom"], head["dom_stats"], a.workload)
    other = "plain" if value_form == "ozaki" else "ozaki"
    if other in w and w[other].get("dom_stats") and w[other]["dom_stats"]["launches"]:
        line["roofline_%s" % other] = roofline(w[other]["dom"], w[other]["dom_stats"], a.workload)
    if not a.no_secondary and a.workload == "theta":
        tmp = os.path.join(tempfile.gettempdir(), "sdpadd_bench")
        data, param25 = workload_files(tmp, "theta", 1e-25)
        prob = sdpadd.Problem(dev, data)
        sec = {}
        for form, path in (("plain", 0), ("ozaki", 1)):
            dev.set_path(path, OZ_SLICES, 2)
            r = prob.solve(param25)
            dev.set_path(0)
            sec[form] = {"phase": r["phase"], "relgap": r["relgap"][0], "iterations": r["iterations"],
                         "solve_s": round(r["seconds"], 4)}
        prob.close()
        line["theta_gap_1e-25"] = sec
        mc = run_workload(dev, "maxcut", 1e-25, 1, 1, {"plain", "ozaki"}, False)
        line["maxcut_n800_gap_1e-25"] = {
            "plain_solve_s": round(mc["plain"]["s"], 4), "ozaki_solve_s": round(mc["ozaki"]["s"], 4),
            "ozaki_speedup": round(mc["plain"]["s"] / mc["ozaki"]["s"], 3),
            "iterations": {k: mc[k]["res"]["iterations"] for k in ("plain", "ozaki")},
            "phase": {k: mc[k]["res"]["phase"] for k in ("plain", "ozaki")},
            "relgap": {k: mc[k]["res"]["relgap"][0] for k in ("plain", "ozaki")}}
    if not a.no_cpu:
        its = line["iterations"].get("plain") or line["iterations"].get("ozaki")
        line["cpu_baseline"] = cpu_baseline(w["data"], w["param"], its, a.workload)
    print(json.dumps(line))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

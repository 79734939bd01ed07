"""End-to-end parity of the GPU Newton-step path against the CPU oracle.

* With the same dot partition (the library's default) the whole trajectory is
  bit-identical to the oracle: every per-iteration scalar and the final
  objectives, residual norms and iterates.
* With the partition set to one chunk (the reference's single-threaded
  order) the trajectory is bit-identical to that order too, so iteration
  count and phase agree, objectives to >= 28 significant digits and
  feasibility norms within 1e-28.
* On SDPLIB problems the reference's own result files pin iteration count,
  phase, the printed per-iteration trace and the objectives' high words.
"""
import os
import re

import numpy as np
import pytest

import sdpadd
from ddgen import same
from oracle_lib import SDPLIB_PARAM, golden_sdplib

pytestmark = pytest.mark.gpu
KEYS = ["mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta"]


def dd_to_mp(v):
    import mpmath
    mpmath.mp.dps = 50
    return mpmath.mpf(v[0]) + mpmath.mpf(v[1])


def digits_agree(a, b):
    import mpmath
    x, y = dd_to_mp(a), dd_to_mp(b)
    if x == y:
        return 99
    return float(-mpmath.log10(abs(x - y) / max(abs(x), abs(y))))


def gpu_solve(gpu, path, param=None):
    p = sdpadd.Problem(gpu, path)
    try:
        return p.solve(param)
    finally:
        p.close()


@pytest.mark.parametrize("name", ["theta1", "mcp100"])
def test_sdplib_bitwise_vs_oracle(gpu, oracle, name):
    path = golden_sdplib(name)
    ref = oracle.solve(path, SDPLIB_PARAM, dot_threads=sdpadd.DEFAULT_DOT_PARTITION)
    got = gpu_solve(gpu, path, SDPLIB_PARAM)
    assert got["iterations"] == ref["iterations"]
    assert got["phase"] == ref["phase"]
    assert len(got["trace"]) == len(ref["trace"])
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k, a[k], b[k])
    for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
        assert same(got[k], ref[k]), k


def _result_file(name):
    txt = open(os.path.join(os.path.dirname(__file__), "golden", "sdplib", name + ".result")).read()
    it = int(re.search(r"Iteration = (\d+)", txt).group(1))
    ph = re.search(r"phase.value = (\S+)", txt).group(1)
    objp = float(re.search(r"objValPrimal = (\S+)", txt).group(1))
    objd = float(re.search(r"objValDual   = (\S+)", txt).group(1))
    return it, ph, objp, objd


@pytest.mark.parametrize("name", ["theta1", "theta2", "mcp100", "mcp124-1", "mcp124-3"])
def test_sdplib_against_reference_results(gpu, name):
    it, ph, objp, objd = _result_file(name)
    got = gpu_solve(gpu, golden_sdplib(name), SDPLIB_PARAM)
    assert got["iterations"] == it and got["phase"] == ph
    # the result file prints -objValDual / -objValPrimal (REVERSE_PRIMAL_DUAL)
    assert "%.16e" % (-got["objValDual"][0]) == "%.16e" % objp
    assert "%.16e" % (-got["objValPrimal"][0]) == "%.16e" % objd


def test_random_dense_single_thread_order_bitwise(gpu, oracle, tmp_path):
    """Workload-0 shape (n=40, m=50, dense A_i) with the dot partition set to
    one chunk, the reference's deterministic single-threaded order: the whole
    trajectory is bit-identical to it, which implies the iteration count and
    phase, objectives to >= 28 digits and feasibility within 1e-28."""
    from sdpadd import gen
    data = gen.random_dense(str(tmp_path / "c0.dat-s"))
    param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-25, epsilon_dash=1e-25)
    ref = oracle.solve(data, param, dot_threads=1)
    gpu.set_dot_partition(1)
    try:
        got = gpu_solve(gpu, data, param)
    finally:
        gpu.set_dot_partition(sdpadd.DEFAULT_DOT_PARTITION)
    assert got["iterations"] == ref["iterations"]
    assert got["phase"] == ref["phase"]
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k)
    for k in ("objValPrimal", "objValDual"):
        assert digits_agree(got[k], ref[k]) >= 28, k
    for k in ("normPrimalVec", "normDualMat"):
        assert abs(dd_to_mp(got[k]) - dd_to_mp(ref[k])) <= 1e-28, k


def test_random_dense_bitwise(gpu, oracle, tmp_path):
    from sdpadd import gen
    data = gen.random_dense(str(tmp_path / "c0.dat-s"))
    param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-25, epsilon_dash=1e-25)
    ref = oracle.solve(data, param, dot_threads=sdpadd.DEFAULT_DOT_PARTITION)
    got = gpu_solve(gpu, data, param)
    assert got["iterations"] == ref["iterations"]
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k)


@pytest.mark.parametrize("builder,kw", [("maxcut_torus", dict(rows=8, cols=5)), ("theta", dict(n=40, p=0.2)),
                                        ("random_sparse", dict(n=30, m=40)), ("control", dict(n1=30, n2=22, m=60))])
def test_small_workloads_bitwise(gpu, oracle, tmp_path, builder, kw):
    from sdpadd import gen
    data = getattr(gen, builder)(str(tmp_path / "w.dat-s"), **kw)
    param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-25, epsilon_dash=1e-25)
    ref = oracle.solve(data, param, dot_threads=sdpadd.DEFAULT_DOT_PARTITION)
    got = gpu_solve(gpu, data, param)
    assert got["iterations"] == ref["iterations"] and got["phase"] == ref["phase"]
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k)


def _mixed(tmp_path):
    from sdpadd import gen
    data = gen.mixed_blocks(str(tmp_path / "mixed.dat-s"))
    return data, gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-20, epsilon_dash=1e-20)


@pytest.mark.parametrize("name", ["truss1", "truss4", "control1", "arch0", "hinf1", "example1", "mixed"])
def test_lp_and_small_blocks_bitwise_vs_oracle(gpu, oracle, tmp_path, name):
    """SDP blocks of size <= 20 (Rsyev step length), LP entries (negative and
    size-1 blocks) and their mix with a Lanczos block: the whole trajectory
    and the final iterate (SDP blocks, then LP entries) are bit-identical to
    the oracle's with the same dot partition.  hinf1 is ill-posed (its end
    depends on the partition) and still identical here."""
    if name == "mixed":
        path, param = _mixed(tmp_path)
    elif name == "example1":
        g = os.path.join(os.path.dirname(__file__), "golden")
        path, param = os.path.join(g, "example1.dat-s"), os.path.join(g, "param_default.sdpa")
    else:
        path, param = golden_sdplib(name), SDPLIB_PARAM
    ref = oracle.solve(path, param, dot_threads=sdpadd.DEFAULT_DOT_PARTITION, dump=True)
    p = sdpadd.Problem(gpu, path)
    try:
        got = p.solve(param)
        X, y, Z = p.solution()
    finally:
        p.close()
    assert got["iterations"] == ref["iterations"] and got["phase"] == ref["phase"]
    assert len(got["trace"]) == len(ref["trace"])
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k, a[k], b[k])
    for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
        assert same(got[k], ref[k]), k
    assert same(X, np.array(ref["xMat"])) and same(Z, np.array(ref["zMat"])) and same(y, np.array(ref["yVec"]))


@pytest.mark.parametrize("name", ["truss1", "truss3", "truss4", "control1", "arch0"])
def test_lp_and_small_blocks_against_reference_results(gpu, name):
    it, ph, objp, objd = _result_file(name)
    got = gpu_solve(gpu, golden_sdplib(name), SDPLIB_PARAM)
    assert got["iterations"] == it and got["phase"] == ph
    assert "%.16e" % (-got["objValDual"][0]) == "%.16e" % objp
    assert "%.16e" % (-got["objValPrimal"][0]) == "%.16e" % objd

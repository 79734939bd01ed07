"""The sparse-Schur path on the GPU (the reference's chordal branch).

The GPU assembles B dense in the reference's pair order and gathers the
pattern (Newton::permuteMat), then factors and solves on the pattern with
level-scheduled kernels; the oracle follows the reference's sparse loops
directly (pair-by-pair accumulation into the pattern, getCholesky's row
loop, solveSystems' storage-order sweeps).  With the same ordering and dot
partition the whole trajectory is bit-identical; the GPU's ordering (its own
minimum-degree code) equals the oracle's.  truss6, the instance the reference
solved on this path, also matches the reference's result file.
"""
import numpy as np
import pytest

import sdpadd
from ddgen import same
from oracle_lib import SDPLIB_PARAM, golden_sdplib
from test_oracle_sdplib import fmt_trace, parse_result

pytestmark = pytest.mark.gpu
KEYS = ["mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta"]


def _gpu(gpu, path, mode):
    p = sdpadd.Problem(gpu, path)
    try:
        p.set_schur_mode(mode)
        info = p.schur_info()
        res = p.solve(SDPLIB_PARAM)
        X, y, Z = p.solution()
        return info, res, (X, y, Z)
    finally:
        p.close()


def _bitwise(got, ref):
    assert got["iterations"] == ref["iterations"]
    assert got["phase"] == ref["phase"]
    assert len(got["trace"]) == len(ref["trace"])
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k, a[k], b[k])
    for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
        assert same(got[k], ref[k]), k


def test_truss6_sparse_path(gpu, oracle):
    """The reference's rule sends truss6 to the sparse path; the GPU takes the
    oracle's ordering and trajectory bit for bit and the reference's printed
    trace, iteration count, phase and objectives."""
    path = golden_sdplib("truss6")
    ref = oracle.solve(path, SDPLIB_PARAM, dot_threads=sdpadd.DEFAULT_DOT_PARTITION, dump=True)
    info, got, (X, y, Z) = _gpu(gpu, path, "auto")
    assert info["sparse"] and ref["schur"] == "sparse"
    assert info["ordering"] == ref["schur_ordering"]
    assert info["nnz"] == ref["schur_nnz"]
    _bitwise(got, ref)
    assert same(X, np.array(ref["xMat"])) and same(Z, np.array(ref["zMat"])) and same(y, np.array(ref["yVec"]))
    trace, fin = parse_result("truss6")
    assert got["iterations"] == int(fin["Iteration"]) and got["phase"] == fin["phase"]
    for rec, line in zip(got["trace"], trace):
        g = fmt_trace(rec)
        line = [x.replace("-0.00e+00", "+0.00e+00") for x in line]
        assert g[:2] + g[4:] == line[:2] + line[4:], (rec["it"], g, line)
    assert "%.16e" % -got["objValDual"][0] == fin["objValPrimal"]
    assert "%.16e" % -got["objValPrimal"][0] == fin["objValDual"]


# instances the reference runs dense, taken through the sparse path as its
# debugging switch does: small SDP blocks with an LP entry (truss1, truss4,
# control1's two blocks), a 161 block with 174 LP entries (arch0), a single
# dense block whose pattern is full (theta1), max-cut (mcp100)
@pytest.mark.parametrize("name", ["truss1", "truss4", "control1", "arch0", "theta1", "mcp100"])
def test_forced_sparse_bitwise_vs_oracle(gpu, oracle, name):
    path = golden_sdplib(name)
    ref = oracle.solve(path, SDPLIB_PARAM, dot_threads=sdpadd.DEFAULT_DOT_PARTITION, schur="sparse")
    info, got, _ = _gpu(gpu, path, "sparse")
    assert info["sparse"] and ref["schur"] == "sparse"
    assert info["ordering"] == ref["schur_ordering"]
    _bitwise(got, ref)


def test_truss6_forced_dense_bitwise_vs_oracle(gpu, oracle):
    """The dense path stays what it was when the rule would pick sparse."""
    path = golden_sdplib("truss6")
    ref = oracle.solve(path, SDPLIB_PARAM, dot_threads=sdpadd.DEFAULT_DOT_PARTITION, schur="dense")
    info, got, _ = _gpu(gpu, path, "dense")
    assert not info["sparse"] and ref["schur"] == "dense"
    _bitwise(got, ref)


def test_schur_mode_switch_on_one_problem(gpu):
    """Changing the mode between solves of one problem re-analyses it."""
    p = sdpadd.Problem(gpu, golden_sdplib("truss1"))
    try:
        assert not p.schur_info()["sparse"]  # m = 6: the reference's rule keeps it dense
        p.set_schur_mode("sparse")
        a = p.schur_info()
        assert a["sparse"] and sorted(a["ordering"]) == list(range(p.m))
        r1 = p.solve(SDPLIB_PARAM)
        p.set_schur_mode("dense")
        assert not p.schur_info()["sparse"]
        r2 = p.solve(SDPLIB_PARAM)
        assert r1["phase"] == r2["phase"] == "pdOPT"
    finally:
        p.close()

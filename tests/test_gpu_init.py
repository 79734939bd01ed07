"""Initial points and full-dd solution files on the GPU path, against the
oracle (whose reading of the reference's example1 initial point and of
exact HI+LO files is pinned in test_oracle_init.py)."""
import os

import numpy as np
import pytest

import sdpadd
from ddgen import same
from oracle_lib import GOLDEN, SDPLIB_PARAM, golden_sdplib
from test_oracle_init import write_pair_init

pytestmark = pytest.mark.gpu
KEYS = ["mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta"]
T = sdpadd.DEFAULT_DOT_PARTITION


def _compare(got, ref):
    assert got["iterations"] == ref["iterations"] and got["phase"] == ref["phase"]
    assert len(got["trace"]) == len(ref["trace"])
    for a, b in zip(got["trace"], ref["trace"]):
        for k in KEYS:
            assert same(a[k], b[k]), (a["it"], k, a[k], b[k])
    for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
        assert same(got[k], ref[k]), k


def _interior_init(oracle, path, param, tmp_path, iters=3):
    """An interior point: the oracle's iterate after a few steps, as a file."""
    r = oracle.solve(path, param, dot_threads=T, max_iter=iters, dump=True)
    _, bs = oracle.load(path)
    f = str(tmp_path / "interior.ini-s")
    write_pair_init(f, r["xMat"], r["yVec"], r["zMat"], bs, oracle.n_lp)
    return f


@pytest.mark.parametrize("case", ["example1-sparse", "example1-dense", "theta1", "mixed"])
def test_initial_point_bitwise_vs_oracle(gpu, oracle, tmp_path, case):
    dense = False
    if case.startswith("example1"):
        path, param = os.path.join(GOLDEN, "example1.dat-s"), os.path.join(GOLDEN, "param_default.sdpa")
        dense = case.endswith("dense")
        init = os.path.join(GOLDEN, "example1.ini" if dense else "example1.ini-s")
    elif case == "theta1":
        path, param = golden_sdplib("theta1"), SDPLIB_PARAM
        init = _interior_init(oracle, path, param, tmp_path)
    else:
        from sdpadd import gen
        path = gen.mixed_blocks(str(tmp_path / "mixed.dat-s"))
        param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=1e-20, epsilon_dash=1e-20)
        init = _interior_init(oracle, path, param, tmp_path)
    ref = oracle.solve(path, param, dot_threads=T, init=init, init_dense=dense)
    p = sdpadd.Problem(gpu, path)
    try:
        p.read_initial_point(init, sparse=not dense)
        got = p.solve(param)
    finally:
        p.close()
    _compare(got, ref)


@pytest.mark.parametrize("name", ["theta1", "arch0"])
def test_written_solution_restores_the_point(gpu, oracle, tmp_path, name):
    """write_solution -> read_initial_point (or set_initial_point with the
    arrays) restores the final point bit for bit: the next solve stops at
    once with the same objectives, mu and point; the oracle reads the file
    the same way; clearing the initial point gives the cold run again."""
    path = golden_sdplib(name)
    f = str(tmp_path / "sol.ini-s")
    p = sdpadd.Problem(gpu, path)
    try:
        cold = p.solve(SDPLIB_PARAM)
        X, y, Z = p.solution()
        p.write_solution(f)
        for how in ("file", "arrays"):
            if how == "file":
                p.read_initial_point(f)
            else:
                p.set_initial_point(X, y, Z)
            warm = p.solve(SDPLIB_PARAM)
            assert warm["iterations"] == 0 and warm["phase"] == cold["phase"] == "pdOPT", how
            for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
                assert same(warm[k], cold[k]), (how, k)
            X2, y2, Z2 = p.solution()
            assert same(X2, X) and same(y2, y) and same(Z2, Z), how
        p.clear_initial_point()
        again = p.solve(SDPLIB_PARAM)
        _compare(again, cold)
    finally:
        p.close()
    r = oracle.solve(path, SDPLIB_PARAM, dot_threads=T, init=f)
    assert r["iterations"] == 0
    for k in ("mu", "objValPrimal", "objValDual"):
        assert same(r[k], cold[k]), k

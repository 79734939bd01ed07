"""Initial points in the oracle (sdpa_main.cpp:237-251, IO::read at
sdpa_io.cpp:152-285, SolveInfo's initial-point form at sdpa_parts.cpp:691-740).

Pinned by: the reference's own example1 initial point in both of its file
formats (sparse .ini-s and dense .ini must give the same run), the closed
form mu0 = X0.Z0 / n of that point, and warm starts from a solution written
as exact HI+LO pairs, which must reproduce the solution bit for bit (the
run stops at once, with the previous objectives and mu).
"""
import os

import mpmath
import numpy as np

from oracle_lib import GOLDEN, SDPLIB_PARAM, golden_sdplib

EX = os.path.join(GOLDEN, "example1.dat-s")
PARAM = os.path.join(GOLDEN, "param_default.sdpa")


def test_example1_initial_point_formats_agree(oracle):
    a = oracle.solve(EX, PARAM, init=os.path.join(GOLDEN, "example1.ini-s"))
    b = oracle.solve(EX, PARAM, init=os.path.join(GOLDEN, "example1.ini"), init_dense=True)
    assert a["trace"] == b["trace"] and a["iterations"] == b["iterations"]
    assert a["phase"] == "pdOPT"
    # mu0 = X0 . Z0 / 2 with X0 = [[5.9, -1.375], [-1.375, 1]], Z0 = diag(11, 9)
    mpmath.mp.dps = 40
    mu0 = mpmath.mpf(a["trace"][0]["mu"][0]) + mpmath.mpf(a["trace"][0]["mu"][1])
    exact = (mpmath.mpf("5.9") * 11 + 9) / 2
    assert abs(mu0 - exact) < 1e-30 * exact
    # the objectives agree with the cold start's to all printed digits
    c = oracle.solve(EX, PARAM)
    assert "%.16e" % a["objValDual"][0] == "%.16e" % c["objValDual"][0]


def write_pair_init(path, X, y, Z, bs, nlp):
    """A sparse initial-point file with every value as the exact HI+LO pair."""
    tok = lambda v: "%.17e%+.17e" % (v[0], v[1])
    with open(path, "w") as f:
        # each value needs a separator before it (the readers' "%*[^0-9+-]")
        f.write(" " + " ".join(tok((-v[0], -v[1])) for v in y) + "\n")
        for target, M in ((1, Z), (2, X)):
            o = 0
            for l, n in enumerate(bs):
                for i in range(n):
                    for j in range(i, n):
                        v = M[o + i + n * j]
                        if i != j and v[0] == 0.0 and v[1] == 0.0:
                            continue
                        f.write("%d %d %d %d %s\n" % (target, l + 1, i + 1, j + 1, tok(v)))
                o += n * n
            for i in range(nlp):
                f.write("%d %d %d %d %s\n" % (target, len(bs) + 1, i + 1, i + 1, tok(M[o + i])))


def test_warm_start_from_a_written_solution_is_exact(oracle, tmp_path):
    for name in ("theta1", "truss1"):
        path = golden_sdplib(name)
        r = oracle.solve(path, SDPLIB_PARAM, dump=True)
        m, bs = oracle.load(path)
        init = str(tmp_path / (name + ".ini-s"))
        write_pair_init(init, r["xMat"], r["yVec"], r["zMat"], bs, oracle.n_lp)
        w = oracle.solve(path, SDPLIB_PARAM, init=init, dump=True)
        assert w["iterations"] == 0 and w["phase"] == r["phase"] == "pdOPT"
        for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat"):
            assert w[k] == r[k], (name, k)
        for k in ("xMat", "zMat", "yVec"):
            assert np.array_equal(np.array(w[k]), np.array(r[k])), (name, k)

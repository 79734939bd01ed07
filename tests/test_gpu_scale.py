"""Parity at the sizes the benchmark workloads have.

* The reference's own SDPLIB results at the workload shapes: maxG11 (the
  n = m = 800 max-cut shape), control11 (blocks 110 + 55, m = 1596), theta3,
  theta4 and theta6 (n = 300, m = 4375: the theta shape).  The plain form
  with SDPLIB/param.sdpa reproduces iteration count, phase, every printed
  trace line and the 17-digit objectives; the Ozaki form reaches the same
  phase and optimal value.
* The generated workloads at full size: the Schur matrix of the control-type
  (m = 1596) and theta (m ~ 4551) instances bit for bit against the oracle
  at a random interior point, and full solves of both forms to pdOPT.
* Every GEMM and Ozaki-product tiling at shapes where the size rule picks
  it, and at small shapes with the tiling forced, against the oracle.
"""
import os
import re
import sys

import numpy as np
import pytest

import sdpadd
from ddgen import rand_dd, rand_spd_cm, same
from oracle_lib import SDPLIB_PARAM, golden_sdplib
from test_oracle_sdplib import fmt_trace, parse_result

pytestmark = pytest.mark.gpu
ONE = np.array([1.0, 0.0])
ZERO = np.array([0.0, 0.0])
M1 = np.array([-1.0, 0.0])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve(gpu, path, param, form=0):
    gpu.set_path(form)
    try:
        p = sdpadd.Problem(gpu, path)
        try:
            return p.solve(param)
        finally:
            p.close()
    finally:
        gpu.set_path(0)


# ---------------------------------------------------------------- SDPLIB --
LARGE = ["maxG11", "theta3", "theta4", "theta6", "control11"]


@pytest.mark.parametrize("name", ["maxG11", "theta3", "theta4", "theta6"])
def test_sdplib_large_plain_vs_reference_results(gpu, name):
    say("plain form on SDPLIB %s" % name)
    trace, fin = parse_result(name)
    r = _solve(gpu, golden_sdplib(name), SDPLIB_PARAM)
    assert r["iterations"] == int(fin["Iteration"]) and r["phase"] == fin["phase"], (r["iterations"], r["phase"])
    assert len(r["trace"]) == len(trace) - 1
    for rec, ref in zip(r["trace"], trace):
        got = fmt_trace(rec)
        ref = [x.replace("-0.00e+00", "+0.00e+00") for x in ref]
        # thetaP/thetaD near 1e-35 are rounding noise; compare the rest exactly
        assert got[:2] + got[4:] == ref[:2] + ref[4:], (rec["it"], got, ref)
    assert "%.16e" % -r["objValDual"][0] == fin["objValPrimal"]
    assert "%.16e" % -r["objValPrimal"][0] == fin["objValDual"]
    # the final mu (~1e-21) moves by ~1e-5 relative between dot partitions
    assert abs(r["mu"][0] - float(fin["mu"])) <= 1e-4 * abs(float(fin["mu"]))


@pytest.mark.parametrize("T", [64, 1])
def test_sdplib_control11_up_to_the_stopping_edge(gpu, T):
    """control11 (blocks 110 + 55, m = 1596, dense F1 rows) ends on the
    1e-20 stopping test, and its last iterations (mu below 5e-20) move with
    the dot partition: with partitions 1, 8, 16, 32 and 64 it stops after
    57-59 iterations (the reference: 59; 59 with the serial partition).
    Every earlier printed trace line, the phase and the 17-digit objectives
    reproduce the reference's result file."""
    say("plain form on SDPLIB control11, dot partition %d" % T)
    trace, fin = parse_result("control11")
    gpu.set_dot_partition(T)
    try:
        r = _solve(gpu, golden_sdplib("control11"), SDPLIB_PARAM)
    finally:
        gpu.set_dot_partition(sdpadd.DEFAULT_DOT_PARTITION)
    assert r["phase"] == fin["phase"] == "pdOPT"
    assert abs(r["iterations"] - int(fin["Iteration"])) <= 2
    stable = [ref for ref in trace[:-1] if float(ref[1]) > 5e-20]
    assert len(stable) == 54
    for rec, ref in zip(r["trace"], stable):
        got = fmt_trace(rec)
        ref = [x.replace("-0.00e+00", "+0.00e+00") for x in ref]
        assert got[:2] + got[4:] == ref[:2] + ref[4:], (rec["it"], got, ref)
    assert "%.16e" % -r["objValDual"][0] == fin["objValPrimal"]
    assert "%.16e" % -r["objValPrimal"][0] == fin["objValDual"]


@pytest.mark.parametrize("name", LARGE)
def test_sdplib_large_ozaki_form(gpu, name):
    """Ozaki Cholesky form (path 1) with the SDPLIB parameters: the
    reference's phase and optimal value (its 17 printed digits, to 1e-14
    relative at gap 1e-20)."""
    say("Ozaki form on SDPLIB %s" % name)
    _, fin = parse_result(name)
    r = _solve(gpu, golden_sdplib(name), SDPLIB_PARAM, form=1)
    assert r["phase"] == fin["phase"]
    objp = float(fin["objValPrimal"])
    assert abs(-r["objValDual"][0] - objp) <= 1e-14 * abs(objp)
    assert abs(r["iterations"] - int(fin["Iteration"])) <= 3


# ------------------------------------------------- generated, full size --
@pytest.fixture(scope="module")
def workloads(tmp_path_factory):
    from sdpadd import gen
    d = tmp_path_factory.mktemp("scale")
    out = {}
    say("generating the control-type (110 + 55, m = 1596) and theta G(300, 0.1) workloads")
    out["control"] = gen.control(str(d / "control.dat-s"))
    out["theta"] = gen.theta(str(d / "theta.dat-s"))
    out["param"] = gen.param_file(str(d / "p20.sdpa"), epsilon_star=1e-20, epsilon_dash=1e-20)
    return out


@pytest.mark.parametrize("which", ["theta", "control"])
def test_schur_full_size_bitwise_vs_oracle(gpu, oracle, workloads, which):
    """sdpadd_schur_build against the oracle's compute_bMat_dense_SDP at a
    random interior point, bit for bit.  theta: F3 pairs for the edges and
    F2 for A_0 = I (300 entries: the multi-chunk path of the F2 kernel);
    control: F1 rows with dense 110 x 110 products."""
    path = workloads[which]
    say("Schur matrix of the %s workload" % which)
    m, bs = oracle.load(path)
    rng = np.random.default_rng(11)
    X = np.concatenate([rand_spd_cm(n, rng) for n in bs])
    Z = np.concatenate([rand_spd_cm(n, rng) for n in bs])
    # F1's inner products with dense A_j follow the dot partition
    oracle.set_dot_threads(sdpadd.DEFAULT_DOT_PARTITION)
    try:
        ref = oracle.schur(m, X, Z)
    finally:
        oracle.set_dot_threads(1)
    p = sdpadd.Problem(gpu, path)
    try:
        assert p.m == m and p.sdp_blocks == bs
        assert np.array_equal(p.formula(), oracle.formula(m, len(bs)))
        got = p.schur(X, Z)
    finally:
        p.close()
    bad = ~(got == ref).all(axis=1)
    assert not bad.any(), (int(bad.sum()), np.argwhere(bad)[:5].ravel())


@pytest.mark.parametrize("which", ["theta", "control"])
def test_full_size_workload_both_forms(gpu, workloads, which):
    """Full-size solves at gap 1e-20: both forms reach pdOPT within three
    iterations of each other, objectives agreeing to 1e-18 relative."""
    import mpmath
    mpmath.mp.dps = 50
    path, param = workloads[which], workloads["param"]
    say("plain form, %s workload" % which)
    plain = _solve(gpu, path, param, 0)
    say("Ozaki form, %s workload" % which)
    ozk = _solve(gpu, path, param, 1)
    assert plain["phase"] == "pdOPT" and ozk["phase"] == "pdOPT", (plain["phase"], ozk["phase"])
    assert abs(plain["iterations"] - ozk["iterations"]) <= 3
    for k in ("objValPrimal", "objValDual"):
        a = mpmath.mpf(plain[k][0]) + mpmath.mpf(plain[k][1])
        b = mpmath.mpf(ozk[k][0]) + mpmath.mpf(ozk[k][1])
        assert abs(a - b) <= 1e-18 * max(1, abs(a)), (k, float(abs(a - b)))


def test_wait_recovery_full_solves(gpu, workloads):
    """Solves whose cross-workgroup triangular solves all hand over to the
    single-workgroup recovery (spin bound of one iteration): the plain form,
    whose recovery keeps Rtrsv's order, reproduces the default solve bit for
    bit; the Ozaki form still reaches pdOPT with the same objectives to
    1e-18."""
    import mpmath
    mpmath.mp.dps = 50
    path, param = workloads["control"], workloads["param"]
    say("control workload, default spin bound, both forms")
    ref = _solve(gpu, path, param, 0)
    ref_oz = _solve(gpu, path, param, 1)
    gpu.set_wait_bound(1)
    try:
        say("control workload, spin bound 1, both forms")
        got = _solve(gpu, path, param, 0)
        got_oz = _solve(gpu, path, param, 1)
    finally:
        gpu.set_wait_bound(0)
    for k in ("iterations", "phase", "mu", "objValPrimal", "objValDual", "relgap", "trace"):
        assert got[k] == ref[k], (k, got[k], ref[k])
    assert got_oz["phase"] == "pdOPT" and abs(got_oz["iterations"] - ref_oz["iterations"]) <= 2
    for k in ("objValPrimal", "objValDual"):
        a = mpmath.mpf(ref_oz[k][0]) + mpmath.mpf(ref_oz[k][1])
        b = mpmath.mpf(got_oz[k][0]) + mpmath.mpf(got_oz[k][1])
        assert abs(a - b) <= 1e-18 * max(1, abs(a)), (k, float(abs(a - b)))


# ----------------------------------------------------------- tilings --
def _rgemm(gpu, tb, m, n, k, alpha, A, B, ldb, beta, C):
    import ctypes
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    dA, dB, dC = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_Rgemm(gpu.h, b"N", tb.encode(), m, n, k, dp(alpha), dA.ptr, m, dB.ptr, ldb, dp(beta),
                                   dC.ptr, m), "Rgemm")
    return dC.get()


@pytest.fixture
def tiles(gpu):
    yield gpu
    gpu.set_tile_policy(0)


@pytest.mark.parametrize("m,n,k", [(7, 5, 3), (65, 63, 17), (130, 70, 200), (200, 129, 64)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("ab", [("one", "zero"), ("m1", "one"), ("gen", "gen")])
def test_rgemm_64_tiles_forced(tiles, oracle, m, n, k, tb, ab):
    """The 64 x 64-tile dd GEMM (chosen by size from 1024 tiles on) at small
    shapes, forced: bit for bit the oracle's Rgemm."""
    gpu = tiles
    rng = np.random.default_rng(m * 1000 + n * 10 + k + 5)
    A, B, C = rand_dd((m * k,), rng), rand_dd((k * n,), rng), rand_dd((m * n,), rng)
    alpha = {"one": ONE, "m1": M1, "gen": rand_dd((1,), rng)[0]}[ab[0]]
    beta = {"zero": ZERO, "one": ONE, "gen": rand_dd((1,), rng)[0]}[ab[1]]
    ldb = k if tb == "N" else n
    ref = oracle.rgemm(tb, m, n, k, alpha, A, m, B, ldb, beta, C, m)
    gpu.set_tile_policy(2)
    assert same(_rgemm(gpu, tb, m, n, k, alpha, A, B, ldb, beta, C), ref)


@pytest.mark.parametrize("tb", ["N", "T"])
def test_rgemm_64_tiles_at_size(gpu, oracle, tb):
    """2048 x 2048 x 16: 1024 64-tiles, the size rule's own choice."""
    m = n = 2048
    k = 16
    rng = np.random.default_rng(2048 + (tb == "T"))
    A, B, C = rand_dd((m * k,), rng), rand_dd((k * n,), rng), rand_dd((m * n,), rng)
    ldb = k if tb == "N" else n
    ref = oracle.rgemm(tb, m, n, k, M1, A, m, B, ldb, ONE, C, m)
    assert same(_rgemm(gpu, tb, m, n, k, M1, A, B, ldb, ONE, C), ref)


def test_rsyrk_64_tiles_forced(tiles, oracle):
    gpu = tiles
    import ctypes
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    n, k = 150, 70
    rng = np.random.default_rng(3)
    A, C = rand_dd((n * k,), rng), rand_dd((n * n,), rng)
    ref = oracle.rsyrk(n, k, M1, A, n, ONE, C, n)
    gpu.set_tile_policy(2)
    dA, dC = gpu.to_device(A), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_Rsyrk(gpu.h, b"L", b"N", n, k, dp(M1), dA.ptr, n, dp(ONE), dC.ptr, n), "Rsyrk")
    assert same(dC.get(), ref)


def _potrf(gpu, A, n):
    import ctypes
    dA = gpu.to_device(A)
    gi = ctypes.c_int64(-7)
    gpu.check(gpu.lib.sdpadd_Rpotrf(gpu.h, b"L", n, dA.ptr, n, ctypes.byref(gi)), "Rpotrf")
    return dA.get(), gi.value


@pytest.mark.parametrize("n", [300, 777])
def test_rpotrf_64_tile_updates_forced(tiles, oracle, n):
    """Rpotrf whose lower-triangle trailing updates run on the 64 x 64 tiles."""
    gpu = tiles
    rng = np.random.default_rng(n + 1)
    A = rand_spd_cm(n, rng)
    ref, info = oracle.rpotrf(n, A, n)
    gpu.set_tile_policy(2)
    got, gi = _potrf(gpu, A, n)
    assert info == 0 and gi == 0
    assert same(got, ref)


def test_rpotrf_at_size_2112(gpu, oracle):
    """m = 2112: the first trailing updates take the 64 x 64 tiles by the
    size rule, the later ones the 32 x 32 tiles; bit for bit the oracle's."""
    n = 2112
    rng = np.random.default_rng(2112)
    A = rand_spd_cm(n, rng)
    say("Rpotrf n = 2112 on the oracle")
    ref, info = oracle.rpotrf(n, A, n)
    got, gi = _potrf(gpu, A, n)
    assert info == 0 and gi == 0
    assert same(got, ref)


def test_schur_batched_products_64_tiles_forced(tiles, oracle, tmp_path):
    """The F1 rows' batched products (gemm over pointer arrays) on 64 x 64
    tiles: the Schur matrix of a dense workload bit for bit."""
    gpu = tiles
    from sdpadd import gen
    path = gen.random_dense(str(tmp_path / "d.dat-s"), n=70, m=30)
    m, bs = oracle.load(path)
    rng = np.random.default_rng(5)
    X = np.concatenate([rand_spd_cm(n, rng) for n in bs])
    Z = np.concatenate([rand_spd_cm(n, rng) for n in bs])
    oracle.set_dot_threads(sdpadd.DEFAULT_DOT_PARTITION)
    try:
        ref = oracle.schur(m, X, Z)
    finally:
        oracle.set_dot_threads(1)
    gpu.set_tile_policy(2)
    p = sdpadd.Problem(gpu, path)
    try:
        got = p.schur(X, Z)
    finally:
        p.close()
    assert same(got, ref)


def _oz(gpu, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, lower):
    dA, dB, dC = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_ozaki_gemm(gpu.h, ta.encode(), tb.encode(), m, n, k, alpha, dA.ptr, lda, dB.ptr, ldb, beta,
                                        dC.ptr, m, lower, 16), "ozaki_gemm")
    return dC.get()


@pytest.mark.parametrize("case", ["full_1600_k128", "lower_2112_k64_self", "full_small_forced"])
def test_ozaki_64_and_32_tiles_identical(tiles, oracle, case):
    """The 64 x 64-tile Ozaki product (chosen from 512 tiles on) and the
    32 x 32-tile one give identical bits (same per-element level sums), and
    both stay within the Ozaki error bound of the oracle's dd product."""
    gpu = tiles
    rng = np.random.default_rng(len(case))
    if case == "full_1600_k128":
        m = n = 1600
        k, lower, tb = 128, 0, "N"
    elif case == "lower_2112_k64_self":
        m = n = 2112
        k, lower, tb = 64, 1, "T"
    else:
        m, n, k, lower, tb = 200, 130, 300, 0, "N"
    A = rand_dd((m * k,), rng)
    if case == "lower_2112_k64_self":
        B, ldb = A, m  # A A^T: one set of planes serves both operands
    else:
        B, ldb = rand_dd((k * n,), rng), k
    C = rand_dd((m * n,), rng)
    outs = {}
    for pol in (1, 2):
        gpu.set_tile_policy(pol)
        outs[pol] = _oz(gpu, "N", tb, m, n, k, -1, A, m, B, ldb, 1, C, lower)
    if case == "full_small_forced":
        pass
    else:
        gpu.set_tile_policy(0)
        outs[0] = _oz(gpu, "N", tb, m, n, k, -1, A, m, B, ldb, 1, C, lower)
        assert same(outs[0], outs[2])  # the size rule picks the 64 x 64 kernel here
    assert same(outs[1], outs[2])
    ref = oracle.rgemm(tb, m, n, k, M1, A, m, B, ldb, ONE, C, m)
    Ah = np.abs(A[:, 0]).reshape(k, m).max(axis=0)
    Bh = (np.abs(B[:, 0]).reshape(n, ldb)[:, :k] if tb == "N" else np.abs(B[:, 0]).reshape(k, ldb)[:, :n].T).max(axis=1)
    bound = k * 2.0 ** (10 - 7 * 16) * np.outer(Ah, Bh).T.reshape(-1)
    got = outs[2]
    err = np.abs((got[:, 0] - ref[:, 0]) + (got[:, 1] - ref[:, 1]))
    if lower:
        r, c = np.meshgrid(np.arange(m), np.arange(n), indexing="xy")  # element (r, c) at c*m + r
        keep = (r >= c).reshape(-1)
        assert same(got[~keep], C[~keep])  # the upper triangle is left alone
        err, bound = err[keep], bound[keep]
    assert np.all(err <= bound + 1e-300), float(np.max(err / (bound + 1e-300)))

"""Ozaki-scheme dd products on INT8 MFMA against the oracle's dd Rgemm.

The Ozaki form is not bit-identical to the reference; its error is stated
against the dd oracle: |C_ij - C_ref_ij| <= k * 2^(10 - 7 S) * rowmax_i(A) *
colmax_j(B), where S is the number of 7-bit digit planes.  Products of small
integers are exact in both forms and must agree bit for bit.
"""
import numpy as np
import pytest

import sdpadd
from ddgen import rand_dd, same

pytestmark = pytest.mark.gpu
ZERO = np.array([0.0, 0.0])
ONE = np.array([1.0, 0.0])


def oz(gpu, tb, m, n, k, alpha, A, B, ldb, beta, C, lower=0, S=16):
    dA, dB, dC = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_ozaki_gemm(gpu.h, b"N", tb.encode(), m, n, k, alpha, dA.ptr, m, dB.ptr, ldb, beta, dC.ptr,
                                        m, lower, S), "ozaki_gemm")
    return dC.get()


def ref_gemm(oracle, tb, m, n, k, alpha, A, B, ldb, beta, C):
    a = ONE if alpha == 1 else np.array([-1.0, 0.0])
    b = ZERO if beta == 0 else ONE
    return oracle.rgemm(tb, m, n, k, a, A, m, B, ldb, b, C, m)


def scale_bound(m, n, k, A, B, tb, ldb):
    Ah = np.abs(A[:, 0]).reshape(k, m).T            # A is m x k column-major
    if tb == "N":
        Bh = np.abs(B[:, 0]).reshape(n, ldb)[:, :k]  # B is k x n: column j -> row j here
    else:
        Bh = np.abs(B[:, 0]).reshape(k, ldb)[:, :n].T
    return np.outer(Ah.max(axis=1), Bh.max(axis=1))


def dd_diff(x, y):
    return np.abs((x[:, 0] - y[:, 0]) + (x[:, 1] - y[:, 1]))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (65, 63, 17), (130, 70, 200), (200, 129, 801)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("alpha,beta", [(1, 0), (-1, 1)])
def test_ozaki_gemm_error_bound(gpu, oracle, m, n, k, tb, alpha, beta):
    rng = np.random.default_rng(7 * m + 11 * n + k)
    A, B, C = rand_dd((m * k,), rng), rand_dd((k * n,), rng), rand_dd((m * n,), rng)
    ldb = k if tb == "N" else n
    ref = ref_gemm(oracle, tb, m, n, k, alpha, A, B, ldb, beta, C)
    got = oz(gpu, tb, m, n, k, alpha, A, B, ldb, beta, C)
    bound = k * 2.0 ** (10 - 7 * 16) * scale_bound(m, n, k, A, B, tb, ldb).T.reshape(-1)
    err = dd_diff(got, ref)
    assert np.all(err <= bound + 1e-300), float(np.max(err / (bound + 1e-300)))
    # and far below double precision: the dd part matters
    assert np.max(err / (bound / (k * 2.0 ** (10 - 7 * 16)))) < 1e-28


@pytest.mark.parametrize("S", [8, 12, 14])
def test_ozaki_gemm_fewer_slices(gpu, oracle, S):
    m, n, k = 96, 80, 150
    rng = np.random.default_rng(S)
    A, B = rand_dd((m * k,), rng), rand_dd((k * n,), rng)
    C = np.zeros((m * n, 2))
    ref = ref_gemm(oracle, "N", m, n, k, 1, A, B, k, 0, C)
    got = oz(gpu, "N", m, n, k, 1, A, B, k, 0, C, S=S)
    sc = scale_bound(m, n, k, A, B, "N", k).T.reshape(-1)
    err = dd_diff(got, ref)
    assert np.all(err <= k * 2.0 ** (10 - 7 * S) * sc)
    # a plane fewer would not meet the bound: the error is not far below it
    assert np.max(err / sc) > 2.0 ** (-7 * S - 20)


def test_ozaki_gemm_exact_on_small_integers(gpu, oracle):
    """Integer entries below 2^20: every digit is captured and all sums are
    exact, so the Ozaki and the plain products agree bit for bit."""
    m, n, k = 100, 90, 300
    rng = np.random.default_rng(3)
    A = np.stack([rng.integers(-2 ** 20, 2 ** 20, m * k).astype(float), np.zeros(m * k)], -1)
    B = np.stack([rng.integers(-2 ** 20, 2 ** 20, k * n).astype(float), np.zeros(k * n)], -1)
    C = np.zeros((m * n, 2))
    ref = ref_gemm(oracle, "N", m, n, k, 1, A, B, k, 0, C)
    got = oz(gpu, "N", m, n, k, 1, A, B, k, 0, C)
    assert same(got, ref)


def test_ozaki_gemm_wide_dynamic_range(gpu, oracle):
    """Rows and columns scaled by 2^(+-300): each row/column has its own
    exponent, so the normwise bound holds per entry."""
    m, n, k = 70, 66, 90
    rng = np.random.default_rng(5)
    A, B = rand_dd((m * k,), rng), rand_dd((k * n,), rng)
    rs = 2.0 ** rng.integers(-300, 300, m)
    cs = 2.0 ** rng.integers(-300, 300, n)
    A = A * np.tile(rs, k)[:, None]              # row i of column-major A: entries i + l*m
    B = B * np.repeat(cs, k)[:, None]            # column j of B: entries l + j*k
    C = np.zeros((m * n, 2))
    ref = ref_gemm(oracle, "N", m, n, k, 1, A, B, k, 0, C)
    got = oz(gpu, "N", m, n, k, 1, A, B, k, 0, C)
    bound = k * 2.0 ** (10 - 7 * 16) * scale_bound(m, n, k, A, B, "N", k).T.reshape(-1)
    assert np.all(dd_diff(got, ref) <= bound)


def test_ozaki_gemm_lower_only(gpu, oracle):
    n, k = 150, 64
    rng = np.random.default_rng(9)
    A = rand_dd((n * k,), rng)
    C = rand_dd((n * n,), rng)
    ref = ref_gemm(oracle, "T", n, n, k, -1, A, A, n, 1, C)
    got = oz(gpu, "T", n, n, k, -1, A, A, n, 1, C, lower=1)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    low = (i >= j).T.reshape(-1)   # column-major flattening
    assert same(got[~low], C[~low])
    bound = k * 2.0 ** (10 - 7 * 16) * scale_bound(n, n, k, A, A, "T", n).T.reshape(-1)
    assert np.all(dd_diff(got[low], ref[low]) <= bound[low])


def test_ozaki_gemm_rejects_bad_arguments(gpu):
    A = np.zeros((4, 2))
    dA = gpu.to_device(A)
    assert gpu.lib.sdpadd_ozaki_gemm(gpu.h, b"N", b"N", 2, 2, 2, 2, dA.ptr, 2, dA.ptr, 2, 0, dA.ptr, 2, 0, 16) == -1
    assert gpu.lib.sdpadd_ozaki_gemm(gpu.h, b"N", b"N", 2, 2, 2, 1, dA.ptr, 2, dA.ptr, 2, 0, dA.ptr, 2, 0, 9) == -1
    assert gpu.lib.sdpadd_ozaki_gemm(gpu.h, b"N", b"N", 2, 2, 9000, 1, dA.ptr, 2, dA.ptr, 9000, 0, dA.ptr, 2, 0,
                                     16) == -1


def test_ozaki_gemm_transposed_a(gpu, oracle):
    """op(A) = A^T: L^T L for a lower triangle, against the oracle's product
    of the explicit transpose."""
    n = 130
    rng = np.random.default_rng(11)
    L = rand_dd((n * n,), rng)
    Lt = L.reshape(n, n, 2).transpose(1, 0, 2).reshape(-1, 2).copy()  # column-major transpose
    C = np.zeros((n * n, 2))
    ref = ref_gemm(oracle, "N", n, n, n, 1, Lt, L, n, 0, C)
    dA, dC = gpu.to_device(L), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_ozaki_gemm(gpu.h, b"T", b"N", n, n, n, 1, dA.ptr, n, dA.ptr, n, 0, dC.ptr, n, 0, 16),
              "ozaki_gemm")
    bound = n * 2.0 ** (10 - 7 * 16) * scale_bound(n, n, n, Lt, L, "N", n).T.reshape(-1)
    assert np.all(dd_diff(dC.get(), ref) <= bound)


# ------------------------------------------------------ FP64-slice variant --
# |S| digits of 18 bits per element, held in doubles and accumulated by FP64
# MFMA: the truncation error is below k 2^(10 - 18|S|) rowmax colmax; at 7
# slices that is below dd rounding, so the bound adds the two dd products'
# own rounding, k 2^-100 rowmax colmax.
@pytest.mark.parametrize("m,n,k", [(7, 5, 3), (65, 63, 17), (130, 70, 200), (200, 129, 801)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("S", [-5, -6, -7])
def test_ozaki_fp64_slices_error_bound(gpu, oracle, m, n, k, tb, S):
    rng = np.random.default_rng(13 * m + n + k - S)
    A, B, C = rand_dd((m * k,), rng), rand_dd((k * n,), rng), rand_dd((m * n,), rng)
    ldb = k if tb == "N" else n
    ref = ref_gemm(oracle, tb, m, n, k, -1, A, B, ldb, 1, C)
    got = oz(gpu, tb, m, n, k, -1, A, B, ldb, 1, C, S=S)
    sc = scale_bound(m, n, k, A, B, tb, ldb).T.reshape(-1)
    err = dd_diff(got, ref)
    assert np.all(err <= k * (2.0 ** (10 - 18 * -S) + 2.0 ** -100) * sc + 1e-300)
    if S == -5:
        # 90 bits: the truncation shows above dd rounding
        assert np.max(err / sc) > 2.0 ** -110


def test_ozaki_fp64_slices_exact_on_small_integers(gpu, oracle):
    m, n, k = 100, 90, 300
    rng = np.random.default_rng(4)
    A = np.stack([rng.integers(-2 ** 20, 2 ** 20, m * k).astype(float), np.zeros(m * k)], -1)
    B = np.stack([rng.integers(-2 ** 20, 2 ** 20, k * n).astype(float), np.zeros(k * n)], -1)
    C = np.zeros((m * n, 2))
    ref = ref_gemm(oracle, "N", m, n, k, 1, A, B, k, 0, C)
    assert same(oz(gpu, "N", m, n, k, 1, A, B, k, 0, C, S=-4), ref)


# ---------------------------------------------------------------- solver --
def _solve(gpu, path, data, param, slices=16, refine=2):
    gpu.set_path(path, slices, refine)
    try:
        p = sdpadd.Problem(gpu, data)
        try:
            return p.solve(param)
        finally:
            p.close()
    finally:
        gpu.set_path(0)


def _mp(v):
    import mpmath
    mpmath.mp.dps = 50
    return mpmath.mpf(v[0]) + mpmath.mpf(v[1])


@pytest.mark.parametrize("builder,kw,eps", [("maxcut_torus", dict(rows=8, cols=5), 1e-25),
                                            ("maxcut_torus", dict(rows=40, cols=20), 1e-25),
                                            ("control", dict(n1=30, n2=22, m=60), 1e-20)])
@pytest.mark.parametrize("path", [1, 2])
def test_ozaki_solve_matches_plain(gpu, tmp_path, builder, kw, eps, path):
    """The Ozaki form reaches pdOPT like the plain form within two
    iterations, with objectives that agree to the stopping tolerance.

    The Ozaki form is not bit-identical, so on instances whose trajectory
    flips on a feasibility threshold it can take another (equally valid)
    branch -- the plain form does the same under another dot partition; the
    instances here are ones whose trajectory is stable."""
    from sdpadd import gen
    data = getattr(gen, builder)(str(tmp_path / "w.dat-s"), **kw)
    param = gen.param_file(str(tmp_path / "p.sdpa"), epsilon_star=eps, epsilon_dash=eps)
    plain = _solve(gpu, 0, data, param)
    assert plain["phase"] == "pdOPT"
    ozk = _solve(gpu, path, data, param)
    assert ozk["phase"] == "pdOPT"
    assert abs(ozk["iterations"] - plain["iterations"]) <= 2
    for k in ("objValPrimal", "objValDual"):
        a, b = _mp(ozk[k]), _mp(plain[k])
        assert abs(a - b) <= 100 * eps * max(1, abs(b)), (k, float(abs(a - b)))
    assert ozk["relgap"][0] <= eps


@pytest.mark.parametrize("name,slices", [("theta1", 16), ("theta2", 16), ("mcp100", 16), ("mcp124-1", 16),
                                         ("mcp124-3", 16), ("theta1", 14), ("mcp100", 14), ("theta1", -7),
                                         ("mcp100", -7)])
def test_ozaki_solve_sdplib(gpu, name, slices):
    """SDPLIB instances with the reference's SDPLIB parameters, Ozaki
    Cholesky form: same phase and optimal values as the reference's result
    files (17 printed digits, compared to 1e-14 relative: the gap is 1e-20)."""
    import re
    from oracle_lib import SDPLIB_PARAM, golden_sdplib
    import os
    txt = open(os.path.join(os.path.dirname(__file__), "golden", "sdplib", name + ".result")).read()
    ph = re.search(r"phase.value = (\S+)", txt).group(1)
    objp = float(re.search(r"objValPrimal = (\S+)", txt).group(1))
    got = _solve(gpu, 1, golden_sdplib(name), SDPLIB_PARAM, slices=slices)  # Ozaki Cholesky (m > 64)
    assert got["phase"] == ph
    assert abs(-got["objValDual"][0] - objp) <= 1e-14 * abs(objp)


def test_ozaki_refinement_with_fewer_planes(gpu):
    """12 digit planes (84 bits): the Schur factor is below dd accuracy, the
    residual of the Schur solve sits above the dd rounding floor and the
    refinement sweeps run and are kept; mcp100 still reaches the reference's
    optimum.  With 16 planes no sweep is needed."""
    import re, os
    from oracle_lib import SDPLIB_PARAM, golden_sdplib
    txt = open(os.path.join(os.path.dirname(__file__), "golden", "sdplib", "mcp100.result")).read()
    objp = float(re.search(r"objValPrimal = (\S+)", txt).group(1))
    got = _solve(gpu, 1, golden_sdplib("mcp100"), SDPLIB_PARAM, slices=12, refine=2)
    assert got["phase"] == "pdOPT" and got["refine_sweeps"] > 0
    assert abs(-got["objValDual"][0] - objp) <= 1e-14 * abs(objp)
    got16 = _solve(gpu, 1, golden_sdplib("mcp100"), SDPLIB_PARAM, slices=16, refine=2)
    assert got16["refine_sweeps"] == 0

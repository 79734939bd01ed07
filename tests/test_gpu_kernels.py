"""Kernel parity: every HIP dd kernel against the CPU oracle, bit for bit.

The kernels keep each output element's dd operation sequence identical to
the MPLAPACK loop the reference runs, so equality (not a tolerance) is the
bar here.
"""
import ctypes

import numpy as np
import pytest

import sdpadd
from ddgen import rand_dd, rand_lower_cm, rand_spd_cm, same

pytestmark = pytest.mark.gpu
ONE = np.array([1.0, 0.0])
ZERO = np.array([0.0, 0.0])


def dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (65, 63, 17), (130, 70, 200), (40, 40, 40)])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("ab", [("one", "zero"), ("m1", "one"), ("gen", "gen")])
def test_rgemm(gpu, oracle, m, n, k, tb, ab):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rand_dd((m * k,), rng)
    B = rand_dd((k * n,), rng)
    C = rand_dd((m * n,), rng)
    alpha = {"one": ONE, "m1": np.array([-1.0, 0.0]), "gen": rand_dd((1,), rng)[0]}[ab[0]]
    beta = {"zero": ZERO, "one": ONE, "gen": rand_dd((1,), rng)[0]}[ab[1]]
    ldb = k if tb == "N" else n
    ref = oracle.rgemm(tb, m, n, k, alpha, A, m, B, ldb, beta, C, m)
    dA, dB, dC = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_Rgemm(gpu.h, b"N", tb.encode(), m, n, k, dp(alpha), dA.ptr, m, dB.ptr, ldb, dp(beta),
                                   dC.ptr, m), "Rgemm")
    assert same(dC.get(), ref)


@pytest.mark.parametrize("n,k", [(5, 3), (64, 64), (100, 37)])
def test_rsyrk(gpu, oracle, n, k):
    rng = np.random.default_rng(n + k)
    A = rand_dd((n * k,), rng)
    C = rand_dd((n * n,), rng)
    alpha = np.array([-1.0, 0.0])
    ref = oracle.rsyrk(n, k, alpha, A, n, ONE, C, n)
    dA, dC = gpu.to_device(A), gpu.to_device(C)
    gpu.check(gpu.lib.sdpadd_Rsyrk(gpu.h, b"L", b"N", n, k, dp(alpha), dA.ptr, n, dp(ONE), dC.ptr, n), "Rsyrk")
    assert same(dC.get(), ref)


@pytest.mark.parametrize("n", [1, 2, 33, 64, 65, 150, 300])
def test_rpotrf(gpu, oracle, n):
    rng = np.random.default_rng(n)
    A = rand_spd_cm(n, rng)
    ref, info = oracle.rpotrf(n, A, n)
    dA = gpu.to_device(A)
    gi = ctypes.c_int64(-7)
    gpu.check(gpu.lib.sdpadd_Rpotrf(gpu.h, b"L", n, dA.ptr, n, ctypes.byref(gi)), "Rpotrf")
    assert info == 0 and gi.value == 0
    assert same(dA.get(), ref)


@pytest.mark.parametrize("n,bad", [(10, 3), (80, 70), (200, 129)])
def test_rpotrf_not_pd(gpu, oracle, n, bad):
    rng = np.random.default_rng(n)
    A = rand_spd_cm(n, rng)
    A[bad + n * bad] = [-5.0, 0.0]  # an indefinite leading minor at index bad
    _, info = oracle.rpotrf(n, A, n)
    dA = gpu.to_device(A)
    gi = ctypes.c_int64(0)
    gpu.check(gpu.lib.sdpadd_Rpotrf(gpu.h, b"L", n, dA.ptr, n, ctypes.byref(gi)), "Rpotrf")
    assert info > 0 and gi.value == info


@pytest.mark.parametrize("n", [300, 1000])
def test_rpotrf_panel_wait_bound(gpu, oracle, n):
    """Each diagonal block and the panel rows below it are one launch whose
    panel workgroups poll for the diagonal workgroup's published columns.
    With a bound of one empty poll (set from the host) a panel workgroup
    gives up as soon as a column is not there yet: Rpotrf then reports -2
    and no panel rows are stored (or, if every poll found its columns, it
    completes bit-identical).  With the default bound it is bit-identical."""
    rng = np.random.default_rng(n + 3)
    A = rand_spd_cm(n, rng)
    ref, info = oracle.rpotrf(n, A, n)
    assert info == 0
    gpu.set_panel_wait_bound(1)
    try:
        dA = gpu.to_device(A)
        gi = ctypes.c_int64(-7)
        gpu.check(gpu.lib.sdpadd_Rpotrf(gpu.h, b"L", n, dA.ptr, n, ctypes.byref(gi)), "Rpotrf")
        print("n = %d, bound 1: info %d" % (n, gi.value))
        assert gi.value in (0, -2)
        if gi.value == 0:
            assert same(dA.get(), ref)
    finally:
        gpu.set_panel_wait_bound(0)
    dA = gpu.to_device(A)
    gi = ctypes.c_int64(-7)
    gpu.check(gpu.lib.sdpadd_Rpotrf(gpu.h, b"L", n, dA.ptr, n, ctypes.byref(gi)), "Rpotrf")
    assert gi.value == 0 and same(dA.get(), ref)


@pytest.mark.parametrize("n", [3, 64, 100, 257])
def test_inverse_cholesky(gpu, oracle, n):
    """chol + zero upper + Rtrsm(L,L,N,N) on I + Rtrmm(L,L,T,N): Jal::getInvCholAndInv."""
    rng = np.random.default_rng(10 + n)
    A = rand_spd_cm(n, rng)
    L, info = oracle.rpotrf(n, A, n)
    Lm = L.reshape(n, n, 2)  # [col][row]
    for j in range(n):
        Lm[j, :j] = 0.0
    L = Lm.reshape(-1, 2)
    I = np.zeros((n * n, 2))
    I[:: n + 1, 0] = 1.0
    inv = oracle.rtrsm("LLNN", n, n, L, n, I, n)
    zinv = oracle.rtrmm(n, n, inv, n, inv.copy(), n)
    dA, dW, dI, dZ = gpu.to_device(A), gpu.empty((n * n, 2)), gpu.empty((n * n, 2)), gpu.empty((n * n, 2))
    rc = gpu.lib.sdpadd_inv_chol(gpu.h, n, dA.ptr, dW.ptr, dI.ptr, dZ.ptr)
    assert rc == 0
    assert same(dW.get(), L)
    assert same(dI.get(), inv)
    assert same(dZ.get(), zinv)


@pytest.mark.parametrize("m,n", [(5, 4), (70, 64), (150, 100)])
def test_rtrsm_rltn(gpu, oracle, m, n):
    rng = np.random.default_rng(m * n)
    A = rand_lower_cm(n, rng)
    B = rand_dd((m * n,), rng)
    ref = oracle.rtrsm("RLTN", m, n, A, n, B, m)
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    gpu.check(gpu.lib.sdpadd_Rtrsm(gpu.h, b"R", b"L", b"T", b"N", m, n, dp(ONE), dA.ptr, n, dB.ptr, m), "Rtrsm")
    assert same(dB.get(), ref)


@pytest.mark.parametrize("m,n", [(6, 3), (100, 40), (130, 130)])
def test_rtrsm_llnn(gpu, oracle, m, n):
    rng = np.random.default_rng(m + 7 * n)
    A = rand_lower_cm(m, rng)
    B = rand_dd((m * n,), rng)
    ref = oracle.rtrsm("LLNN", m, n, A, m, B, m)
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    gpu.check(gpu.lib.sdpadd_Rtrsm(gpu.h, b"L", b"L", b"N", b"N", m, n, dp(ONE), dA.ptr, m, dB.ptr, m), "Rtrsm")
    assert same(dB.get(), ref)


@pytest.mark.parametrize("m,n", [(6, 3), (100, 70)])
def test_rtrmm(gpu, oracle, m, n):
    rng = np.random.default_rng(m + n)
    A = rand_lower_cm(m, rng)
    B = rand_dd((m * n,), rng)
    ref = oracle.rtrmm(m, n, A, m, B, m)
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    gpu.check(gpu.lib.sdpadd_Rtrmm(gpu.h, b"L", b"L", b"T", b"N", m, n, dp(ONE), dA.ptr, m, dB.ptr, m), "Rtrmm")
    assert same(dB.get(), ref)


@pytest.mark.parametrize("n", [1, 2, 17, 50, 64, 65, 257, 511, 512, 1100, 2049, 4551])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_rtrsv_rtrmv(gpu, oracle, n, trans):
    rng = np.random.default_rng(n)
    A = rand_lower_cm(n, rng)
    x = rand_dd((n,), rng)
    x[n // 2] = 0.0  # the zero-skip branch
    ref = oracle.rtrsv(trans, n, A, n, x)
    dA, dx = gpu.to_device(A), gpu.to_device(x)
    gpu.check(gpu.lib.sdpadd_Rtrsv(gpu.h, b"L", trans.encode(), b"N", n, dA.ptr, n, dx.ptr, 1), "Rtrsv")
    assert same(dx.get(), ref)
    ref = oracle.rtrmv(trans, n, A, n, x)
    dx = gpu.to_device(x)
    gpu.check(gpu.lib.sdpadd_Rtrmv(gpu.h, b"L", trans.encode(), b"N", n, dA.ptr, n, dx.ptr, 1), "Rtrmv")
    assert same(dx.get(), ref)


@pytest.mark.parametrize("n", [1, 31, 33, 100, 130, 511, 512, 1100, 2049, 4551])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_ozaki_trsv(gpu, oracle, n, trans):
    """The Ozaki form's block-inverse solve: not Rtrsv's operation order, so
    compared with the oracle's dd substitution to a tolerance (a
    well-conditioned factor: unit-scale diagonal, off-diagonal 0.5/sqrt(n))."""
    from ddgen import dd_from_hi
    rng = np.random.default_rng(n + 7)
    a = np.tril(rng.standard_normal((n, n))) * (0.5 / np.sqrt(n)) + np.diag(1.0 + rng.uniform(0, 1, n))
    A = dd_from_hi(a, rng)
    A[np.triu_indices(n, 1)] = 0.0
    A = np.ascontiguousarray(A.transpose(1, 0, 2)).reshape(-1, 2)
    x = rand_dd((n,), rng)
    ref = oracle.rtrsv(trans, n, A, n, x)
    dA, dx = gpu.to_device(A), gpu.to_device(x)
    gpu.check(gpu.lib.sdpadd_ozaki_trsv(gpu.h, trans.encode(), n, dA.ptr, n, dx.ptr), "ozaki_trsv")
    got = dx.get()
    err = np.abs((got[:, 0] - ref[:, 0]) + (got[:, 1] - ref[:, 1]))
    assert np.max(err) <= 1e-26 * np.max(np.abs(ref[:, 0])), float(np.max(err) / np.max(np.abs(ref[:, 0])))


@pytest.mark.parametrize("n", [512, 2049, 4551])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_trsv_wait_recovery(gpu, oracle, n, trans):
    """The cross-workgroup solves with a spin bound of one iteration (set
    from the host): waits run out as soon as a flag is not already up, every
    workgroup leaves, and the single-workgroup solve enqueued behind redoes
    the solve from the saved right-hand side.  Rtrsv stays bit-identical to
    the oracle; the Ozaki form's block-inverse solve, which otherwise
    differs from Rtrsv's order in the last bits, now equals it bit for bit,
    which shows that the recovery ran."""
    from ddgen import dd_from_hi
    rng = np.random.default_rng(n + 7)
    a = np.tril(rng.standard_normal((n, n))) * (0.5 / np.sqrt(n)) + np.diag(1.0 + rng.uniform(0, 1, n))
    A = dd_from_hi(a, rng)
    A[np.triu_indices(n, 1)] = 0.0
    A = np.ascontiguousarray(A.transpose(1, 0, 2)).reshape(-1, 2)
    x = rand_dd((n,), rng)
    x[n // 3] = 0.0
    ref = oracle.rtrsv(trans, n, A, n, x)
    dA = gpu.to_device(A)
    gpu.set_wait_bound(1)
    try:
        dx = gpu.to_device(x)
        gpu.check(gpu.lib.sdpadd_Rtrsv(gpu.h, b"L", trans.encode(), b"N", n, dA.ptr, n, dx.ptr, 1), "Rtrsv")
        assert same(dx.get(), ref)
        dx = gpu.to_device(x)
        gpu.check(gpu.lib.sdpadd_ozaki_trsv(gpu.h, trans.encode(), n, dA.ptr, n, dx.ptr), "ozaki_trsv")
        assert same(dx.get(), ref)
    finally:
        gpu.set_wait_bound(0)
    # the default bound: the block-inverse order again
    dx = gpu.to_device(x)
    gpu.check(gpu.lib.sdpadd_ozaki_trsv(gpu.h, trans.encode(), n, dA.ptr, n, dx.ptr), "ozaki_trsv")
    got = dx.get()
    assert np.max(_dd_err(got, ref)) <= 1e-26 * np.max(np.abs(ref[:, 0]))


def _dd_err(got, ref):
    return np.abs((got[:, 0] - ref[:, 0]) + (got[:, 1] - ref[:, 1]))


@pytest.mark.parametrize("n", [1, 15, 16, 17, 63, 64, 65, 300, 801, 2000])
def test_ozaki_matvecs_and_dot(gpu, oracle, n):
    """The Ozaki form's split-order mat-vecs and dot against the oracle's
    reference-order results: |diff| <= 4 n 2^-104 sum |terms| (each order's
    dd rounding)."""
    rng = np.random.default_rng(n + 100)
    A = rand_dd((n * n,), rng)
    L = rand_lower_cm(n, rng)
    x = rand_dd((n,), rng)
    y = rand_dd((n,), rng)
    alpha, beta = rand_dd((1,), rng)[0], rand_dd((1,), rng)[0]
    tol = (4 * n + 16) * 2.0 ** -104
    absA = np.abs(A[:, 0]).reshape(n, n).T   # row-major view of column-major A
    absL = np.abs(L[:, 0]).reshape(n, n).T
    ax = np.abs(x[:, 0])
    # y = alpha A x + beta y
    ref = oracle.rgemv(n, n, alpha, A, n, x, beta, y)
    dA, dx, dy = gpu.to_device(A), gpu.to_device(x), gpu.to_device(y)
    gpu.check(gpu.lib.sdpadd_ozaki_matvec(gpu.h, b"G", n, n, dp(alpha), dA.ptr, n, dx.ptr, dp(beta), dy.ptr), "mv G")
    scale = abs(alpha[0]) * (absA @ ax) + abs(beta[0]) * np.abs(y[:, 0])
    assert np.all(_dd_err(dy.get(), ref) <= tol * scale + 1e-300)
    for op, tr in ((b"L", "N"), (b"T", "T")):
        ref = oracle.rtrmv(tr, n, L, n, x)
        dL, dx2, dout = gpu.to_device(L), gpu.to_device(x), gpu.empty((n, 2))
        gpu.check(gpu.lib.sdpadd_ozaki_matvec(gpu.h, op, n, n, None, dL.ptr, n, dx2.ptr, None, dout.ptr), "mv")
        sc = (absL @ ax) if tr == "N" else (absL.T @ ax)
        assert np.all(_dd_err(dout.get(), ref) <= tol * sc + 1e-300), op
    res = np.zeros(2)
    gpu.check(gpu.lib.sdpadd_ozaki_dot(gpu.h, n, dx.ptr, gpu.to_device(y).ptr, dp(res)), "dot")
    ref = oracle.rdot(x, y)
    assert abs((res[0] - ref[0]) + (res[1] - ref[1])) <= tol * float(np.sum(np.abs(x[:, 0] * y[:, 0]))) + 1e-300


@pytest.mark.parametrize("m,n", [(1, 1), (33, 17), (64, 16), (65, 33), (130, 801), (300, 300)])
@pytest.mark.parametrize("ab", [("gen", "zero"), ("m1", "one"), ("one", "gen")])
def test_rgemv(gpu, oracle, m, n, ab):
    rng = np.random.default_rng(m + n)
    A = rand_dd((m * n,), rng)
    x = rand_dd((n,), rng)
    y = rand_dd((m,), rng)
    alpha = {"gen": rand_dd((1,), rng)[0], "m1": np.array([-1.0, 0.0]), "one": ONE}[ab[0]]
    beta = {"zero": ZERO, "one": ONE, "gen": rand_dd((1,), rng)[0]}[ab[1]]
    ref = oracle.rgemv(m, n, alpha, A, m, x, beta, y)
    dA, dx, dy = gpu.to_device(A), gpu.to_device(x), gpu.to_device(y)
    gpu.check(gpu.lib.sdpadd_Rgemv(gpu.h, b"N", m, n, dp(alpha), dA.ptr, m, dx.ptr, 1, dp(beta), dy.ptr, 1), "Rgemv")
    assert same(dy.get(), ref)


@pytest.mark.parametrize("n", [0, 1, 5, 1023, 1024, 1025, 40000, 640000])
@pytest.mark.parametrize("T", [1, 7, 64, 1024, 2048])
def test_rdot_partition(gpu, oracle, n, T):
    """Rdot_omp's static partition into T chunks, partials added in order."""
    rng = np.random.default_rng(n + T)
    x = rand_dd((max(n, 1),), rng)
    y = rand_dd((max(n, 1),), rng)
    oracle.set_dot_threads(T)
    try:
        ref = oracle.rdot(x, y, n=n)
    finally:
        oracle.set_dot_threads(1)
    gpu.set_dot_partition(T)
    try:
        dx, dy = gpu.to_device(x), gpu.to_device(y)
        out = np.zeros(2)
        gpu.check(gpu.lib.sdpadd_Rdot(gpu.h, n, dx.ptr, 1, dy.ptr, 1, dp(out)), "Rdot")
    finally:
        gpu.set_dot_partition(sdpadd.DEFAULT_DOT_PARTITION)
    assert same(out, ref)


@pytest.mark.parametrize("n", [21, 40, 200, 800])
def test_lanczos(gpu, oracle, n):
    rng = np.random.default_rng(n)
    L = rand_lower_cm(n, rng)
    g = rng.standard_normal((n, n))
    Mh = (g + g.T) / 2
    M = np.ascontiguousarray(np.stack([Mh, np.zeros_like(Mh)], -1).transpose(1, 0, 2)).reshape(-1, 2)
    oracle.set_dot_threads(sdpadd.DEFAULT_DOT_PARTITION)
    try:
        ref = oracle.lanczos(n, L, M)
    finally:
        oracle.set_dot_threads(1)
    dL, dM = gpu.to_device(L), gpu.to_device(M)
    out = np.zeros(2)
    gpu.check(gpu.lib.sdpadd_lanczos_min_eig(gpu.h, n, dL.ptr, dM.ptr, dp(out)), "lanczos")
    assert same(out, ref)


@pytest.mark.parametrize("n,scale,threads", [(2, 1.0, None), (3, 1.0, None), (7, 1.0, None), (13, 1.0, None),
                                             (20, 1.0, None), (20, 1.0, 1), (9, 1e-140, None), (9, 1e140, None)])
def test_small_min_eig(gpu, oracle, n, scale, threads):
    """Blocks of size <= 20: lambda_min(L M L^T) by the Rsyev path, bit for
    bit with the oracle (which mpmath pins), including matrices scaled below
    rmin and above rmax (Rsyev's Rlascl path) and the serial dot partition."""
    T = threads or sdpadd.DEFAULT_DOT_PARTITION
    rng = np.random.default_rng(100 + n)
    L = rand_lower_cm(n, rng)
    g = rng.standard_normal((n, n)) * scale
    Mh = (g + g.T) / 2
    M = np.ascontiguousarray(np.stack([Mh, np.zeros_like(Mh)], -1).transpose(1, 0, 2)).reshape(-1, 2)
    oracle.set_dot_threads(T)
    gpu.set_dot_partition(T)
    try:
        ref, rinfo = oracle.qr_min_eigen(n, L, M)
        dL, dM = gpu.to_device(L), gpu.to_device(M)
        out = np.zeros(2)
        info = ctypes.c_int(-1)
        gpu.check(gpu.lib.sdpadd_small_min_eig(gpu.h, n, dL.ptr, dM.ptr, dp(out), ctypes.byref(info)), "small_min_eig")
    finally:
        oracle.set_dot_threads(1)
        gpu.set_dot_partition(sdpadd.DEFAULT_DOT_PARTITION)
    assert info.value == 0 and rinfo == 0
    assert same(out, ref), (out, ref)

"""sdpadd_Rtrmm_noalloc (the allocation-free Rtrmm, scratch in a fixed
device array) against sdpadd_Rtrmm through the C ABI: bit for bit on two
calls in a row, the second smaller than the first, and the size limit.
Named to be collected last: it must pass alone and after every other test."""
import ctypes

import numpy as np
import pytest

from ddgen import rand_dd, rand_lower_cm, same

pytestmark = pytest.mark.gpu
ONE = np.array([1.0, 0.0])


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _trmm(gpu, fn, m, n, A, B, ldb):
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    rc = fn(gpu.h, b"L", b"L", b"T", b"N", m, n, _dp(ONE), dA.ptr, m, dB.ptr, ldb)
    return rc, dB.get()


def test_noalloc_rtrmm_matches_rtrmm_twice(gpu, oracle):
    rng = np.random.default_rng(77)
    for m, n, ldb in [(700, 900, 703), (130, 75, 130)]:  # second call smaller
        A = rand_lower_cm(m, rng)
        B = rand_dd((ldb * n,), rng)
        rc1, ref = _trmm(gpu, gpu.lib.sdpadd_Rtrmm, m, n, A, B, ldb)
        rc2, got = _trmm(gpu, gpu.lib.sdpadd_Rtrmm_noalloc, m, n, A, B, ldb)
        assert rc1 == 0 and rc2 == 0, gpu.lib.sdpadd_last_error(gpu.h)
        assert same(got, ref)
        # rows ldb > m of each column are left alone
        Bm, Gm = B.reshape(n, ldb, 2), got.reshape(n, ldb, 2)
        assert same(Gm[:, m:], Bm[:, m:])
    # and the reference's own values on the smaller call
    Bc = B.reshape(n, ldb, 2)[:, :m].reshape(-1, 2)
    assert same(got.reshape(n, ldb, 2)[:, :m].reshape(-1, 2), oracle.rtrmm(m, n, A, m, Bc, m))


def test_noalloc_rtrmm_rejects_oversize(gpu):
    m, n = 1025, 1024  # m * n > 2^20
    A = np.zeros((m * m, 2))
    B = np.zeros((m * n, 2))
    dA, dB = gpu.to_device(A), gpu.to_device(B)
    rc = gpu.lib.sdpadd_Rtrmm_noalloc(gpu.h, b"L", b"L", b"T", b"N", m, n, _dp(ONE), dA.ptr, m, dB.ptr, m)
    assert rc == -1
    assert b"scratch" in gpu.lib.sdpadd_last_error(gpu.h)

"""The device dd operators (csrc/dd.hpp) against the oracle's restatement
of QD 2.3.23 (oracle/dd_oracle.hpp: Dekker split, sloppy add and divide,
Karp square root), bit for bit, on 2^20 operand pairs per operator.

The pairs (tests/ddgen.py special_dd_pairs) cover N(0,1) values, exponents
over 2^+-1000, lo = 0, signed zeros, exact and near cancellation and
half-ulp ties.  The device forms TwoProduct's error with an FMA, QD with
Dekker's split; both give the exact error of a product p whenever that error
is representable, i.e. |p| >= 2^-969 (and no overflow).  Below that the
error term falls into gradual underflow, where the two round differently and
both lose dd accuracy: those pairs are checked to agree in the high word
(products, whose results are themselves subnormal there, to 2^-16 relative)
and counted, not held to bit equality.
"""
import numpy as np
import pytest

from ddgen import special_dd_pairs

pytestmark = pytest.mark.gpu

N = 1 << 20
OPS = {0: "a+b", 1: "a-b", 2: "a*b", 3: "a/b", 4: "sqrt(a)", 5: "a+=b", 6: "a-=b", 7: "a*=b", 8: "a*hi(b)",
       9: "a/hi(b)", 10: "a+hi(b)", 11: "hi(a)/b"}


def _product_scale(op, a, b):
    """Magnitude of the TwoProducts the operator forms (None: it forms none)."""
    with np.errstate(all="ignore"):
        if op in (2, 7, 8):
            return np.abs(a[:, 0] * b[:, 0])
        if op in (3, 9, 11, 4):
            return np.abs(a[:, 0])  # q*b ~ a (division), (sqrt a)^2 ~ a
    return None


@pytest.fixture(scope="module")
def pairs():
    return special_dd_pairs(N, np.random.default_rng(20261203))


@pytest.mark.parametrize("op", sorted(OPS))
def test_dd_operator_bitwise_vs_qd(gpu, oracle, pairs, op):
    a, b = pairs
    if op == 4:
        a = np.abs(a)
    ref = oracle.dd_op(op, a, b)
    got = gpu.dd_vec_op(op, a, b)
    same = ((got.view(np.int64) == ref.view(np.int64)) | (np.isnan(got) & np.isnan(ref))).all(axis=1)
    p = _product_scale(op, a, b)
    if p is None:
        assert same.all(), (OPS[op], int((~same).sum()), a[~same][:3], b[~same][:3])
        return
    inside = (p >= 2.0 ** -969) & (p <= 2.0 ** 1000)
    assert inside.sum() > N // 2
    bad = inside & ~same
    assert not bad.any(), (OPS[op], int(bad.sum()), a[bad][:3], b[bad][:3], got[bad][:3], ref[bad][:3])
    # outside: the error term underflows; finite results agree closely
    out = ~inside & np.isfinite(ref).all(axis=1) & np.isfinite(got).all(axis=1) & (ref[:, 0] != 0)
    rel = np.abs((got[out, 0] - ref[out, 0]) + (got[out, 1] - ref[out, 1])) / np.abs(ref[out, 0])
    bound = 2.0 ** -16 if op in (2, 7, 8) else 2.0 ** -60
    assert rel.size == 0 or rel.max() <= bound, (OPS[op], rel.max())


def test_dd_signed_zero_and_lo_zero_cases(gpu, oracle):
    """Hand-picked operands: signed zeros, lo = 0, exact cancellation,
    sqrt(0), division of zero, powers of two."""
    vals = [(0.0, 0.0), (-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0), (2.0 ** -1000, 0.0),
            (2.0 ** 1000, 0.0), (1.0, 2.0 ** -53), (1.0, -2.0 ** -54), (3.0, 1e-17), (-3.0, -1e-17)]
    A = np.array([x for x in vals for _ in vals])
    B = np.array([y for _ in vals for y in vals])
    for op in sorted(OPS):
        a = np.abs(A) if op == 4 else A
        keep = np.ones(len(a), bool) if op not in (3, 9, 11) else (B[:, 0] != 0)
        ref = oracle.dd_op(op, a[keep], B[keep])
        got = gpu.dd_vec_op(op, a[keep], B[keep])
        # bit patterns, so that the signs of zeros are compared too
        same = (got.view(np.int64) == ref.view(np.int64)) | (np.isnan(got) & np.isnan(ref))
        assert same.all(), (OPS[op], a[keep][~same.all(axis=1)][:3], B[keep][~same.all(axis=1)][:3])

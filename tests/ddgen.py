"""Random dd test data (normalized pairs: |lo| < ulp(hi)/2)."""
import numpy as np


def dd_from_hi(hi, rng):
    hi = np.asarray(hi, dtype=np.float64)
    u = rng.uniform(-0.49, 0.49, size=hi.shape)
    lo = u * np.spacing(np.abs(hi))
    return np.stack([hi, lo], axis=-1)


def rand_dd(shape, rng, scale=1.0):
    return dd_from_hi(rng.standard_normal(shape) * scale, rng)


def rand_spd_cm(n, rng, shift=None):
    """Column-major (n*n, 2) symmetric positive definite dd matrix."""
    g = rng.standard_normal((n, n))
    a = g @ g.T / n + (shift if shift is not None else 1.0) * np.eye(n)
    a = (a + a.T) / 2
    lo = rng.uniform(-0.49, 0.49, (n, n)) * np.spacing(np.abs(a))
    lo = np.triu(lo) + np.triu(lo, 1).T
    m = np.stack([a, lo], axis=-1)
    return np.ascontiguousarray(m.transpose(1, 0, 2)).reshape(-1, 2)


def rand_lower_cm(n, rng):
    """Column-major lower-triangular dd matrix with a dominant diagonal."""
    a = np.tril(rng.standard_normal((n, n))) * 0.1 + np.diag(1.0 + rng.uniform(0, 1, n))
    d = dd_from_hi(a, rng)
    d[np.triu_indices(n, 1)] = 0.0
    return np.ascontiguousarray(d.transpose(1, 0, 2)).reshape(-1, 2)


def same(a, b):
    """Bitwise equality of dd arrays (-0 == +0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))

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


def _dd_normalize(hi, lo):
    """(hi, lo) -> the normalized pair of the exact sum hi + lo (quick two-sum)."""
    s = hi + lo
    with np.errstate(invalid="ignore"):
        e = lo - (s - hi)
    e = np.where(np.isfinite(s), e, 0.0)
    return np.stack([s, e], axis=-1)


def special_dd_pairs(n, rng):
    """n pairs (a, b) of dd operands covering the cases QD's operators
    branch or round on: N(0,1) values, exponents spread over 2^+-1000,
    lo = 0, signed zeros, exact and near cancellation (a = +-b, a = b(1+u)),
    powers of two with lo at half an ulp, and mixed scales whose product or
    quotient lands near the ends of the range."""
    k = n // 8
    parts_a, parts_b = [], []

    def spread(cnt, emin, emax):
        m = rng.uniform(1.0, 2.0, cnt) * rng.choice([-1.0, 1.0], cnt)
        e = rng.integers(emin, emax + 1, cnt)
        hi = np.ldexp(m, e)
        lo = rng.uniform(-0.5, 0.5, cnt) * np.spacing(np.abs(hi))
        return _dd_normalize(hi, lo)

    parts_a.append(rand_dd((k,), rng)); parts_b.append(rand_dd((k,), rng))
    # both operands anywhere in [2^-1000, 2^1000]: the comparison keeps
    # pairs whose results stay finite and normal (checked by the caller)
    parts_a.append(spread(k, -1000, 1000)); parts_b.append(spread(k, -1000, 1000))
    # one operand near either end, the other near 1
    big = spread(k, 990, 1010 - 20)
    parts_a.append(big); parts_b.append(spread(k, -8, 8))
    parts_a.append(spread(k, -1000, -960)); parts_b.append(spread(k, -8, 8))
    # lo = 0 on either or both sides
    a = rand_dd((k,), rng); b = rand_dd((k,), rng)
    a[: k // 2, 1] = 0.0; b[k // 4: 3 * k // 4, 1] = 0.0
    parts_a.append(a); parts_b.append(b)
    # signed zeros against values and zeros
    z = np.zeros((k, 2))
    sgn = rng.choice([-1.0, 1.0], (k, 2))
    z = z * sgn
    b = rand_dd((k,), rng)
    b[::3] = np.zeros(2) * rng.choice([-1.0, 1.0])
    parts_a.append(z); parts_b.append(b)
    # exact and near cancellation
    a = rand_dd((k,), rng)
    b = a.copy()
    b[::2] *= -1.0
    u = rng.integers(-3, 4, k) * 2.0 ** -100
    b[1::2, 1] += (np.abs(b[1::2, 0]) * u[1::2])
    parts_a.append(a); parts_b.append(_dd_normalize(b[:, 0], b[:, 1]))
    # powers of two with lo at +-half an ulp of hi (ties of the final add)
    e = rng.integers(-60, 60, n - 7 * k)
    hi = np.ldexp(1.0, e) * rng.choice([-1.0, 1.0], n - 7 * k)
    lo = 0.5 * np.spacing(np.abs(hi)) * rng.choice([-1.0, 1.0], n - 7 * k)
    lo = np.where(rng.uniform(size=n - 7 * k) < 0.5, lo, lo / 2)
    parts_a.append(np.stack([hi, lo], -1)); parts_b.append(rand_dd((n - 7 * k,), rng))
    A = np.concatenate(parts_a)
    B = np.concatenate(parts_b)
    p = rng.permutation(n)
    return np.ascontiguousarray(A[p]), np.ascontiguousarray(B[p])

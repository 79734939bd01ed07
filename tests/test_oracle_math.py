"""Pins the oracle's arithmetic and linear algebra against mathematics the
reference's results must satisfy: exact error-free transformations, dd
accuracy against 60-digit mpmath, QD's decimal reader, BLAS identities
computed exactly with fractions, Cholesky and eigenvalue invariants, the
Schur complement closed form B_ij = (X A_i Z^-1) . A_j, and agreement of the
three Schur formulas F1/F2/F3 that the reference chooses between."""
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from ddgen import rand_dd, rand_spd_cm

mpmath.mp.dps = 40
EPS = 2.0 ** -104


def F(x):
    return Fraction(float(x))


def ddF(v):
    return F(v[0]) + F(v[1])


def mp(v):
    return mpmath.mpf(float(v[0])) + mpmath.mpf(float(v[1]))


def test_two_sum_two_prod_exact(oracle):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(2000) * 10.0 ** rng.integers(-20, 20, 2000)
    b = rng.standard_normal(2000) * 10.0 ** rng.integers(-20, 20, 2000)
    A = np.stack([a, np.zeros_like(a)], -1)
    B = np.stack([b, np.zeros_like(b)], -1)
    s = oracle.dd_op(0, A, B)  # sloppy add of two doubles = TwoSum, exact
    p = oracle.dd_op(2, A, B)  # product of two doubles = TwoProduct, exact
    for i in range(2000):
        assert ddF(s[i]) == F(a[i]) + F(b[i])
        assert ddF(p[i]) == F(a[i]) * F(b[i])
        assert abs(s[i][1]) <= np.spacing(abs(s[i][0])) / 2


@pytest.mark.parametrize("op,fn,bound", [(0, lambda x, y: x + y, 4), (5, lambda x, y: x + y, 4), (1, lambda x, y: x - y, 4),
                                         (2, lambda x, y: x * y, 8), (7, lambda x, y: x * y, 8),
                                         (3, lambda x, y: x / y, 16)])
def test_dd_ops_accuracy(oracle, op, fn, bound):
    rng = np.random.default_rng(op)
    A = rand_dd((500,), rng)
    B = rand_dd((500,), rng)
    if op in (0, 5, 1):  # same signs keep the sloppy add's bound relative to the result
        B[:, 0] = np.abs(B[:, 0]) * np.sign(A[:, 0]) * (1 if op != 1 else -1)
        B[:, 1] = np.abs(B[:, 1]) * np.sign(B[:, 0]) * np.sign(B[:, 1] + 1e-300)
    R = oracle.dd_op(op, A, B)
    for i in range(500):
        exact = fn(mp(A[i]), mp(B[i]))
        assert abs(mp(R[i]) - exact) <= bound * EPS * abs(exact)


def test_dd_sqrt_accuracy(oracle):
    rng = np.random.default_rng(5)
    A = rand_dd((300,), rng)
    A[:, 0] = np.abs(A[:, 0])
    R = oracle.dd_op(4, A, A)
    for i in range(300):
        exact = mpmath.sqrt(mp(A[i]))
        assert abs(mp(R[i]) - exact) <= 8 * EPS * exact


def test_qd_reader_known_answers(oracle):
    rc, v = oracle.read_dd("3.14159265358979323846264338327950288")
    assert rc == 0 and abs(mp(v) - mpmath.pi) < 1e-31
    rc, v = oracle.read_dd("-11")
    assert rc == 0 and v[0] == -11.0 and v[1] == 0.0
    rc, v = oracle.read_dd("1.1")
    assert abs(mp(v) - mpmath.mpf("1.1")) < 2 * EPS
    rc, v = oracle.read_dd("2.5e-30")
    assert abs(mp(v) - mpmath.mpf("2.5e-30")) < 4 * EPS * 2.5e-30
    rc, v = oracle.read_dd("1.0E4")
    assert v[0] == 1e4 and v[1] == 0.0
    assert oracle.read_dd("1.2.3")[0] == -1
    assert oracle.read_dd("12x")[0] == -1


def test_rgemm_exact_rational(oracle):
    rng = np.random.default_rng(2)
    m, n, k = 5, 4, 7
    A = rand_dd((m * k,), rng)
    B = rand_dd((k * n,), rng)
    C = oracle.rgemm("N", m, n, k, [1.0, 0.0], A, m, B, k, [0.0, 0.0], np.zeros((m * n, 2)), m)
    Ct = oracle.rgemm("T", m, n, k, [1.0, 0.0], A, m, B.reshape(n, k, 2).transpose(1, 0, 2).reshape(-1, 2).copy(), n,
                      [0.0, 0.0], np.zeros((m * n, 2)), m)
    for j in range(n):
        for i in range(m):
            ex = sum(ddF(A[i + l * m]) * ddF(B[l + j * k]) for l in range(k))
            mag = sum(abs(ddF(A[i + l * m]) * ddF(B[l + j * k])) for l in range(k))
            assert abs(ddF(C[i + j * m]) - ex) <= 4 * k * F(EPS) * mag
    assert np.array_equal(C, Ct) or np.allclose(C[:, 0], Ct[:, 0], rtol=0, atol=0)


def test_rpotrf_reconstructs(oracle):
    rng = np.random.default_rng(3)
    n = 90  # crosses the 64-column block boundary of Rpotrf
    A = rand_spd_cm(n, rng)
    L, info = oracle.rpotrf(n, A, n)
    assert info == 0
    Lm = np.zeros((n, n), dtype=object)
    Am = np.zeros((n, n), dtype=object)
    for j in range(n):
        for i in range(n):
            Am[i, j] = mp(A[i + j * n])
            if i >= j:
                Lm[i, j] = mp(L[i + j * n])
            else:
                Lm[i, j] = mpmath.mpf(0)
    for (i, j) in [(0, 0), (5, 3), (70, 70), (89, 10), (64, 63), (89, 88)]:
        s = sum(Lm[i, t] * Lm[j, t] for t in range(min(i, j) + 1))
        assert abs(s - Am[i, j]) <= 64 * n * EPS * abs(Am[i, i] * Am[j, j]) ** 0.5
    # a non-positive pivot is reported at its 1-based index
    B = A.copy()
    B[30 + n * 30] = [-1.0, 0.0]
    _, info = oracle.rpotrf(n, B, n)
    assert info == 31


def test_rsteqr_eigenpairs(oracle):
    rng = np.random.default_rng(4)
    n = 12
    d = rand_dd((n,), rng)
    e = rand_dd((n,), rng)
    e[-1] = 0.0
    T = mpmath.matrix(n, n)
    for i in range(n):
        T[i, i] = mp(d[i])
        if i + 1 < n:
            T[i, i + 1] = T[i + 1, i] = mp(e[i])
    ev = sorted(mpmath.eigsy(T)[0])
    w, _, z, info = oracle.rsteqr(d, e)
    assert info == 0
    for i in range(n):
        assert abs(mp(w[i]) - ev[i]) < 1e-28 * max(1, abs(ev[i]))
    # columns of z are the eigenvectors in the same (ascending) order
    for c in range(n):
        v = mpmath.matrix([mp(z[r + c * n]) for r in range(n)])
        r = T * v - mp(w[c]) * v
        assert mpmath.norm(r) < 1e-27


def test_lanczos_estimate_tracks_lambda_min(oracle):
    rng = np.random.default_rng(6)
    n = 30
    g = rng.standard_normal((n, n))
    M = (g + g.T) / 2
    L = np.eye(n)
    Lc = np.stack([L.T.reshape(-1), np.zeros(n * n)], -1)
    Mc = np.stack([M.T.reshape(-1), np.zeros(n * n)], -1)
    est = oracle.lanczos(n, Lc, Mc)
    lmin = np.linalg.eigvalsh(M)[0]
    assert abs(est[0] - lmin) <= 1e-2 * abs(lmin) + 1e-4 + 1e-8


def _schur_closed_form(oracle, path):
    m, bs = oracle.load(path)
    rng = np.random.default_rng(7)
    Xs, Zs = [], []
    for n in bs:
        Xs.append(rand_spd_cm(n, rng))
        Zs.append(rand_spd_cm(n, rng))
    X = np.concatenate(Xs)
    Z = np.concatenate(Zs)
    return m, bs, X, Z


@pytest.mark.parametrize("builder,kw", [("random_dense", dict(n=22, m=8)), ("random_sparse", dict(n=24, m=9)),
                                        ("theta", dict(n=25, p=0.15)), ("maxcut_torus", dict(rows=5, cols=5))])
def test_schur_formulas_agree_with_closed_form(oracle, tmp_path, builder, kw):
    from sdpadd import gen
    path = getattr(gen, builder)(str(tmp_path / "s.dat-s"), **kw)
    m, bs, X, Z = _schur_closed_form(oracle, path)
    B = oracle.schur(m, X, Z)
    import ctypes
    c_dp = ctypes.POINTER(ctypes.c_double)
    oracle.lib.orc_schur_forced.argtypes = [c_dp, c_dp, c_dp, ctypes.c_int]
    Bf = []
    for f in (0, 1, 2):
        out = np.zeros((m * m, 2))
        oracle.lib.orc_schur_forced(X.ctypes.data_as(c_dp), Z.ctypes.data_as(c_dp), out.ctypes.data_as(c_dp), f)
        Bf.append(out)
    # dense A_i/A_j from the file, closed form in float64-exact rationals is too slow;
    # use 60-digit mpmath on the first block
    n = bs[0]
    import lzma  # noqa: F401  (keep import list explicit)
    A = _dense_constraints(path, m, n)
    Xm = _to_mp(X[: n * n], n)
    Zm = _to_mp(Z[: n * n], n)
    W = {}
    for (i, j) in [(0, 0), (1, 0), (m - 1, m - 2), (m // 2, 1)]:
        if i not in W:
            W[i] = Xm * A[i] * Zm
        ref = sum((W[i][r, c] * A[j][r, c] for r in range(n) for c in range(n)), mpmath.mpf(0))
        for Bx in [B] + Bf:
            v = mp(Bx[i + m * j])
            assert abs(v - ref) <= 1e-25 * (1 + abs(ref)), (i, j)
    # B symmetric bit for bit (the reference writes both triangles from one value)
    Bm = B.reshape(m, m, 2)
    assert np.array_equal(Bm, Bm.transpose(1, 0, 2))


def _to_mp(buf, n):
    M = mpmath.matrix(n, n)
    for j in range(n):
        for i in range(n):
            M[i, j] = mp(buf[i + j * n])
    return M


def _dense_constraints(path, m, n):
    A = [mpmath.matrix(n, n) for _ in range(m)]
    lines = open(path).read().splitlines()
    for ln in lines[5:]:
        p = ln.split()
        if len(p) != 5:
            continue
        k, blk, i, j = map(int, p[:4])
        if k == 0 or blk != 1:
            continue
        v = mpmath.mpf(p[4])
        A[k - 1][i - 1, j - 1] += v
        if i != j:
            A[k - 1][j - 1, i - 1] += v
    return A


def _mp_sym(A, n):
    T = mpmath.matrix(n, n)
    for j in range(n):
        for i in range(j, n):
            T[i, j] = T[j, i] = mp(A[i + n * j])
    return T


@pytest.mark.parametrize("n,scale", [(2, 1.0), (5, 1.0), (13, 1.0), (20, 1.0), (9, 1e-140), (9, 1e140)])
def test_rsyev_eigenvalues(oracle, n, scale):
    """Rsyev ('N', 'L'): ascending eigenvalues to dd accuracy against
    40-digit mpmath, including matrices scaled below rmin (~1e-130) and
    above rmax, which take the Rlascl path; only the lower triangle is read."""
    rng = np.random.default_rng(n)
    A = rand_dd((n * n,), rng)
    A[:, 0] *= scale
    A[:, 1] *= scale
    for j in range(n):
        for i in range(j + 1, n):
            A[j + n * i] = [7.0, 0.0]  # upper triangle must be ignored
    ev = sorted(mpmath.eigsy(_mp_sym(A, n))[0])
    w, info = oracle.rsyev(A)
    assert info == 0
    norm = max(abs(x) for x in ev)
    for i in range(n):
        assert abs(mp(w[i]) - ev[i]) < 1e-29 * n * norm, (i, w[i], ev[i])


@pytest.mark.parametrize("scale", [1.0, 1e150, 1e-90])
def test_rsterf_tridiagonal(oracle, scale):
    """Rsterf on a tridiagonal with a zero off-diagonal (two unreduced
    blocks), plain and scaled beyond ssfmax (~1e145) and below ssfmin
    (~6e-84): eigenvalues against mpmath."""
    rng = np.random.default_rng(11)
    n = 14
    d = rand_dd((n,), rng) * scale
    e = rand_dd((n,), rng) * scale
    e[6] = 0.0
    T = mpmath.matrix(n, n)
    for i in range(n):
        T[i, i] = mp(d[i])
        if i + 1 < n:
            T[i, i + 1] = T[i + 1, i] = mp(e[i])
    ev = sorted(mpmath.eigsy(T)[0])
    w, info = oracle.rsterf(d, e[: n - 1].copy())
    assert info == 0
    norm = max(abs(x) for x in ev)
    for i in range(n):
        assert abs(mp(w[i]) - ev[i]) < 1e-29 * n * norm


def test_qr_min_eigen_is_lambda_min_of_congruence(oracle):
    """The small-block step length's eigenvalue: lambda_min(L M L^T) for a
    lower-triangular L and symmetric M, against mpmath."""
    rng = np.random.default_rng(13)
    n = 11
    L = rand_dd((n * n,), rng)
    for j in range(n):
        for i in range(j):
            L[i + n * j] = 0.0
    M = rand_dd((n * n,), rng)
    for j in range(n):
        for i in range(j):
            M[i + n * j] = M[j + n * i]
    Lm = mpmath.matrix(n, n)
    for j in range(n):
        for i in range(n):
            Lm[i, j] = mp(L[i + n * j])
    S = Lm * _mp_sym(M, n) * Lm.T
    lmin = min(mpmath.eigsy(S)[0])
    got, info = oracle.qr_min_eigen(n, L, M)
    assert info == 0
    assert abs(mp(got) - lmin) < 1e-28 * mpmath.norm(S)


# ---- the remaining MPLAPACK routines of the path, against exact products --
# Each check is an identity of the routine's definition evaluated exactly
# (Fractions of the dd inputs) or at 60 digits; a transposed operand, the
# wrong triangle, a dropped term or a sign slip misses it by far more than
# the dd rounding bound.
def _exact_cm(buf, rows, cols):
    """column-major (rows*cols, 2) dd -> list of rows of Fractions."""
    return [[ddF(buf[i + rows * j]) for j in range(cols)] for i in range(rows)]


def _close(got, exact, scale):
    """|got - exact| <= 8 (terms) 2^-104 scale, elementwise."""
    for g, e, s in zip(got, exact, scale):
        assert abs(ddF(g) - e) <= Fraction(8) * Fraction(2) ** -104 * s + Fraction(1, 10 ** 300), (float(ddF(g) - e), float(s))


def _lower(n, rng):
    from ddgen import rand_lower_cm
    return rand_lower_cm(n, rng)


@pytest.mark.parametrize("tb", ["N", "T"])
def test_rgemv_exact(oracle, tb):
    """Rgemv 'N': y := alpha A x + beta y (mplapack/Rgemv.cpp:31-204)."""
    rng = np.random.default_rng(41)
    m, n = 13, 9
    A, x, y = rand_dd((m * n,), rng), rand_dd((n,), rng), rand_dd((m,), rng)
    al, be = rand_dd((1,), rng)[0], rand_dd((1,), rng)[0]
    got = oracle.rgemv(m, n, al, A, m, x, be, y)
    Ae = _exact_cm(A, m, n)
    a, b = ddF(al), ddF(be)
    exact = [a * sum(Ae[i][j] * ddF(x[j]) for j in range(n)) + b * ddF(y[i]) for i in range(m)]
    scale = [abs(a) * sum(abs(Ae[i][j] * ddF(x[j])) for j in range(n)) + abs(b * ddF(y[i])) for i in range(m)]
    _close(got, exact, [s * n for s in scale])


@pytest.mark.parametrize("trans", ["N", "T"])
def test_rtrmv_exact(oracle, trans):
    """Rtrmv L: x := L x ('N') or L^T x ('T') (mplapack/Rtrmv.cpp:31-234)."""
    rng = np.random.default_rng(42 + (trans == "T"))
    n = 14
    L, x = _lower(n, rng), rand_dd((n,), rng)
    got = oracle.rtrmv(trans, n, L, n, x)
    Le = _exact_cm(L, n, n)
    op = (lambda i, k: Le[i][k]) if trans == "N" else (lambda i, k: Le[k][i])
    exact = [sum(op(i, k) * ddF(x[k]) for k in range(n)) for i in range(n)]
    scale = [n * sum(abs(op(i, k) * ddF(x[k])) for k in range(n)) for i in range(n)]
    _close(got, exact, scale)
    # strictly upper entries of L are never read: garbage there changes nothing
    Lg = L.copy()
    for j in range(n):
        for i in range(j):
            Lg[i + n * j] = [1e300, 0.0]
    assert np.array_equal(oracle.rtrmv(trans, n, Lg, n, x), got)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_rtrsv_residual(oracle, trans):
    """Rtrsv L: solves L x = b ('N') or L^T x = b ('T'); the exact residual
    of the computed x is at dd rounding level (mplapack/Rtrsv.cpp:31-234)."""
    rng = np.random.default_rng(44 + (trans == "T"))
    n = 16
    L, b = _lower(n, rng), rand_dd((n,), rng)
    x = oracle.rtrsv(trans, n, L, n, b)
    Le = _exact_cm(L, n, n)
    op = (lambda i, k: Le[i][k]) if trans == "N" else (lambda i, k: Le[k][i])
    for i in range(n):
        r = sum(op(i, k) * ddF(x[k]) for k in range(n)) - ddF(b[i])
        s = sum(abs(op(i, k) * ddF(x[k])) for k in range(n)) + abs(ddF(b[i]))
        assert abs(r) <= 16 * n * Fraction(2) ** -104 * s, (i, float(r))


def test_rtrsm_both_sides_residual(oracle):
    """Rtrsm (R,L,T,N): X L^T = B, and (L,L,N,N): L X = B
    (mplapack/Rtrsm.cpp:128-142, 256-283), residuals at dd level."""
    rng = np.random.default_rng(46)
    m, n = 7, 11
    Ln, B = _lower(n, rng), rand_dd((m * n,), rng)
    X = oracle.rtrsm("RLTN", m, n, Ln, n, B, m)
    Le, Xe, Be = _exact_cm(Ln, n, n), _exact_cm(X, m, n), _exact_cm(B, m, n)
    for i in range(m):
        for j in range(n):
            r = sum(Xe[i][k] * Le[j][k] for k in range(n)) - Be[i][j]  # (X L^T)_ij
            s = sum(abs(Xe[i][k] * Le[j][k]) for k in range(n)) + abs(Be[i][j])
            assert abs(r) <= 16 * n * Fraction(2) ** -104 * s, ("RLTN", i, j, float(r))
    Lm = _lower(m, rng)
    X = oracle.rtrsm("LLNN", m, n, Lm, m, B, m)
    Le, Xe = _exact_cm(Lm, m, m), _exact_cm(X, m, n)
    for i in range(m):
        for j in range(n):
            r = sum(Le[i][k] * Xe[k][j] for k in range(m)) - Be[i][j]  # (L X)_ij
            s = sum(abs(Le[i][k] * Xe[k][j]) for k in range(m)) + abs(Be[i][j])
            assert abs(r) <= 16 * m * Fraction(2) ** -104 * s, ("LLNN", i, j, float(r))


def test_rtrmm_exact(oracle):
    """Rtrmm (L,L,T,N): B := L^T B (mplapack/Rtrmm.cpp:157-169)."""
    rng = np.random.default_rng(47)
    m, n = 9, 6
    L, B = _lower(m, rng), rand_dd((m * n,), rng)
    got = oracle.rtrmm(m, n, L, m, B, m)
    Le, Be = _exact_cm(L, m, m), _exact_cm(B, m, n)
    exact, scale = [], []
    for j in range(n):
        for i in range(m):
            exact.append(sum(Le[k][i] * Be[k][j] for k in range(m)))
            scale.append(m * sum(abs(Le[k][i] * Be[k][j]) for k in range(m)))
    _close(got, exact, scale)


def test_rsyrk_exact_lower_only(oracle):
    """Rsyrk (L,N): the lower triangle of C := alpha A A^T + beta C, the
    strict upper triangle untouched (mplapack/Rsyrk.cpp:158-176)."""
    rng = np.random.default_rng(48)
    n, k = 10, 7
    A, C = rand_dd((n * k,), rng), rand_dd((n * n,), rng)
    al, be = rand_dd((1,), rng)[0], rand_dd((1,), rng)[0]
    got = oracle.rsyrk(n, k, al, A, n, be, C, n)
    Ae, Ce = _exact_cm(A, n, k), _exact_cm(C, n, n)
    a, b = ddF(al), ddF(be)
    for j in range(n):
        for i in range(n):
            g = got[i + n * j]
            if i < j:
                assert np.array_equal(g, C[i + n * j]), (i, j)
                continue
            e = a * sum(Ae[i][l] * Ae[j][l] for l in range(k)) + b * Ce[i][j]
            s = k * (abs(a) * sum(abs(Ae[i][l] * Ae[j][l]) for l in range(k)) + abs(b * Ce[i][j]))
            assert abs(ddF(g) - e) <= 8 * Fraction(2) ** -104 * s + Fraction(1, 10 ** 300), (i, j)

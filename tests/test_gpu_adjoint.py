"""Adjoint of the interior-point solution (sdpadd_adjoint_setup/_adjoint).

Pins: the exact duality sensitivities at the optimum (dp*/db = y*,
dp*/dC = X*, dp*/dA_i = -y*_i X*, with p* = b.y*), dd central finite
differences of re-solved perturbed problems (step 2^-30 ~ 9.3e-10, exact in
dd), linearity and batch consistency of the vector-Jacobian products, and the
torch.autograd.Function over the same C ABI.
"""
import os
import tempfile

import mpmath
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 2.0 ** -30  # finite-difference step; perturbations are dyadic, so b + t*H*d is exact in dd


def _arrays(kind, **kw):
    from sdpadd import gen
    path = os.path.join(tempfile.gettempdir(), "adj_%s_%s.dat-s" % (kind, "_".join("%s%s" % i for i in kw.items())))
    {"dense": gen.random_dense, "sparse": gen.random_sparse, "theta": gen.theta,
     "maxcut": gen.maxcut_torus}[kind](path, **kw)
    return gen.read_arrays(path)


def _dd(v):
    v = np.asarray(v, dtype=np.float64)
    return np.stack([v, np.zeros_like(v)], axis=-1)


def _two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _perturbed(vals, d, t):
    """dd values vals + t*H*d, exact (vals float64, d dyadic)."""
    hi, lo = _two_sum(np.asarray(vals, dtype=np.float64), t * H * d)
    return np.stack([hi, lo], axis=-1)


def _build(dev, arr, b=None, cvals=None, avals=None):
    import sdpadd
    m, bs, b0, ents = arr
    out = []
    ci = ai = 0
    for k, l, r, c, v in ents:
        if k == 0:
            vv = _dd(v) if cvals is None else cvals[ci:ci + len(v)]
            ci += len(v)
        else:
            vv = _dd(v) if avals is None else avals[ai:ai + len(v)]
            ai += len(v)
        out.append((k, l, r, c, vv))
    return sdpadd.Problem.from_arrays(dev, m, bs, _dd(b0) if b is None else b, out)


def _flat(arr):
    m, bs, b, ents = arr
    cv = np.concatenate([v for k, l, r, c, v in ents if k == 0])
    av = np.concatenate([v for k, l, r, c, v in ents if k > 0])
    return b, cv, av


def _solve(prob, eps=1e-30):
    res = prob.solve({"epsilonStar": eps, "epsilonDash": eps})
    return res


def _mp(a):
    return mpmath.mpf(float(a[0])) + mpmath.mpf(float(a[1]))


def _phi(X, y, p, Xbar, ybar, pbar):
    """<Xbar, X> + <ybar, y> + pbar p in 200-bit arithmetic."""
    mpmath.mp.prec = 200
    s = mpmath.mpf(0)
    for xb, x in zip(Xbar, X):
        s += mpmath.mpf(float(xb)) * _mp(x)
    for yb, yy in zip(ybar, y):
        s += mpmath.mpf(float(yb)) * _mp(yy)
    return s + mpmath.mpf(float(pbar)) * _mp(p)


def _entry_values(arr, prob):
    """(A-gradient index, factor, stored (k, l, r, c)) per stored A value, as in
    sdpadd.autograd.Structure.a_map."""
    from sdpadd.autograd import Structure
    m, bs, b, ents = arr
    st = Structure(m, bs, ents)
    return st, st.a_map(prob)


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def _dd_sub(a, b):
    """(a - b) of dd arrays to about 1e-32 relative (hi/lo differences)."""
    return (a[..., 0] - b[..., 0]) + (a[..., 1] - b[..., 1])


# (kind, generator sizes, gap/feasibility tolerance of the solve, bounds on
# the centrality reached, on the relative deviation from the duality
# sensitivities and on the finite-difference mismatch).  The random instances
# stall near gap 1e-23 (the CPU oracle, which follows the reference's
# algorithm, as well), so they are solved to 1e-20; the 8x5 max-cut instance
# reaches 1e-25.  The Schur matrix at the final iterate has a condition
# number near 1/mu^2, beyond dd's 1e32 once mu < 1e-16: that limits the
# centring (to ~1e-6 at mu ~ 1e-25) and through it the agreement.
CASES = [("maxcut", {"rows": 8, "cols": 5}, 1e-25, 1e-4, 1e-16, 1e-4),
         ("maxcut", {"rows": 8, "cols": 5}, 1e-20, 1e-8, 1e-13, 1e-8),
         ("dense", {"n": 24, "m": 12}, 1e-20, 1e-8, 1e-13, 1e-8),
         ("sparse", {"n": 25, "m": 15, "density": 0.2}, 1e-20, 1e-8, 1e-13, 1e-8)]


@pytest.mark.parametrize("kind,kw,eps,tcen,tol,tfd", CASES)
def test_duality_sensitivities(gpu, kind, kw, eps, tcen, tol, tfd):
    """p* = b.y*: bbar = y*, Cbar = X*, Abar_i = -y*_i X* (per stored value)."""
    arr = _arrays(kind, **kw)
    prob = _build(gpu, arr)
    res = _solve(prob, eps)
    assert res["phase"] == "pdOPT"
    X, y, Z = prob.solution()
    prob.adjoint_setup()
    bbar, Cbar, Abar = prob.adjoint(pbar=np.array([[1.0, 0.0]]))
    ny, nx = np.max(np.abs(y[:, 0])), np.max(np.abs(X[:, 0]))
    eb = np.max(np.abs(_dd_sub(bbar[0], y))) / ny
    ec = np.max(np.abs(_dd_sub(Cbar[0], X))) / nx
    con, blk, row, col = prob.adjoint_entries()
    bs = arr[1]
    boff = np.concatenate([[0], np.cumsum([n * n for n in bs])])
    n_of = np.array(bs)[blk]
    xi = boff[blk] + row + n_of * col
    # expected -y_i X_rc, doubled for stored off-diagonal entries of sparse blocks
    exp_hi = -(y[con, 0] * X[xi, 0])
    twice = np.ones(len(con))
    cnt = {}
    for c_, b_ in zip(con, blk):
        cnt[(c_, b_)] = cnt.get((c_, b_), 0) + 1
    sparse = np.array([cnt[(c_, b_)] != bs[b_] ** 2 for c_, b_ in zip(con, blk)])
    twice[sparse & (row != col)] = 2.0
    ea = np.max(np.abs(Abar[0, :, 0] + Abar[0, :, 1] - twice * exp_hi)) / (ny * nx)
    assert prob.centrality < tcen
    assert eb < tol and ec < tol and ea < tol, (eb, ec, ea)
    prob.close()


# At gap 1e-25 a perturbed max-cut problem can end a step short of the gap
# (pFEAS), so the finite differences use the instances solved to 1e-20; the
# last case is the adjoint workload at its full size (n = 200, m = 300).
FD_CASES = CASES[1:] + [("sparse", {"n": 200, "m": 300, "density": 0.2}, 1e-20, 1e-8, 1e-13, 1e-8)]


@pytest.mark.parametrize("kind,kw,eps,tcen,tol,tfd", FD_CASES)
def test_finite_differences(gpu, kind, kw, eps, tcen, tol, tfd):
    """Directional derivatives of phi = <Xbar, X*> + <ybar, y*> + pbar p*
    along dyadic directions in b, C and the A_k agree with dd central
    differences of re-solved problems."""
    arr = _arrays(kind, **kw)
    rng = np.random.default_rng(7)
    m, bs, b0, ents = arr
    nn = sum(n * n for n in bs)
    b, cv, av = _flat(arr)
    prob = _build(gpu, arr)
    assert _solve(prob, eps)["phase"] == "pdOPT"
    # a symmetric random cotangent
    Xbar = np.zeros(nn)
    o = 0
    for n in bs:
        g = rng.standard_normal((n, n))
        Xbar[o:o + n * n] = ((g + g.T) / 2).reshape(-1)
        o += n * n
    ybar = rng.standard_normal(m)
    pbar = 0.7
    bbar, Cbar, Abar = prob.adjoint(Xbar=_dd(Xbar)[None], ybar=_dd(ybar)[None], pbar=np.array([[pbar, 0.0]]))
    st, (idx, fac) = _entry_values(arr, prob)
    dirs = {"b": np.round(rng.standard_normal(m) * 2 ** 10) / 2 ** 10,
            "C": np.round(rng.standard_normal(len(cv)) * 2 ** 10) / 2 ** 10,
            "A": np.round(rng.standard_normal(len(av)) * 2 ** 10) / 2 ** 10}
    # predicted directional derivatives
    c_entries = [(l, r, c) for k, l, r, c, v in ents if k == 0]
    boff = np.concatenate([[0], np.cumsum([n * n for n in bs])])
    gC = np.concatenate([(Cbar[0, boff[l] + r + bs[l] * c, 0] + Cbar[0, boff[l] + r + bs[l] * c, 1])
                         * np.where(r != c, 2.0, 1.0) for l, r, c in c_entries])
    gA = (Abar[0, idx, 0] + Abar[0, idx, 1]) * fac
    pred = {"b": float(np.dot(bbar[0, :, 0] + bbar[0, :, 1], dirs["b"])),
            "C": float(np.dot(gC, dirs["C"])), "A": float(np.dot(gA, dirs["A"]))}
    for what, d in dirs.items():
        vals = []
        for t in (1.0, -1.0):
            kwp = {}
            if what == "b":
                kwp["b"] = _perturbed(b, d, t)
            elif what == "C":
                kwp["cvals"] = _perturbed(cv, d, t)
            else:
                kwp["avals"] = _perturbed(av, d, t)
            pp = _build(gpu, arr, **kwp)
            r = _solve(pp, eps)
            assert r["phase"] == "pdOPT"
            X, y, _ = pp.solution()
            vals.append(_phi(X, y, r["objValDual"], Xbar, ybar, pbar))
            pp.close()
        fd = float((vals[0] - vals[1]) / (2 * H))
        # the finite difference itself carries (solve-to-solve noise ~ mu) / H
        err = abs(fd - pred[what]) / max(1.0, abs(fd))
        print("fd %s %s %s: fd %.17g adjoint %.17g rel %.3g" % (kind, kw, what, fd, pred[what], err))
        assert err < tfd, (what, fd, pred[what])
    prob.close()


def test_batch_and_linearity(gpu):
    """64 cotangents in one call equal one call each; the products are linear."""
    arr = _arrays("sparse", n=30, m=25, density=0.2)
    prob = _build(gpu, arr)
    assert _solve(prob, 1e-20)["phase"] == "pdOPT"
    m, bs = arr[0], arr[1]
    nn = sum(n * n for n in bs)
    rng = np.random.default_rng(3)
    K = 64
    Xb = _dd(rng.standard_normal((K, nn)))
    yb = _dd(rng.standard_normal((K, m)))
    pb = _dd(rng.standard_normal(K))
    bb, Cb, Ab = prob.adjoint(Xbar=Xb, ybar=yb, pbar=pb)
    for k in (0, 17, 63):
        b1, C1, A1 = prob.adjoint(Xbar=Xb[k:k + 1], ybar=yb[k:k + 1], pbar=pb[k:k + 1])
        assert np.array_equal(b1[0], bb[k]) and np.array_equal(C1[0], Cb[k]) and np.array_equal(A1[0], Ab[k])
    # linearity: the sum of cotangents 0 and 1
    Xs = _dd(Xb[0, :, 0] + Xb[1, :, 0])[None]
    ys = _dd(yb[0, :, 0] + yb[1, :, 0])[None]
    ps = _dd(pb[0:1, 0] + pb[1:2, 0])
    bs_, Cs_, As_ = prob.adjoint(Xbar=Xs, ybar=ys, pbar=ps)
    for full, one in ((bb, bs_), (Cb, Cs_), (Ab, As_)):
        ref = full[0, :, 0] + full[1, :, 0]
        assert _rel(one[0, :, 0], ref) < 1e-10
    prob.close()


def test_autograd_function(gpu):
    """torch.autograd over the C ABI: d(p*)/d(b, C, A) = (y*, X*, -y_i X*)."""
    import torch
    from sdpadd import autograd
    arr = _arrays("sparse", n=25, m=15, density=0.2)
    st, b, cv, av = autograd.Structure.from_arrays(*arr)
    # CPU tensors: the solve and the adjoint run on the GPU through the C ABI
    b, cv, av = (t.requires_grad_(True) for t in (b, cv, av))
    p, X, y = autograd.solve(b, cv, av, st, gpu, {"epsilonStar": 1e-20, "epsilonDash": 1e-20})
    p.backward()
    yh = y.detach().cpu().numpy()
    Xh = X.detach().cpu().numpy()
    assert _rel(b.grad.cpu().numpy(), yh) < 1e-13
    gc = []
    for k, l, r, c in st.c_groups:
        n, o = st.bs[l], st.boff[l]
        gc.append(Xh[o + r + n * c] * np.where(r != c, 2.0, 1.0))
    assert _rel(cv.grad.cpu().numpy(), np.concatenate(gc)) < 1e-13
    ga = []
    for k, l, r, c in st.a_groups:
        n, o = st.bs[l], st.boff[l]
        ga.append(-yh[k - 1] * Xh[o + r + n * c] * np.where(r != c, 2.0, 1.0))
    assert _rel(av.grad.cpu().numpy(), np.concatenate(ga)) < 1e-13
    # a second output through autograd: y* and X* cotangents run too
    b2, cv2, av2 = (t.detach().clone().requires_grad_(True) for t in (b, cv, av))
    p2, X2, y2 = autograd.solve(b2, cv2, av2, st, gpu, {"epsilonStar": 1e-20, "epsilonDash": 1e-20})
    (X2.sum() + (y2 * y2).sum()).backward()
    assert torch.isfinite(b2.grad).all() and torch.isfinite(av2.grad).all()


def test_random_sparse_200_batch64(gpu):
    """The adjoint workload at full size: n = 200, m = 300, A_i at 20%
    density, 64 N(0,1) cotangents in one call; the p* rows match y* and X*."""
    arr = _arrays("sparse", n=200, m=300, density=0.2)
    prob = _build(gpu, arr)
    assert _solve(prob, 1e-20)["phase"] == "pdOPT"
    X, y, _ = prob.solution()
    prob.adjoint_setup()
    assert prob.centrality < 1e-8
    m, nn, K = arr[0], 200 * 200, 64
    rng = np.random.default_rng(64)
    pb = np.zeros((K, 2))
    pb[0, 0] = 1.0  # cotangent 0: p* alone
    Xb = _dd(rng.standard_normal((K, nn)))
    yb = _dd(rng.standard_normal((K, m)))
    Xb[0] = 0.0
    yb[0] = 0.0
    bb, Cb, Ab = prob.adjoint(Xbar=Xb, ybar=yb, pbar=pb)
    assert np.all(np.isfinite(bb)) and np.all(np.isfinite(Cb)) and np.all(np.isfinite(Ab))
    assert np.max(np.abs(_dd_sub(bb[0], y))) / np.max(np.abs(y[:, 0])) < 1e-13
    assert np.max(np.abs(_dd_sub(Cb[0], X))) / np.max(np.abs(X[:, 0])) < 1e-13
    prob.close()


def test_autograd_rejects_non_optimal(gpu):
    """The derivative is of the optimality conditions: a solve that stops
    before pdOPT (here at an iteration limit) raises instead of returning
    gradients of a point that is not optimal."""
    import sdpadd
    from sdpadd import autograd
    arr = _arrays("sparse", n=25, m=15, density=0.2)
    st, b, cv, av = autograd.Structure.from_arrays(*arr)
    with pytest.raises(sdpadd.SdpaddError, match="pdOPT"):
        autograd.solve(b, cv, av, st, gpu, {"maxIteration": 3, "epsilonStar": 1e-20, "epsilonDash": 1e-20})
    # the same problem solved to the end goes through
    p, X, y = autograd.solve(b, cv, av, st, gpu, {"epsilonStar": 1e-20, "epsilonDash": 1e-20})
    assert np.isfinite(float(p))


def test_autograd_dd_values(gpu):
    """dd (hi, lo) inputs and outputs through SDPSolve: the outputs are the
    library's dd solution bit for bit, the high words equal the float64
    interface's outputs for the same values, the gradient of p* reaches both
    words of a dd input, and vjp_dd gives the full-dd products."""
    import torch
    from sdpadd import autograd
    arr = _arrays("sparse", n=25, m=15, density=0.2)
    st, b, cv, av = autograd.Structure.from_arrays(*arr)
    eps = {"epsilonStar": 1e-20, "epsilonDash": 1e-20}
    rng = np.random.default_rng(11)
    # dd inputs with non-zero low words (dyadic, far below the high words' ulp)
    d = np.round(rng.standard_normal(len(b)) * 2 ** 10) / 2 ** 10
    bdd = torch.from_numpy(_perturbed(b.numpy(), d, 2.0 ** -30)).requires_grad_(True)
    cdd = torch.from_numpy(_dd(cv.numpy())).requires_grad_(True)
    add = torch.from_numpy(_dd(av.numpy())).requires_grad_(True)
    p, X, y = autograd.solve(bdd, cdd, add, st, gpu, eps, dd=True)
    assert p.shape == (2,) and X.shape == (st.nn, 2) and y.shape == (st.m, 2)
    ref = _build(gpu, arr, b=bdd.detach().numpy())
    res = _solve(ref, 1e-20)
    Xr, yr, _ = ref.solution()
    assert np.array_equal(X.detach().numpy(), Xr) and np.array_equal(y.detach().numpy(), yr)
    assert np.array_equal(p.detach().numpy(), np.array(res["objValDual"]))
    p[0].backward()
    g = bdd.grad.numpy()
    assert np.array_equal(g[:, 0], g[:, 1])
    assert _rel(g[:, 0], yr[:, 0] + yr[:, 1]) < 1e-13
    assert cdd.grad.shape == (len(cv), 2) and add.grad.shape == (len(av), 2)
    # vjp_dd: the same cotangent (pbar = 1) in dd.  dp*/db = y* holds at the
    # optimum; at gap 1e-20 the products come from the Schur matrix at the
    # centred final iterate (condition ~ 1/mu^2, beyond dd's 1e32), which
    # bounds the agreement near 1e-15 (5.2e-16 measured here)
    gb, gc, ga, res2 = autograd.vjp_dd(bdd, cdd, add, st, gpu, pbar=[1.0, 0.0], params=eps)
    assert res2["phase"] == "pdOPT"
    e = np.max(np.abs(_dd_sub(gb, yr))) / np.max(np.abs(yr[:, 0]))
    print("dd vjp sparse 25x15: dp*/db vs y* %.3g" % e)
    assert e < 1e-14
    ref.close()
    # float64 inputs: the high words of the dd outputs
    p1, X1, y1 = autograd.solve(torch.from_numpy(bdd.detach().numpy()[:, 0].copy()), cv, av, st, gpu, eps)
    assert np.isfinite(float(p1)) and X1.shape == (st.nn,)


def test_theta_duality_sensitivities(gpu):
    """The headline workload (Lovasz theta of G(300, 0.1), m = 4551) solved
    to gap 1e-20: dp*/db = y* and dp*/dC = X* on the stored values of C
    through the torch Function, and the full-dd products of vjp_dd."""
    import torch
    from sdpadd import autograd
    arr = _arrays("theta", n=300, p=0.1)
    st, b, cv, av = autograd.Structure.from_arrays(*arr)
    assert st.m == 4551
    eps = {"epsilonStar": 1e-20, "epsilonDash": 1e-20}
    b, cv = b.requires_grad_(True), cv.requires_grad_(True)
    p, X, y = autograd.solve(b, cv, av, st, gpu, eps)
    p.backward()
    Xh, yh = X.detach().numpy(), y.detach().numpy()
    eb = _rel(b.grad.numpy(), yh)
    gc = []
    for k, l, r, c in st.c_groups:
        n, o = st.bs[l], st.boff[l]
        gc.append(Xh[o + r + n * c] * np.where(r != c, 2.0, 1.0))
    ec = _rel(cv.grad.numpy(), np.concatenate(gc))
    print("theta m=4551 gap 1e-20: dp*/db vs y* %.3g, dp*/dC vs X* %.3g" % (eb, ec))
    # measured 9.0e-14 and 7.2e-16: the Schur matrix at the centred final
    # iterate (condition ~ 1/mu^2) bounds the agreement, not the word size
    assert eb < 5e-13 and ec < 1e-13, (eb, ec)
    gb, gcd, ga, res = autograd.vjp_dd(b, cv, av, st, gpu, pbar=[1.0, 0.0], params=eps)
    prob = _build(gpu, arr)
    _solve(prob, 1e-20)
    Xd, yd, _ = prob.solution()
    prob.close()
    ebd = np.max(np.abs(_dd_sub(gb, yd))) / np.max(np.abs(yd[:, 0]))
    print("theta dd: dp*/db vs y* %.3g" % ebd)
    assert ebd < 5e-13

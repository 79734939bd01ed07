"""The oracle's sparse-Schur path (the reference's chordal branch).

Pins: the reference's own Schur decision ("SPARSE computation" / "DENSE
computation") in every committed SDPLIB result file; truss6, the instance the
reference solved on the sparse path, in test_oracle_sdplib (every printed
trace line, objectives, final mu); the minimum-degree ordering on graphs
whose elimination is worked out by hand; the factorisation against exact
rational arithmetic (R^T R = P B P^T with R the factor the pattern stores),
the solve by its residual, and the negative-pivot rule of Lal::getCholesky.
The ordering itself is not the reference's (Spooles 2.2 is absent): any
elimination order gives a valid factorisation, and the trajectories agree
with the reference's to the printed digits (truss6).
"""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle_lib import SDPLIB_PARAM, golden_sdplib

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sdplib")


def _names():
    return sorted(os.path.basename(p)[:-7] for p in glob.glob(os.path.join(GOLDEN, "*.result")))


@pytest.mark.parametrize("name", _names())
def test_schur_decision_matches_reference(oracle, name):
    """Chordal::ordering_bMat's thresholds give the reference's choice."""
    txt = open(os.path.join(GOLDEN, name + ".result")).read()
    ref = "sparse" if "SPARSE computation" in txt else "dense"
    assert ("DENSE computation" in txt) != (ref == "sparse")
    r = oracle.solve(golden_sdplib(name), SDPLIB_PARAM, max_iter=0)
    assert r["schur"] == ref


# graphs and their elimination worked out by hand: (adjacency, newToOld,
# factor-pattern entries).  Path: an end first, no fill.  Star: the leaves
# (degree 1) in index order, then the centre and the last leaf tie at
# degree 1 and the smaller index goes first.  5-cycle: vertex 0 first
# (fill 1-4), then 1 (fill 2-4), no fill after.  K4: index order, dense.
# Two components and an isolated vertex: the isolated one (degree 0) first.
HAND = [
    ([[1], [0, 2], [1, 3], [2, 4], [3]], [0, 1, 2, 3, 4], 9),
    ([[1, 2, 3, 4], [0], [0], [0], [0]], [1, 2, 3, 0, 4], 9),
    ([[1, 4], [0, 2], [1, 3], [2, 4], [3, 0]], [0, 1, 2, 3, 4], 12),
    ([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]], [0, 1, 2, 3], 10),
    ([[1], [0], [], [4], [3]], [2, 0, 1, 3, 4], 7),
]


@pytest.mark.parametrize("adj,order,nnz", HAND)
def test_min_degree_by_hand(oracle, adj, order, nnz):
    assert oracle.min_degree(adj) == (order, nnz)


def _fr(v):
    return Fraction(float(v[0])) + Fraction(float(v[1]))


def _sparse_spd(m, rng, density=0.15):
    """A symmetric positive definite matrix with a random sparsity pattern and
    small dyadic entries (exact in double)."""
    a = np.zeros((m, m))
    mask = np.triu(rng.random((m, m)) < density, 1)
    a[mask] = np.round(rng.standard_normal(mask.sum()) * 64) / 64
    a = a + a.T
    a += np.diag(np.abs(a).sum(axis=1) + 1 + np.round(rng.random(m) * 8) / 8)
    return a


def _dd(a):
    return np.stack([a.T.reshape(-1), np.zeros(a.size)], axis=-1)


@pytest.mark.parametrize("m,seed", [(12, 1), (30, 2), (45, 3)])
def test_sparse_factor_reconstructs(oracle, m, seed):
    """R^T R = P B P^T to dd accuracy, R the pattern with its stored diagonal
    1/sqrt(pivot) inverted; entries outside the pattern are exact zeros."""
    rng = np.random.default_rng(seed)
    a = _sparse_spd(m, rng)
    x, U, order, nnz = oracle.sparse_factor_solve(_dd(a), np.zeros((m, 2)))
    assert sorted(order) == list(range(m))
    R = [[Fraction(0)] * m for _ in range(m)]
    count = 0
    for c in range(m):
        for r in range(m):
            v = U[r + m * c]
            if v[0] != 0 or v[1] != 0:
                assert r <= c
                count += 1
                R[r][c] = 1 / _fr(v) if r == c else _fr(v)
    assert count <= nnz
    scale = Fraction(float(np.abs(a).max()))
    worst = Fraction(0)
    for i in range(m):
        for j in range(i, m):
            s = sum(R[k][i] * R[k][j] for k in range(min(i, j) + 1))
            worst = max(worst, abs(s - Fraction(float(a[order[i], order[j]]))))
    assert worst <= Fraction(1, 10 ** 28) * scale, float(worst / scale)


@pytest.mark.parametrize("m,seed", [(20, 4), (60, 5)])
def test_sparse_solve_residual(oracle, m, seed):
    """B x = g to dd accuracy (exact rational residual)."""
    rng = np.random.default_rng(seed)
    a = _sparse_spd(m, rng)
    g = np.round(rng.standard_normal(m) * 1024) / 1024
    x, U, order, nnz = oracle.sparse_factor_solve(_dd(a), np.stack([g, np.zeros(m)], axis=-1))
    xs = [_fr(v) for v in x]
    worst = max(abs(sum(Fraction(float(a[i, j])) * xs[j] for j in range(m)) - Fraction(float(g[i]))) for i in range(m))
    assert worst <= Fraction(1, 10 ** 27) * max(1, max(abs(Fraction(float(v))) for v in g)), float(worst)


def test_negative_pivot_becomes_zero(oracle):
    """Lal::getCholesky keeps going on a negative pivot: its stored inverse
    square root is 0, so the row's other entries become 0 and the rows after
    it are factored as if it had no effect on them."""
    a = np.array([[4.0, 0, 2], [0, -1.0, 1], [2, 1, 9.0]])
    x, U, order, nnz = oracle.sparse_factor_solve(_dd(a), np.zeros((3, 2)))
    m = 3
    pos = order.index(1)  # the row of constraint 1 (pivot -1)
    assert U[pos + m * pos][0] == 0 and U[pos + m * pos][1] == 0
    for c in range(pos + 1, m):
        assert U[pos + m * c][0] == 0 and U[pos + m * c][1] == 0
    # the other pivots are positive and stored as 1/sqrt
    for p in range(m):
        if p != pos:
            assert U[p + m * p][0] > 0


@pytest.mark.parametrize("name", ["truss1", "control1"])
def test_forced_sparse_agrees_with_dense(oracle, name):
    """On instances the reference runs dense, the sparse path (forced as the
    reference's debugging switch does) takes the same iterations to the same
    optimum."""
    d = oracle.solve(golden_sdplib(name), SDPLIB_PARAM, schur="dense")
    s = oracle.solve(golden_sdplib(name), SDPLIB_PARAM, schur="sparse")
    assert s["schur"] == "sparse" and d["schur"] == "dense"
    assert s["phase"] == d["phase"] == "pdOPT"
    assert abs(s["iterations"] - d["iterations"]) <= 1
    for k in ("objValPrimal", "objValDual"):
        a, b = s[k][0] + s[k][1], d[k][0] + d[k][1]
        assert abs(a - b) <= 1e-15 * max(1.0, abs(b)), (k, a, b)

"""Synthetic SDP instances in SDPA sparse format (.dat-s) for the five
workloads the Newton step is measured on.

Internal (SDPA-DD) convention: min C.X s.t. A_i.X = b_i, X psd, and its dual
max b.y s.t. sum y_i A_i + Z = C, Z psd.  The file stores F0 = -C, F_i = A_i
and c = b (sdpa_io.cpp:692).  Seeds are 20261203 + workload index.

  0  random feasible dense SDP: n=40, m=50, dense A_i
  1  max-cut on a 40x20 toroidal grid, weights +-1: n=m=800
  2  control-type blocks 110 and 55, m=1596
  3  random feasible SDP: n=200, m=300, A_i at 20% density
  4  Lovasz theta on G(300, 0.1): m = |E|+1
Sizes can be scaled down for tests through the keyword arguments.
"""
import os

import numpy as np

SEED0 = 20261203


def _haar(n, rng):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def _pd(n, rng, lo=0.5, hi=2.0):
    q = _haar(n, rng)
    return (q * rng.uniform(lo, hi, n)) @ q.T


class Writer:
    """Collects (k, block, i, j, value) entries; 1-based as in the format."""

    def __init__(self, m, blocks):
        self.m, self.blocks = m, list(blocks)
        self.b = np.zeros(m)
        self.lines = []

    def add_dense(self, k, blk, mat, tol=0.0):
        n = mat.shape[0]
        iu, ju = np.triu_indices(n)
        v = mat[iu, ju]
        keep = np.abs(v) > tol
        for i, j, x in zip(iu[keep], ju[keep], v[keep]):
            self.lines.append("%d %d %d %d %.17e" % (k, blk, i + 1, j + 1, x))

    def add(self, k, blk, i, j, x):
        self.lines.append("%d %d %d %d %.17e" % (k, blk, i + 1, j + 1, x))

    def write(self, path, title):
        with open(path, "w") as f:
            f.write('"%s"\n' % title)
            f.write("%d = mDIM\n%d = nBLOCK\n" % (self.m, len(self.blocks)))
            f.write(" ".join(str(b) for b in self.blocks) + " = bLOCKsTRUCT\n")
            f.write(" ".join("%.17e" % x for x in self.b) + "\n")
            f.write("\n".join(self.lines) + "\n")
        return path


def random_dense(path, n=40, m=50, seed=SEED0 + 0):
    """Workload 0: dense symmetric N(0,1) A_i, strictly feasible by construction."""
    rng = np.random.default_rng(seed)
    A = []
    for _ in range(m):
        g = np.triu(rng.standard_normal((n, n)))
        A.append(g + np.triu(g, 1).T)
    X0, Z0 = _pd(n, rng), _pd(n, rng)
    y0 = rng.standard_normal(m)
    w = Writer(m, [n])
    w.b = np.array([np.sum(a * X0) for a in A])
    C = sum(y * a for y, a in zip(y0, A)) + Z0
    w.add_dense(0, 1, -C)
    for k, a in enumerate(A):
        w.add_dense(k + 1, 1, a)
    return w.write(path, "random dense SDP n=%d m=%d seed=%d" % (n, m, seed))


def maxcut_torus(path, rows=40, cols=20, seed=SEED0 + 1):
    """Workload 1: C = -L/4 for a +-1 weighted toroidal grid, A_i = e_i e_i^T, b = 1."""
    rng = np.random.default_rng(seed)
    n = rows * cols
    L = np.zeros((n, n))
    node = lambda r, c: (r % rows) * cols + (c % cols)
    for r in range(rows):
        for c in range(cols):
            for (i, j) in ((node(r, c), node(r + 1, c)), (node(r, c), node(r, c + 1))):
                wgt = rng.choice([-1.0, 1.0])
                L[i, j] -= wgt
                L[j, i] -= wgt
                L[i, i] += wgt
                L[j, j] += wgt
    w = Writer(n, [n])
    w.b = np.ones(n)
    w.add_dense(0, 1, L / 4.0, tol=0.0)  # F0 = -C = L/4 (its nonzeros only)
    for i in range(n):
        w.add(i + 1, 1, i, i, 1.0)
    return w.write(path, "max-cut torus %dx%d seed=%d" % (rows, cols, seed))


def _sprand(shape, density, rng):
    a = rng.standard_normal(shape)
    return a * (rng.uniform(size=shape) < density)


def control(path, n1=110, n2=55, m=1596, density=0.10, seed=SEED0 + 2):
    """Workload 2: block 1 A_i = P_i^T E + E^T P_i, block 2 a rank-2 term."""
    rng = np.random.default_rng(seed)
    E = _sprand((n1, n1), density, rng)
    A1, A2 = [], []
    for _ in range(m):
        P = _sprand((n1, n1), density, rng)
        A1.append(P.T @ E + E.T @ P)
        u, v = _sprand(n2, density * 2, rng), _sprand(n2, density * 2, rng)
        if not u.any():
            u[rng.integers(n2)] = 1.0
        if not v.any():
            v[rng.integers(n2)] = 1.0
        A2.append(np.outer(u, v) + np.outer(v, u))
    X1, X2, Z1, Z2 = _pd(n1, rng), _pd(n2, rng), _pd(n1, rng), _pd(n2, rng)
    y0 = rng.standard_normal(m)
    w = Writer(m, [n1, n2])
    w.b = np.array([np.sum(a1 * X1) + np.sum(a2 * X2) for a1, a2 in zip(A1, A2)])
    C1 = sum(y * a for y, a in zip(y0, A1)) + Z1
    C2 = sum(y * a for y, a in zip(y0, A2)) + Z2
    w.add_dense(0, 1, -C1)
    w.add_dense(0, 2, -C2)
    for k in range(m):
        w.add_dense(k + 1, 1, A1[k])
        w.add_dense(k + 1, 2, A2[k])
    return w.write(path, "control-type n=%d+%d m=%d seed=%d" % (n1, n2, m, seed))


def random_sparse(path, n=200, m=300, density=0.20, seed=SEED0 + 3):
    """Workload 3: A_i with about `density` of their upper entries N(0,1)."""
    rng = np.random.default_rng(seed)
    A = []
    for _ in range(m):
        g = np.triu(_sprand((n, n), density, rng))
        A.append(g + np.triu(g, 1).T)
    X0, Z0 = _pd(n, rng), _pd(n, rng)
    y0 = rng.standard_normal(m)
    w = Writer(m, [n])
    w.b = np.array([np.sum(a * X0) for a in A])
    C = sum(y * a for y, a in zip(y0, A)) + Z0
    w.add_dense(0, 1, -C)
    for k, a in enumerate(A):
        w.add_dense(k + 1, 1, a)
    return w.write(path, "random sparse SDP n=%d m=%d seed=%d" % (n, m, seed))


def mixed_blocks(path, sdp=(30, 8, 3), lp=12, m=25, density=0.3, seed=SEED0 + 5):
    """SDP blocks of both step-length kinds (> 20: Lanczos, <= 20: Rsyev), a
    size-1 block and an LP block of `lp` entries; each A_i touches a random
    subset of blocks.  Strictly feasible by construction."""
    rng = np.random.default_rng(seed)
    nb = len(sdp)
    blocks = list(sdp) + [1, -lp]
    X0 = [_pd(n, rng) for n in sdp]
    Z0 = [_pd(n, rng) for n in sdp]
    x0, z0 = rng.uniform(0.5, 2.0, lp + 1), rng.uniform(0.5, 2.0, lp + 1)
    y0 = rng.standard_normal(m)
    w = Writer(m, blocks)
    Csdp = [z.copy() for z in Z0]
    clp = z0.copy()
    for k in range(m):
        bk = 0.0
        for l, n in enumerate(sdp):
            if rng.random() < 0.6 or l == k % nb:
                g = np.triu(_sprand((n, n), density, rng))
                a = g + np.triu(g, 1).T
                if not a.any():
                    a[0, 0] = 1.0
                w.add_dense(k + 1, l + 1, a)
                bk += np.sum(a * X0[l])
                Csdp[l] += y0[k] * a
        v = np.where(rng.random(lp + 1) < density, rng.standard_normal(lp + 1), 0.0)
        for i in np.nonzero(v)[0]:
            blk, row = (nb + 1, 0) if i == 0 else (nb + 2, i - 1)
            w.add(k + 1, blk, row, row, v[i])
        bk += np.sum(v * x0)
        clp += y0[k] * v
        w.b[k] = bk
    for l in range(nb):
        w.add_dense(0, l + 1, -Csdp[l])
    w.add(0, nb + 1, 0, 0, -clp[0])
    for i in range(lp):
        w.add(0, nb + 2, i, i, -clp[i + 1])
    return w.write(path, "mixed SDP/LP blocks %s + 1 + LP %d, m=%d seed=%d" % (sdp, lp, m, seed))


def theta(path, n=300, p=0.1, seed=SEED0 + 4):
    """Workload 4: Lovasz theta of G(n, p): C = -J, tr X = 1, X_ij = 0 on edges."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.uniform(size=iu.shape) < p
    edges = list(zip(iu[keep], ju[keep]))
    m = len(edges) + 1
    w = Writer(m, [n])
    w.b = np.zeros(m)
    w.b[0] = 1.0
    w.add_dense(0, 1, np.ones((n, n)))  # F0 = -C = J
    for i in range(n):
        w.add(1, 1, i, i, 1.0)
    for k, (i, j) in enumerate(edges):
        w.add(k + 2, 1, i, j, 1.0)
    return w.write(path, "theta G(%d,%.2f) m=%d seed=%d" % (n, p, m, seed))


def read_arrays(path):
    """Parses an SDPA sparse-format file into SDPA's internal form as float64
    arrays: (m, block_struct, b, entries) with entries a list of
    (k, l, rows, cols, vals), k = 0 for C (= -F0), 1..m for A_k, 0-based
    block l, rows and cols.  Entries of one (k, l) keep file order."""
    import re
    toks = []
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t[0] in '"*':
                continue
            toks.extend(x for x in re.split(r"[\s,{}()=]+", t) if x and _isnum(x))
    pos = 0

    def take():
        nonlocal pos
        pos += 1
        return toks[pos - 1]

    m = int(take())
    nblock = int(take())
    bs = [abs(int(take())) for _ in range(nblock)]
    b = np.array([float(take()) for _ in range(m)])
    groups = {}
    while pos + 5 <= len(toks):
        k, l, i, j, v = int(take()), int(take()), int(take()), int(take()), float(take())
        if k == 0:
            v = -v
        g = groups.setdefault((k, l - 1), ([], [], []))
        g[0].append(i - 1)
        g[1].append(j - 1)
        g[2].append(v)
    entries = [(k, l, np.array(r), np.array(c), np.array(v)) for (k, l), (r, c, v) in sorted(groups.items())]
    return m, bs, b, entries


def _isnum(t):
    try:
        float(t)
        return True
    except ValueError:
        return False


BUILDERS = {0: random_dense, 1: maxcut_torus, 2: control, 3: random_sparse, 4: theta}


def write_config(idx, path, **kw):
    return BUILDERS[idx](path, **kw)


def param_file(path, epsilon_star=None, epsilon_dash=None, gamma_star=None):
    """A param.sdpa with the reference defaults (sdpa_parts.cpp:160-170), some overridden."""
    vals = [200, 1.0e-30, 1.0e4, 2.0, -1.0e5, 1.0e5, 0.1, 0.3, 0.9, 1.0e-30]
    if epsilon_star is not None:
        vals[1] = epsilon_star
    if gamma_star is not None:
        vals[8] = gamma_star
    if epsilon_dash is not None:
        vals[9] = epsilon_dash
    with open(path, "w") as f:
        f.write("%d\tunsigned int maxIteration;\n" % vals[0])
        for v in vals[1:]:
            f.write("%r\n" % v)
    return path


if __name__ == "__main__":
    import sys
    idx = int(sys.argv[1])
    out = sys.argv[2] if len(sys.argv) > 2 else "config%d.dat-s" % idx
    print(write_config(idx, out))

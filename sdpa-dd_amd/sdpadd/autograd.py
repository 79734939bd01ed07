"""torch.autograd.Function over the C ABI: an SDP solve whose outputs
(p* = b.y*, X*, y*) are differentiable with respect to b, C and the stored
entries of the A_k.

The forward pass builds the problem through ``sdpadd_problem_create`` and
solves it with ``sdpadd_solve``; the backward pass calls
``sdpadd_adjoint_setup`` (the Schur matrix at the final iterate, factored
once) and ``sdpadd_adjoint`` (implicit differentiation of the central-path
equations).

Values are float64 tensors of shape (n,), entering the dd solver with a
zero low word, or dd tensors of shape (n, 2) holding (hi, lo) pairs; with
``dd=True`` the outputs are dd pairs too (p* of shape (2,), X* (nn, 2), y*
(m, 2)), so the solution keeps its ~32 digits through torch.  A dd pair
represents the value hi + lo; for autograd the high word carries the
derivative and the low word none, so a cotangent is read from the high
word and a dd input receives the same gradient in both words (the chain
rule of v = hi + lo).  Gradients are float64 in torch's type; the full-dd
vector-Jacobian products are ``vjp_dd``.

A solve that does not end pdOPT (or one of the phases passed as ``allow``)
raises: the implicit differentiation is of the optimality conditions.
"""
import numpy as np
import torch

from . import Problem, SdpaddError


class Structure:
    """Block sizes and the stored (row, col) pattern of C and of each A_k per
    block, in SDPA's internal form (C = -F0).  Values are supplied per call."""

    def __init__(self, m, block_struct, entries):
        self.m = int(m)
        self.bs = [int(n) for n in block_struct]
        self.boff = np.concatenate([[0], np.cumsum([n * n for n in self.bs])]).astype(np.int64)
        self.nn = int(self.boff[-1])
        groups = sorted((int(k), int(l), np.asarray(r, dtype=np.int32), np.asarray(c, dtype=np.int32))
                        for k, l, r, c, *_ in entries)
        self.c_groups = [g for g in groups if g[0] == 0]
        self.a_groups = [g for g in groups if g[0] > 0]
        self.nc = sum(len(g[2]) for g in self.c_groups)
        self.na = sum(len(g[2]) for g in self.a_groups)
        self._amap = None

    @classmethod
    def from_arrays(cls, m, block_struct, b, entries):
        """Structure plus the value tensors (b, cvals, avals) of parsed arrays
        (as returned by ``sdpadd.gen.read_arrays``)."""
        st = cls(m, block_struct, entries)
        vals = {(int(k), int(l)): np.asarray(v, dtype=np.float64) for k, l, r, c, v in entries}
        cv = np.concatenate([vals[(k, l)] for k, l, _, _ in st.c_groups]) if st.c_groups else np.zeros(0)
        av = np.concatenate([vals[(k, l)] for k, l, _, _ in st.a_groups]) if st.a_groups else np.zeros(0)
        t = lambda a: torch.tensor(np.asarray(a, dtype=np.float64))
        return st, t(b), t(cv), t(av)

    def problem(self, dev, b, cvals, avals):
        def dd(v):
            v = np.asarray(v, dtype=np.float64)
            return v if v.ndim == 2 else np.stack([v, np.zeros(len(v))], axis=-1)
        ents, pos = [], 0
        for k, l, r, c in self.c_groups:
            ents.append((k, l, r, c, dd(cvals[pos:pos + len(r)])))
            pos += len(r)
        pos = 0
        for k, l, r, c in self.a_groups:
            ents.append((k, l, r, c, dd(avals[pos:pos + len(r)])))
            pos += len(r)
        return Problem.from_arrays(dev, self.m, self.bs, dd(b), ents)

    def a_map(self, prob):
        """Index into the adjoint's A-gradient entries and factor per stored
        A value: sparse blocks keep their stored entries (already the
        derivative with respect to the stored value); a block the library
        holds dense (>= 20% stored) contributes 2 A_rc off the diagonal."""
        if self._amap is not None:
            return self._amap
        con, blk, row, col = prob.adjoint_entries()
        start = {}
        for q in range(len(con) - 1, -1, -1):
            start[(int(con[q]), int(blk[q]))] = q
        idx, fac = [], []
        for k, l, r, c in self.a_groups:
            n = self.bs[l]
            q0 = start[(k - 1, l)]
            if len(r) >= 0.2 * n * n:
                idx.append(q0 + r + n * c)
                fac.append(np.where(r != c, 2.0, 1.0))
            else:
                idx.append(q0 + np.arange(len(r)))
                fac.append(np.ones(len(r)))
        self._amap = (np.concatenate(idx), np.concatenate(fac)) if idx else (np.zeros(0, int), np.zeros(0))
        return self._amap


def _grads(prob, st, Xbar, ybar, pbar):
    """Full-dd gradients with respect to b (m, 2), the stored C values
    (nc, 2) and the stored A values (na, 2) for one dd cotangent."""
    bbar, Cbar, Abar = prob.adjoint(Xbar=Xbar[None], ybar=ybar[None], pbar=pbar[None])
    gc = []
    for _, l, r, c in st.c_groups:
        n, o = st.bs[l], st.boff[l]
        gc.append(Cbar[0, o + r + n * c] * np.where(r != c, 2.0, 1.0)[:, None])
    gc = np.concatenate(gc) if gc else np.zeros((0, 2))
    idx, fac = st.a_map(prob)
    return bbar[0], gc, Abar[0, idx] * fac[:, None]


class SDPSolve(torch.autograd.Function):
    """(p*, X*, y*) = SDPSolve.apply(b, cvals, avals, structure, device, params, dd, allow)."""

    @staticmethod
    def forward(ctx, b, cvals, avals, structure, dev, params, dd=False, allow=("pdOPT",)):
        prob = structure.problem(dev, b.detach().cpu().numpy(), cvals.detach().cpu().numpy(),
                                 avals.detach().cpu().numpy())
        res = prob.solve(params)
        if res["phase"] not in allow:
            prob.close()
            raise SdpaddError("SDPSolve: the solve ended %s after %d iterations (gap %.3g); the derivative is of "
                              "the optimality conditions and needs one of %s" %
                              (res["phase"], res["iterations"], res["relgap"][0], ", ".join(allow)))
        X, y, _ = prob.solution()
        ctx.prob, ctx.st, ctx.res, ctx.dd = prob, structure, res, dd
        ctx.shapes = (tuple(b.shape), tuple(cvals.shape), tuple(avals.shape))
        dv = b.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
        if dd:
            return t(np.array(res["objValDual"], dtype=np.float64)), t(X), t(y)
        return (torch.tensor(res["objValDual"][0], dtype=torch.float64, device=dv), t(X[:, 0].copy()),
                t(y[:, 0].copy()))

    @staticmethod
    def backward(ctx, gp, gX, gy):
        prob, st = ctx.prob, ctx.st

        def cot(g, n):
            # the cotangent of a value: the high word's (the low word carries no derivative)
            v = np.zeros(n) if g is None else g.detach().cpu().numpy().reshape(n, -1)[:, 0]
            return np.stack([v, np.zeros(n)], axis=-1)

        gb, gc, ga = _grads(prob, st, cot(gX, st.nn), cot(gy, st.m), cot(gp, 1)[0])
        dv = next(g.device for g in (gX, gy, gp) if g is not None)

        def out(g, shape):
            v = g[:, 0] + g[:, 1]  # float64 value of the dd gradient
            if len(shape) == 2:  # a dd input: v = hi + lo, the same gradient for both words
                v = np.stack([v, v], axis=-1)
            return torch.from_numpy(np.ascontiguousarray(v)).to(dv)

        bs, cs, as_ = ctx.shapes
        return out(gb, bs), out(gc, cs), out(ga, as_), None, None, None, None, None


def solve(b, cvals, avals, structure, dev, params=None, dd=False, allow=("pdOPT",)):
    """Differentiable (p*, X*, y*) of the SDP with these values (float64
    tensors, or dd (hi, lo) pair tensors of shape (n, 2)); dd=True returns
    the solution as dd pairs."""
    return SDPSolve.apply(b, cvals, avals, structure, dev, params, dd, tuple(allow))


def vjp_dd(b, cvals, avals, structure, dev, Xbar=None, ybar=None, pbar=None, params=None, allow=("pdOPT",)):
    """Full-dd vector-Jacobian products of (X*, y*, p*) for one dd cotangent
    (Xbar (nn, 2), ybar (m, 2), pbar (2,); None = zero): the gradients with
    respect to b (m, 2), the stored C values (nc, 2) and the stored A values
    (na, 2), each a dd pair.  Returns (gb, gc, ga, solve result)."""
    as_np = lambda t: t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)
    prob = structure.problem(dev, as_np(b), as_np(cvals), as_np(avals))
    try:
        res = prob.solve(params)
        if res["phase"] not in allow:
            raise SdpaddError("vjp_dd: the solve ended %s" % res["phase"])
        z = lambda n: np.zeros((n, 2))
        Xb = z(structure.nn) if Xbar is None else np.asarray(Xbar, dtype=np.float64).reshape(structure.nn, 2)
        yb = z(structure.m) if ybar is None else np.asarray(ybar, dtype=np.float64).reshape(structure.m, 2)
        pb = np.zeros(2) if pbar is None else np.asarray(pbar, dtype=np.float64).reshape(2)
        gb, gc, ga = _grads(prob, structure, Xb, yb, pb)
        return gb, gc, ga, res
    finally:
        prob.close()

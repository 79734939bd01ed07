"""torch.autograd.Function over the C ABI: an SDP solve whose outputs
(p* = b.y*, X*, y*) are differentiable with respect to b, C and the stored
entries of the A_k.

The forward pass builds the problem through ``sdpadd_problem_create`` and
solves it with ``sdpadd_solve``; the backward pass calls
``sdpadd_adjoint_setup`` (the Schur matrix at the final iterate, factored
once) and ``sdpadd_adjoint`` (implicit differentiation of the central-path
equations).  Tensors are float64; the values enter the dd solver exactly
(low word zero) and the high words of the dd results come back.  The full
dd gradients are available from ``Problem.adjoint``.
"""
import numpy as np
import torch

from . import Problem


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
        dd = lambda v: np.stack([np.asarray(v, dtype=np.float64), np.zeros(len(v))], axis=-1)
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


class SDPSolve(torch.autograd.Function):
    """(p*, X*, y*) = SDPSolve.apply(b, cvals, avals, structure, device, params)."""

    @staticmethod
    def forward(ctx, b, cvals, avals, structure, dev, params):
        prob = structure.problem(dev, b.detach().cpu().numpy(), cvals.detach().cpu().numpy(),
                                 avals.detach().cpu().numpy())
        res = prob.solve(params)
        X, y, _ = prob.solution()
        ctx.prob, ctx.st, ctx.res = prob, structure, res
        dv = b.device
        p = torch.tensor(res["objValDual"][0], dtype=torch.float64, device=dv)
        return p, torch.from_numpy(X[:, 0].copy()).to(dv), torch.from_numpy(y[:, 0].copy()).to(dv)

    @staticmethod
    def backward(ctx, gp, gX, gy):
        prob, st = ctx.prob, ctx.st
        dd = lambda t, n: np.stack([np.zeros(n) if t is None else t.detach().cpu().numpy().reshape(n),
                                    np.zeros(n)], axis=-1)[None]
        bbar, Cbar, Abar = prob.adjoint(Xbar=dd(gX, st.nn), ybar=dd(gy, st.m), pbar=dd(gp, 1)[0])
        dv = gX.device if gX is not None else gy.device
        gb = torch.from_numpy(bbar[0, :, 0].copy()).to(dv)
        gc = []
        for _, l, r, c in st.c_groups:
            n, o = st.bs[l], st.boff[l]
            gc.append(Cbar[0, o + r + n * c, 0] * np.where(r != c, 2.0, 1.0))
        gc = torch.from_numpy(np.concatenate(gc) if gc else np.zeros(0)).to(dv)
        idx, fac = st.a_map(prob)
        ga = torch.from_numpy(Abar[0, idx, 0] * fac).to(dv)
        return gb, gc, ga, None, None, None


def solve(b, cvals, avals, structure, dev, params=None):
    """Differentiable (p*, X*, y*) of the SDP with these values."""
    return SDPSolve.apply(b, cvals, avals, structure, dev, params)

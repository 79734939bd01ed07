"""ctypes/subprocess access to the CPU oracle (oracle/_build).  Used only by
tests, smoke() and the bench's CPU-baseline leg, as the checker."""
import ctypes
import json
import lzma
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
CLI = os.path.join(ROOT, "oracle", "_build", "sdpa_oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    def __init__(self):
        self.lib = ctypes.CDLL(LIB)
        L = self.lib
        c_dp = ctypes.POINTER(ctypes.c_double)
        L.orc_set_dot_threads.argtypes = [ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_min_degree.argtypes = [ctypes.c_int, ip, ip, ip]
        L.orc_min_degree.restype = ctypes.c_long
        L.orc_sparse_factor_solve.argtypes = [ctypes.c_int, c_dp, c_dp, c_dp, c_dp, ip]
        L.orc_sparse_factor_solve.restype = ctypes.c_long
        L.orc_dd_op.argtypes = [ctypes.c_int, ctypes.c_long, c_dp, c_dp, c_dp]
        L.orc_read_dd.argtypes = [ctypes.c_char_p, c_dp]
        L.orc_read_dd.restype = ctypes.c_int
        L.orc_Rgemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_long, ctypes.c_long, ctypes.c_long, c_dp, c_dp,
                                ctypes.c_long, c_dp, ctypes.c_long, c_dp, c_dp, ctypes.c_long]
        L.orc_Rsyrk_LN.argtypes = [ctypes.c_long, ctypes.c_long, c_dp, c_dp, ctypes.c_long, c_dp, c_dp, ctypes.c_long]
        L.orc_Rtrsm_RLTN.argtypes = [ctypes.c_long, ctypes.c_long, c_dp, c_dp, ctypes.c_long, c_dp, ctypes.c_long]
        L.orc_Rtrsm_LLNN.argtypes = [ctypes.c_long, ctypes.c_long, c_dp, c_dp, ctypes.c_long, c_dp, ctypes.c_long]
        L.orc_Rtrmm_LLTN.argtypes = [ctypes.c_long, ctypes.c_long, c_dp, c_dp, ctypes.c_long, c_dp, ctypes.c_long]
        L.orc_Rtrsv.argtypes = [ctypes.c_char, ctypes.c_long, c_dp, ctypes.c_long, c_dp]
        L.orc_Rtrmv.argtypes = [ctypes.c_char, ctypes.c_long, c_dp, ctypes.c_long, c_dp]
        L.orc_Rgemv_N.argtypes = [ctypes.c_long, ctypes.c_long, c_dp, c_dp, ctypes.c_long, c_dp, c_dp, c_dp]
        L.orc_Rdot.argtypes = [ctypes.c_long, c_dp, ctypes.c_long, c_dp, ctypes.c_long, c_dp]
        L.orc_Rpotrf_L.argtypes = [ctypes.c_long, c_dp, ctypes.c_long]
        L.orc_Rpotrf_L.restype = ctypes.c_long
        L.orc_Rsteqr_I.argtypes = [ctypes.c_long, c_dp, c_dp, c_dp, ctypes.c_long]
        L.orc_Rsteqr_I.restype = ctypes.c_long
        L.orc_lanczos.argtypes = [ctypes.c_long, c_dp, c_dp, c_dp]
        L.orc_Rsyev_NL.argtypes = [ctypes.c_long, c_dp, ctypes.c_long, c_dp]
        L.orc_Rsyev_NL.restype = ctypes.c_long
        L.orc_Rsterf.argtypes = [ctypes.c_long, c_dp, c_dp]
        L.orc_Rsterf.restype = ctypes.c_long
        L.orc_qr_min_eigen.argtypes = [ctypes.c_long, c_dp, c_dp, c_dp]
        L.orc_qr_min_eigen.restype = ctypes.c_long
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_load.restype = ctypes.c_int
        L.orc_formula.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.orc_schur.argtypes = [c_dp, c_dp, c_dp]

    def min_degree(self, adj):
        """The sparse-Schur path's minimum-degree ordering of an adjacency
        (list of neighbour lists): (newToOld, factor-pattern entries)."""
        m = len(adj)
        ptr = np.zeros(m + 1, dtype=np.int32)
        for i, a in enumerate(adj):
            ptr[i + 1] = ptr[i] + len(a)
        idx = np.array([j for a in adj for j in a] or [0], dtype=np.int32)
        order = np.zeros(max(1, m), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        nnz = self.lib.orc_min_degree(m, ptr.ctypes.data_as(ip), idx.ctypes.data_as(ip), order.ctypes.data_as(ip))
        return order[:m].tolist(), int(nnz)

    def sparse_factor_solve(self, B, g):
        """The sparse path on a symmetric dd matrix B (m*m x 2, column-major):
        (x, U, order, nnz) with U the factor pattern (new indices, upper,
        1/sqrt(pivot) on its diagonal) as an (m*m, 2) array."""
        m = g.shape[0]
        x = np.zeros((m, 2))
        U = np.zeros((m * m, 2))
        order = np.zeros(m, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        nnz = self.lib.orc_sparse_factor_solve(m, _dp(np.ascontiguousarray(B)), _dp(np.ascontiguousarray(g)), _dp(x),
                                               _dp(U), order.ctypes.data_as(ip))
        return x, U, order.tolist(), int(nnz)

    def set_dot_threads(self, t):
        self.lib.orc_set_dot_threads(int(t))

    def dd_op(self, op, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros_like(a)
        self.lib.orc_dd_op(op, a.shape[0], _dp(a), _dp(b), _dp(out))
        return out

    def read_dd(self, s):
        out = np.zeros(2)
        rc = self.lib.orc_read_dd(s.encode(), _dp(out))
        return rc, out

    # BLAS on column-major (len, 2) buffers ------------------------------
    def rgemm(self, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        C = np.array(C, dtype=np.float64, copy=True)
        self.lib.orc_Rgemm(b"N", transb.encode(), m, n, k, _dp(np.asarray(alpha, float)), _dp(np.ascontiguousarray(A)),
                           lda, _dp(np.ascontiguousarray(B)), ldb, _dp(np.asarray(beta, float)), _dp(C), ldc)
        return C

    def rsyrk(self, n, k, alpha, A, lda, beta, C, ldc):
        C = np.array(C, dtype=np.float64, copy=True)
        self.lib.orc_Rsyrk_LN(n, k, _dp(np.asarray(alpha, float)), _dp(np.ascontiguousarray(A)), lda,
                              _dp(np.asarray(beta, float)), _dp(C), ldc)
        return C

    def rtrsm(self, kind, m, n, A, lda, B, ldb):
        B = np.array(B, dtype=np.float64, copy=True)
        f = self.lib.orc_Rtrsm_RLTN if kind == "RLTN" else self.lib.orc_Rtrsm_LLNN
        f(m, n, _dp(np.array([1.0, 0.0])), _dp(np.ascontiguousarray(A)), lda, _dp(B), ldb)
        return B

    def rtrmm(self, m, n, A, lda, B, ldb):
        B = np.array(B, dtype=np.float64, copy=True)
        self.lib.orc_Rtrmm_LLTN(m, n, _dp(np.array([1.0, 0.0])), _dp(np.ascontiguousarray(A)), lda, _dp(B), ldb)
        return B

    def rtrsv(self, trans, n, A, lda, x):
        x = np.array(x, dtype=np.float64, copy=True)
        self.lib.orc_Rtrsv(trans.encode(), n, _dp(np.ascontiguousarray(A)), lda, _dp(x))
        return x

    def rtrmv(self, trans, n, A, lda, x):
        x = np.array(x, dtype=np.float64, copy=True)
        self.lib.orc_Rtrmv(trans.encode(), n, _dp(np.ascontiguousarray(A)), lda, _dp(x))
        return x

    def rgemv(self, m, n, alpha, A, lda, x, beta, y):
        y = np.array(y, dtype=np.float64, copy=True)
        self.lib.orc_Rgemv_N(m, n, _dp(np.asarray(alpha, float)), _dp(np.ascontiguousarray(A)), lda,
                             _dp(np.ascontiguousarray(x)), _dp(np.asarray(beta, float)), _dp(y))
        return y

    def rdot(self, x, y, incx=1, incy=1, n=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if n is None:
            n = x.shape[0]
        out = np.zeros(2)
        self.lib.orc_Rdot(n, _dp(x), incx, _dp(y), incy, _dp(out))
        return out

    def rpotrf(self, n, A, lda):
        A = np.array(A, dtype=np.float64, copy=True)
        info = self.lib.orc_Rpotrf_L(n, _dp(A), lda)
        return A, info

    def rsteqr(self, d, e):
        n = d.shape[0]
        d = np.array(d, dtype=np.float64, copy=True)
        e = np.array(e, dtype=np.float64, copy=True)
        z = np.zeros((n * n, 2))
        info = self.lib.orc_Rsteqr_I(n, _dp(d), _dp(e), _dp(z), n)
        return d, e, z, info

    def lanczos(self, n, L, M):
        out = np.zeros(2)
        self.lib.orc_lanczos(n, _dp(np.ascontiguousarray(L)), _dp(np.ascontiguousarray(M)), _dp(out))
        return out

    def rsyev(self, A):
        """Eigenvalues (ascending) from the lower triangle of the n x n dd
        matrix A (column-major, shape (n*n, 2)); returns (w, info)."""
        a = np.array(A, dtype=np.float64)
        n = int(round(np.sqrt(a.shape[0])))
        w = np.zeros((n, 2))
        info = self.lib.orc_Rsyev_NL(n, _dp(a), n, _dp(w))
        return w, info

    def rsterf(self, d, e):
        d = np.array(d, dtype=np.float64)
        e = np.array(e, dtype=np.float64)
        info = self.lib.orc_Rsterf(d.shape[0], _dp(d), _dp(e))
        return d, info

    def qr_min_eigen(self, n, L, M):
        out = np.zeros(2)
        info = self.lib.orc_qr_min_eigen(n, _dp(np.ascontiguousarray(L)), _dp(np.ascontiguousarray(M)), _dp(out))
        return out, info

    def load(self, path):
        """(m, SDP block sizes); self.n_lp holds the LP entry count."""
        m, nb = ctypes.c_int(), ctypes.c_int()
        bs = (ctypes.c_int * 4096)()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.orc_load(path.encode(), ctypes.byref(m), ctypes.byref(nb), bs, 4096, err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self.n_lp = -bs[nb.value]
        return m.value, [bs[i] for i in range(nb.value)]

    def formula(self, m, nb):
        out = (ctypes.c_int * (m * nb))()
        self.lib.orc_formula(out)
        return np.array(out[:], dtype=np.int32)

    def schur(self, m, X, Zinv):
        B = np.zeros((m * m, 2))
        self.lib.orc_schur(_dp(np.ascontiguousarray(X)), _dp(np.ascontiguousarray(Zinv)), _dp(B))
        return B

    # whole solver -------------------------------------------------------
    def solve(self, data, param=None, dot_threads=1, max_iter=None, dump=False, timeout=3000, init=None,
              init_dense=False, schur=None):
        """schur: None (the reference's rule), "dense" or "sparse" (forced)."""
        cmd = [CLI, data]
        if param:
            cmd.append(param)
        if schur:
            cmd += ["--schur", schur]
        if init:
            cmd += ["--init", init] + (["--init-dense"] if init_dense else [])
        cmd += ["--dot-threads", str(dot_threads)]
        if max_iter is not None:
            cmd += ["--max-iter", str(max_iter)]
        if dump:
            cmd.append("--dump-solution")
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        out = json.loads(r.stdout)
        # wall-clock stamps differ from run to run: keep them out of the trace
        out["iter_seconds"] = [rec.pop("seconds", None) for rec in out.get("trace", [])]
        return out


def golden_sdplib(name):
    """Decompresses tests/golden/sdplib/<name>.dat-s.xz to a temp file."""
    src = os.path.join(GOLDEN, "sdplib", name + ".dat-s.xz")
    d = os.path.join(tempfile.gettempdir(), "sdpadd_golden")
    os.makedirs(d, exist_ok=True)
    dst = os.path.join(d, name + ".dat-s")
    if not os.path.exists(dst):
        with lzma.open(src, "rb") as f, open(dst, "wb") as g:
            g.write(f.read())
    return dst


SDPLIB_PARAM = os.path.join(GOLDEN, "sdplib", "param.sdpa")

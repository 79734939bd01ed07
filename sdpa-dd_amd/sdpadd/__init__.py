"""Python front end of the MI355X double-double Newton-step library.

The compute path is the C ABI of ``libsdpadd_hip.so`` (``include/sdpadd.h``);
this module only loads it with ctypes and moves numpy arrays in and out.
A dd array is a float64 array whose last axis has length 2 (hi, lo), the
memory layout of QD's ``dd_real``.  Matrices are column-major.

There is no CPU fallback: if the library is missing or no GPU is present,
calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libsdpadd_hip.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "sdpa_dd_gpu")

PHASES = ["noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF", "pFEAS_dINF", "pINF_dFEAS", "pdOPT", "pUNBD", "dUNBD"]

_lib = None


# Chunks of the static dot-product partition (the reference's Rdot_omp with
# this many OpenMP threads); the library's default.
DEFAULT_DOT_PARTITION = 64


class SdpaddError(RuntimeError):
    pass


def load_library(path=None):
    """Loads libsdpadd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise SdpaddError("libsdpadd_hip.so not built (run sdpa-dd_amd/build.py): %s" % path)
    lib = ctypes.CDLL(path)
    c_dp = ctypes.POINTER(ctypes.c_double)
    c_i64 = ctypes.c_int64
    vp = ctypes.c_void_p
    sig = {
        "sdpadd_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "sdpadd_destroy": (ctypes.c_int, [vp]),
        "sdpadd_last_error": (ctypes.c_char_p, [vp]),
        "sdpadd_set_dot_partition": (ctypes.c_int, [vp, ctypes.c_int]),
        "sdpadd_set_path": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "sdpadd_malloc": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
        "sdpadd_set_tile_policy": (ctypes.c_int, [vp, ctypes.c_int]),
        "sdpadd_set_wait_bound": (ctypes.c_int, [vp, ctypes.c_uint]),
        "sdpadd_set_panel_wait_bound": (ctypes.c_int, [vp, ctypes.c_uint]),
        "sdpadd_dd_vec_op": (ctypes.c_int, [vp, ctypes.c_int, c_i64, vp, vp, vp]),
        "sdpadd_stats_enable": (ctypes.c_int, [vp, ctypes.c_int]),
        "sdpadd_stats_reset": (ctypes.c_int, [vp]),
        "sdpadd_stats_filter": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sdpadd_stats_get": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(c_i64), c_dp, c_dp]),
        "sdpadd_free": (ctypes.c_int, [vp, vp]),
        "sdpadd_upload": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t]),
        "sdpadd_download": (ctypes.c_int, [vp, vp, vp, ctypes.c_size_t]),
        "sdpadd_Rgemm": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_i64, c_dp, vp, c_i64, vp, c_i64,
                                        c_dp, vp, c_i64]),
        "sdpadd_ozaki_gemm": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_i64, ctypes.c_int, vp, c_i64,
                                             vp, c_i64,
                                             ctypes.c_int, vp, c_i64, ctypes.c_int, ctypes.c_int]),
        "sdpadd_Rsyrk": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_dp, vp, c_i64, c_dp, vp, c_i64]),
        "sdpadd_Rtrsm": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_dp,
                                        vp, c_i64, vp, c_i64]),
        "sdpadd_Rtrmm": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, c_i64, c_i64, c_dp,
                                        vp, c_i64, vp, c_i64]),
        "sdpadd_Rtrmm_noalloc": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, c_i64, c_i64,
                                                c_dp, vp, c_i64, vp, c_i64]),
        "sdpadd_Rtrsv": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, c_i64, vp, c_i64, vp, c_i64]),
        "sdpadd_ozaki_trsv": (ctypes.c_int, [vp, ctypes.c_char, c_i64, vp, c_i64, vp]),
        "sdpadd_ozaki_matvec": (ctypes.c_int, [vp, ctypes.c_char, c_i64, c_i64, c_dp, vp, c_i64, vp, c_dp, vp]),
        "sdpadd_ozaki_dot": (ctypes.c_int, [vp, c_i64, vp, vp, c_dp]),
        "sdpadd_Rtrmv": (ctypes.c_int, [vp, ctypes.c_char, ctypes.c_char, ctypes.c_char, c_i64, vp, c_i64, vp, c_i64]),
        "sdpadd_Rgemv": (ctypes.c_int, [vp, ctypes.c_char, c_i64, c_i64, c_dp, vp, c_i64, vp, c_i64, c_dp, vp, c_i64]),
        "sdpadd_Rdot": (ctypes.c_int, [vp, c_i64, vp, c_i64, vp, c_i64, c_dp]),
        "sdpadd_Raxpy": (ctypes.c_int, [vp, c_i64, c_dp, vp, c_i64, vp, c_i64]),
        "sdpadd_Rscal": (ctypes.c_int, [vp, c_i64, c_dp, vp, c_i64]),
        "sdpadd_Rcopy": (ctypes.c_int, [vp, c_i64, vp, c_i64, vp, c_i64]),
        "sdpadd_Rpotrf": (ctypes.c_int, [vp, ctypes.c_char, c_i64, vp, c_i64, ctypes.POINTER(c_i64)]),
        "sdpadd_Rsteqr": (ctypes.c_int, [vp, ctypes.c_char, c_i64, c_dp, c_dp, c_dp, c_i64, c_dp, ctypes.POINTER(c_i64)]),
        "sdpadd_lanczos_min_eig": (ctypes.c_int, [vp, c_i64, vp, vp, c_dp]),
        "sdpadd_small_min_eig": (ctypes.c_int, [vp, c_i64, vp, vp, c_dp, ctypes.POINTER(ctypes.c_int)]),
        "sdpadd_read_initial_point": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int]),
        "sdpadd_set_initial_point": (ctypes.c_int, [vp, c_dp, c_dp, c_dp]),
        "sdpadd_clear_initial_point": (ctypes.c_int, [vp]),
        "sdpadd_write_solution": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "sdpadd_inv_chol": (ctypes.c_int, [vp, c_i64, vp, vp, vp, vp]),
        "sdpadd_problem_load": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "sdpadd_problem_create": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(vp)]),
        "sdpadd_problem_set_entries": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, c_i64, ctypes.POINTER(ctypes.c_int),
                                                      ctypes.POINTER(ctypes.c_int), c_dp]),
        "sdpadd_problem_set_b": (ctypes.c_int, [vp, c_dp]),
        "sdpadd_problem_finalize": (ctypes.c_int, [vp]),
        "sdpadd_problem_destroy": (ctypes.c_int, [vp]),
        "sdpadd_problem_dims": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
        "sdpadd_problem_formula": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
        "sdpadd_schur_build": (ctypes.c_int, [vp, vp, vp, vp]),
        "sdpadd_problem_schur_ddflops": (ctypes.c_int, [vp, c_dp]),
        "sdpadd_problem_set_schur_mode": (ctypes.c_int, [vp, ctypes.c_int]),
        "sdpadd_problem_schur_info": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(c_i64),
                                                     ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_int)]),
        "sdpadd_inner_A": (ctypes.c_int, [vp, vp, vp]),
        "sdpadd_plus_A": (ctypes.c_int, [vp, vp, vp]),
        "sdpadd_read_params": (ctypes.c_int, [ctypes.c_char_p, vp]),
        "sdpadd_default_params": (None, [vp]),
        "sdpadd_solve": (ctypes.c_int, [vp, vp, vp]),
        "sdpadd_trace": (ctypes.c_int, [vp, c_dp, ctypes.c_int]),
        "sdpadd_solution": (ctypes.c_int, [vp, c_dp, c_dp, c_dp]),
        "sdpadd_adjoint_setup": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(c_i64), c_dp]),
        "sdpadd_adjoint_entries": (ctypes.c_int, [vp] + [ctypes.POINTER(ctypes.c_int)] * 4),
        "sdpadd_adjoint": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, c_dp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class Params(ctypes.Structure):
    _fields_ = [("maxIteration", ctypes.c_int)] + [(n, ctypes.c_double) for n in (
        "epsilonStar", "lambdaStar", "omegaStar", "lowerBound", "upperBound", "betaStar", "betaBar", "gammaStar",
        "epsilonDash")]


class Result(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("phase", ctypes.c_int), ("phase_internal", ctypes.c_int)] + [
        (n, ctypes.c_double * 2) for n in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat", "relgap")
    ] + [(n, ctypes.c_double) for n in ("seconds", "t_schur", "t_chol", "t_lanczos", "t_inverse")] + [
        ("refine_sweeps", ctypes.c_int)]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dd_scalar(hi, lo=0.0):
    return np.array([hi, lo], dtype=np.float64)


class Device:
    """A library handle (one GPU, one stream)."""

    def __init__(self, device=0, dot_partition=None):
        self.lib = load_library()
        self.h = ctypes.c_void_p()
        rc = self.lib.sdpadd_create(device, ctypes.byref(self.h))
        if rc != 0:
            raise SdpaddError("sdpadd_create failed (%d): no usable GPU" % rc)
        self.set_dot_partition(DEFAULT_DOT_PARTITION if dot_partition is None else dot_partition)

    def close(self):
        if self.h:
            self.lib.sdpadd_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def check(self, rc, what):
        if rc < 0:
            raise SdpaddError("%s failed (%d): %s" % (what, rc, self.lib.sdpadd_last_error(self.h).decode()))
        return rc

    def set_path(self, path, slices=16, refine_steps=2):
        """path 0: plain dd (the reference's operation order); 1: Ozaki INT8-MFMA
        Cholesky trailing updates; 2: as 1 plus the n x n GEMMs."""
        self.check(self.lib.sdpadd_set_path(self.h, path, slices, refine_steps), "set_path")

    def set_tile_policy(self, policy):
        """0: GEMM/Ozaki tile shape by grid size; 1: always 32 x 32; 2: always 64 x 64."""
        self.check(self.lib.sdpadd_set_tile_policy(self.h, int(policy)), "set_tile_policy")

    def set_wait_bound(self, spins):
        """Bound on each spin-wait of the cross-workgroup triangular solves
        (0 = default); a wait that runs out hands the solve to the
        single-workgroup recovery (same results)."""
        self.check(self.lib.sdpadd_set_wait_bound(self.h, int(spins)), "set_wait_bound")

    def set_panel_wait_bound(self, polls):
        """Bound on a Cholesky panel workgroup's consecutive empty polls for
        its diagonal block's columns (0 = default); one that runs out makes
        Rpotrf report info -2 and a solve fail with an error."""
        self.check(self.lib.sdpadd_set_panel_wait_bound(self.h, int(polls)), "set_panel_wait_bound")

    def dd_vec_op(self, op, a, b):
        """QD operator `op` (0..11, see sdpadd_dd_vec_op) elementwise on (n, 2) dd arrays."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        da, db, do = self.to_device(a), self.to_device(b), self.empty(a.shape)
        self.check(self.lib.sdpadd_dd_vec_op(self.h, int(op), a.shape[0], da.ptr, db.ptr, do.ptr), "dd_vec_op")
        return do.get()

    def set_dot_partition(self, t):
        self.dot_partition = t
        self.check(self.lib.sdpadd_set_dot_partition(self.h, t), "set_dot_partition")

    def stats(self, on=True, families=None):
        """Per-family kernel timing on/off (and reset); families: an
        iterable of family names to time, None for all."""
        self.check(self.lib.sdpadd_stats_enable(self.h, 1 if on else 0), "stats_enable")
        self.check(self.lib.sdpadd_stats_filter(self.h, ",".join(families).encode() if families else None),
                   "stats_filter")
        self.check(self.lib.sdpadd_stats_reset(self.h), "stats_reset")

    def stats_get(self, family):
        n = ctypes.c_int64()
        ms = np.zeros(1)
        fl = np.zeros(1)
        self.check(self.lib.sdpadd_stats_get(self.h, family.encode(), ctypes.byref(n), _dp(ms), _dp(fl)), "stats_get")
        return {"launches": n.value, "ms": float(ms[0]), "ddflops": float(fl[0])}

    # device memory -----------------------------------------------------
    def to_device(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self.check(self.lib.sdpadd_malloc(self.h, max(a.nbytes, 16), ctypes.byref(p)), "malloc")
        if a.nbytes:
            self.check(self.lib.sdpadd_upload(self.h, p, a.ctypes.data, a.nbytes), "upload")
        return DeviceArray(self, p, a.shape)

    def empty(self, shape):
        return self.to_device(np.zeros(shape))


class DeviceArray:
    def __init__(self, dev, ptr, shape):
        self.dev, self.ptr, self.shape = dev, ptr, tuple(shape)

    @property
    def nbytes(self):
        return int(np.prod(self.shape)) * 8

    def get(self):
        out = np.empty(self.shape, dtype=np.float64)
        if out.nbytes:
            self.dev.check(self.dev.lib.sdpadd_download(self.dev.h, out.ctypes.data, self.ptr, out.nbytes), "download")
        return out

    def free(self):
        if self.ptr:
            self.dev.lib.sdpadd_free(self.dev.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- BLAS-level helpers on host arrays (upload, run, download) ------------
def cm(a):
    """(rows, cols, 2) dd matrix -> column-major (rows*cols, 2) buffer."""
    a = np.asarray(a, dtype=np.float64)
    return np.ascontiguousarray(a.transpose(1, 0, 2)).reshape(-1, 2)


def from_cm(buf, rows, cols):
    return np.ascontiguousarray(np.asarray(buf).reshape(cols, rows, 2).transpose(1, 0, 2))


def rgemm(dev, transb, A, B, C, alpha=(1.0, 0.0), beta=(0.0, 0.0)):
    """C := alpha*A*op(B) + beta*C on dd matrices given as (r, c, 2) arrays."""
    m, k = A.shape[:2]
    n = B.shape[1] if transb == "N" else B.shape[0]
    dA, dB, dC = dev.to_device(cm(A)), dev.to_device(cm(B)), dev.to_device(cm(C))
    al, be = np.array(alpha, dtype=np.float64), np.array(beta, dtype=np.float64)
    dev.check(dev.lib.sdpadd_Rgemm(dev.h, b"N", transb.encode(), m, n, k, _dp(al), dA.ptr, max(1, m), dB.ptr,
                                   max(1, B.shape[0]), _dp(be), dC.ptr, max(1, m)), "Rgemm")
    return from_cm(dC.get(), m, n)


class Problem:
    """A problem on the device, loaded from a .dat-s file or built in memory."""

    def __init__(self, dev, path=None, _handle=None):
        self.dev = dev
        self.p = ctypes.c_void_p()
        self.na = None
        if _handle is not None:
            self.p = _handle
        elif path is not None:
            dev.check(dev.lib.sdpadd_problem_load(dev.h, path.encode(), ctypes.byref(self.p)), "problem_load")
        m, nb = ctypes.c_int(), ctypes.c_int()
        bs = (ctypes.c_int * 65536)()
        dev.check(dev.lib.sdpadd_problem_dims(self.p, ctypes.byref(m), ctypes.byref(nb), bs, 65536), "dims")
        self.m, self.block_struct = m.value, [bs[i] for i in range(nb.value)]

    @property
    def sdp_blocks(self):
        """Sizes of the SDP blocks (block sizes >= 2), in order."""
        return [n for n in self.block_struct if n >= 2]

    @property
    def n_lp(self):
        """LP entries (negative blocks and blocks of size 1)."""
        return sum(-n if n < 0 else 1 for n in self.block_struct if n < 2)

    @property
    def nn(self):
        """Length of an iterate: the SDP blocks (n^2 each), then the LP entries."""
        return sum(n * n for n in self.sdp_blocks) + self.n_lp

    @classmethod
    def from_arrays(cls, dev, m, block_struct, b, entries):
        """Builds a problem in SDPA's internal form.  b: (m, 2) dd; entries:
        iterable of (k, l, rows, cols, vals) with k = 0 for C, 1..m for A_k,
        l the 0-based block, 0-based rows/cols and (nnz, 2) dd values."""
        h = ctypes.c_void_p()
        bs = (ctypes.c_int * len(block_struct))(*block_struct)
        dev.check(dev.lib.sdpadd_problem_create(dev.h, m, len(block_struct), bs, ctypes.byref(h)), "problem_create")
        try:
            for k, l, rows, cols, vals in entries:
                r = np.ascontiguousarray(rows, dtype=np.int32)
                c = np.ascontiguousarray(cols, dtype=np.int32)
                v = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1, 2)
                ip = ctypes.POINTER(ctypes.c_int)
                dev.check(dev.lib.sdpadd_problem_set_entries(h, int(k), int(l), len(r), r.ctypes.data_as(ip),
                                                             c.ctypes.data_as(ip), _dp(v)), "problem_set_entries")
            bb = np.ascontiguousarray(b, dtype=np.float64).reshape(m, 2)
            dev.check(dev.lib.sdpadd_problem_set_b(h, _dp(bb)), "problem_set_b")
            dev.check(dev.lib.sdpadd_problem_finalize(h), "problem_finalize")
        except Exception:
            dev.lib.sdpadd_problem_destroy(h)
            raise
        return cls(dev, _handle=h)

    def formula(self):
        out = (ctypes.c_int * max(1, self.m * len(self.sdp_blocks)))()
        self.dev.check(self.dev.lib.sdpadd_problem_formula(self.p, out), "formula")
        return np.array(out[:], dtype=np.int32)

    def schur_ddflops(self):
        """Algorithmic dd operations of one Schur assembly."""
        out = np.zeros(1)
        self.dev.check(self.dev.lib.sdpadd_problem_schur_ddflops(self.p, _dp(out)), "schur_ddflops")
        return float(out[0])

    SCHUR_MODES = {"auto": 0, "dense": 1, "sparse": 2}

    def set_schur_mode(self, mode):
        """The Schur path of the next solves: "auto" (the reference's rule),
        "dense", or "sparse" (whenever the structure allows)."""
        self.dev.check(self.dev.lib.sdpadd_problem_set_schur_mode(self.p, self.SCHUR_MODES[mode]), "set_schur_mode")

    def schur_info(self):
        """{"sparse": bool, "nnz": factor-pattern entries, "aggregate":
        aggregate-pattern entries, "ordering": newToOld (sparse only)}."""
        sp = ctypes.c_int()
        nnz, agg = ctypes.c_int64(), ctypes.c_int64()
        order = (ctypes.c_int * max(1, self.m))()
        self.dev.check(self.dev.lib.sdpadd_problem_schur_info(self.p, ctypes.byref(sp), ctypes.byref(nnz),
                                                              ctypes.byref(agg), order), "schur_info")
        out = {"sparse": bool(sp.value), "nnz": nnz.value, "aggregate": agg.value}
        if sp.value:
            out["ordering"] = [order[i] for i in range(self.m)]
        return out

    def schur(self, X, Zinv):
        dX, dZ = self.dev.to_device(X), self.dev.to_device(Zinv)
        dB = self.dev.empty((self.m * self.m, 2))
        self.dev.check(self.dev.lib.sdpadd_schur_build(self.p, dX.ptr, dZ.ptr, dB.ptr), "schur_build")
        return dB.get()

    def solve(self, params=None):
        prm = Params()
        self.dev.lib.sdpadd_default_params(ctypes.byref(prm))
        if isinstance(params, str):
            if self.dev.lib.sdpadd_read_params(params.encode(), ctypes.byref(prm)) != 0:
                raise SdpaddError("cannot read parameter file %s" % params)
        elif isinstance(params, dict):
            for k, v in params.items():
                setattr(prm, k, v)
        res = Result()
        self.dev.check(self.dev.lib.sdpadd_solve(self.p, ctypes.byref(prm), ctypes.byref(res)), "solve")
        out = {
            "iterations": res.iterations, "phase": PHASES[res.phase], "phase_internal": PHASES[res.phase_internal],
            "seconds": res.seconds, "t_schur": res.t_schur, "t_chol": res.t_chol, "t_lanczos": res.t_lanczos,
            "t_inverse": res.t_inverse, "refine_sweeps": res.refine_sweeps,
        }
        for k in ("mu", "objValPrimal", "objValDual", "normPrimalVec", "normDualMat", "relgap"):
            out[k] = [getattr(res, k)[0], getattr(res, k)[1]]
        cap = 2000
        buf = np.zeros(17 * cap)
        n = self.dev.lib.sdpadd_trace(self.p, _dp(buf), cap)
        tr = []
        keys = ["mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta"]
        for i in range(min(n, cap)):
            rec = {"it": int(buf[17 * i])}
            for q, k in enumerate(keys):
                rec[k] = [buf[17 * i + 1 + 2 * q], buf[17 * i + 2 + 2 * q]]
            tr.append(rec)
        out["trace"] = tr
        return out

    def read_initial_point(self, path, sparse=None):
        """Next solves start from the point in an initial-point file
        (sparse format if the name ends in '-s', as sdpa_dd's -is/-id)."""
        if sparse is None:
            sparse = path.endswith("-s")
        if self.dev.lib.sdpadd_read_initial_point(self.p, path.encode(), int(bool(sparse))) != 0:
            raise SdpaddError("read_initial_point: %s" % self.dev.lib.sdpadd_last_error(self.dev.h).decode())

    def set_initial_point(self, X, y, Z):
        """Next solves start from (X, y, Z): dd host arrays in the layout and
        sign of solution()."""
        X, y, Z = (np.ascontiguousarray(a, dtype=np.float64) for a in (X, y, Z))
        self.dev.check(self.dev.lib.sdpadd_set_initial_point(self.p, _dp(X), _dp(y), _dp(Z)), "set_initial_point")

    def clear_initial_point(self):
        self.dev.check(self.dev.lib.sdpadd_clear_initial_point(self.p), "clear_initial_point")

    def write_solution(self, path):
        """The last solve's final point as a full-dd sparse initial-point file."""
        self.dev.check(self.dev.lib.sdpadd_write_solution(self.p, path.encode()), "write_solution")

    def solution(self):
        X = np.zeros((self.nn, 2))
        Z = np.zeros((self.nn, 2))
        y = np.zeros((self.m, 2))
        self.dev.check(self.dev.lib.sdpadd_solution(self.p, _dp(X), _dp(y), _dp(Z)), "solution")
        return X, y, Z

    # ---- adjoint (implicit differentiation at the final iterate) --------
    def adjoint_setup(self, center_steps=8):
        """Centres the final iterate of the last solve (center_steps Newton
        steps at mu = X.Z/n) and factors the Schur matrix there once; returns
        the number of A-gradient entries per cotangent.  self.centrality:
        max|XZ - mu I| / mu after centring."""
        na = ctypes.c_int64()
        cen = np.zeros(1)
        rc = self.dev.lib.sdpadd_adjoint_setup(self.p, center_steps, ctypes.byref(na), _dp(cen))
        if rc != 0:
            raise SdpaddError("adjoint_setup failed (%d): %s" % (rc, self.dev.lib.sdpadd_last_error(self.dev.h).decode()))
        self.na = na.value
        self.centrality = float(cen[0])
        return self.na

    def adjoint_entries(self):
        """(constraint, block, row, col) of every A-gradient entry (0-based)."""
        if self.na is None:
            self.adjoint_setup()
        out = [np.zeros(self.na, dtype=np.int32) for _ in range(4)]
        ip = ctypes.POINTER(ctypes.c_int)
        self.dev.check(self.dev.lib.sdpadd_adjoint_entries(self.p, *[o.ctypes.data_as(ip) for o in out]),
                       "adjoint_entries")
        return tuple(out)

    def adjoint(self, Xbar=None, ybar=None, pbar=None, nrhs=None):
        """Vector-Jacobian products of (X*, y*, p* = b.y*) for nrhs cotangents.
        Xbar (nrhs, nn, 2), ybar (nrhs, m, 2), pbar (nrhs, 2) dd host arrays
        (None = zero).  Returns (bbar (nrhs, m, 2), Cbar (nrhs, nn, 2),
        Abar (nrhs, na, 2)) as host arrays."""
        if self.na is None:
            self.adjoint_setup()
        K = nrhs or next(len(a) for a in (Xbar, ybar, pbar) if a is not None)
        dX = self.dev.to_device(np.asarray(Xbar).reshape(K, self.nn, 2)) if Xbar is not None else None
        dy = self.dev.to_device(np.asarray(ybar).reshape(K, self.m, 2)) if ybar is not None else None
        pb = np.ascontiguousarray(pbar, dtype=np.float64).reshape(K, 2) if pbar is not None else None
        db, dC, dA = (self.dev.empty((K, self.m, 2)), self.dev.empty((K, self.nn, 2)),
                      self.dev.empty((K, max(1, self.na), 2)))
        rc = self.dev.lib.sdpadd_adjoint(self.p, K, dX.ptr if dX else None, dy.ptr if dy else None,
                                         _dp(pb) if pb is not None else None, db.ptr, dC.ptr, dA.ptr)
        if rc != 0:
            raise SdpaddError("adjoint failed (%d): %s" % (rc, self.dev.lib.sdpadd_last_error(self.dev.h).decode()))
        return db.get(), dC.get(), dA.get()[:, :self.na]

    def close(self):
        if self.p:
            self.dev.lib.sdpadd_problem_destroy(self.p)
            self.p = ctypes.c_void_p()

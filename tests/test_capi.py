"""CPU checks of the C ABI: the library loads and exports every symbol the
header declares; the host-side parts (tridiagonal eigen-solver, argument
validation) behave without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import sdpadd
from ddgen import rand_dd, same

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "sdpadd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sdpadd_[a-zA-Z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(sdpadd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    lib = sdpadd.load_library()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, n


@pytest.mark.parametrize("n", [1, 2, 5, 17, 40])
def test_host_steqr_matches_oracle(oracle, n):
    """The product's tridiagonal QL/QR (used by the Lanczos step length) is
    bit-identical to the oracle's Rsteqr, eigenvectors included."""
    rng = np.random.default_rng(n)
    d = rand_dd((n,), rng)
    e = rand_dd((n,), rng)
    if n > 3:
        e[1] = 0.0  # a split point
    e[-1] = 0.0
    wd, we, wz, winfo = oracle.rsteqr(d, e)
    lib = sdpadd.load_library()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    gd, ge = d.copy(), e.copy()
    gz = np.zeros((n * n, 2))
    work = np.zeros((2 * n + 2, 2))
    info = ctypes.c_int64(-1)
    rc = lib.sdpadd_Rsteqr(None, b"I", n, dp(gd), dp(ge), dp(gz), n, dp(work), ctypes.byref(info))
    assert rc == 0 and info.value == winfo
    assert same(gd, wd)
    assert same(gz, wz)


def test_unsupported_options_rejected():
    lib = sdpadd.load_library()
    one = np.array([1.0, 0.0])
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    # transa 'T' is not a combination the Newton step uses: refused before any GPU work
    assert lib.sdpadd_Rgemm(None, b"T", b"N", 2, 2, 2, dp(one), None, 2, None, 2, dp(one), None, 2) == -1
    assert lib.sdpadd_Rsyrk(None, b"U", b"N", 2, 2, dp(one), None, 2, dp(one), None, 2) == -1
    z = np.zeros((4, 2))
    info = ctypes.c_int64(0)
    assert lib.sdpadd_Rsteqr(None, b"N", 2, dp(z), dp(z), dp(z), 2, dp(z), ctypes.byref(info)) == -1


def test_generated_workloads_have_the_stated_shapes(oracle, tmp_path):
    from sdpadd import gen
    p = gen.maxcut_torus(str(tmp_path / "c1.dat-s"))
    assert oracle.load(p) == (800, [800])
    p = gen.random_dense(str(tmp_path / "c0.dat-s"))
    assert oracle.load(p) == (50, [40])
    p = gen.theta(str(tmp_path / "c4.dat-s"))
    m, bs = oracle.load(p)
    assert bs == [300] and 4000 < m < 5000
    f = oracle.formula(m, 1)
    assert f[0] == 1 and set(f[1:].tolist()) == {2}  # A_0 = I uses F2, edges F3


# MPLAPACK's dd signatures (mplapack/mpblas_dd.h:38-56, mplapack_dd.h:40-41)
# as the demangler prints them for mplapackint = int64_t = long and QD's
# dd_real passed by value.
MPLAPACK_SIGNATURES = [
    "Rgemm(char const*, char const*, long, long, long, dd_real, dd_real*, long, dd_real*, long, dd_real, dd_real*, long)",
    "Rsyrk(char const*, char const*, long, long, dd_real, dd_real*, long, dd_real, dd_real*, long)",
    "Rtrsm(char const*, char const*, char const*, char const*, long, long, dd_real, dd_real*, long, dd_real*, long)",
    "Rtrmm(char const*, char const*, char const*, char const*, long, long, dd_real, dd_real*, long, dd_real*, long)",
    "Rtrsv(char const*, char const*, char const*, long, dd_real*, long, dd_real*, long)",
    "Rtrmv(char const*, char const*, char const*, long, dd_real*, long, dd_real*, long)",
    "Rgemv(char const*, long, long, dd_real, dd_real*, long, dd_real*, long, dd_real, dd_real*, long)",
    "Rdot(long, dd_real*, long, dd_real*, long)",
    "Raxpy(long, dd_real, dd_real*, long, dd_real*, long)",
    "Rscal(long, dd_real, dd_real*, long)",
    "Rcopy(long, dd_real*, long, dd_real*, long)",
    "Rpotrf(char const*, long, dd_real*, long, long&)",
    "Rsteqr(char const*, long, dd_real*, dd_real*, dd_real*, long, dd_real*, long&)",
]


def test_mplapack_shim_exports_the_reference_signatures():
    """libmpblas_dd_hip.so carries MPLAPACK's C++ symbols (the library seam
    of configure.ac:47-50), so SDPA-DD links it unchanged."""
    import subprocess
    shim = os.path.join(ROOT, "sdpa-dd_amd", "libmpblas_dd_hip.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", shim], capture_output=True, text=True, check=True).stdout
    exported = {ln.split(" ", 2)[2] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in MPLAPACK_SIGNATURES if s not in exported]
    assert not missing, missing

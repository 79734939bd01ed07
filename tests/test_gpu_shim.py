"""The MPLAPACK-symbol library (libmpblas_dd_hip.so) driven the way
SDPA-DD's host code drives MPLAPACK: a C++ program (tests/shim/shim_driver.cpp)
calls Rgemm, Rsyrk, Rtrsm, Rtrmm, Rtrsv, Rtrmv, Rgemv, Rdot, Raxpy, Rscal,
Rcopy, Rpotrf and Rsteqr with dd_real by value and host arrays (leading
dimensions above the row count, negative strides, a non-positive-definite
matrix), and every result is compared bit for bit with the oracle's
restatement of the reference routine."""
import os
import subprocess

import numpy as np
import pytest

import sdpadd
from ddgen import rand_dd, rand_lower_cm, rand_spd_cm, same

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sdpa-dd_amd")
ONE = np.array([1.0, 0.0])


def _hex(v):
    return "%s %s" % (float(v[0]).hex(), float(v[1]).hex())


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("shim")
    exe = str(d / "shim_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "shim", "shim_driver.cpp"), "-o", exe,
                    "-L" + PKG, "-lmpblas_dd_hip", "-Wl,-rpath," + PKG], check=True)
    return exe, d


def _pad(buf, rows, cols, ld, rng):
    """(rows*cols, 2) column-major -> (ld*cols, 2) with random padding rows."""
    out = rand_dd((ld * cols,), rng)
    for j in range(cols):
        out[j * ld:j * ld + rows] = buf[j * rows:(j + 1) * rows]
    return out


def test_shim_routines_bitwise_vs_oracle(driver, oracle):
    exe, base = driver
    rng = np.random.default_rng(31)
    cases = []

    def case(name, args, ins):
        d = base / ("c%02d_%s" % (len(cases), name))
        d.mkdir()
        (d / "args.txt").write_text(name + " " + " ".join(str(a) for a in args) + "\n")
        for i, a in enumerate(ins):
            if a is not None:
                np.ascontiguousarray(a, dtype=np.float64).tofile(str(d / ("in%d.bin" % i)))
        cases.append(d)
        return d

    expect = {}
    # Rgemm N/T with padded leading dimensions and general alpha, beta
    for tb in ("N", "T"):
        m, n, k = 70, 45, 33
        lda, ldb, ldc = 75, (k if tb == "N" else n) + 3, 72
        A = rand_dd((lda * k,), rng)
        B = rand_dd((ldb * (n if tb == "N" else k),), rng)
        C = rand_dd((ldc * n,), rng)
        al, be = rand_dd((1,), rng)[0], rand_dd((1,), rng)[0]
        d = case("Rgemm", [tb, m, n, k, lda, ldb, ldc, _hex(al), _hex(be)], [A, B, C])
        expect[d] = oracle.rgemm(tb, m, n, k, al, A, lda, B, ldb, be, C, ldc)
    n, k = 50, 20
    A, C = rand_dd((n * k,), rng), rand_dd((n * n,), rng)
    al, be = np.array([-1.0, 0.0]), ONE
    d = case("Rsyrk", [n, k, n, n, _hex(al), _hex(be)], [A, None, C])
    expect[d] = oracle.rsyrk(n, k, al, A, n, be, C, n)
    m, n = 60, 40
    L = rand_lower_cm(n, rng)
    B = rand_dd((m * n,), rng)
    d = case("Rtrsm", ["R", "T", m, n, n, m], [L, B])
    expect[d] = oracle.rtrsm("RLTN", m, n, L, n, B, m)
    L = rand_lower_cm(m, rng)
    d = case("Rtrsm", ["L", "N", m, n, m, m], [L, B])
    expect[d] = oracle.rtrsm("LLNN", m, n, L, m, B, m)
    d = case("Rtrmm", [m, n, m, m], [L, B])
    expect[d] = oracle.rtrmm(m, n, L, m, B, m)
    for tr in ("N", "T"):
        x = rand_dd((m,), rng)
        d = case("Rtrsv", [tr, m, m], [L, x])
        expect[d] = oracle.rtrsv(tr, m, L, m, x)
        d = case("Rtrmv", [tr, m, m], [L, x])
        expect[d] = oracle.rtrmv(tr, m, L, m, x)
    A, x, y = rand_dd((m * n,), rng), rand_dd((n,), rng), rand_dd((m,), rng)
    al, be = rand_dd((1,), rng)[0], rand_dd((1,), rng)[0]
    d = case("Rgemv", [m, n, m, _hex(al), _hex(be)], [A, x, y])
    expect[d] = oracle.rgemv(m, n, al, A, m, x, be, y)
    # Rdot: unit strides (the library's dot partition) and a strided dot
    x, y = rand_dd((300,), rng), rand_dd((300,), rng)
    d = case("Rdot", [300, 1, 1], [x, y])
    oracle.set_dot_threads(sdpadd.DEFAULT_DOT_PARTITION)
    expect[d] = ("dot", oracle.rdot(x, y))
    oracle.set_dot_threads(1)
    X = rand_dd((40 * 40,), rng)
    d = case("Rdot", [40, 40, 1], [X[3:], rand_dd((40,), rng)])
    xin = np.fromfile(str(d / "in0.bin")).reshape(-1, 2)
    yin = np.fromfile(str(d / "in1.bin")).reshape(-1, 2)
    expect[d] = ("dot", oracle.rdot(xin, yin, incx=40, incy=1, n=40))
    # Raxpy with a negative stride, Rscal, Rcopy
    nn = 33
    a = rand_dd((1,), rng)[0]
    x, y = rand_dd((2 * nn,), rng), rand_dd((nn,), rng)
    d = case("Raxpy", [nn, -2, 1, _hex(a)], [x, y])
    xs = x[(nn - 1) * 2::-2]  # element i of a stride -2 vector: x[(nn-1)*2 - 2i]
    prod = oracle.dd_op(2, np.repeat(a[None, :], nn, 0), xs)
    expect[d] = oracle.dd_op(5, y, prod)
    x = rand_dd((3 * nn,), rng)
    d = case("Rscal", [nn, 3, _hex(a)], [x])
    e = x.copy()
    e[::3][:nn] = oracle.dd_op(2, np.repeat(a[None, :], nn, 0), x[::3][:nn])
    expect[d] = e
    x, y = rand_dd((nn,), rng), rand_dd((2 * nn,), rng)
    d = case("Rcopy", [nn, 1, 2], [x, y])
    e = y.copy()
    e[::2][:nn] = x
    expect[d] = e
    # Rpotrf: padded leading dimension, and a matrix that is not PD
    n, ld = 150, 157
    A = _pad(rand_spd_cm(n, rng), n, n, ld, rng)
    d = case("Rpotrf", [n, ld], [A])
    expect[d] = ("potrf",) + oracle.rpotrf(n, A, ld)
    B = rand_spd_cm(n, rng)
    B[77 + n * 77] = [-3.0, 0.0]
    d = case("Rpotrf", [n, n], [B])
    expect[d] = ("potrf",) + oracle.rpotrf(n, B, n)
    # Rsteqr
    dd_, ee = rand_dd((12,), rng), rand_dd((12,), rng)
    ee[-1] = 0.0
    d = case("Rsteqr", [12], [dd_, ee])
    wd, we, wz, winfo = oracle.rsteqr(dd_, ee)
    expect[d] = ("steqr", np.concatenate([wd, wz]), winfo)

    out = subprocess.run([exe] + [str(c) for c in cases], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    for d in cases:
        e = expect[d]
        if isinstance(e, tuple) and e[0] == "dot":
            hi, lo = (float.fromhex(v) for v in (d / "res.txt").read_text().split())
            assert same(np.array([hi, lo]), e[1]), d.name
            continue
        got = np.fromfile(str(d / "out.bin")).reshape(-1, 2)
        if isinstance(e, tuple) and e[0] == "potrf":
            info = int((d / "res.txt").read_text())
            assert info == e[2], (d.name, info, e[2])
            if info == 0:
                assert same(got, e[1]), d.name
            continue
        if isinstance(e, tuple) and e[0] == "steqr":
            assert int((d / "res.txt").read_text()) == e[2]
            assert same(got, e[1]), d.name
            continue
        assert same(got, e), d.name


def test_shim_unsupported_option_stops_loudly(driver, tmp_path):
    """An option combination with no GPU kernel is reported like an illegal
    argument (Mxerbla_dd's message) and stops the process."""
    exe, _ = driver
    d = tmp_path / "bad"
    d.mkdir()
    (d / "args.txt").write_text("Rtrsm R N 4 4 4 4\n")  # side R with transa N: no kernel
    np.zeros((16, 2)).tofile(str(d / "in0.bin"))
    np.zeros((16, 2)).tofile(str(d / "in1.bin"))
    out = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "On entry to Rtrsm parameter number" in out.stderr

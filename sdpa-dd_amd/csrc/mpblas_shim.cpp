// libmpblas_dd_hip.so: the dd BLAS/LAPACK routines of SDPA-DD's Newton step
// under the C++ symbols of MPLAPACK's dd library (mplapack/mpblas_dd.h:38-56,
// mplapack/mplapack_dd.h:38-92), computed on the GPU through
// libsdpadd_hip.so.  Linked in place of -lmpblas_dd_opt -lmplapack_dd
// (--with-system-mplapack, configure.ac:47-50, 196-225; Makefile.am:3-9),
// it lets the reference's own host code drive the GPU kernels unchanged.
//
// The contract is MPLAPACK's: C++ linkage, dd_real by value, caller-owned
// host arrays, column-major with leading dimensions, mplapackint = int64_t,
// option strings read by their first letter, LAPACK info through a
// reference.  Each call stages its operands to device buffers (kept and
// grown across calls), runs the kernels and copies the outputs back; the
// results are bit-identical to the reference's routines (with Rdot's unit
// stride summed in the library's dot partition, sdpadd_set_dot_partition).
// Option combinations outside what the Newton step uses have no GPU kernel:
// they are reported like an illegal argument (Mxerbla_dd's message) and the
// process stops, as the reference's own argument errors do.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/sdpadd.h"

typedef int64_t mplapackint;
// QD's dd_real (qd/include/qd/dd_real.h:62-68): two doubles, no user-declared
// copy constructor or destructor, so it is passed by value in two FP
// registers; this struct has the same layout, name and passing.
struct dd_real {
    double x[2];
};

namespace {

std::mutex g_mu;
sdpadd_handle g_h = nullptr;
int g_T = 64;  // dot partition of unit-stride dots (SDPADD_DOT_PARTITION, default 64)
void *g_buf[4] = {nullptr, nullptr, nullptr, nullptr};
size_t g_cap[4] = {0, 0, 0, 0};

[[noreturn]] void fail(const char *name, int arg, const char *why) {
    std::fprintf(stderr, " ** On entry to %s parameter number %2d had an illegal value\n", name, arg);
    if (why) std::fprintf(stderr, " ** %s: %s\n", name, why);
    std::exit(arg > 0 ? arg : 1);
}
[[noreturn]] void gpu_fail(const char *name) {
    std::fprintf(stderr, " ** %s: GPU call failed: %s\n", name, g_h ? sdpadd_last_error(g_h) : "no device");
    std::exit(1);
}
sdpadd_handle H(const char *name) {
    if (!g_h) {
        const char *dev = std::getenv("SDPADD_DEVICE");
        if (sdpadd_create(dev ? std::atoi(dev) : 0, &g_h) != 0) {
            g_h = nullptr;
            gpu_fail(name);
        }
        const char *t = std::getenv("SDPADD_DOT_PARTITION");
        if (t && std::atoi(t) >= 1) g_T = std::atoi(t);
        if (sdpadd_set_dot_partition(g_h, g_T) != 0) gpu_fail(name);
    }
    return g_h;
}
// Device buffer `slot` of at least `bytes`.
double *buf(int slot, size_t bytes, const char *name) {
    if (bytes > g_cap[slot]) {
        if (g_buf[slot]) sdpadd_free(g_h, g_buf[slot]);
        g_buf[slot] = nullptr;
        g_cap[slot] = 0;
        if (sdpadd_malloc(g_h, bytes, &g_buf[slot]) != 0) gpu_fail(name);
        g_cap[slot] = bytes;
    }
    return static_cast<double *>(g_buf[slot]);
}
// Host column-major ld x cols array -> device slot (the whole leading
// dimension travels, so the device array keeps the caller's ld).
double *up(int slot, const dd_real *a, mplapackint ld, mplapackint cols, const char *name) {
    const size_t bytes = sizeof(dd_real) * (size_t)(ld > 0 ? ld : 1) * (size_t)(cols > 0 ? cols : 1);
    double *d = buf(slot, bytes, name);
    if (a && ld > 0 && cols > 0 && sdpadd_upload(g_h, d, a, sizeof(dd_real) * (size_t)ld * (size_t)cols) != 0)
        gpu_fail(name);
    return d;
}
void down(dd_real *a, const double *d, mplapackint ld, mplapackint cols, const char *name) {
    if (ld > 0 && cols > 0 && sdpadd_download(g_h, a, d, sizeof(dd_real) * (size_t)ld * (size_t)cols) != 0)
        gpu_fail(name);
}
bool is(const char *opt, char c) { return opt && (opt[0] == c || opt[0] == (char)(c + 32)); }
// elements a strided vector of n entries spans
mplapackint span(mplapackint n, mplapackint inc) { return n > 0 ? 1 + (n - 1) * (inc < 0 ? -inc : inc) : 0; }
void check(int rc, const char *name) {
    if (rc < 0) gpu_fail(name);
}

}  // namespace

// ---- mpblas_dd.h ---------------------------------------------------------------
void Rgemm(const char *transa, const char *transb, mplapackint const m, mplapackint const n, mplapackint const k,
           dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb, dd_real const beta,
           dd_real *c, mplapackint const ldc) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rgemm";
    if (!is(transa, 'N')) fail(name, 1, "only transa = 'N' runs on the GPU");
    if (!is(transb, 'N') && !is(transb, 'T')) fail(name, 2, "transb must be 'N' or 'T'");
    if (m < 0) fail(name, 3, nullptr);
    if (n < 0) fail(name, 4, nullptr);
    if (k < 0) fail(name, 5, nullptr);
    if (lda < (m > 1 ? m : 1)) fail(name, 8, nullptr);
    if (ldb < ((is(transb, 'N') ? k : n) > 1 ? (is(transb, 'N') ? k : n) : 1)) fail(name, 10, nullptr);
    if (ldc < (m > 1 ? m : 1)) fail(name, 13, nullptr);
    if (m == 0 || n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, k, name);
    double *dB = up(1, b, ldb, is(transb, 'N') ? n : k, name);
    double *dC = up(2, c, ldc, n, name);
    check(sdpadd_Rgemm(h, 'N', is(transb, 'N') ? 'N' : 'T', m, n, k, alpha.x, dA, lda, dB, ldb, beta.x, dC, ldc), name);
    down(c, dC, ldc, n, name);
}

void Rsyrk(const char *uplo, const char *trans, mplapackint const n, mplapackint const k, dd_real const alpha,
           dd_real *a, mplapackint const lda, dd_real const beta, dd_real *c, mplapackint const ldc) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rsyrk";
    if (!is(uplo, 'L')) fail(name, 1, "only uplo = 'L' runs on the GPU");
    if (!is(trans, 'N')) fail(name, 2, "only trans = 'N' runs on the GPU");
    if (n < 0) fail(name, 3, nullptr);
    if (k < 0) fail(name, 4, nullptr);
    if (lda < (n > 1 ? n : 1)) fail(name, 7, nullptr);
    if (ldc < (n > 1 ? n : 1)) fail(name, 10, nullptr);
    if (n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, k, name);
    double *dC = up(2, c, ldc, n, name);
    check(sdpadd_Rsyrk(h, 'L', 'N', n, k, alpha.x, dA, lda, beta.x, dC, ldc), name);
    down(c, dC, ldc, n, name);
}

void Rtrsm(const char *side, const char *uplo, const char *transa, const char *diag, mplapackint const m,
           mplapackint const n, dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rtrsm";
    const bool left = is(side, 'L');
    if (!left && !is(side, 'R')) fail(name, 1, nullptr);
    if (!is(uplo, 'L')) fail(name, 2, "only uplo = 'L' runs on the GPU");
    if (!(left ? is(transa, 'N') : is(transa, 'T'))) fail(name, 3, "only (L,L,N,N) and (R,L,T,N) run on the GPU");
    if (!is(diag, 'N')) fail(name, 4, "only diag = 'N' runs on the GPU");
    if (!(alpha.x[0] == 1.0 && alpha.x[1] == 0.0)) fail(name, 7, "only alpha = 1 runs on the GPU");
    if (m < 0) fail(name, 5, nullptr);
    if (n < 0) fail(name, 6, nullptr);
    const mplapackint na = left ? m : n;
    if (lda < (na > 1 ? na : 1)) fail(name, 9, nullptr);
    if (ldb < (m > 1 ? m : 1)) fail(name, 11, nullptr);
    if (m == 0 || n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, na, name);
    double *dB = up(1, b, ldb, n, name);
    check(sdpadd_Rtrsm(h, left ? 'L' : 'R', 'L', left ? 'N' : 'T', 'N', m, n, alpha.x, dA, lda, dB, ldb), name);
    down(b, dB, ldb, n, name);
}

void Rtrmm(const char *side, const char *uplo, const char *transa, const char *diag, mplapackint const m,
           mplapackint const n, dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rtrmm";
    if (!is(side, 'L') || !is(uplo, 'L') || !is(transa, 'T') || !is(diag, 'N'))
        fail(name, 1, "only (L,L,T,N) runs on the GPU");
    if (!(alpha.x[0] == 1.0 && alpha.x[1] == 0.0)) fail(name, 7, "only alpha = 1 runs on the GPU");
    if (m < 0) fail(name, 5, nullptr);
    if (n < 0) fail(name, 6, nullptr);
    if (lda < (m > 1 ? m : 1)) fail(name, 9, nullptr);
    if (ldb < (m > 1 ? m : 1)) fail(name, 11, nullptr);
    if (m == 0 || n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, m, name);
    double *dB = up(1, b, ldb, n, name);
    check(sdpadd_Rtrmm(h, 'L', 'L', 'T', 'N', m, n, alpha.x, dA, lda, dB, ldb), name);
    down(b, dB, ldb, n, name);
}

static void trv(const char *name, bool solve, const char *uplo, const char *trans, const char *diag, mplapackint n,
                dd_real *a, mplapackint lda, dd_real *x, mplapackint incx) {
    if (!is(uplo, 'L')) fail(name, 1, "only uplo = 'L' runs on the GPU");
    if (!is(trans, 'N') && !is(trans, 'T')) fail(name, 2, nullptr);
    if (!is(diag, 'N')) fail(name, 3, "only diag = 'N' runs on the GPU");
    if (n < 0) fail(name, 4, nullptr);
    if (lda < (n > 1 ? n : 1)) fail(name, 6, nullptr);
    if (incx != 1) fail(name, 8, "only incx = 1 runs on the GPU");
    if (n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, n, name);
    double *dx = up(1, x, n, 1, name);
    const char t = is(trans, 'N') ? 'N' : 'T';
    check(solve ? sdpadd_Rtrsv(h, 'L', t, 'N', n, dA, lda, dx, 1) : sdpadd_Rtrmv(h, 'L', t, 'N', n, dA, lda, dx, 1), name);
    down(x, dx, n, 1, name);
}
void Rtrsv(const char *uplo, const char *trans, const char *diag, mplapackint const n, dd_real *a, mplapackint const lda,
           dd_real *x, mplapackint const incx) {
    std::lock_guard<std::mutex> lk(g_mu);
    trv("Rtrsv", true, uplo, trans, diag, n, a, lda, x, incx);
}
void Rtrmv(const char *uplo, const char *trans, const char *diag, mplapackint const n, dd_real *a, mplapackint const lda,
           dd_real *x, mplapackint const incx) {
    std::lock_guard<std::mutex> lk(g_mu);
    trv("Rtrmv", false, uplo, trans, diag, n, a, lda, x, incx);
}

void Rgemv(const char *trans, mplapackint const m, mplapackint const n, dd_real const alpha, dd_real *a,
           mplapackint const lda, dd_real *x, mplapackint const incx, dd_real const beta, dd_real *y,
           mplapackint const incy) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rgemv";
    if (!is(trans, 'N')) fail(name, 1, "only trans = 'N' runs on the GPU");
    if (m < 0) fail(name, 2, nullptr);
    if (n < 0) fail(name, 3, nullptr);
    if (lda < (m > 1 ? m : 1)) fail(name, 6, nullptr);
    if (incx != 1) fail(name, 8, "only incx = 1 runs on the GPU");
    if (incy != 1) fail(name, 11, "only incy = 1 runs on the GPU");
    if (m == 0 || n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, n, name);
    double *dx = up(1, x, n, 1, name);
    double *dy = up(2, y, m, 1, name);
    check(sdpadd_Rgemv(h, 'N', m, n, alpha.x, dA, lda, dx, 1, beta.x, dy, 1), name);
    down(y, dy, m, 1, name);
}

dd_real Rdot(mplapackint const n, dd_real *dx, mplapackint const incx, dd_real *dy, mplapackint const incy) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rdot";
    dd_real r{{0.0, 0.0}};
    if (n <= 0) return r;
    sdpadd_handle h = H(name);
    if (incx == 1 && incy == 1) {
        double *a = up(0, dx, n, 1, name);
        double *b = up(1, dy, n, 1, name);
        check(sdpadd_Rdot(h, n, a, 1, b, 1, r.x), name);
        return r;
    }
    // a strided dot is one sequential sum in the reference (Rdot_omp.cpp:112-117):
    // gather both vectors in their visiting order and sum them as one chunk
    std::vector<dd_real> gx((size_t)n), gy((size_t)n);
    mplapackint ix = incx < 0 ? (1 - n) * incx : 0, iy = incy < 0 ? (1 - n) * incy : 0;
    for (mplapackint i = 0; i < n; i++, ix += incx, iy += incy) {
        gx[(size_t)i] = dx[ix];
        gy[(size_t)i] = dy[iy];
    }
    double *a = up(0, gx.data(), n, 1, name);
    double *b = up(1, gy.data(), n, 1, name);
    check(sdpadd_set_dot_partition(h, 1), name);
    const int rc = sdpadd_Rdot(h, n, a, 1, b, 1, r.x);
    (void)sdpadd_set_dot_partition(h, g_T);
    check(rc, name);
    return r;
}

void Raxpy(mplapackint const n, dd_real const da, dd_real *dx, mplapackint const incx, dd_real *dy,
           mplapackint const incy) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Raxpy";
    if (n <= 0 || (da.x[0] == 0.0 && da.x[1] == 0.0)) return;  // Raxpy.cpp:42-45
    if (incx == 0 || incy == 0) fail(name, incx == 0 ? 4 : 6, "zero strides do not run on the GPU");
    sdpadd_handle h = H(name);
    const mplapackint sx = span(n, incx), sy = span(n, incy);
    double *a = up(0, dx, sx, 1, name);
    double *b = up(1, dy, sy, 1, name);
    check(sdpadd_Raxpy(h, n, da.x, a, incx, b, incy), name);
    down(dy, b, sy, 1, name);
}

void Rscal(mplapackint const n, dd_real const da, dd_real *dx, mplapackint const incx) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rscal";
    if (n <= 0 || incx <= 0) return;  // Rscal.cpp: nothing for incx <= 0
    sdpadd_handle h = H(name);
    const mplapackint sx = span(n, incx);
    double *a = up(0, dx, sx, 1, name);
    check(sdpadd_Rscal(h, n, da.x, a, incx), name);
    down(dx, a, sx, 1, name);
}

void Rcopy(mplapackint const n, dd_real *dx, mplapackint const incx, dd_real *dy, mplapackint const incy) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rcopy";
    if (n <= 0) return;
    if (incx == 0 || incy == 0) fail(name, incx == 0 ? 3 : 5, "zero strides do not run on the GPU");
    sdpadd_handle h = H(name);
    const mplapackint sx = span(n, incx), sy = span(n, incy);
    double *a = up(0, dx, sx, 1, name);
    double *b = up(1, dy, sy, 1, name);
    check(sdpadd_Rcopy(h, n, a, incx, b, incy), name);
    down(dy, b, sy, 1, name);
}

// ---- mplapack_dd.h -------------------------------------------------------------
void Rpotrf(const char *uplo, mplapackint const n, dd_real *a, mplapackint const lda, mplapackint &info) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rpotrf";
    info = 0;
    if (!is(uplo, 'L')) fail(name, 1, "only uplo = 'L' runs on the GPU");
    if (n < 0) {
        info = -2;
        fail(name, 2, nullptr);
    }
    if (lda < (n > 1 ? n : 1)) {
        info = -4;
        fail(name, 4, nullptr);
    }
    if (n == 0) return;
    sdpadd_handle h = H(name);
    double *dA = up(0, a, lda, n, name);
    int64_t inf = 0;
    check(sdpadd_Rpotrf(h, 'L', n, dA, lda, &inf), name);
    down(a, dA, lda, n, name);
    info = inf;
}

void Rsteqr(const char *compz, mplapackint const n, dd_real *d, dd_real *e, dd_real *z, mplapackint const ldz,
            dd_real *work, mplapackint &info) {
    std::lock_guard<std::mutex> lk(g_mu);
    const char *name = "Rsteqr";
    info = 0;
    if (!is(compz, 'I')) fail(name, 1, "only compz = 'I' is provided");
    if (n < 0) fail(name, 2, nullptr);
    if (ldz < (n > 1 ? n : 1)) fail(name, 6, nullptr);
    int64_t inf = 0;
    // the (k+1)-tridiagonal of the Lanczos estimate is tiny: the library
    // runs it on the host arrays it is given
    check(sdpadd_Rsteqr(H(name), 'I', n, reinterpret_cast<double *>(d), reinterpret_cast<double *>(e),
                        reinterpret_cast<double *>(z), ldz, reinterpret_cast<double *>(work), &inf),
          name);
    info = inf;
}

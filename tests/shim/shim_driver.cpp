// Test driver for libmpblas_dd_hip.so: calls the MPLAPACK dd symbols the way
// SDPA-DD's host code does (C++ linkage, dd_real by value, host arrays) on
// cases written by tests/test_gpu_shim.py, and writes the results back.
//   shim_driver <case directory>...
// Each directory holds args.txt ("<routine> <ints...> <dd scalars as two
// %a doubles>"), in0.bin .. in2.bin (raw dd arrays); the driver writes
// out.bin (the array the routine modifies) and res.txt (a dd result or info).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

typedef int64_t mplapackint;
struct dd_real {
    double x[2];
};
// mplapack/mpblas_dd.h:38-56, mplapack/mplapack_dd.h:40-41
dd_real Rdot(mplapackint const n, dd_real *dx, mplapackint const incx, dd_real *dy, mplapackint const incy);
void Rcopy(mplapackint const n, dd_real *dx, mplapackint const incx, dd_real *dy, mplapackint const incy);
void Raxpy(mplapackint const n, dd_real const da, dd_real *dx, mplapackint const incx, dd_real *dy, mplapackint const incy);
void Rscal(mplapackint const n, dd_real const da, dd_real *dx, mplapackint const incx);
void Rtrmv(const char *uplo, const char *trans, const char *diag, mplapackint const n, dd_real *a, mplapackint const lda,
           dd_real *x, mplapackint const incx);
void Rtrsv(const char *uplo, const char *trans, const char *diag, mplapackint const n, dd_real *a, mplapackint const lda,
           dd_real *x, mplapackint const incx);
void Rgemv(const char *trans, mplapackint const m, mplapackint const n, dd_real const alpha, dd_real *a,
           mplapackint const lda, dd_real *x, mplapackint const incx, dd_real const beta, dd_real *y, mplapackint const incy);
void Rtrmm(const char *side, const char *uplo, const char *transa, const char *diag, mplapackint const m,
           mplapackint const n, dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb);
void Rtrsm(const char *side, const char *uplo, const char *transa, const char *diag, mplapackint const m,
           mplapackint const n, dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb);
void Rgemm(const char *transa, const char *transb, mplapackint const m, mplapackint const n, mplapackint const k,
           dd_real const alpha, dd_real *a, mplapackint const lda, dd_real *b, mplapackint const ldb, dd_real const beta,
           dd_real *c, mplapackint const ldc);
void Rsyrk(const char *uplo, const char *trans, mplapackint const n, mplapackint const k, dd_real const alpha,
           dd_real *a, mplapackint const lda, dd_real const beta, dd_real *c, mplapackint const ldc);
void Rsteqr(const char *compz, mplapackint const n, dd_real *d, dd_real *e, dd_real *z, mplapackint const ldz,
            dd_real *work, mplapackint &info);
void Rpotrf(const char *uplo, mplapackint const n, dd_real *a, mplapackint const lda, mplapackint &info);

static std::vector<dd_real> load(const std::string &p) {
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    std::vector<dd_real> v;
    if (!f) return v;
    v.resize((size_t)f.tellg() / sizeof(dd_real));
    f.seekg(0);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(dd_real)));
    return v;
}
static void save(const std::string &p, const std::vector<dd_real> &v) {
    std::ofstream f(p, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(dd_real)));
}
static dd_real dd(const std::string &hi, const std::string &lo) {
    return dd_real{{std::strtod(hi.c_str(), nullptr), std::strtod(lo.c_str(), nullptr)}};
}

int main(int argc, char **argv) {
    for (int c = 1; c < argc; c++) {
        const std::string dir = argv[c];
        std::ifstream af(dir + "/args.txt");
        std::string op;
        af >> op;
        std::vector<std::string> t;
        for (std::string w; af >> w;) t.push_back(w);
        auto I = [&](size_t i) { return (mplapackint)std::atoll(t[i].c_str()); };
        std::vector<dd_real> a = load(dir + "/in0.bin"), b = load(dir + "/in1.bin"), x = load(dir + "/in2.bin");
        std::ofstream res(dir + "/res.txt");
        res.precision(17);
        std::vector<dd_real> *out = nullptr;
        if (op == "Rgemm") {  // transb m n k lda ldb ldc alpha beta
            Rgemm("N", t[0].c_str(), I(1), I(2), I(3), dd(t[7], t[8]), a.data(), I(4), b.data(), I(5), dd(t[9], t[10]),
                  x.data(), I(6));
            out = &x;
        } else if (op == "Rsyrk") {  // n k lda ldc alpha beta
            Rsyrk("Lower", "NoTranspose", I(0), I(1), dd(t[4], t[5]), a.data(), I(2), dd(t[6], t[7]), x.data(), I(3));
            out = &x;
        } else if (op == "Rtrsm") {  // side transa m n lda ldb
            Rtrsm(t[0].c_str(), "Lower", t[1].c_str(), "Nonunit", I(2), I(3), dd_real{{1.0, 0.0}}, a.data(), I(4),
                  b.data(), I(5));
            out = &b;
        } else if (op == "Rtrmm") {  // m n lda ldb
            Rtrmm("Left", "Lower", "Transpose", "Nonunit", I(0), I(1), dd_real{{1.0, 0.0}}, a.data(), I(2), b.data(),
                  I(3));
            out = &b;
        } else if (op == "Rtrsv" || op == "Rtrmv") {  // trans n lda
            (op == "Rtrsv" ? Rtrsv : Rtrmv)("Lower", t[0].c_str(), "Nonunit", I(1), a.data(), I(2), b.data(), 1);
            out = &b;
        } else if (op == "Rgemv") {  // m n lda alpha beta
            Rgemv("NoTranspose", I(0), I(1), dd(t[3], t[4]), a.data(), I(2), b.data(), 1, dd(t[5], t[6]), x.data(), 1);
            out = &x;
        } else if (op == "Rdot") {  // n incx incy
            dd_real r = Rdot(I(0), a.data(), I(1), b.data(), I(2));
            res << std::hexfloat << r.x[0] << " " << r.x[1] << "\n";
        } else if (op == "Raxpy") {  // n incx incy alpha
            Raxpy(I(0), dd(t[3], t[4]), a.data(), I(1), b.data(), I(2));
            out = &b;
        } else if (op == "Rscal") {  // n incx alpha
            Rscal(I(0), dd(t[2], t[3]), a.data(), I(1));
            out = &a;
        } else if (op == "Rcopy") {  // n incx incy
            Rcopy(I(0), a.data(), I(1), b.data(), I(2));
            out = &b;
        } else if (op == "Rpotrf") {  // n lda
            mplapackint info = -99;
            Rpotrf("L", I(0), a.data(), I(1), info);
            res << info << "\n";
            out = &a;
        } else if (op == "Rsteqr") {  // n  (a = d, b = e, z out)
            const mplapackint n = I(0);
            std::vector<dd_real> z((size_t)(n * n)), work((size_t)(2 * n));
            mplapackint info = -99;
            Rsteqr("I", n, a.data(), b.data(), z.data(), n, work.data(), info);
            res << info << "\n";
            z.insert(z.begin(), a.begin(), a.end());  // eigenvalues, then the eigenvectors
            save(dir + "/out.bin", z);
        } else {
            std::fprintf(stderr, "unknown routine %s\n", op.c_str());
            return 2;
        }
        if (out) save(dir + "/out.bin", *out);
        std::printf("%s done\n", op.c_str());
        std::fflush(stdout);
    }
    return 0;
}

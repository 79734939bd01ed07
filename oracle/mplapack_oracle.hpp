// ORACLE — test infrastructure only.
//
// CPU restatement of the dd BLAS/LAPACK subset (MPLAPACK 2.0.1) that the
// SDPA-DD Newton step calls.  Only the variants reached on the path are
// restated; each keeps the reference routine's per-element operation order.
#pragma once
#include <cstdint>
#include "dd_oracle.hpp"

namespace orc {
typedef int64_t mint;

// Number of "threads" in the OpenMP static partition Rdot sums over
// (mplapack/Rdot_omp.cpp:80-120).  1 reproduces a single-threaded reference
// run; T>1 reproduces a T-thread run whose critical sections complete in
// thread order.
extern int g_dot_threads;

void Rgemm(char ta, char tb, mint m, mint n, mint k, dd_real alpha, const dd_real *A, mint lda, const dd_real *B,
           mint ldb, dd_real beta, dd_real *C, mint ldc);
void Rsyrk_LN(mint n, mint k, dd_real alpha, const dd_real *A, mint lda, dd_real beta, dd_real *C, mint ldc);
void Rtrsm_RLTN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb);
void Rtrsm_LLNN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb);
void Rtrmm_LLTN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb);
void Rtrsv_LN(mint n, const dd_real *A, mint lda, dd_real *x);
void Rtrsv_LT(mint n, const dd_real *A, mint lda, dd_real *x);
void Rtrmv_LN(mint n, const dd_real *A, mint lda, dd_real *x);
void Rtrmv_LT(mint n, const dd_real *A, mint lda, dd_real *x);
void Rgemv_N(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, const dd_real *x, dd_real beta, dd_real *y);
dd_real Rdot(mint n, const dd_real *x, mint incx, const dd_real *y, mint incy);
void Raxpy(mint n, dd_real a, const dd_real *x, mint incx, dd_real *y, mint incy);
void Rscal(mint n, dd_real a, dd_real *x, mint incx);
void Rcopy(mint n, const dd_real *x, mint incx, dd_real *y, mint incy);
void Rswap(mint n, dd_real *x, mint incx, dd_real *y, mint incy);

// Cholesky (Lower only).  info: 0 ok, k>0 leading minor k not positive.
void Rpotrf_L(mint n, dd_real *A, mint lda, mint &info);
void Rpotrf2_L(mint n, dd_real *A, mint lda, mint &info);

// Implicit QL/QR on a symmetric tridiagonal; compz = 'I' (eigenvectors of
// the tridiagonal into Z) or 'N'.
void Rsteqr(char compz, mint n, dd_real *d, dd_real *e, dd_real *z, mint ldz, dd_real *work, mint &info);
void Rlaev2(dd_real a, dd_real b, dd_real c, dd_real &rt1, dd_real &rt2, dd_real &cs1, dd_real &sn1);
void Rlae2(dd_real a, dd_real b, dd_real c, dd_real &rt1, dd_real &rt2);
dd_real Rlapy2(dd_real x, dd_real y);
// Root-free QL/QR eigenvalues of a symmetric tridiagonal, sorted ascending.
void Rsterf(mint n, dd_real *d, dd_real *e, mint &info);
// Eigenvalues (ascending) of a symmetric matrix from its lower triangle
// (Rsyev with jobz 'N', uplo 'L'; a is overwritten).
void Rsyev_NL(mint n, dd_real *a, mint lda, dd_real *w, mint &info);
void Rlartg(dd_real f, dd_real g, dd_real &cs, dd_real &sn, dd_real &r);
}  // namespace orc

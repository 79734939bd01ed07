// Step-length eigenvalues of SDP blocks of size <= 20 and the elementwise
// LP-entry operations of the Newton step.
//
// For a block of size n <= 20 the reference takes lambda_min(L M L^T) from
// all eigenvalues (Jal::getMinEigen, sdpa_jordan.cpp:62-66, and
// Lal::getMinEigenValue, sdpa_linear.cpp:124-156): Rsyev with jobz 'N' and
// uplo 'L', i.e. Rsytrd (unblocked Rsytd2 for these sizes) and the
// root-free QR of Rsterf.  Here one wave takes one block: the two products
// and the Rsymv/Rsyr2 updates of the reduction run one lane per row (each
// element keeps the reference's operation order), the Householder scalars
// and Rsterf run on lane 0.
#pragma once
#include <cstdint>

#include "kernels.hpp"

namespace sdk {

struct SmallEigTask {
    const dd *L;  // n x n lower-triangular factor (column-major, lda n)
    const dd *M;  // n x n symmetric
    int n;        // 1 <= n <= 20
};
constexpr int SMALL_EIG_MAXN = 20;

// out[t] = lambda_min(L_t M_t L_t^T); info[t] = Rsyev's info (0 ok, > 0
// Rsterf did not converge).  T: the Rdot partition of Rsytd2's dot.
void small_min_eig(hipStream_t s, const SmallEigTask *d_tasks, int ntasks, int T, dd *out, int *info);

// LP entries (length n, elementwise):
// ret = a + b s  (Lal::plus's LP loop, sdpa_linear.cpp:1724-1732)
void lp_plus(hipStream_t s, idx_t n, const dd *a, const dd *b, dd sc, dd *ret);
// ret = (a b) c  (Jal::ns_jordan_triple_product's LP part, sdpa_jordan.cpp:200-210)
void lp_triple(hipStream_t s, idx_t n, const dd *a, const dd *b, const dd *c, dd *ret);
// out = a b  (the LP terms of an inner product, before their ordered sum)
void lp_prod(hipStream_t s, idx_t n, const dd *a, const dd *b, dd *out);
// out = (m l) l  (the LP terms of Jal::getMinEigen, sdpa_jordan.cpp:80-85)
void lp_eig(hipStream_t s, idx_t n, const dd *m, const dd *l, dd *out);
// invChol = 1/sqrt(a) and, if inv, inv = 1/a; *info = 1 if some a < 0, else 0
// (Jal::getInvChol / getInvCholAndInv, sdpa_jordan.cpp:140-170).
void lp_inv_chol(hipStream_t s, idx_t n, const dd *a, dd *invChol, dd *inv, int64_t *info);

}  // namespace sdk

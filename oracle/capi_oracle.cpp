// ORACLE — test infrastructure only.
// extern "C" wrappers so tests can call the oracle through ctypes.  Arrays of
// dd values are passed as interleaved (hi, lo) doubles, the dd_real layout.
#include <cstring>
#include <string>
#include <vector>

#include "sdpa_oracle.hpp"

using namespace orc;

static dd_real *D(double *p) { return reinterpret_cast<dd_real *>(p); }
static const dd_real *D(const double *p) { return reinterpret_cast<const dd_real *>(p); }

extern "C" {

void orc_set_dot_threads(int t) { g_dot_threads = t < 1 ? 1 : t; }

// The sparse-Schur path's exact minimum degree on a CSR adjacency; order
// (m ints) gets newToOld.  Returns the factor pattern's entries, diagonal
// included.
long orc_min_degree(int m, const int *ptr, const int *idx, int *order) {
    std::vector<std::vector<int>> adj(m), rows;
    for (int i = 0; i < m; i++) adj[i].assign(idx + ptr[i], idx + ptr[i + 1]);
    std::vector<int> o = min_degree_order(m, adj, rows);
    long nnz = 0;
    for (int i = 0; i < m; i++) {
        order[i] = o[i];
        nnz += 1 + (long)rows[i].size();
    }
    return nnz;
}

// The sparse path's factorisation and solve on a symmetric dd matrix B
// (m x m, column-major): the pattern of B's off-diagonal nonzeros ordered by
// min_degree_order, the values gathered from B as Newton::permuteMat does,
// Lal::getCholesky, then solve_sparse for g.  Writes x (m), order (newToOld)
// and U (m x m, column-major, new indices): the factor pattern's entries
// at (row, column), row <= column, its diagonal holding 1/sqrt(pivot) as
// the reference stores it.  Returns the pattern's entries.
long orc_sparse_factor_solve(int m, const double *B, const double *g, double *x, double *U, int *order) {
    std::vector<std::vector<int>> adj(m), rows;
    const dd_real *b = D(B);
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++)
            if (i != j && (b[i + (size_t)m * j].x[0] != 0.0 || b[i + (size_t)m * j].x[1] != 0.0)) adj[i].push_back(j);
    SparseSchur S;
    S.on = true;
    S.m = m;
    S.ordering = min_degree_order(m, adj, rows);
    S.reverse_ordering.assign(m, 0);
    for (int i = 0; i < m; i++) S.reverse_ordering[S.ordering[i]] = i;
    S.diagonalIndex.assign(m + 1, 0);
    for (int p = 0; p < m; p++) S.diagonalIndex[p + 1] = S.diagonalIndex[p] + 1 + (int)rows[p].size();
    for (int p = 0; p < m; p++) {
        S.row_index.push_back(p);
        S.column_index.push_back(p);
        for (int c : rows[p]) {
            S.row_index.push_back(p);
            S.column_index.push_back(c);
        }
    }
    for (size_t k = 0; k < S.row_index.size(); k++)
        S.sp_ele.push_back(b[S.ordering[S.row_index[k]] + (size_t)m * S.ordering[S.column_index[k]]]);
    chol_sparse(S);
    std::vector<dd_real> gv(D(g), D(g) + m), xv;
    solve_sparse(S, gv, xv);
    for (int i = 0; i < m; i++) {
        D(x)[i] = xv[i];
        order[i] = S.ordering[i];
    }
    for (size_t q = 0; q < (size_t)m * m; q++) D(U)[q] = dd_real(0.0);
    for (size_t k = 0; k < S.row_index.size(); k++) D(U)[S.row_index[k] + (size_t)m * S.column_index[k]] = S.sp_ele[k];
    return (long)S.row_index.size();
}
int orc_get_dot_threads(void) { return g_dot_threads; }

// Elementwise dd ops on n values: op 0 add (a+b), 1 sub, 2 mul, 3 div, 4 sqrt(a),
// 5 in-place add (a+=b), 6 in-place sub, 7 in-place mul, 8 a*double(b.hi),
// 9 a/double(b.hi), 10 a+double(b.hi), 11 double(a.hi)/b.
void orc_dd_op(int op, long n, const double *a, const double *b, double *out) {
    for (long i = 0; i < n; i++) {
        dd_real x = D(a)[i], y = D(b)[i], r;
        switch (op) {
        case 0: r = x + y; break;
        case 1: r = x - y; break;
        case 2: r = x * y; break;
        case 3: r = x / y; break;
        case 4: r = sqrt(x); break;
        case 5: r = x; r += y; break;
        case 6: r = x; r -= y; break;
        case 7: r = x; r *= y; break;
        case 8: r = x * y.x[0]; break;
        case 9: r = x / y.x[0]; break;
        case 10: r = x + y.x[0]; break;
        case 11: r = x.x[0] / y; break;
        default: r = dd_real(0.0);
        }
        D(out)[i] = r;
    }
}

int orc_read_dd(const char *s, double *out) {
    dd_real v;
    int rc = read_dd(s, v);
    out[0] = v.x[0];
    out[1] = v.x[1];
    return rc;
}

void orc_Rgemm(char ta, char tb, long m, long n, long k, const double *alpha, const double *A, long lda,
               const double *B, long ldb, const double *beta, double *C, long ldc) {
    Rgemm(ta, tb, m, n, k, D(alpha)[0], D(A), lda, D(B), ldb, D(beta)[0], D(C), ldc);
}
void orc_Rsyrk_LN(long n, long k, const double *alpha, const double *A, long lda, const double *beta, double *C, long ldc) {
    Rsyrk_LN(n, k, D(alpha)[0], D(A), lda, D(beta)[0], D(C), ldc);
}
void orc_Rtrsm_RLTN(long m, long n, const double *alpha, const double *A, long lda, double *B, long ldb) {
    Rtrsm_RLTN(m, n, D(alpha)[0], D(A), lda, D(B), ldb);
}
void orc_Rtrsm_LLNN(long m, long n, const double *alpha, const double *A, long lda, double *B, long ldb) {
    Rtrsm_LLNN(m, n, D(alpha)[0], D(A), lda, D(B), ldb);
}
void orc_Rtrmm_LLTN(long m, long n, const double *alpha, const double *A, long lda, double *B, long ldb) {
    Rtrmm_LLTN(m, n, D(alpha)[0], D(A), lda, D(B), ldb);
}
void orc_Rtrsv(char trans, long n, const double *A, long lda, double *x) {
    if (trans == 'N') Rtrsv_LN(n, D(A), lda, D(x));
    else Rtrsv_LT(n, D(A), lda, D(x));
}
void orc_Rtrmv(char trans, long n, const double *A, long lda, double *x) {
    if (trans == 'N') Rtrmv_LN(n, D(A), lda, D(x));
    else Rtrmv_LT(n, D(A), lda, D(x));
}
void orc_Rgemv_N(long m, long n, const double *alpha, const double *A, long lda, const double *x, const double *beta, double *y) {
    Rgemv_N(m, n, D(alpha)[0], D(A), lda, D(x), D(beta)[0], D(y));
}
void orc_Rdot(long n, const double *x, long incx, const double *y, long incy, double *out) {
    dd_real r = Rdot(n, D(x), incx, D(y), incy);
    out[0] = r.x[0];
    out[1] = r.x[1];
}
long orc_Rpotrf_L(long n, double *A, long lda) {
    mint info = 0;
    Rpotrf_L(n, D(A), lda, info);
    return (long)info;
}
long orc_Rsteqr_I(long n, double *d, double *e, double *z, long ldz) {
    std::vector<dd_real> work(2 * (n > 1 ? n : 1));
    mint info = 0;
    Rsteqr('I', n, D(d), D(e), D(z), ldz, work.data(), info);
    return (long)info;
}
// Eigenvalues (ascending) of the symmetric matrix whose lower triangle is A.
long orc_Rsyev_NL(long n, double *A, long lda, double *w) {
    mint info = 0;
    Rsyev_NL(n, D(A), lda, D(w), info);
    return (long)info;
}
long orc_Rsterf(long n, double *d, double *e) {
    mint info = 0;
    Rsterf(n, D(d), D(e), info);
    return (long)info;
}
long orc_qr_min_eigen(long n, const double *L, const double *M, double *out) {
    DenseMatrix l, mm;
    l.nRow = l.nCol = (int)n;
    mm.nRow = mm.nCol = (int)n;
    l.de_ele.assign(D(L), D(L) + n * n);
    mm.de_ele.assign(D(M), D(M) + n * n);
    mint info = 0;
    dd_real r = qr_min_eigen(l, mm, info);
    out[0] = r.x[0];
    out[1] = r.x[1];
    return (long)info;
}
void orc_lanczos(long n, const double *L, const double *M, double *out) {
    DenseMatrix l, mm;
    l.nRow = l.nCol = (int)n;
    mm.nRow = mm.nCol = (int)n;
    l.de_ele.assign(D(L), D(L) + n * n);
    mm.de_ele.assign(D(M), D(M) + n * n);
    dd_real r = lanczos_min_eigen(l, mm);
    out[0] = r.x[0];
    out[1] = r.x[1];
}

// Loads a .dat-s file, returns m and the number of blocks; block sizes into bs.
static InputData g_data;
int orc_load(const char *path, int *m, int *nblock, int *bs, int bs_cap, char *err, int errcap) {
    g_data = InputData();
    std::string e;
    if (!read_dat_s(path, g_data, e)) {
        std::strncpy(err, e.c_str(), errcap - 1);
        err[errcap - 1] = 0;
        return -1;
    }
    *m = g_data.m;
    *nblock = (int)g_data.blockStruct.size();
    if (*nblock < bs_cap) bs[*nblock] = -g_data.LP_nBlock;  // LP entry count after the SDP sizes
    for (int i = 0; i < *nblock && i < bs_cap; i++) bs[i] = g_data.blockStruct[i];
    return 0;
}
// useFormula table (m*nBlock, index i*nBlock+l) of the loaded problem.
void orc_formula(int *out) {
    auto f = compute_formula(g_data, dd_real(1.2));
    for (size_t i = 0; i < f.size(); i++) out[i] = f[i];
}
// Schur matrix for the loaded problem with given X and invZ (blocks concatenated).
// forced >= 0 overrides the formula table (F3 only where A_i is sparse).
void orc_schur_forced(const double *X, const double *invZ, double *B, int forced);
void orc_schur(const double *X, const double *invZ, double *B) { orc_schur_forced(X, invZ, B, -1); }
void orc_schur_forced(const double *X, const double *invZ, double *B, int forced) {
    DenseLinearSpace x, z;
    x.initialize(g_data.blockStruct, g_data.LP_nBlock);
    z.initialize(g_data.blockStruct, g_data.LP_nBlock);
    size_t off = 0;
    for (size_t l = 0; l < g_data.blockStruct.size(); l++) {
        size_t len = x.SDP_block[l].de_ele.size();
        std::memcpy(x.SDP_block[l].de_ele.data(), X + 2 * off, len * 16);
        std::memcpy(z.SDP_block[l].de_ele.data(), invZ + 2 * off, len * 16);
        off += len;
    }
    if (g_data.LP_nBlock > 0) {
        std::memcpy(x.LP_block.data(), X + 2 * off, (size_t)g_data.LP_nBlock * 16);
        std::memcpy(z.LP_block.data(), invZ + 2 * off, (size_t)g_data.LP_nBlock * 16);
    }
    std::vector<dd_real> b((size_t)g_data.m * g_data.m, dd_real(0.0));
    std::vector<int> f = compute_formula(g_data, dd_real(1.2));
    if (forced >= 0) {
        const int nB = (int)g_data.blockStruct.size();
        for (int l = 0; l < nB; l++)
            for (size_t k = 0; k < g_data.SDP_constraint[l].size(); k++) {
                int i = g_data.SDP_constraint[l][k];
                bool dense = g_data.A[i].SDP_sp_block[g_data.SDP_blockIndex[l][k]].type == SparseMatrix::DENSE;
                f[i * nB + l] = (forced == 2 && dense) ? 1 : forced;
            }
    }
    schur_dense(g_data, x, z, b, f);
    std::memcpy(B, b.data(), b.size() * 16);
}
}

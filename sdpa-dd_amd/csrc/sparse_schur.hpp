// The sparse-Schur path (the reference's chordal branch): B held as the
// upper-triangular pattern of its factor under a fill-reducing ordering,
// factored by the reference's row Cholesky and solved in its order.
//
// Reference: Chordal::ordering_bMat (sdpa_chordal.cpp:348-440) decides and
// orders (Spooles 2.2's minimum degree there; an exact minimum degree of
// our own here), Newton::initialize_sparse_bMat (sdpa_newton.cpp:225-336)
// lays out the pattern, compute_bMat_sparse_SDP/_LP (1070-1142, 1199-1219)
// fill it, Lal::getCholesky (sdpa_linear.cpp:351-395) factors it and
// Lal::solveSystems (654-790) solves between permuteVec and
// reverse_permuteVec (sdpa_newton.cpp:1257-1275).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "problem.hpp"

namespace sps {

typedef sdk::idx_t idx_t;

// Host analysis of one problem (structure only; values come per solve).
struct Analysis {
    bool on = false;
    int m = 0;
    long long aggregate = 0;  // entries of the aggregate sparsity pattern (self included)
    std::vector<int> ordering, reverse;  // newToOld, oldToNew
    std::vector<int> diag;               // m + 1: start of row p (its diagonal entry)
    std::vector<int> col, row;           // column and row of each entry (rows compressed, columns ascending)
    std::vector<idx_t> gather;           // entry k <- dense B[ordering[row] + m ordering[col]]
    // factorisation, left-looking per entry: levels of rows (a row waits for
    // every row whose pattern holds its index), the entries of each level,
    // and per entry the positions (a, b) of U_{i,c1}, U_{i,c2} of every row
    // i that updates it, i ascending (the order the row loop applies them)
    std::vector<int> lev_ptr, lev_rows, lev_ent_ptr, lev_ents;
    std::vector<int> con_ptr, con_a, con_b;
    // forward solve: per row j the rows i < j with U_ij in the pattern (i
    // ascending) and the position of U_ij; backward levels (a row waits for
    // the rows of its pattern)
    std::vector<int> fw_ptr, fw_row, fw_pos;
    std::vector<int> bw_ptr, bw_rows;
    long long contributions() const { return (long long)con_a.size(); }
};

// mode 0: the reference's rule (m > 100, more than 5 blocks counting LP
// entries, every block's constraint count <= m sqrt(0.25), aggregate
// pattern <= 0.25 m^2, factor pattern 2 nnz - m <= 0.4 m^2); 1: dense;
// 2: the thresholds of the reference's debugging switch (0, 0, 2.0, 2.0,
// sdpa_chordal.cpp:53-58), i.e. sparse whenever the structure allows.
// Returns on = false for the dense path.
void analyse(const sdp::HostProblem &P, int mode, Analysis &S);

// Exact minimum degree on an adjacency (lists may include the vertex
// itself): repeatedly eliminate the vertex with the fewest remaining
// neighbours, ties to the smaller index; its neighbours become a clique.
// Returns newToOld; rows[p] = new indices of the p-th vertex's neighbours
// at its elimination, ascending.
std::vector<int> min_degree(int m, const std::vector<std::vector<int>> &adj, std::vector<std::vector<int>> &rows);

// Device image of an analysis.
struct Device {
    bool on = false;
    int m = 0, nnz = 0, nlev = 0, nblev = 0;
    int *d_ordering = nullptr, *d_diag = nullptr, *d_col = nullptr, *d_row = nullptr;
    idx_t *d_gather = nullptr;
    int *d_lev_ptr = nullptr, *d_lev_rows = nullptr, *d_lev_ent_ptr = nullptr, *d_lev_ents = nullptr;
    int *d_con_ptr = nullptr, *d_con_a = nullptr, *d_con_b = nullptr;
    int *d_fw_ptr = nullptr, *d_fw_row = nullptr, *d_fw_pos = nullptr;
    int *d_bw_ptr = nullptr, *d_bw_rows = nullptr;
    std::vector<void *> allocs;
};
int upload(const Analysis &S, Device &D, hipStream_t s, std::string &err);
void release(Device &D);

// sp[k] = B[gather[k]]  (Newton::permuteMat: B assembled dense in the
// reference's pair order holds the values the sparse assembly adds)
void gather(hipStream_t s, const Device &D, const dd *B, dd *sp);
// Lal::getCholesky on the pattern (a negative pivot becomes 0, as there)
void factor(hipStream_t s, const Device &D, dd *sp);
// dy = B^{-1} g through permuteVec, the two sweeps and reverse_permuteVec;
// x: m dd of workspace
void solve(hipStream_t s, const Device &D, const dd *sp, const dd *g, dd *dy, dd *x);

}  // namespace sps

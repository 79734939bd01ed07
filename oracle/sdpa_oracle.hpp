// ORACLE — test infrastructure only.
//
// CPU restatement of SDPA-DD 7.1.3's SDP path: the .dat-s reader, the
// parameter file, and the Mehrotra predictor-corrector loop with its HKM
// Newton step, Schur assembly (F1/F2/F3 for SDP blocks, the diagonal form
// for LP entries) into a dense B or into the sparse pattern of its factor,
// Cholesky solve (dense, or the sparse row Cholesky of the chordal path),
// step length (Lanczos for SDP blocks of size > 20, Rsyev's tridiagonal
// reduction and root-free QR for smaller ones, ratios for LP entries) and
// phase logic.  The sparse path's ordering is an exact minimum degree in
// place of the reference's Spooles 2.2.  SOCP blocks are not restated;
// inputs with them are rejected.
#pragma once
#include <string>
#include <vector>

#include "mplapack_oracle.hpp"

namespace orc {

// sdpa_struct.h:71-108
struct SparseMatrix {
    int nRow = 0, nCol = 0;
    enum Type { SPARSE, DENSE } type = SPARSE;
    int NonZeroNumber = 0, NonZeroCount = 0, NonZeroEffect = 0;
    std::vector<dd_real> de_ele;
    std::vector<int> row_index, column_index;
    std::vector<dd_real> sp_ele;
    void changeToDense();
};
struct DenseMatrix {
    int nRow = 0, nCol = 0;
    std::vector<dd_real> de_ele;
    void initialize(int n) {
        nRow = nCol = n;
        de_ele.assign((size_t)n * n, dd_real(0.0));
    }
    void setIdentity(dd_real s) {
        for (auto &v : de_ele) v = 0.0;
        for (int i = 0; i < nRow; i++) de_ele[(size_t)i * nRow + i] = s;
    }
};
// sdpa_struct.h: SDP blocks plus the LP part (one scalar per LP entry; the
// LP entries of a matrix are stored at ascending LP_sp_index).
struct SparseLinearSpace {
    std::vector<int> SDP_sp_index;
    std::vector<SparseMatrix> SDP_sp_block;
    std::vector<int> LP_sp_index;
    std::vector<dd_real> LP_sp_block;
};
struct DenseLinearSpace {
    std::vector<DenseMatrix> SDP_block;
    std::vector<dd_real> LP_block;
    void initialize(const std::vector<int> &bs, int nLP) {
        SDP_block.resize(bs.size());
        for (size_t l = 0; l < bs.size(); l++) SDP_block[l].initialize(bs[l]);
        LP_block.assign(nLP, dd_real(0.0));
    }
    void setZero() {
        for (auto &b : SDP_block)
            for (auto &v : b.de_ele) v = 0.0;
        for (auto &v : LP_block) v = 0.0;
    }
    void setIdentity(dd_real s) {
        for (auto &b : SDP_block) b.setIdentity(s);
        for (auto &v : LP_block) v = s;
    }
};

struct Parameter {
    int maxIteration = 200;
    double epsilonStar = 1.0e-30, lambdaStar = 1.0e+4, omegaStar = 2.0, lowerBound = -1.0e+5, upperBound = 1.0e+5,
           betaStar = 0.1, betaBar = 0.3, gammaStar = 0.9, epsilonDash = 1.0e-30;
    bool readFile(const std::string &path);
};

struct InputData {
    int m = 0;
    std::vector<int> blockStruct;  // SDP blocks only
    int LP_nBlock = 0;             // LP entries (all LP blocks and size-1 blocks)
    std::vector<dd_real> b;
    SparseLinearSpace C;
    std::vector<SparseLinearSpace> A;
    std::vector<std::vector<int>> SDP_constraint, SDP_blockIndex;
    std::vector<std::vector<int>> LP_constraint, LP_blockIndex;
    void initialize_index();
};

// Reads an SDPA sparse-format (.dat-s) file (sdpa_io.cpp:124-150, 429-786).
// Returns false with a message in err on unsupported or malformed input.
bool read_dat_s(const std::string &path, InputData &data, std::string &err);

enum PhaseType { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

struct IterRecord {
    int it;
    dd_real mu, thetaP, thetaD, objP, objD, alphaP, alphaD, beta;
    double seconds = 0.0;  // wall time since the solve started, when the record was taken
};

struct SolveResult {
    int iterations = 0;
    int phase = noINFO;  // internal phase before the primal/dual swap
    int phase_reported = noINFO;  // after Phase::reverse (what the result file prints)
    dd_real mu, objValPrimal, objValDual, normPrimalVec, normDualMat, relgap;
    std::vector<dd_real> xMat, zMat, yVec;  // concatenated SDP blocks (column-major), then the LP entries
    std::vector<IterRecord> trace;
    std::string message;
    bool schur_sparse = false;          // the sparse-Schur path ran
    std::vector<int> schur_ordering;    // its newToOld ordering
    long long schur_nnz = 0;            // entries of its factor pattern
};

// An initial point (sdpa_main.cpp:237-251): X and Z over the whole space
// (SDP blocks concatenated column-major, then the LP entries), y in the
// solver's sign (the file's y negated).
struct InitPoint {
    std::vector<dd_real> x, z, y;
};
// Reads an initial-point file the way IO::read does (sdpa_io.cpp:152-285):
// X = Z = lambdaStar I first, then the file's y (negated) and entries
// (sparse: "target block i j value", target 1 = Z, else X; dense: Z's
// blocks then X's, full n x n each, then the LP entries).  A value token
// may also be the exact pair HI+LO / HI-LO of two doubles (the full-dd dump
// format; QD's reader stops after HI's exponent).
bool read_init(const std::string &path, const InputData &data, const Parameter &param, bool sparse, InitPoint &ip,
               std::string &err);

// The sparse-Schur path (the chordal branch of the reference): B in the
// upper-triangular pattern of its factor under the ordering newToOld
// (ordering), rows compressed (diagonalIndex), each row's diagonal first
// and its columns ascending.
struct SchurPair {
    int i, j, k1, k2, loc;  // constraints, their positions in the block's constraint list, entry of the pattern
};
struct SparseSchur {
    bool on = false;
    int m = 0;
    std::vector<int> ordering, reverse_ordering, diagonalIndex, row_index, column_index;
    std::vector<dd_real> sp_ele;
    std::vector<std::vector<SchurPair>> sdp, lp;  // per SDP block, per LP entry
};
// 0: the reference's rule; 1: always dense; 2: sparse whenever the
// pattern allows (the thresholds of the reference's debugging switch).
extern int g_schur_mode;
bool sparse_schur_setup(const InputData &d, SparseSchur &S, int mode);
std::vector<int> min_degree_order(int m, const std::vector<std::vector<int>> &adj, std::vector<std::vector<int>> &rows);
void schur_sparse(const InputData &data, const DenseLinearSpace &X, const DenseLinearSpace &invZ, SparseSchur &S,
                  const std::vector<int> &useFormula);
void chol_sparse(SparseSchur &S);
void solve_sparse(const SparseSchur &S, const std::vector<dd_real> &g, std::vector<dd_real> &dy);

// Runs the solver from lambdaStar I, or from *init.
bool solve(const InputData &data, const Parameter &param, SolveResult &res, const InitPoint *init = nullptr);

// Pieces exposed to tests.
void schur_dense(const InputData &data, const DenseLinearSpace &X, const DenseLinearSpace &invZ,
                 std::vector<dd_real> &B, const std::vector<int> &useFormula);
std::vector<int> compute_formula(const InputData &data, dd_real kappa);
dd_real lanczos_min_eigen(const DenseMatrix &L, const DenseMatrix &M);
// Smallest eigenvalue of L M L^T by Rsyev ('N', 'L'); info as Rsyev's.
dd_real qr_min_eigen(const DenseMatrix &L, const DenseMatrix &M, mint &info);

}  // namespace orc

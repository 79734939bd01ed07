// Host-side problem representation (the reference's InputData in SDPA's
// internal form) and its device image used by the Newton-step kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "ddr.hpp"
#include "kernels.hpp"

namespace sdp {

typedef sdk::idx_t idx_t;

// One SDP block of one matrix (SparseMatrix, sdpa_struct.h:71-108).
struct Block {
    int n = 0;
    bool dense = false;
    int nnz_count = 0;   // NonZeroCount
    int nnz_effect = 0;  // NonZeroEffect
    std::vector<int> row, col;
    std::vector<dd> val;   // sparse values, file order
    std::vector<dd> dval;  // dense values n*n column-major
    void change_to_dense();
};
struct Space {
    std::vector<int> index;  // SDP block numbers present, ascending
    std::vector<Block> blk;
    // LP entries (SparseLinearSpace::LP_sp_index / LP_sp_block): ascending
    // indices into the LP part; raw (index, value) pairs in input order until
    // finish_lp() sorts them (a repeated index keeps its last value).
    std::vector<int> lp_index;
    std::vector<dd> lp_val;
    std::vector<std::pair<int, dd>> lp_raw;
    void finish_lp();
    int find(int l) const {
        for (size_t k = 0; k < index.size(); k++)
            if (index[k] == l) return (int)k;
        return -1;
    }
};

struct Params {
    int maxIteration = 200;
    double epsilonStar = 1.0e-30, lambdaStar = 1.0e+4, omegaStar = 2.0, lowerBound = -1.0e+5, upperBound = 1.0e+5,
           betaStar = 0.1, betaBar = 0.3, gammaStar = 0.9, epsilonDash = 1.0e-30;
};
bool read_params(const std::string &path, Params &p);

struct HostProblem {
    int m = 0;
    // Block structure as given (sdpa_main.cpp:111-130): a size >= 2 is an SDP
    // block, a negative size -k (or the size 1) adds k (or 1) LP entries.
    std::vector<int> given;
    std::vector<int> gtype;     // 1 SDP, 3 LP
    std::vector<int> gnum;      // SDP block number, or the first LP index
    std::vector<int> bs;        // SDP block sizes
    std::vector<idx_t> boff;    // offset of block l in a concatenated dense space
    idx_t nn = 0;               // sum of n_l^2
    int nlp = 0;                // LP entries, stored after the SDP blocks (at nn)
    idx_t nt() const { return nn + nlp; }
    std::vector<dd> b;
    Space C;
    std::vector<Space> A;
    std::vector<std::vector<int>> cons, cidx;  // SDP_constraint / SDP_blockIndex per block
    std::vector<std::vector<int>> lp_cons, lp_cidx;  // LP_constraint / LP_blockIndex per LP entry
    std::vector<int> formula;                  // useFormula[i*nBlock + l]
    // Sets given/gtype/gnum/bs/nlp from a given block structure; false on a 0.
    bool set_structure(const std::vector<int> &given_bs);
    void layout();
    void index();
};

bool read_dat_s(const std::string &path, HostProblem &p, std::string &err);

// An initial point (sdpa_main.cpp:237-251).  Either a full point (X and Z
// over the whole space -- SDP blocks, then LP entries -- and y in the
// solver's sign), or the settings an initial-point file makes over
// X = Z = lambdaStar I, applied when a solve knows its lambdaStar.
struct InitPoint {
    bool full = false;
    std::vector<dd> x, z, y;
    struct Set {
        int target;  // 1: Z, otherwise X
        idx_t a, b;  // positions set (b == a on the diagonal and for LP entries)
        dd v;
    };
    std::vector<Set> sets;
    void build(const HostProblem &P, double lambda, std::vector<dd> &X, std::vector<dd> &Y, std::vector<dd> &Z) const;
};
// IO::read for initial points (sdpa_io.cpp:152-285): y (negated), then
// sparse entries "target block i j value" (block <= the SDP block count: an
// SDP block; beyond it: LP entry i) or dense Z then X.  A value may also be
// the exact pair HI+LO / HI-LO of two doubles, as write_point writes it
// (QD's reader stops after HI's exponent, so the reference reads HI).
bool read_init(const std::string &path, const HostProblem &P, bool sparse, InitPoint &ip, std::string &err);
// Writes (X, y, Z) as a sparse initial-point file with full-dd values:
// every diagonal entry, the nonzero entries above it and every LP entry.
bool write_point(const std::string &path, const HostProblem &P, const std::vector<dd> &X, const std::vector<dd> &y,
                 const std::vector<dd> &Z, std::string &err);
bool qd_read(const char *s, dd &out);
std::vector<int> compute_formula(const HostProblem &p, ddr kappa);
// Algorithmic dd operations (multiplies and adds) of one Schur assembly.
double schur_ddflops(const HostProblem &p);

// ---------------------------------------------------------------- device --
// Pairs of the Schur assembly for one block and their formula data.
struct SchurBlock {
    int n = 0;
    // constraints of this block in k1 order
    std::vector<int> ci;             // constraint index i
    std::vector<int> form;           // formula of (i, l)
    // pairs (i, j) evaluated in row i's loop, B[i,j] (+ B[j,i]) += value
    int npairs = 0;
    int *d_pair_i = nullptr, *d_pair_j = nullptr;  // constraint ids
    int *d_pair_slot_i = nullptr;                  // position of i in ci (work matrix slot)
    int *d_pair_slot_j = nullptr;                  // position of j in ci
    int *d_pair_form = nullptr;
    std::vector<int> h_pair_form;
    // pair ids with a sparse A_j, by kernel: F2 rows (one wave per pair) and
    // F1/F3 rows (one thread per pair)
    int n_f2 = 0, n_other = 0;
    int *d_list_f2 = nullptr, *d_list_other = nullptr;
    // per-constraint block data on device (index = slot)
    int *d_type = nullptr;       // 0 sparse, 1 dense
    int *d_nnz = nullptr;        // NonZeroCount
    idx_t *d_off = nullptr;      // sparse entry offset or dense value offset
    int *d_row = nullptr, *d_col = nullptr;
    dd *d_val = nullptr;         // sparse values
    dd *d_dense = nullptr;       // dense values
    // A_i * Zinv row lists for sparse constraints using F1/F2: per slot,
    // per result row r, entries (source row, value) in application order.
    idx_t *d_rl_ptr = nullptr;   // [nslots*(n+1)] (row pointer per slot)
    int *d_rl_src = nullptr;
    dd *d_rl_val = nullptr;
    int *d_wslot = nullptr;      // work-matrix slot of each constraint slot (-1: none)
    // slot lists per formula
    std::vector<int> slots_f1, slots_f2, slots_f3;
    // constraint slots whose A_i*Zinv is built from row lists (sparse) or by
    // GEMM (dense), and slots that need X*work1 (F1, and F2 with dense pairs)
    std::vector<int> sp_slots, de_slots, g_slots;
    int *d_sp_slots = nullptr, *d_de_slots = nullptr, *d_g_slots = nullptr;
    // dot tasks for pairs whose A_j is dense
    int ndense_pairs = 0;
    sdk::DotTask *d_pair_tasks = nullptr;
};

struct DeviceProblem {
    int m = 0;
    std::vector<int> bs;
    std::vector<idx_t> boff;
    idx_t nn = 0;
    int nlp = 0;  // LP entries at [nn, nn + nlp) of every iterate
    // LP entries of each A_i (CSR over constraints, ascending LP index) and
    // of C, added after the SDP blocks of an inner product
    idx_t *d_lp_ptr = nullptr, *d_clp_ptr = nullptr;
    int *d_lp_idx = nullptr, *d_clp_idx = nullptr;
    dd *d_lp_val = nullptr, *d_clp_val = nullptr;
    // LP part of the Schur matrix: per constraint pair (i <= j), the LP
    // entries both touch, ascending, with A_i's and A_j's values
    int n_lp_pairs = 0;
    int *d_lpp_i = nullptr, *d_lpp_j = nullptr;
    idx_t *d_lpp_ptr = nullptr;
    int *d_lpp_l = nullptr;
    dd *d_lpp_ai = nullptr, *d_lpp_aj = nullptr;
    dd *d_b = nullptr;
    // inner products A_i . W: tasks per constraint in block order
    int ntask = 0;
    idx_t *d_task_ptr = nullptr;        // [m+1] tasks of constraint i
    int *d_task_type = nullptr;         // 0 sparse, 1 dense
    int *d_task_block = nullptr;
    int *d_task_nnz = nullptr;
    idx_t *d_task_off = nullptr;
    std::vector<int> h_task_type;
    // C . W tasks (same format, one "constraint")
    int nctask = 0;
    idx_t *d_ctask_ptr = nullptr;
    int *d_ctask_type = nullptr, *d_ctask_block = nullptr, *d_ctask_nnz = nullptr;
    idx_t *d_ctask_off = nullptr;
    std::vector<int> h_ctask_type;
    int *d_sp_row = nullptr, *d_sp_col = nullptr;  // all sparse entries (A then C)
    dd *d_sp_val = nullptr;
    dd *d_dense = nullptr;                         // all dense blocks (A then C)
    // element update lists for sum_i y_i A_i (per element of the concatenated space)
    idx_t *d_el_ptr = nullptr;
    int *d_el_idx = nullptr;
    dd *d_el_val = nullptr;
    idx_t *d_cel_ptr = nullptr;  // element lists of C
    dd *d_cel_val = nullptr;
    idx_t *d_boff = nullptr;
    int *d_bsize = nullptr;
    // dense inner-product tasks (y offset relative to the W passed in)
    int n_dense_tasks = 0, n_dense_ctasks = 0;
    idx_t dense_task_maxn = 0, dense_ctask_maxn = 0;  // longest dense task
    sdk::DotTaskOff *d_dense_tasks = nullptr, *d_dense_ctasks = nullptr;
    dd *d_dot_partial = nullptr;  // chunk sums of long dense tasks (grown on demand)
    size_t dot_partial_cap = 0;
    std::vector<SchurBlock> schur;
    // pointer arrays for batched GEMMs
    void *d_ptrs = nullptr;
    size_t ptr_cap = 0;
    // scratch for batched work matrices
    size_t work_elems = 0;
    dd *d_work1 = nullptr, *d_work2 = nullptr;
    dd *d_pair_val = nullptr;
    // one SDP block: pair id of each element of B (-1: none), column-major
    int *d_pidx = nullptr;
    dd *d_task_val = nullptr;
    dd *d_ctask_val = nullptr;
    std::vector<void *> allocs;
};

int upload_problem(const HostProblem &hp, DeviceProblem &dp, hipStream_t s, std::string &err);
void free_problem(DeviceProblem &dp);

// out[i] = sum over blocks of A_i . W (the reference's multi_InnerProductToA).
void inner_A(hipStream_t s, DeviceProblem &dp, const dd *W, dd *out, int T, bool fast = false);
// *out = C . W
void inner_C(hipStream_t s, DeviceProblem &dp, const dd *W, dd *out, int T, bool fast = false);
// R = sum_i y_i A_i (multi_plusToA)
void plus_A(hipStream_t s, DeviceProblem &dp, const dd *y, dd *R);
// R += C
void plus_C(hipStream_t s, DeviceProblem &dp, dd *R);
// B = Schur complement for (X, invZ) (B zeroed first).  lp_larger_first:
// the LP terms x z^-1 a_i a_j with i the later constraint of each pair (the
// sparse-Schur path's order) instead of the earlier one.
void schur(hipStream_t s, DeviceProblem &dp, const dd *X, const dd *invZ, dd *B, int T, bool lp_larger_first = false);

}  // namespace sdp

// Library handle internals: device, stream and reusable scratch.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.hpp"
#include "ozaki.hpp"

namespace sdp {

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int dot_T = 64;             // chunks of the static dot partition
    dd *d_dot_work = nullptr;   // >= dot_T dd
    unsigned *d_counter = nullptr;
    dd *d_scal = nullptr;       // scalar slots for dot results
    int64_t *d_info = nullptr;
    dd *d_rinv = nullptr;       // Cholesky scratch (sdk::POTRF_SCRATCH_DD dd)
    dd *d_max_work = nullptr;   // max-abs partials (sdk::MAX_ABS_BLOCKS)
    oz::Work ozw;               // digit planes of the Ozaki products
    // Solver form: 0 = plain (the reference's dd operation order); 1 = Ozaki
    // Cholesky (trailing updates of the Schur and X/Z Cholesky on INT8 MFMA
    // with oz_slices digit planes, up to refine_steps sweeps of dd iterative
    // refinement of dy); 2 = as 1 plus the n x n GEMMs of the direction.
    int path = 0;
    // workspace of the cooperative triangular solves (grown on demand)
    void *d_tsw = nullptr;
    size_t tsw_cap = 0;
    // Rtrsv on stream s, stream-ordered: a grid of workgroups for n >= 512,
    // one workgroup otherwise or when the workspace or the launch is
    // unavailable (the same operation order either way).
    void trsv(hipStream_t s, char trans, sdk::idx_t n, const dd *L, sdk::idx_t lda, dd *x);
    // The Ozaki form's block-inverse solves: operators of the factor L (both
    // directions) formed by binv_prepare, used by trsv_binv for that n.
    dd *d_binv = nullptr;
    size_t binv_cap = 0;
    sdk::idx_t binv_n = -1;
    dd *d_binv_lt = nullptr;  // L^T of the factor the block operators belong to
    size_t binv_lt_cap = 0;
    int binv_prepare(hipStream_t s, sdk::idx_t n, const dd *L, sdk::idx_t lda);
    void trsv_binv(hipStream_t s, char trans, sdk::idx_t n, const dd *L, sdk::idx_t lda, dd *x);
    int oz_slices = 16;
    int refine_steps = 2;
    // general workspace of the BLAS-level entry points (grown on demand,
    // kept until the handle is destroyed)
    void *d_ws = nullptr;
    size_t ws_cap = 0;
    void *workspace(size_t bytes);
    std::string last_error;

    int init(int dev);
    void destroy();
    // Fixed-partition dot into scalar slot `slot` (device).
    void dot(sdk::idx_t n, const dd *x, const dd *y, int slot) {
        sdk::dot(stream, n, x, y, d_scal + slot, dot_T, d_dot_work, d_counter);
    }
    void set_error(const std::string &e) { last_error = e; }
};

}  // namespace sdp

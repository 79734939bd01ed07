// extern "C" entry points of libsdpadd_hip.so (declared in include/sdpadd.h).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include <mutex>

#include "../../include/sdpadd.h"
#include "adjoint.hpp"
#include "context.hpp"
#include "smallblk.hpp"
#include "solver.hpp"
#include "steqr.hpp"

using sdk::idx_t;

struct sdpadd_ctx {
    sdp::Context c;
};
struct sdpadd_problem_s {
    sdpadd_ctx *h = nullptr;
    sdp::HostProblem hp;
    sdp::DeviceProblem dp;
    bool finalized = false;
    std::vector<sdp::IterRecord> trace;
    std::vector<dd> X, y, Z;
    sdp::Adjoint adj;
    sdp::InitPoint init;
    bool has_init = false;
    // the Schur path: mode (0 the reference's rule, 1 dense, 2 sparse when
    // the structure allows), its analysis and device image (for sps_mode)
    int schur_mode = 0, sps_mode = -1;
    sps::Analysis spa;
    sps::Device spd;
};

// Scratch of the allocation-free entry points: a fixed device array
// (each GPU holds its own copy, kept until the process ends).  Host code
// uses its address from hipGetSymbolAddress, taken once per device at
// sdpadd_create with that device current, never the array's name.
constexpr size_t NOALLOC_DD = size_t(1) << 20;  // 16 MiB of dd
static __device__ __align__(256) dd scratch[NOALLOC_DD];
constexpr int NOALLOC_MAX_DEVICES = 64;
static dd *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

namespace {
inline dd D(const double *p) { return dd{p[0], p[1]}; }
inline dd *DP(double *p) { return reinterpret_cast<dd *>(p); }
inline const dd *DP(const double *p) { return reinterpret_cast<const dd *>(p); }
inline bool up(char c, char u) { return c == u || c == (char)(u + 32); }
int check(sdpadd_handle h) {
    hipError_t e = hipStreamSynchronize(h->c.stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        h->c.set_error(hipGetErrorString(e));
        return -3;
    }
    return 0;
}
int bad(sdpadd_handle h, const char *msg) {
    if (h) h->c.set_error(msg);
    return -1;
}
}  // namespace

extern "C" {

int sdpadd_create(int device, sdpadd_handle *h) {
    if (!h) return -1;
    auto *ctx = new sdpadd_ctx();
    if (ctx->c.init(device) != 0) {
        delete ctx;
        *h = nullptr;
        return -3;
    }
    if (device >= 0 && device < NOALLOC_MAX_DEVICES) {
        // init made `device` current
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess)
                g_scratch_ptr[device] = static_cast<dd *>(ptr);
            else
                (void)hipGetLastError();  // the allocation-free entries then report it
        }
    }
    *h = ctx;
    return 0;
}
int sdpadd_destroy(sdpadd_handle h) {
    if (!h) return -1;
    h->c.destroy();
    delete h;
    return 0;
}
const char *sdpadd_last_error(sdpadd_handle h) { return h ? h->c.last_error.c_str() : "null handle"; }
int sdpadd_set_dot_partition(sdpadd_handle h, int chunks) {
    if (!h || chunks < 1 || chunks > 2048) return bad(h, "dot partition must be in [1, 2048]");
    h->c.dot_T = chunks;
    return 0;
}
int sdpadd_set_path(sdpadd_handle h, int path, int slices, int refine_steps) {
    if (!h) return -1;
    if (path < 0 || path > 2) return bad(h, "path must be 0 (plain), 1 (Ozaki Cholesky) or 2 (Ozaki Cholesky and GEMMs)");
    if (slices != 8 && slices != 12 && slices != 14 && slices != 16 && (slices < -7 || slices > -4))
        return bad(h, "slices must be 8, 12, 14 or 16 (INT8 planes) or -4 .. -7 (FP64 slices)");
    if (refine_steps < 0 || refine_steps > 8) return bad(h, "refine_steps must be in [0, 8]");
    h->c.path = path;
    h->c.oz_slices = slices;
    h->c.refine_steps = refine_steps;
    return 0;
}
int sdpadd_set_tile_policy(sdpadd_handle h, int policy) {
    if (!h || policy < 0 || policy > 2) return bad(h, "tile policy must be 0 (by size), 1 (32 x 32) or 2 (64 x 64)");
    sdk::g_tile_policy = policy;
    return 0;
}
int sdpadd_set_wait_bound(sdpadd_handle h, unsigned spins) {
    if (!h) return bad(h, "set_wait_bound: null handle");
    sdk::g_trsv_spin = spins;
    return 0;
}
int sdpadd_set_panel_wait_bound(sdpadd_handle h, unsigned polls) {
    if (!h) return bad(h, "set_panel_wait_bound: null handle");
    sdk::g_panel_spin = polls;
    return 0;
}
int sdpadd_dd_vec_op(sdpadd_handle h, int op, int64_t n, const double *d_a, const double *d_b, double *d_out) {
    if (!h || op < 0 || op > 11 || n < 0 || (n > 0 && (!d_a || !d_b || !d_out))) return bad(h, "dd_vec_op: bad argument");
    sdk::dd_vec_op(h->c.stream, op, n, DP(d_a), DP(d_b), DP(d_out));
    return check(h);
}
int sdpadd_stats_enable(sdpadd_handle h, int on) {
    if (!h) return -1;
    sdk::stats_enable(on != 0);
    return 0;
}
int sdpadd_stats_filter(sdpadd_handle h, const char *families) {
    if (!h) return -1;
    sdk::stats_filter(families);
    return 0;
}
int sdpadd_stats_reset(sdpadd_handle h) {
    if (!h) return -1;
    sdk::stats_reset();
    return 0;
}
int sdpadd_stats_get(sdpadd_handle h, const char *family, int64_t *launches, double *ms, double *ddflops) {
    if (!h || !family) return -1;
    sdk::stats_get(family, launches, ms, ddflops);
    return 0;
}
int sdpadd_malloc(sdpadd_handle h, size_t bytes, void **d_ptr) {
    if (!h || !d_ptr) return -1;
    if (hipMalloc(d_ptr, bytes ? bytes : 16) != hipSuccess) return bad(h, "hipMalloc failed");
    return 0;
}
int sdpadd_free(sdpadd_handle h, void *d_ptr) {
    if (!h) return -1;
    return hipFree(d_ptr) == hipSuccess ? 0 : bad(h, "hipFree failed");
}
int sdpadd_upload(sdpadd_handle h, void *d_dst, const void *src, size_t bytes) {
    if (!h) return -1;
    if (hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return bad(h, "upload failed");
    return 0;
}
int sdpadd_download(sdpadd_handle h, void *dst, const void *d_src, size_t bytes) {
    if (!h) return -1;
    if (hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost) != hipSuccess) return bad(h, "download failed");
    return 0;
}

int sdpadd_Rgemm(sdpadd_handle h, char transa, char transb, int64_t m, int64_t n, int64_t k, const double *alpha,
                 const double *d_A, int64_t lda, const double *d_B, int64_t ldb, const double *beta, double *d_C,
                 int64_t ldc) {
    if (!h || !up(transa, 'N') || !(up(transb, 'N') || up(transb, 'T'))) return bad(h, "Rgemm: unsupported op");
    if (!alpha || !beta) return bad(h, "Rgemm: alpha and beta must not be null");
    if (m < 0 || n < 0 || k < 0 || lda < std::max<int64_t>(1, m) || ldc < std::max<int64_t>(1, m) ||
        ldb < std::max<int64_t>(1, up(transb, 'N') ? k : n))
        return bad(h, "Rgemm: bad dimension");
    sdk::gemm(h->c.stream, up(transb, 'T') ? 'T' : 'N', m, n, k, D(alpha), DP(d_A), lda, DP(d_B), ldb, D(beta), DP(d_C), ldc);
    return check(h);
}
int sdpadd_ozaki_gemm(sdpadd_handle h, char transa, char transb, int64_t m, int64_t n, int64_t k, int alpha,
                      const double *d_A, int64_t lda, const double *d_B, int64_t ldb, int beta, double *d_C, int64_t ldc,
                      int lower, int slices) {
    if (!h) return -1;
    if (m < 0 || n < 0 || k < 0 || lda < std::max<int64_t>(1, up(transa, 'N') ? m : k) || ldc < std::max<int64_t>(1, m) ||
        ldb < std::max<int64_t>(1, up(transb, 'N') ? k : n))
        return bad(h, "ozaki_gemm: bad dimension");
    int rc = oz::gemm(h->c.stream, h->c.ozw, transa, transb, m, n, k, alpha, DP(d_A), lda, DP(d_B), ldb, beta, DP(d_C),
                      ldc, lower != 0, slices);
    if (rc == -1)
        return bad(h, "ozaki_gemm: unsupported argument (alpha +-1, beta 0/1, k <= 8192, slices 8/12/14/16 or -4..-7)");
    if (rc != 0) return bad(h, "ozaki_gemm: out of device memory");
    return check(h);
}
int sdpadd_Rsyrk(sdpadd_handle h, char uplo, char trans, int64_t n, int64_t k, const double *alpha, const double *d_A,
                 int64_t lda, const double *beta, double *d_C, int64_t ldc) {
    if (!h || !up(uplo, 'L') || !up(trans, 'N')) return bad(h, "Rsyrk: unsupported op");
    if (!alpha || !beta) return bad(h, "Rsyrk: alpha and beta must not be null");
    if (n < 0 || k < 0 || lda < std::max<int64_t>(1, n) || ldc < std::max<int64_t>(1, n)) return bad(h, "Rsyrk: bad dimension");
    // Rsyrk skips columns with a[j,l] == 0; adding those zero products leaves
    // every value unchanged, so the GEMM order gives the same bits.
    sdk::syrk_ln(h->c.stream, n, k, D(alpha), DP(d_A), lda, D(beta), DP(d_C), ldc);
    return check(h);
}
int sdpadd_Rtrsm(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n, const double *alpha,
                 const double *d_A, int64_t lda, double *d_B, int64_t ldb) {
    if (!h) return -1;
    if (!alpha) return bad(h, "Rtrsm: alpha must not be null");
    if (!(alpha[0] == 1.0 && alpha[1] == 0.0) || !up(uplo, 'L') || !up(diag, 'N')) return bad(h, "Rtrsm: unsupported op");
    // A is m x m for side 'L', n x n for side 'R' (Rtrsm.cpp:95-112)
    if (m < 0 || n < 0 || ldb < std::max<int64_t>(1, m) || lda < std::max<int64_t>(1, up(side, 'L') ? m : n))
        return bad(h, "Rtrsm: bad dimension");
    if (up(side, 'R') && up(transa, 'T'))
        sdk::trsm_rltn(h->c.stream, m, n, DP(d_A), lda, DP(d_B), ldb);
    else if (up(side, 'L') && up(transa, 'N'))
        sdk::trsm_llnn(h->c.stream, m, n, DP(d_A), lda, DP(d_B), ldb);
    else
        return bad(h, "Rtrsm: unsupported op");
    return check(h);
}
int sdpadd_Rtrmm(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n, const double *alpha,
                 const double *d_A, int64_t lda, double *d_B, int64_t ldb) {
    if (!h) return -1;
    if (!alpha) return bad(h, "Rtrmm: alpha must not be null");
    if (!(alpha[0] == 1.0 && alpha[1] == 0.0) || !up(side, 'L') || !up(uplo, 'L') || !up(transa, 'T') || !up(diag, 'N'))
        return bad(h, "Rtrmm: unsupported op");
    if (m < 0 || n < 0 || ldb < std::max<int64_t>(1, m) || lda < std::max<int64_t>(1, m)) return bad(h, "Rtrmm: bad dimension");
    if (m == 0 || n == 0) return 0;
    dd *tmp = static_cast<dd *>(h->c.workspace(sizeof(dd) * (size_t)m * (size_t)n));
    if (!tmp) return bad(h, "Rtrmm: workspace allocation failed");
    sdk::trmm_lltn(h->c.stream, m, n, DP(d_A), lda, DP(d_B), ldb, tmp, m);
    (void)hipMemcpy2DAsync(d_B, sizeof(dd) * ldb, tmp, sizeof(dd) * m, sizeof(dd) * m, n, hipMemcpyDeviceToDevice,
                           h->c.stream);
    return check(h);
}
int sdpadd_Rtrmm_noalloc(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                         const double *alpha, const double *d_A, int64_t lda, double *d_B, int64_t ldb) {
    if (!h) return -1;
    if (!alpha) return bad(h, "Rtrmm_noalloc: alpha must not be null");
    if (!(alpha[0] == 1.0 && alpha[1] == 0.0) || !up(side, 'L') || !up(uplo, 'L') || !up(transa, 'T') || !up(diag, 'N'))
        return bad(h, "Rtrmm_noalloc: unsupported op");
    if (m < 0 || n < 0 || ldb < std::max<int64_t>(1, m) || lda < std::max<int64_t>(1, m))
        return bad(h, "Rtrmm_noalloc: bad dimension");
    if ((uint64_t)m * (uint64_t)n > NOALLOC_DD) return bad(h, "Rtrmm_noalloc: m*n exceeds the 1048576-element scratch");
    if (m == 0 || n == 0) return 0;
    dd *tmp = (h->c.device >= 0 && h->c.device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[h->c.device] : nullptr;
    if (!tmp) return bad(h, "Rtrmm_noalloc: no scratch address for this device");
    // the product is formed in the scratch (every element written before the
    // copy reads it, so nothing needs clearing) and copied over B
    sdk::trmm_lltn(h->c.stream, m, n, DP(d_A), lda, DP(d_B), ldb, tmp, m);
    (void)hipMemcpy2DAsync(d_B, sizeof(dd) * ldb, tmp, sizeof(dd) * m, sizeof(dd) * m, n, hipMemcpyDeviceToDevice,
                           h->c.stream);
    return check(h);
}
int sdpadd_Rtrsv(sdpadd_handle h, char uplo, char trans, char diag, int64_t n, const double *d_A, int64_t lda, double *d_x,
                 int64_t incx) {
    if (!h || !up(uplo, 'L') || !up(diag, 'N') || incx != 1 || !(up(trans, 'N') || up(trans, 'T')))
        return bad(h, "Rtrsv: unsupported op");
    if (n < 0 || lda < std::max<int64_t>(1, n)) return bad(h, "Rtrsv: bad dimension");
    if (n == 0) return 0;
    h->c.trsv(h->c.stream, up(trans, 'N') ? 'N' : 'T', n, DP(d_A), lda, DP(d_x));
    return check(h);
}
int sdpadd_ozaki_trsv(sdpadd_handle h, char trans, int64_t n, const double *d_L, int64_t lda, double *d_x) {
    if (!h || n < 0 || lda < std::max<int64_t>(1, n) || !(up(trans, 'N') || up(trans, 'T')))
        return bad(h, "ozaki_trsv: bad argument");
    if (n == 0) return 0;
    hipStream_t s = h->c.stream;
    if (h->c.binv_prepare(s, n, DP(d_L), lda) != 0) return bad(h, "ozaki_trsv: workspace allocation failed");
    h->c.trsv_binv(s, up(trans, 'N') ? 'N' : 'T', n, DP(d_L), lda, DP(d_x));
    return check(h);
}
int sdpadd_ozaki_matvec(sdpadd_handle h, char op, int64_t m, int64_t n, const double *alpha, const double *d_A,
                        int64_t lda, const double *d_x, const double *beta, double *d_y) {
    if (!h || m < 0 || n < 0 || lda < std::max<int64_t>(1, m)) return bad(h, "ozaki_matvec: bad argument");
    hipStream_t s = h->c.stream;
    if (up(op, 'G')) {
        if (!alpha || !beta) return bad(h, "ozaki_matvec: alpha and beta needed");
        sdk::gemv_n_fast(s, m, n, D(alpha), DP(d_A), lda, DP(d_x), D(beta), DP(d_y));
    } else if (up(op, 'L') || up(op, 'T')) {
        if (m != n) return bad(h, "ozaki_matvec: triangular operand must be square");
        sdk::trmv_lower_fast(s, up(op, 'T') ? 'T' : 'N', n, DP(d_A), lda, DP(d_x), DP(d_y));
    } else {
        return bad(h, "ozaki_matvec: op must be G, L or T");
    }
    return check(h);
}
int sdpadd_ozaki_dot(sdpadd_handle h, int64_t n, const double *d_x, const double *d_y, double *result) {
    if (!h || n < 0 || !result) return bad(h, "ozaki_dot: bad argument");
    hipStream_t s = h->c.stream;
    sdk::dot_fast(s, n, DP(d_x), DP(d_y), h->c.d_scal, h->c.d_dot_work, h->c.d_counter + 2);
    (void)hipMemcpyAsync(result, h->c.d_scal, sizeof(dd), hipMemcpyDeviceToHost, s);
    return check(h);
}
int sdpadd_Rtrmv(sdpadd_handle h, char uplo, char trans, char diag, int64_t n, const double *d_A, int64_t lda, double *d_x,
                 int64_t incx) {
    if (!h || !up(uplo, 'L') || !up(diag, 'N') || incx != 1 || !(up(trans, 'N') || up(trans, 'T')))
        return bad(h, "Rtrmv: unsupported op");
    if (n < 0 || lda < std::max<int64_t>(1, n)) return bad(h, "Rtrmv: bad dimension");
    if (n == 0) return 0;
    dd *tmp = static_cast<dd *>(h->c.workspace(sizeof(dd) * (size_t)n));
    if (!tmp) return bad(h, "Rtrmv: workspace allocation failed");
    sdk::trmv_lower(h->c.stream, up(trans, 'N') ? 'N' : 'T', n, DP(d_A), lda, DP(d_x), tmp);
    sdk::copy(h->c.stream, n, tmp, DP(d_x));
    return check(h);
}
int sdpadd_Rgemv(sdpadd_handle h, char trans, int64_t m, int64_t n, const double *alpha, const double *d_A, int64_t lda,
                 const double *d_x, int64_t incx, const double *beta, double *d_y, int64_t incy) {
    if (!h || !up(trans, 'N') || incx != 1 || incy != 1) return bad(h, "Rgemv: unsupported op");
    if (!alpha || !beta) return bad(h, "Rgemv: alpha and beta must not be null");
    if (m < 0 || n < 0 || lda < std::max<int64_t>(1, m)) return bad(h, "Rgemv: bad dimension");
    sdk::gemv_n(h->c.stream, m, n, D(alpha), DP(d_A), lda, DP(d_x), D(beta), DP(d_y));
    return check(h);
}
int sdpadd_Rdot(sdpadd_handle h, int64_t n, const double *d_x, int64_t incx, const double *d_y, int64_t incy,
                double *result) {
    if (!h || incx != 1 || incy != 1) return bad(h, "Rdot: unsupported op");
    if (n < 0 || !result) return bad(h, "Rdot: bad argument");
    h->c.dot(n, DP(d_x), DP(d_y), 0);
    (void)hipMemcpyAsync(result, h->c.d_scal, sizeof(dd), hipMemcpyDeviceToHost, h->c.stream);
    return check(h);
}
int sdpadd_Raxpy(sdpadd_handle h, int64_t n, const double *alpha, const double *d_x, int64_t incx, double *d_y,
                 int64_t incy) {
    if (!h) return -1;
    if (!alpha || incx == 0 || incy == 0) return bad(h, "Raxpy: bad argument");
    sdk::axpy_strided(h->c.stream, n, D(alpha), DP(d_x), incx, DP(d_y), incy);
    return check(h);
}
int sdpadd_Rscal(sdpadd_handle h, int64_t n, const double *alpha, double *d_x, int64_t incx) {
    if (!h) return -1;
    if (!alpha) return bad(h, "Rscal: alpha must not be null");
    sdk::scal_strided(h->c.stream, n, D(alpha), DP(d_x), incx);
    return check(h);
}
int sdpadd_Rcopy(sdpadd_handle h, int64_t n, const double *d_x, int64_t incx, double *d_y, int64_t incy) {
    if (!h) return -1;
    if (incx == 0 || incy == 0) return bad(h, "Rcopy: bad argument");
    sdk::copy_strided(h->c.stream, n, DP(d_x), incx, DP(d_y), incy);
    return check(h);
}
int sdpadd_Rpotrf(sdpadd_handle h, char uplo, int64_t n, double *d_A, int64_t lda, int64_t *info) {
    if (!h || !up(uplo, 'L')) return bad(h, "Rpotrf: unsupported op");
    if (n < 0 || lda < std::max<int64_t>(1, n)) return bad(h, "Rpotrf: bad dimension");
    sdk::potrf_lower(h->c.stream, n, DP(d_A), lda, h->c.d_info, h->c.d_rinv);
    (void)hipMemcpyAsync(info, h->c.d_info, sizeof(int64_t), hipMemcpyDeviceToHost, h->c.stream);
    return check(h);
}
int sdpadd_Rsteqr(sdpadd_handle h, char compz, int64_t n, double *d, double *e, double *z, int64_t ldz, double *work,
                  int64_t *info) {
    if (!up(compz, 'I') || n < 0 || ldz < std::max<int64_t>(1, n)) return bad(h, "Rsteqr: unsupported op");
    static_assert(sizeof(ddr) == 16, "ddr layout");
    *info = sdp::steqr_rows(n, reinterpret_cast<ddr *>(d), reinterpret_cast<ddr *>(e), reinterpret_cast<ddr *>(z), ldz,
                            reinterpret_cast<ddr *>(work), 0, n);
    return 0;
}

int sdpadd_lanczos_min_eig(sdpadd_handle h, int64_t n, const double *d_L, const double *d_M, double *out) {
    if (!h || n < 1) return bad(h, "lanczos: bad dimension");
    dd r = sdp::lanczos_min_eigen(h->c, (int)n, DP(d_L), DP(d_M));
    out[0] = r.hi;
    out[1] = r.lo;
    return check(h);
}
int sdpadd_small_min_eig(sdpadd_handle h, int64_t n, const double *d_L, const double *d_M, double *out, int *info) {
    if (!h || n < 1 || n > sdk::SMALL_EIG_MAXN || !out || !info) return bad(h, "small_min_eig: bad arguments");
    hipStream_t s = h->c.stream;
    sdk::SmallEigTask tk{DP(d_L), DP(d_M), (int)n};
    // task, result and info in the handle's workspace (256-byte slots)
    char *ws = static_cast<char *>(h->c.workspace(768));
    if (!ws) return bad(h, "small_min_eig: workspace allocation failed");
    static_assert(sizeof(sdk::SmallEigTask) <= 256, "task slot");
    auto *d_tk = reinterpret_cast<sdk::SmallEigTask *>(ws);
    dd *d_out = reinterpret_cast<dd *>(ws + 256);
    int *d_info = reinterpret_cast<int *>(ws + 512);
    (void)hipMemcpyAsync(d_tk, &tk, sizeof(tk), hipMemcpyHostToDevice, s);
    sdk::small_min_eig(s, d_tk, 1, h->c.dot_T, d_out, d_info);
    dd r{0.0, 0.0};
    (void)hipMemcpyAsync(&r, d_out, sizeof(dd), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(info, d_info, sizeof(int), hipMemcpyDeviceToHost, s);
    int rc = check(h);
    out[0] = r.hi;
    out[1] = r.lo;
    return rc;
}
int sdpadd_inv_chol(sdpadd_handle h, int64_t n, const double *d_A, double *d_work, double *d_invChol, double *d_inv) {
    if (!h || n < 1) return bad(h, "inv_chol: bad dimension");
    hipStream_t s = h->c.stream;
    sdk::copy(s, n * n, DP(d_A), DP(d_work));
    sdk::potrf_lower(s, n, DP(d_work), n, h->c.d_info, h->c.d_rinv);
    int64_t info = 0;
    (void)hipMemcpyAsync(&info, h->c.d_info, sizeof(info), hipMemcpyDeviceToHost, s);
    int rc = check(h);
    if (rc) return rc;
    if (info != 0) return (int)info;
    sdk::zero_upper(s, n, DP(d_work));
    sdk::inv_lower(s, n, DP(d_work), n, DP(d_invChol), n);
    if (d_inv) sdk::trmm_lltn_self(s, n, DP(d_invChol), n, DP(d_inv), n);
    return check(h);
}

// ---- problems ----
int sdpadd_problem_load(sdpadd_handle h, const char *path, sdpadd_problem *p) {
    if (!h || !path || !p) return -1;
    auto *pr = new sdpadd_problem_s();
    pr->h = h;
    std::string err;
    if (!sdp::read_dat_s(path, pr->hp, err)) {
        h->c.set_error(err);
        delete pr;
        return -1;
    }
    if (sdp::upload_problem(pr->hp, pr->dp, h->c.stream, err) != 0) {
        h->c.set_error(err);
        delete pr;
        return -3;
    }
    pr->finalized = true;
    *p = pr;
    return 0;
}
int sdpadd_problem_create(sdpadd_handle h, int m, int nblock, const int *bs, sdpadd_problem *p) {
    if (!h || m < 1 || nblock < 1 || !bs || !p) return bad(h, "problem_create: bad arguments");
    auto *pr = new sdpadd_problem_s();
    pr->h = h;
    pr->hp.m = m;
    if (!pr->hp.set_structure(std::vector<int>(bs, bs + nblock))) {
        delete pr;
        return bad(h, "problem_create: a block size is 0");
    }
    pr->hp.b.assign(m, mk(0.0));
    pr->hp.A.assign(m, sdp::Space());
    *p = pr;
    return 0;
}
int sdpadd_problem_set_entries(sdpadd_problem p, int k, int l, int64_t nnz, const int *rows, const int *cols,
                               const double *vals) {
    if (!p || p->finalized || k < 0 || k > p->hp.m || l < 0 || l >= (int)p->hp.given.size() || nnz < 0)
        return bad(p ? p->h : nullptr, "set_entries: bad arguments");
    sdp::Space &sp = k ? p->hp.A[k - 1] : p->hp.C;
    if (p->hp.gtype[l] == 3) {  // LP block: diagonal entries (row == col) of a diagonal block
        const int len = std::abs(p->hp.given[l]);
        for (int64_t q = 0; q < nnz; q++) {
            if (rows[q] != cols[q] || rows[q] < 0 || rows[q] >= len)
                return bad(p->h, "set_entries: LP entries must be diagonal and in range");
            sp.lp_raw.push_back({p->hp.gnum[l] + rows[q], D(vals + 2 * q)});
        }
        return 0;
    }
    l = p->hp.gnum[l];
    int slot = sp.find(l);
    if (slot < 0) {
        // keep the block list sorted by block number
        size_t pos = 0;
        while (pos < sp.index.size() && sp.index[pos] < l) pos++;
        sp.index.insert(sp.index.begin() + pos, l);
        sdp::Block b;
        b.n = p->hp.bs[l];
        sp.blk.insert(sp.blk.begin() + pos, b);
        slot = (int)pos;
    }
    sdp::Block &b = sp.blk[slot];
    for (int64_t q = 0; q < nnz; q++) {
        if (rows[q] < 0 || cols[q] < 0 || rows[q] >= b.n || cols[q] >= b.n) return bad(p->h, "set_entries: index out of range");
        b.row.push_back(rows[q]);
        b.col.push_back(cols[q]);
        b.val.push_back(D(vals + 2 * q));
        b.nnz_count++;
        b.nnz_effect += (rows[q] == cols[q]) ? 1 : 2;
    }
    return 0;
}
int sdpadd_problem_set_b(sdpadd_problem p, const double *b) {
    if (!p || p->finalized) return -1;
    for (int i = 0; i < p->hp.m; i++) p->hp.b[i] = D(b + 2 * i);
    return 0;
}
int sdpadd_problem_finalize(sdpadd_problem p) {
    if (!p || p->finalized) return -1;
    for (auto &b : p->hp.C.blk) b.change_to_dense();
    for (auto &a : p->hp.A)
        for (auto &b : a.blk) b.change_to_dense();
    p->hp.C.finish_lp();
    for (auto &a : p->hp.A) a.finish_lp();
    p->hp.layout();
    p->hp.index();
    p->hp.formula = sdp::compute_formula(p->hp, ddr(1.2));
    std::string err;
    if (sdp::upload_problem(p->hp, p->dp, p->h->c.stream, err) != 0) {
        p->h->c.set_error(err);
        return -3;
    }
    p->finalized = true;
    return 0;
}
int sdpadd_problem_schur_ddflops(sdpadd_problem p, double *ddflops) {
    if (!p || !p->finalized || !ddflops) return -1;
    *ddflops = sdp::schur_ddflops(p->hp);
    return 0;
}
// Analyses the Schur path for the problem's current mode (once per mode).
static int ensure_schur_path(sdpadd_problem p) {
    if (p->sps_mode == p->schur_mode) return 0;
    sps::analyse(p->hp, p->schur_mode, p->spa);
    std::string err;
    if (sps::upload(p->spa, p->spd, p->h->c.stream, err) != 0) {
        p->h->c.set_error(err);
        return -2;
    }
    p->sps_mode = p->schur_mode;
    return 0;
}
int sdpadd_problem_set_schur_mode(sdpadd_problem p, int mode) {
    if (!p || mode < 0 || mode > 2) return -1;
    p->schur_mode = mode;
    return 0;
}
int sdpadd_problem_schur_info(sdpadd_problem p, int *sparse, int64_t *nnz, int64_t *aggregate, int *ordering) {
    if (!p || !p->finalized) return -1;
    int rc = ensure_schur_path(p);
    if (rc) return rc;
    if (sparse) *sparse = p->spa.on ? 1 : 0;
    if (nnz) *nnz = p->spa.on ? p->spa.diag[p->spa.m] : 0;
    if (aggregate) *aggregate = p->spa.aggregate;
    if (ordering && p->spa.on) std::copy(p->spa.ordering.begin(), p->spa.ordering.end(), ordering);
    return 0;
}
int sdpadd_problem_destroy(sdpadd_problem p) {
    if (!p) return -1;
    sps::release(p->spd);
    p->adj.release();
    sdp::free_problem(p->dp);
    delete p;
    return 0;
}
int sdpadd_problem_dims(sdpadd_problem p, int *m, int *nblock, int *bs, int cap) {
    if (!p) return -1;
    *m = p->hp.m;
    *nblock = (int)p->hp.given.size();
    for (int l = 0; l < *nblock && l < cap; l++) bs[l] = p->hp.given[l];
    return 0;
}
int sdpadd_problem_formula(sdpadd_problem p, int *out) {
    if (!p || !p->finalized) return -1;
    std::memcpy(out, p->hp.formula.data(), sizeof(int) * p->hp.formula.size());
    return 0;
}
int sdpadd_schur_build(sdpadd_problem p, const double *d_X, const double *d_Zinv, double *d_B) {
    if (!p || !p->finalized) return -1;
    sdp::schur(p->h->c.stream, p->dp, DP(d_X), DP(d_Zinv), DP(d_B), p->h->c.dot_T);
    return check(p->h);
}
int sdpadd_inner_A(sdpadd_problem p, const double *d_W, double *d_out) {
    if (!p || !p->finalized) return -1;
    sdp::inner_A(p->h->c.stream, p->dp, DP(d_W), DP(d_out), p->h->c.dot_T);
    return check(p->h);
}
int sdpadd_plus_A(sdpadd_problem p, const double *d_y, double *d_R) {
    if (!p || !p->finalized) return -1;
    sdp::plus_A(p->h->c.stream, p->dp, DP(d_y), DP(d_R));
    return check(p->h);
}

void sdpadd_default_params(sdpadd_params *prm) {
    sdp::Params d;
    prm->maxIteration = d.maxIteration;
    prm->epsilonStar = d.epsilonStar;
    prm->lambdaStar = d.lambdaStar;
    prm->omegaStar = d.omegaStar;
    prm->lowerBound = d.lowerBound;
    prm->upperBound = d.upperBound;
    prm->betaStar = d.betaStar;
    prm->betaBar = d.betaBar;
    prm->gammaStar = d.gammaStar;
    prm->epsilonDash = d.epsilonDash;
}
int sdpadd_read_params(const char *path, sdpadd_params *prm) {
    sdp::Params d;
    if (!sdp::read_params(path, d)) return -1;
    prm->maxIteration = d.maxIteration;
    prm->epsilonStar = d.epsilonStar;
    prm->lambdaStar = d.lambdaStar;
    prm->omegaStar = d.omegaStar;
    prm->lowerBound = d.lowerBound;
    prm->upperBound = d.upperBound;
    prm->betaStar = d.betaStar;
    prm->betaBar = d.betaBar;
    prm->gammaStar = d.gammaStar;
    prm->epsilonDash = d.epsilonDash;
    return 0;
}
int sdpadd_solve(sdpadd_problem p, const sdpadd_params *prm, sdpadd_result *res) {
    if (!p || !p->finalized || !prm || !res) return -1;
    sdp::Params d;
    d.maxIteration = prm->maxIteration;
    d.epsilonStar = prm->epsilonStar;
    d.lambdaStar = prm->lambdaStar;
    d.omegaStar = prm->omegaStar;
    d.lowerBound = prm->lowerBound;
    d.upperBound = prm->upperBound;
    d.betaStar = prm->betaStar;
    d.betaBar = prm->betaBar;
    d.gammaStar = prm->gammaStar;
    d.epsilonDash = prm->epsilonDash;
    sdp::SolveOutput out;
    int rc = ensure_schur_path(p);
    if (rc) return rc;
    rc = sdp::solve(p->h->c, p->hp, p->dp, d, out, &p->X, &p->y, &p->Z, p->has_init ? &p->init : nullptr, &p->spd);
    if (rc != 0) return rc;
    p->trace = out.trace;
    res->iterations = out.iterations;
    res->phase = out.phase_reported;
    res->phase_internal = out.phase;
    auto put = [](double *dst, dd v) {
        dst[0] = v.hi;
        dst[1] = v.lo;
    };
    put(res->mu, out.mu);
    put(res->objValPrimal, out.objValPrimal);
    put(res->objValDual, out.objValDual);
    put(res->normPrimalVec, out.normPrimalVec);
    put(res->normDualMat, out.normDualMat);
    put(res->relgap, out.relgap);
    res->seconds = out.t.total;
    res->t_schur = out.t.schur;
    res->t_chol = out.t.chol;
    res->t_lanczos = out.t.lanczos;
    res->t_inverse = out.t.inverse;
    res->refine_sweeps = out.refine_sweeps;
    return 0;
}
int sdpadd_trace(sdpadd_problem p, double *out, int cap) {
    if (!p) return -1;
    int n = (int)p->trace.size();
    for (int r = 0; r < n && r < cap; r++) {
        const sdp::IterRecord &t = p->trace[r];
        double *o = out + 17 * r;
        o[0] = t.it;
        const dd v[8] = {t.mu, t.thetaP, t.thetaD, t.objP, t.objD, t.alphaP, t.alphaD, t.beta};
        for (int q = 0; q < 8; q++) {
            o[1 + 2 * q] = v[q].hi;
            o[2 + 2 * q] = v[q].lo;
        }
    }
    return n;
}
int sdpadd_adjoint_setup(sdpadd_problem p, int center_steps, int64_t *na, double *centrality) {
    if (!p || !p->finalized || center_steps < 0 || center_steps > 16) return -1;
    std::string err;
    int rc = p->adj.setup(p->h->c, p->hp, p->dp, p->X, p->y, p->Z, center_steps, err);
    if (rc != 0) p->h->c.set_error(err);
    if (rc == 0 && na) *na = p->adj.na;
    if (rc == 0 && centrality) *centrality = p->adj.centrality;
    return rc;
}
int sdpadd_adjoint_entries(sdpadd_problem p, int *con, int *blk, int *row, int *col) {
    if (!p || !p->finalized) return -1;
    idx_t q = 0;
    for (int i = 0; i < p->hp.m; i++)
        for (size_t k = 0; k < p->hp.A[i].index.size(); k++) {
            const sdp::Block &b = p->hp.A[i].blk[k];
            const idx_t cnt = b.dense ? (idx_t)b.n * b.n : (idx_t)b.nnz_count;
            for (idx_t e = 0; e < cnt; e++, q++) {
                if (con) con[q] = i;
                if (blk) blk[q] = p->hp.A[i].index[k];
                if (row) row[q] = b.dense ? (int)(e % b.n) : b.row[e];
                if (col) col[q] = b.dense ? (int)(e / b.n) : b.col[e];
            }
        }
    return 0;
}
int sdpadd_adjoint(sdpadd_problem p, int nrhs, const double *d_Xbar, const double *d_ybar, const double *pbar,
                   double *d_bbar, double *d_Cbar, double *d_Abar) {
    if (!p || !p->finalized || nrhs < 0) return -1;
    std::vector<dd> pb;
    if (pbar)
        for (int k = 0; k < nrhs; k++) pb.push_back(D(pbar + 2 * k));
    std::string err;
    int rc = p->adj.apply(p->h->c, p->dp, nrhs, DP(d_Xbar), DP(d_ybar), pbar ? pb.data() : nullptr, DP(d_bbar),
                          DP(d_Cbar), DP(d_Abar), err);
    if (rc != 0) p->h->c.set_error(err);
    return rc;
}
int sdpadd_read_initial_point(sdpadd_problem p, const char *path, int sparse) {
    if (!p || !p->finalized || !path) return -1;
    std::string err;
    sdp::InitPoint ip;
    if (!sdp::read_init(path, p->hp, sparse != 0, ip, err)) {
        p->h->c.set_error(err);
        return -1;
    }
    p->init = std::move(ip);
    p->has_init = true;
    return 0;
}
int sdpadd_set_initial_point(sdpadd_problem p, const double *X, const double *y, const double *Z) {
    if (!p || !p->finalized || !X || !y || !Z) return -1;
    const sdp::idx_t nt = p->hp.nt();
    sdp::InitPoint ip;
    ip.full = true;
    ip.x.assign(DP(X), DP(X) + nt);
    ip.z.assign(DP(Z), DP(Z) + nt);
    ip.y.assign(DP(y), DP(y) + p->hp.m);
    p->init = std::move(ip);
    p->has_init = true;
    return 0;
}
int sdpadd_clear_initial_point(sdpadd_problem p) {
    if (!p) return -1;
    p->init = sdp::InitPoint();
    p->has_init = false;
    return 0;
}
int sdpadd_write_solution(sdpadd_problem p, const char *path) {
    if (!p || !path) return -1;
    std::string err;
    if (!sdp::write_point(path, p->hp, p->X, p->y, p->Z, err)) {
        p->h->c.set_error(err);
        return -1;
    }
    return 0;
}
int sdpadd_solution(sdpadd_problem p, double *X, double *y, double *Z) {
    if (!p || p->X.empty()) return -1;
    if (X) std::memcpy(X, p->X.data(), sizeof(dd) * p->X.size());
    if (y) std::memcpy(y, p->y.data(), sizeof(dd) * p->y.size());
    if (Z) std::memcpy(Z, p->Z.data(), sizeof(dd) * p->Z.size());
    return 0;
}

}  // extern "C"

// Internal launcher API for the dd kernels (host side).  All pointers are
// device pointers to dd values (16 bytes, {hi, lo}), matrices column-major.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "dd.hpp"

namespace sdk {

typedef int64_t idx_t;

// C := alpha*op(A)*op(B) + beta*C with op(B) = B ('N') or B**T ('T'); A is
// never transposed.  Batched with element strides sA, sB, sC over `batch`.
// Each C element accumulates its k terms in ascending order as
// c += (alpha*b_lj) * a_il, the order of Rgemm_NN/NT.
void gemm(hipStream_t s, char tb, idx_t m, idx_t n, idx_t k, dd alpha, const dd *A, idx_t lda, const dd *B, idx_t ldb,
          dd beta, dd *C, idx_t ldc, int batch = 1, idx_t sA = 0, idx_t sB = 0, idx_t sC = 0,
          bool lower_only = false, idx_t diag_offset = 0);

// C := A*B (alpha 1, beta 0, 'N' 'N') in GEMM_KS parts over consecutive k
// ranges, each summed in ascending k, added in order (not Rgemm's order:
// the Ozaki form's n x n products).  work: GEMM_KS m n dd.
constexpr int GEMM_KS = 4;
void gemm_split(hipStream_t s, idx_t m, idx_t n, idx_t k, const dd *A, idx_t lda, const dd *B, idx_t ldb, dd *C,
                idx_t ldc, dd *work);

// Lower Cholesky in place.  info (device int64) receives 0, the 1-based
// index of the first non-positive pivot, or POTRF_WAIT_FAILED when a panel
// workgroup's bounded wait for the diagonal block ran out (nothing is
// stored then; the factor is unusable).  scratch: POTRF_SCRATCH_DD dd of
// device memory, one per stream that factors concurrently (pivot
// reciprocals, two publication buffers and a ticket counter).
constexpr int64_t POTRF_WAIT_FAILED = -2;
constexpr size_t POTRF_SCRATCH_DD = 64 + 2 * (64 * 64 + 64) + 1;
void potrf_lower(hipStream_t s, idx_t n, dd *A, idx_t lda, int64_t *d_info, dd *scratch);
// Cholesky of the precision path: dd panels, Ozaki INT8-MFMA trailing
// updates with S digit planes.  Returns 0 or the Ozaki product's error code.
}  // namespace sdk
namespace oz { struct Work; }
namespace sdk {
// Look-ahead for potrf_lower_oz: a helper stream with its own Ozaki
// workspace and two events (all default priority).
struct OzLookAhead {
    hipStream_t hs = nullptr;
    oz::Work *w2 = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    // optional second workspace: with it, w2 and wp hold the panels' digit
    // planes in turn (written by each panel's last diagonal-block launch)
    oz::Work *wp = nullptr;
};
// la: when given and n > 1024, each panel's trailing update is split into
// the next panel's columns (on s, first) and the rest (on la->hs), so that
// the next panel's diagonal and panel solves overlap the bulk of the update.
int potrf_lower_oz(hipStream_t s, idx_t n, dd *A, idx_t lda, int64_t *d_info, dd *scratch, oz::Work &w, int S,
                   const OzLookAhead *la = nullptr);
// The Ozaki form's B := L^{-1} (rows below each 256-row panel updated with
// the Ozaki product).  Returns 0 or the product's error code.
int inv_lower_oz(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *B, idx_t ldb, oz::Work &w, int S);

// x := L^{-1} x ('N') or L^{-T} x ('T'), L lower non-unit.
void trsv_lower(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x);
// The same solve spread over a grid of workgroups that wait on each other
// (large n), stream-ordered with no host synchronisation.  work: at least
// trsv_coop_bytes(n) bytes of device memory.  Should the grid not be
// resident at once, its bounded waits run out and a single-workgroup solve
// enqueued behind it redoes the solve from a saved copy of x, so x is
// always complete.  Returns 0, or 1 when the launch fails (nothing ran).
size_t trsv_coop_bytes(idx_t n);
int trsv_lower_coop(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x, void *work);
// The Ozaki form's solves: block-inverse cooperative L^{-1} x / L^{-T} x
// (not Rtrsv's operation order).  trsv_binv_prepare forms the per-block
// operators of both directions once per factor (ops_n, ops_t:
// binv_ops_elems(n) dd each; inv: binv_ops_elems(n) dd of scratch);
// trsv_binv uses the trsv_coop_bytes workspace and the recovery of
// trsv_lower_coop and returns 0, or 1 when the launch fails.  lt (n x n,
// leading dimension n, may be null): prepare stores L^T there (the upper
// triangle of lt is the lower triangle of L, transposed) and the L^T x
// solve then reads L by columns of lt instead of along rows of L.
size_t binv_ops_elems(idx_t n);
void trsv_binv_prepare(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *ops_n, dd *ops_t, dd *inv, dd *lt);
int trsv_binv(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x, const dd *ops, void *work,
              const dd *lt);
// p[0..cnt) := 0 with a kernel (no memset node in captured graphs).
void zero_ints(hipStream_t s, int *p, int cnt);

// B := L^{-1} (B enters as anything; set to identity then solved like
// Rtrsm('L','L','N','N') on the identity).  fast (the Ozaki form): each
// solved value is multiplied by the reciprocal of its pivot instead of
// divided by the pivot.
void inv_lower(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *B, idx_t ldb, bool fast = false);

// C := A**T * A for lower triangular A (Rtrmm('L','L','T','N') applied to a
// copy of A); C full n x n.
void trmm_lltn_self(hipStream_t s, idx_t n, const dd *A, idx_t lda, dd *C, idx_t ldc);

// General Rtrsm('R','L','T','N', alpha = 1) and Rtrsm('L','L','N','N', alpha = 1)
// and Rtrmm('L','L','T','N', alpha = 1) on arbitrary operands.
void trsm_rltn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, dd *B, idx_t ldb);
void trsm_llnn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, dd *B, idx_t ldb);
void trmm_lltn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, const dd *B, idx_t ldb, dd *C, idx_t ldc);
void syrk_ln(hipStream_t s, idx_t n, idx_t k, dd alpha, const dd *A, idx_t lda, dd beta, dd *C, idx_t ldc);

// x := L**T x ('T') or L x ('N'), Rtrmv lower non-unit.  out may differ from x.
void trmv_lower(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, const dd *x, dd *out);
// y := alpha*A*x (+ beta*y), Rgemv 'N'.
void gemv_n(hipStream_t s, idx_t m, idx_t n, dd alpha, const dd *A, idx_t lda, const dd *x, dd beta, dd *y);

// The Ozaki form's mat-vecs and dot (sums split into parts and merged, not
// the reference's order).  dot_fast: work >= 256 dd, *counter 0 on entry
// (left 0).
void trmv_lower_fast(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, const dd *x, dd *out);
void gemv_n_fast(hipStream_t s, idx_t m, idx_t n, dd alpha, const dd *A, idx_t lda, const dd *x, dd beta, dd *y);
void dot_fast(hipStream_t s, idx_t n, const dd *x, const dd *y, dd *out, dd *work, unsigned *counter);
struct DotTask;
// Many dots in that order; partial: >= G * ntasks dd (G <= the partial
// buffer's per-task share).
void dot_tasks_fast(hipStream_t s, const DotTask *d_tasks, int ntasks, idx_t maxn, int G, dd *partial);

// Level-1 BLAS with any strides (negative strides start at the far end, as
// in the reference): y += a*x (skipped when a == 0), x = a*x (nothing for
// incx <= 0), y = x.  x and y must not overlap.
void axpy_strided(hipStream_t s, idx_t n, dd a, const dd *x, idx_t incx, dd *y, idx_t incy);
void scal_strided(hipStream_t s, idx_t n, dd a, dd *x, idx_t incx);
void copy_strided(hipStream_t s, idx_t n, const dd *x, idx_t incx, dd *y, idx_t incy);
// Elementwise QD operator `op` (0..11: + - * / sqrt += -= *= *double
// /double +double double/) on n dd pairs.
void dd_vec_op(hipStream_t s, int op, idx_t n, const dd *a, const dd *b, dd *out);
// Elementwise (length len, flat).
void fill(hipStream_t s, idx_t len, dd v, dd *x);
void copy(hipStream_t s, idx_t len, const dd *x, dd *y);
void scal(hipStream_t s, idx_t len, dd a, dd *x);                    // x = a*x
void scal_dev(hipStream_t s, idx_t len, const dd *a, dd *x);         // x = (*a)*x, scalar on device
void axpy(hipStream_t s, idx_t len, dd a, const dd *x, dd *y);        // y += a*x (skipped if a == 0)
void axpy_dev(hipStream_t s, idx_t len, const dd *a, bool negate, const dd *x, dd *y);
void set_identity(hipStream_t s, idx_t n, dd v, dd *A);
void zero_upper(hipStream_t s, idx_t n, dd *A);
void symmetrize(hipStream_t s, idx_t n, dd *A);

// Dot products with the static T-way partition (Rdot_omp order).
// Result written to *out (device).  work: >= T dd + counter space.
struct DotTask {
    const dd *x;
    const dd *y;
    idx_t n;
    dd *out;
};
void dot(hipStream_t s, idx_t n, const dd *x, const dd *y, dd *out, int T, dd *work, unsigned *counter);
// Many independent dots, one workgroup each (tasks on device).
void dot_batch(hipStream_t s, const DotTask *d_tasks, int ntasks, int T);
// Many dots in one launch (maxn = longest task); partial: >= T * ntasks dd.
void dot_tasks(hipStream_t s, const DotTask *d_tasks, int ntasks, idx_t maxn, int T, dd *partial);
// Same with y = ybase + yoff per task.
struct DotTaskOff {
    const dd *x;
    idx_t yoff;
    idx_t n;
    dd *out;
};
void dot_batch_off(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T);
// Same result for long tasks: one workgroup per (chunk, task); partial:
// >= T * ntasks dd.
void dot_tasks_off(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T, dd *partial);
// The same with each chunk summed by a pairwise tree (the Ozaki form).
void dot_tasks_off_fast(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T, dd *partial);
// out_i = A_ii of an n x n column-major block.
void gather_diag(hipStream_t s, idx_t n, const dd *A, dd *out);
// Lanczos step pieces (Lal::getMinEigen): qold = q, q = v*r; and the tail
// alpha = q.w, r = w - alpha q - beta qold, rr = r.r with out = {alpha, rr}.
void lanczos_q(hipStream_t s, idx_t n, dd v, const dd *r, dd *q, dd *qold);
void lanczos_tail(hipStream_t s, idx_t n, int T, const dd *q, const dd *qold, const dd *w, dd *r, dd nbeta, dd *out);
// The same steps with the scalars on the device (sc[2] = beta, sc[3] =
// 1/beta of the previous step; the tail leaves the next step's pair there),
// so that several steps run without the host.
void lanczos_q_dev(hipStream_t s, idx_t n, const dd *sc, const dd *r, dd *q, dd *qold);
void lanczos_tail_dev(hipStream_t s, idx_t n, int T, const dd *q, const dd *qold, const dd *w, dd *r, dd *sc, dd *out);
// The same step with its two dots in a fixed split order (the Ozaki form).
void lanczos_tail_fast(hipStream_t s, idx_t n, const dd *q, const dd *qold, const dd *w, dd *r, dd *sc, dd *out);
// The whole step of the Ozaki form (lanczos_q_dev, the three fast mat-vecs
// with L and M of leading dimension n, lanczos_tail_fast) in three launches
// with the same results; ctr: a zeroed device counter, left zeroed.
// Returns false (nothing enqueued) for n > 4096.
bool lanczos_step_fast(hipStream_t s, idx_t n, const dd *L, const dd *M, dd *r, dd *q, dd *qold, dd *w, dd *tmp,
                       dd *sc, dd *out, unsigned *ctr);
// Batched GEMM over pointer arrays (device arrays of batch pointers):
// C[b] := A[b]*B[b] (alpha = 1, beta = 0), order as gemm().
void gemm_ptrs(hipStream_t s, idx_t m, idx_t n, idx_t k, const dd *const *A, idx_t lda, const dd *const *B, idx_t ldb,
               dd *const *C, idx_t ldc, int batch);
// Sequential strided dot (Rdot with non-unit stride): one thread per task.
// Max |x_i| over len values into *out.
// *out = max |x_i| (device); work: MAX_ABS_BLOCKS dd; copy_to (may be null):
// also copies x there in the same pass.
constexpr int MAX_ABS_BLOCKS = 1024;
void max_abs(hipStream_t s, idx_t len, const dd *x, dd *out, dd *work, unsigned *counter, dd *copy_to = nullptr);

// Tile choice of the dd GEMM and the Ozaki product: 0 = by grid size
// (default), 1 = always the 32 x 32 tiles, 2 = always the 64 x 64 tiles.
// Both tilings give every output element the same operation sequence.
extern int g_tile_policy;
// Bound on each spin-wait of the cross-workgroup triangular solves
// (iterations of about 0.25 us; 0 = the default of 2^22).  A wait that runs
// out hands the solve to the single-workgroup recovery.
extern unsigned g_trsv_spin;
extern unsigned g_panel_spin;  // bound of the Cholesky panel workgroups' polls (0 = default)

// Per-kernel-family timing with HIP events on the launch stream (off by
// default).  Families: "gemm", "ozaki", "potrf", "panel", "trsm", "trmm",
// "trsv", "trsv_prep", "mv", "dot".  flops are algorithmic dd operations.
void stats_enable(bool on);
// Times only the listed families (comma-separated); null or "" for all.
void stats_filter(const char *families);
void stats_reset();
bool stats_get(const char *family, int64_t *launches, double *ms, double *ddflops);
struct StatScope {
    int slot = -1;
    hipStream_t s = nullptr;
    StatScope(const char *family, hipStream_t stream, double ddflops);
    ~StatScope();
};

}  // namespace sdk

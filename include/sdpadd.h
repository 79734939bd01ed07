/* sdpadd.h — C ABI of the MI355X double-double Newton-step library
 * (libsdpadd_hip.so).
 *
 * Conventions
 *   - A dd value is two doubles {hi, lo}: the memory layout of QD's dd_real,
 *     so a dd_real* from SDPA-DD's host code can be passed as double*.
 *   - Matrices are column-major with a leading dimension, as in MPLAPACK.
 *   - Arguments named d_* are device pointers (from sdpadd_malloc).
 *   - Every function returns int: 0 = success, < 0 = argument or HIP error
 *     (message from sdpadd_last_error), > 0 = LAPACK-style info where
 *     documented.  Nothing in the library calls exit().
 *   - Calls on one handle are stream-ordered and return after completion.
 *     One host thread per handle.
 */
#ifndef SDPADD_H
#define SDPADD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdpadd_ctx *sdpadd_handle;
typedef struct sdpadd_problem_s *sdpadd_problem;

/* ---- context and memory ------------------------------------------------ */
int sdpadd_create(int device, sdpadd_handle *h);
int sdpadd_destroy(sdpadd_handle h);
const char *sdpadd_last_error(sdpadd_handle h);
/* Number of chunks of the static partition that unit-stride dot products
 * sum over (mplapack/Rdot_omp.cpp:80-120 with that many OpenMP threads,
 * partial sums added in thread order).  1 <= chunks <= 2048; default 64. */
int sdpadd_set_dot_partition(sdpadd_handle h, int chunks);
/* Form of the Newton step used by sdpadd_solve (no counterpart in the
 * reference, which has only the first form): path 0 = plain, every dd
 * operation in the reference's order (bit-identical to it for the same dot
 * partition); path 1 = Ozaki Cholesky, the trailing updates of the Schur
 * and X/Z Cholesky factorisations run on INT8 MFMA with `slices` 7-bit digit
 * planes (8, 12, 14 or 16) and the Schur solve is followed by up to
 * `refine_steps` (0..8) sweeps of dd iterative refinement against the
 * unfactored Schur matrix; path 2 = as 1, and the n x n GEMMs of the
 * direction run on INT8 MFMA too (digit planes are scaled per row, so
 * entries far below their row's largest lose relative precision: fine for
 * dense iterates such as max-cut's, not for theta-type iterates whose
 * structurally zero entries shrink with mu).  A negative `slices` (-4 .. -7)
 * selects the FP64-slice variant: |slices| 18-bit digits held in doubles,
 * accumulated by FP64 MFMA (-5: about 2^-80 normwise, -6 and -7: dd).
 * Default: 0, 16, 2. */
int sdpadd_set_path(sdpadd_handle h, int path, int slices, int refine_steps);
/* Tile shape of the dd GEMM and Ozaki-product kernels (no counterpart in
 * the reference): 0 = chosen by grid size (default), 1 = always 32 x 32
 * output tiles, 2 = always 64 x 64.  Both tilings give every output element
 * the same operation sequence, so results do not depend on it; process-wide
 * (it exists so tests and tuning can reach either kernel at any size). */
int sdpadd_set_tile_policy(sdpadd_handle h, int policy);
/* Bound on each spin-wait of the cross-workgroup triangular solves (no
 * counterpart in the reference): `spins` iterations of about 0.25 us, 0 =
 * the default 2^22 (about a second).  The solves launch their workgroups
 * plainly when the grid fits the CU count and assume they are resident
 * together; a wait that runs out makes every workgroup leave and a
 * single-workgroup solve, enqueued behind each of them, redo the solve from
 * a saved copy of the right-hand side in Rtrsv's operation order, so results
 * never depend on co-residency.  Process-wide; tests set a tiny bound to
 * reach that recovery. */
int sdpadd_set_wait_bound(sdpadd_handle h, unsigned spins);
/* Bound on the consecutive empty polls of a Cholesky panel workgroup waiting
 * for the columns of its diagonal block (no counterpart in the reference):
 * 0 = the default 2^22 (seconds).  The diagonal workgroup is the first of
 * its launch to start and waits for nothing, so the wait always ends; a
 * bound that runs out anyway stores nothing and reports info = -2 (Rpotrf)
 * or a solve error.  Process-wide; tests set a tiny bound to reach it. */
int sdpadd_set_panel_wait_bound(sdpadd_handle h, unsigned polls);
int sdpadd_malloc(sdpadd_handle h, size_t bytes, void **d_ptr);
int sdpadd_free(sdpadd_handle h, void *d_ptr);
int sdpadd_upload(sdpadd_handle h, void *d_dst, const void *src, size_t bytes);
int sdpadd_download(sdpadd_handle h, void *dst, const void *d_src, size_t bytes);

/* Kernel timing with HIP events on the stream each kernel runs on, per
 * family ("gemm", "ozaki" (Ozaki products), "potrf" (diagonal Cholesky
 * blocks), "panel" (Cholesky panel solves), "trsm", "trmm", "trsv",
 * "trsv_prep" (block operators of the Ozaki form's solves), "mv", "dot"):
 * launches, summed
 * kernel milliseconds and algorithmic dd-flops since the last reset.  No
 * counterpart in the reference (its ComputeTime buckets are host wall
 * clocks, sdpa_parts.h:46-83). */
int sdpadd_stats_enable(sdpadd_handle h, int on);
/* Limits the timed families to a comma-separated list (NULL or "": all);
 * a family outside it costs no events. */
int sdpadd_stats_filter(sdpadd_handle h, const char *families);
int sdpadd_stats_reset(sdpadd_handle h);
int sdpadd_stats_get(sdpadd_handle h, const char *family, int64_t *launches, double *ms, double *ddflops);

/* ---- dd BLAS/LAPACK on device arrays ------------------------------------
 * Each replaces the MPLAPACK routine of the same name (C++ signatures in
 * mplapack/mpblas_dd.h:38-56 and mplapack/mplapack_dd.h:38-92) for the
 * option combinations SDPA-DD's Newton step uses; other combinations return
 * -1.  alpha and beta point to host dd scalars. */

/* QD's dd_real operators elementwise on n device dd pairs: out_i = a_i op
 * b_i with op 0 a+b (sloppy add), 1 a-b, 2 a*b, 3 a/b (sloppy divide),
 * 4 sqrt(a) (Karp), 5 a+=b, 6 a-=b, 7 a*=b, 8 a*hi(b), 9 a/hi(b),
 * 10 a+hi(b), 11 hi(a)/b (qd/include/qd/dd_inline.h:26-393,
 * qd/src/dd_real.cpp:47-69 with QD 2.3.23's default configuration).  The
 * scalar arithmetic every kernel is built from. */
int sdpadd_dd_vec_op(sdpadd_handle h, int op, int64_t n, const double *d_a, const double *d_b, double *d_out);

/* Rgemm: transa 'N'; transb 'N' or 'T' (mplapack/Rgemm_omp.cpp:41-127). */
int sdpadd_Rgemm(sdpadd_handle h, char transa, char transb, int64_t m, int64_t n, int64_t k, const double *alpha,
                 const double *d_A, int64_t lda, const double *d_B, int64_t ldb, const double *beta, double *d_C,
                 int64_t ldc);
/* Ozaki form of Rgemm (the products of mplapack/Rgemm_omp.cpp:41-127 on
 * INT8 MFMA): C = beta*C + alpha*(op(A) op(B)) with alpha = +1 or -1, beta
 * = 0 or 1, op(X) = X ('N') or X^T ('T'); lower != 0 writes only entries
 * i >= j.  Each row of A and column of op(B) is split into `slices`
 * (8, 12, 14 or 16) signed 7-bit digit planes under its own power-of-two
 * scale; plane products accumulate exactly in int32 and are recombined in
 * dd.  The error of C_ij is about k * 2^(-7*slices) * max_l |A_il| *
 * max_l |B_lj| (16 slices: dd accuracy normwise).  k <= 8192.  slices
 * -4 .. -7: the FP64-slice variant (|slices| 18-bit digits in doubles,
 * accumulated exactly by FP64 MFMA; error about k * 2^(10-18|slices|) *
 * max_l |A_il| * max_l |B_lj| plus dd rounding). */
int sdpadd_ozaki_gemm(sdpadd_handle h, char transa, char transb, int64_t m, int64_t n, int64_t k, int alpha,
                      const double *d_A, int64_t lda, const double *d_B, int64_t ldb, int beta, double *d_C, int64_t ldc,
                      int lower, int slices);
/* Rsyrk: uplo 'L', trans 'N' (mplapack/Rsyrk.cpp:31-215). */
int sdpadd_Rsyrk(sdpadd_handle h, char uplo, char trans, int64_t n, int64_t k, const double *alpha, const double *d_A,
                 int64_t lda, const double *beta, double *d_C, int64_t ldc);
/* Rtrsm: (side,uplo,transa,diag) = (R,L,T,N) or (L,L,N,N), alpha = 1
 * (mplapack/Rtrsm.cpp:31-291). */
int sdpadd_Rtrsm(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                 const double *alpha, const double *d_A, int64_t lda, double *d_B, int64_t ldb);
/* Rtrmm: (L,L,T,N), alpha = 1 (mplapack/Rtrmm.cpp:31-275). */
int sdpadd_Rtrmm(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                 const double *alpha, const double *d_A, int64_t lda, double *d_B, int64_t ldb);
/* Rtrmm (L,L,T,N) as sdpadd_Rtrmm, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Its m x n product buffer is a fixed device array of 2^20 dd
 * (16 MiB) per GPU that the library holds until the process ends;
 * m * n > 1048576 is rejected (-1).  Calls on one GPU share that array:
 * calls that overlap on one GPU (other streams, other threads, other
 * handles) race.  Results are identical to sdpadd_Rtrmm's. */
int sdpadd_Rtrmm_noalloc(sdpadd_handle h, char side, char uplo, char transa, char diag, int64_t m, int64_t n,
                         const double *alpha, const double *d_A, int64_t lda, double *d_B, int64_t ldb);
/* The Ozaki form's triangular solve (no counterpart in the reference): x :=
 * L^{-1} x ('N') or L^{-T} x ('T') for lower non-unit L, by 32-row blocks
 * with inverted diagonal blocks in a cooperative grid; not Rtrsv's operation
 * order (the Ozaki form refines its Schur solves against B).  Each call
 * forms the block operators of L; sdpadd_solve forms them once per factor. */
int sdpadd_ozaki_trsv(sdpadd_handle h, char trans, int64_t n, const double *d_L, int64_t lda, double *d_x);
/* The Ozaki form's mat-vecs and dot (no counterpart in the reference): sums
 * split into independent parts and merged (not Rgemv/Rtrmv/Rdot's order).
 * op 'G': y = alpha*A*x + beta*y (A m x n); 'L': y = L*x, 'T': y = L^T*x
 * (L lower n x n, m = n; alpha, beta unused).  result: dd of x.y (host). */
int sdpadd_ozaki_matvec(sdpadd_handle h, char op, int64_t m, int64_t n, const double *alpha, const double *d_A,
                        int64_t lda, const double *d_x, const double *beta, double *d_y);
int sdpadd_ozaki_dot(sdpadd_handle h, int64_t n, const double *d_x, const double *d_y, double *result);
/* Rtrsv / Rtrmv: uplo 'L', trans 'N' or 'T', diag 'N', incx 1
 * (mplapack/Rtrsv.cpp:31-234, Rtrmv.cpp:31-234). */
int sdpadd_Rtrsv(sdpadd_handle h, char uplo, char trans, char diag, int64_t n, const double *d_A, int64_t lda,
                 double *d_x, int64_t incx);
int sdpadd_Rtrmv(sdpadd_handle h, char uplo, char trans, char diag, int64_t n, const double *d_A, int64_t lda,
                 double *d_x, int64_t incx);
/* Rgemv: trans 'N', unit strides (mplapack/Rgemv.cpp:31-204). */
int sdpadd_Rgemv(sdpadd_handle h, char trans, int64_t m, int64_t n, const double *alpha, const double *d_A,
                 int64_t lda, const double *d_x, int64_t incx, const double *beta, double *d_y, int64_t incy);
/* Rdot: unit strides; result to host dd *result (mplapack/Rdot.cpp:38-47). */
int sdpadd_Rdot(sdpadd_handle h, int64_t n, const double *d_x, int64_t incx, const double *d_y, int64_t incy,
                double *result);
/* Raxpy (y += alpha*x), Rscal (x = alpha*x; nothing for incx <= 0) and
 * Rcopy (y = x) with any nonzero strides, negative strides starting at the
 * far end (mplapack/Raxpy_omp.cpp:79-112, Rscal.cpp, Rcopy.cpp); x and y
 * must not overlap. */
int sdpadd_Raxpy(sdpadd_handle h, int64_t n, const double *alpha, const double *d_x, int64_t incx, double *d_y,
                 int64_t incy);
int sdpadd_Rscal(sdpadd_handle h, int64_t n, const double *alpha, double *d_x, int64_t incx);
int sdpadd_Rcopy(sdpadd_handle h, int64_t n, const double *d_x, int64_t incx, double *d_y, int64_t incy);
/* Rpotrf: uplo 'L'; *info = 0 or the first non-positive leading minor
 * (mplapack/Rpotrf.cpp:32-154); -2 (no counterpart in the reference) when a
 * panel wait ran out of the bound of sdpadd_set_panel_wait_bound. */
int sdpadd_Rpotrf(sdpadd_handle h, char uplo, int64_t n, double *d_A, int64_t lda, int64_t *info);
/* Rsteqr on HOST arrays, compz 'I' (mplapack/Rsteqr.cpp:32-503).
 * work: 2*n dd. */
int sdpadd_Rsteqr(sdpadd_handle h, char compz, int64_t n, double *d, double *e, double *z, int64_t ldz, double *work,
                  int64_t *info);

/* ---- Newton-step pieces --------------------------------------------------- */
/* Truncated-Lanczos estimate of lambda_min(L M L^T) for one block, as
 * Lal::getMinEigen computes it (sdpa_linear.cpp:27-122).  d_L, d_M device
 * n x n (lda = n); result to host dd *out. */
int sdpadd_lanczos_min_eig(sdpadd_handle h, int64_t n, const double *d_L, const double *d_M, double *out);
/* lambda_min(L M L^T) for a block of size n <= 20 as the reference takes it
 * there: M L^T and L (M L^T) by Rgemm, then all eigenvalues by Rsyev ('N',
 * 'L') (Jal::getMinEigen, sdpa_jordan.cpp:62-66; Lal::getMinEigenValue,
 * sdpa_linear.cpp:124-156).  d_L, d_M device n x n; result to host dd *out;
 * *info as Rsyev's (> 0: Rsterf did not converge). */
int sdpadd_small_min_eig(sdpadd_handle h, int64_t n, const double *d_L, const double *d_M, double *out, int *info);
/* Inverse Cholesky pieces of Solutions::computeInverse (sdpa_dataset.cpp:96-118,
 * sdpa_jordan.cpp:123-170): d_invChol = chol(A)^{-1}; if d_inv != NULL also
 * d_inv = L^{-T} L^{-1}.  d_work: n*n dd.  Returns > 0 (failing pivot) if A is
 * not positive definite. */
int sdpadd_inv_chol(sdpadd_handle h, int64_t n, const double *d_A, double *d_work, double *d_invChol, double *d_inv);

/* ---- Problems ------------------------------------------------------------- */
/* Reads an SDPA sparse-format file (.dat-s) the way IO::read does
 * (sdpa_io.cpp:124-150, 289-426, 429-786): C = -F0, A_k = F_k, b = c.  Block
 * sizes as sdpa_main.cpp:111-130 reads them: >= 2 an SDP block, -k (or 1) a
 * diagonal block of k (or 1) LP entries, which must be given on the
 * diagonal. */
int sdpadd_problem_load(sdpadd_handle h, const char *path, sdpadd_problem *p);
/* Builds a problem in SDPA's internal form: the block structure as in a
 * data file (>= 2 SDP, -k or 1 LP), then entries of matrix k (0 = C, 1..m =
 * A_k) block l (0-based, in that structure), given as 0-based (row, col)
 * pairs with dd values in file order (SparseLinearSpace::setElement_SDP and
 * setElement_LP, sdpa_struct.cpp:1217-1265; LP entries need row == col, a
 * repeated LP entry keeps its last value). */
int sdpadd_problem_create(sdpadd_handle h, int m, int nblock, const int *block_struct, sdpadd_problem *p);
int sdpadd_problem_set_entries(sdpadd_problem p, int k, int l, int64_t nnz, const int *rows, const int *cols,
                               const double *vals);
int sdpadd_problem_set_b(sdpadd_problem p, const double *b);
/* Dense conversion (>= 20% stored entries), formula table, device image. */
int sdpadd_problem_finalize(sdpadd_problem p);
int sdpadd_problem_destroy(sdpadd_problem p);
/* m, number of blocks, block structure as given (up to cap; LP blocks
 * negative).  Iterates hold the SDP blocks (each n x n, in order) followed by
 * the LP entries. */
int sdpadd_problem_dims(sdpadd_problem p, int *m, int *nblock, int *block_struct, int cap);
/* useFormula table of Newton::computeFormula_SDP (sdpa_newton.cpp:694-819):
 * out[i*nsdp + l] in {0: F1, 1: F2, 2: F3}, l over the SDP blocks only. */
int sdpadd_problem_formula(sdpadd_problem p, int *out);
/* Algorithmic dd operations (multiplies and adds) of one Schur assembly
 * with this problem's formula table: per row, A_i Z^{-1} (2 nnz_i n, or
 * 2 n^3 dense) and F1's X (A_i Z^{-1}) (2 n^3); per visited pair, 2 nnz_j
 * (F1), 2 n nnz_j or 2 n^2 (F2), 3 nnz_i nnz_j (F3), with nnz =
 * NonZeroEffect (sdpa_newton.cpp:895-1068).  No counterpart in the
 * reference; it is the numerator of Schur dd-GFLOP/s. */
int sdpadd_problem_schur_ddflops(sdpadd_problem p, double *ddflops);
/* The Schur path of the next solves: 0 = the reference's rule
 * (Chordal::ordering_bMat, sdpa_chordal.cpp:348-440: the sparse path when
 * m > 100, more than 5 blocks counting LP entries, every block's constraint
 * count <= m/2, aggregate sparsity pattern <= m^2/4 and factor pattern
 * 2 nnz - m <= 0.4 m^2), 1 = dense, 2 = sparse whenever the structure
 * allows (the thresholds 0, 0, 2.0, 2.0 of the reference's debugging switch,
 * sdpa_chordal.cpp:53-58).  The sparse path replaces Newton::Make_bMat's
 * dense B and Rpotrf/Rtrsv by the pattern of Newton::initialize_sparse_bMat
 * (sdpa_newton.cpp:225-336), Lal::getCholesky (sdpa_linear.cpp:351-395) and
 * Lal::solveSystems (654-790) with permuteVec (sdpa_newton.cpp:1257-1275);
 * its ordering is an exact minimum degree (ties to the smaller constraint)
 * in place of the reference's Spooles 2.2.  The Ozaki forms apply to the
 * dense path only. */
int sdpadd_problem_set_schur_mode(sdpadd_problem p, int mode);
/* The path the current mode gives: *sparse 1 or 0, *nnz entries of the
 * factor pattern (0 when dense), *aggregate entries of the aggregate
 * sparsity pattern (0 when the rule stopped before forming it), ordering
 * (m ints, may be null; written when sparse): newToOld. */
int sdpadd_problem_schur_info(sdpadd_problem p, int *sparse, int64_t *nnz, int64_t *aggregate, int *ordering);
/* Schur complement of Newton::compute_bMat_dense_SDP (sdpa_newton.cpp:979-1068)
 * for device X and Zinv (blocks concatenated, each n_l x n_l): d_B (m x m)
 * is overwritten. */
int sdpadd_schur_build(sdpadd_problem p, const double *d_X, const double *d_Zinv, double *d_B);
/* g_i = A_i . W (InputData::multi_InnerProductToA, sdpa_dataset.cpp:477-485). */
int sdpadd_inner_A(sdpadd_problem p, const double *d_W, double *d_out);
/* R = sum_i y_i A_i (InputData::multi_plusToA, sdpa_dataset.cpp:488-493). */
int sdpadd_plus_A(sdpadd_problem p, const double *d_y, double *d_R);

/* ---- Solver --------------------------------------------------------------- */
/* param.sdpa (sdpa_parts.cpp:136-183). */
typedef struct {
    int maxIteration;
    double epsilonStar, lambdaStar, omegaStar, lowerBound, upperBound, betaStar, betaBar, gammaStar, epsilonDash;
} sdpadd_params;
void sdpadd_default_params(sdpadd_params *prm);
int sdpadd_read_params(const char *path, sdpadd_params *prm);

typedef struct {
    int iterations;
    int phase;          /* SolveInfo::phaseType after Phase::reverse (what the result file prints) */
    int phase_internal; /* before the primal/dual swap */
    double mu[2], objValPrimal[2], objValDual[2], normPrimalVec[2], normDualMat[2], relgap[2];
    double seconds;     /* main loop wall time */
    double t_schur, t_chol, t_lanczos, t_inverse;
    int refine_sweeps;  /* Ozaki forms: refinement sweeps kept over the solve */
} sdpadd_result;

/* Runs the predictor-corrector loop of sdpa_main.cpp:294-377 on the device.
 * Internal convention: objValPrimal = C.X, objValDual = b.y. */
int sdpadd_solve(sdpadd_problem p, const sdpadd_params *prm, sdpadd_result *res);
/* Per-iteration records of the last solve: 17 doubles each
 * (it, then dd pairs mu, thetaP, thetaD, objP, objD, alphaP, alphaD, beta).
 * Returns the number of records (copies at most cap). */
int sdpadd_trace(sdpadd_problem p, double *out, int cap);
/* Final iterate of the last solve (host arrays: X and Z nn + nlp dd -- the
 * SDP blocks, then the LP entries -- and y m dd). */
int sdpadd_solution(sdpadd_problem p, double *X, double *y, double *Z);

/* ---- Initial points and full-dd solution files ---------------------------- */
/* Reads an initial-point file for the next solves the way sdpa_dd's -is
 * (sparse != 0) and -id options do (sdpa_main.cpp:237-251, IO::read at
 * sdpa_io.cpp:152-285): the file's y negated, its entries set over
 * X = Z = lambdaStar I (sparse: "target block i j value", target 1 = Z,
 * otherwise X, blocks numbered over the SDP blocks and one past them for LP
 * entry i; dense: Z then X in full, each SDP block then the LP entries).  A
 * value is a decimal (read as QD reads it) or an exact pair HI+LO of two
 * doubles as sdpadd_write_solution writes.  The solve then starts from
 * that point: mu0 = X.Z / n and SolveInfo's initial-point form. */
int sdpadd_read_initial_point(sdpadd_problem p, const char *path, int sparse);
/* The same from host arrays in the layout and sign of sdpadd_solution
 * (exact: a solution set back reproduces it bit for bit). */
int sdpadd_set_initial_point(sdpadd_problem p, const double *X, const double *y, const double *Z);
/* Back to the default start X = Z = lambdaStar I, y = 0. */
int sdpadd_clear_initial_point(sdpadd_problem p);
/* Writes the last solve's final point as a sparse initial-point file with
 * every value as the exact pair HI+LO (%.17e each): read back with
 * sdpadd_read_initial_point it restores the point bit for bit; the
 * reference's reader (QD's, which stops after HI's exponent) reads HI.
 * This extends the reference's 17-digit text output (sdpa_io.cpp:799-831,
 * 1015-1203), which cannot carry a dd point. */
int sdpadd_write_solution(sdpadd_problem p, const char *path);

/* ---- Adjoint -------------------------------------------------------------- */
/* No counterpart in the reference (which has no derivatives): implicit
 * differentiation of the central-path equations at the final iterate of the
 * last sdpadd_solve, with the HKM Newton system the solver factors
 * (Newton::compute_bMat_dense_SDP, sdpa_newton.cpp:979-1068; compute_DxMat,
 * :1329-1343).  Internal convention: min C.X s.t. A_i.X = b_i; p* = b.y*.
 * The linearisation is exact only on the central path, so
 * sdpadd_adjoint_setup first takes up to center_steps (0..16) full Newton
 * steps toward the central point at mu = X.Z/n, stopping once
 * max|XZ - mu I| <= 1e-10 mu (8 steps suffice from the solver's last
 * iterate), then assembles the
 * Schur matrix there and factors it once.  *na receives the number of
 * A-gradient entries per cotangent, *centrality max|XZ - mu I| / mu after
 * centring (either may be NULL).  Returns > 0 when Z or the Schur matrix is
 * not positive definite. */
int sdpadd_adjoint_setup(sdpadd_problem p, int center_steps, int64_t *na, double *centrality);
/* Which entry each A-gradient value belongs to (host arrays of na ints, any
 * may be NULL): constraint (0-based), block (0-based), row, column.  Sparse
 * blocks list their stored entries in file order, dense blocks all n*n
 * entries column-major. */
int sdpadd_adjoint_entries(sdpadd_problem p, int *con, int *blk, int *row, int *col);
/* Vector-Jacobian products for nrhs cotangents.  Device arrays (dd pairs):
 * d_Xbar nrhs x nn (cotangent of X*), d_ybar nrhs x m (of y*); host pbar
 * nrhs dd (of p*).  Outputs (device): d_bbar nrhs x m, d_Cbar nrhs x nn (the
 * gradient with respect to the symmetric matrix C), d_Abar nrhs x na (for a
 * stored sparse entry the derivative with respect to its value, which sets
 * both (r, c) and (c, r); for a dense block the gradient with respect to the
 * symmetric matrix).  Inputs and outputs may be NULL (a NULL cotangent is
 * zero).  Synchronous. */
int sdpadd_adjoint(sdpadd_problem p, int nrhs, const double *d_Xbar, const double *d_ybar, const double *pbar,
                   double *d_bbar, double *d_Cbar, double *d_Abar);

#ifdef __cplusplus
}
#endif
#endif /* SDPADD_H */

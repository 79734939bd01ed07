// Dense dd kernels for CDNA4 (gfx950): GEMM, Cholesky, triangular solves,
// inverse of a lower triangle, TRMM, mat-vecs, dots and elementwise ops.
//
// Bit-exactness contract: every output element goes through the same
// sequence of dd operations as in the column-major MPLAPACK loops the CPU
// solver runs (mplapack/Rgemm_NN_omp.cpp, Rsyrk.cpp, Rtrsm.cpp, Rtrmm.cpp,
// Rtrsv.cpp, Rtrmv.cpp, Rgemv.cpp, Rpotrf.cpp, Rpotrf2.cpp, Rdot_omp.cpp).
// Parallelism comes from distributing independent elements; the k order of
// each element's accumulation is never split or reassociated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "ozaki.hpp"

namespace sdk {

// ------------------------------------------------------------- statistics --
// Per-family launch counts, algorithmic dd-flops and HIP-event time.  Family
// names are string literals; entries are found by name in a short vector.
// An optional filter (comma-separated names) limits the timed families, so
// that a timed run pays for the events of the kernel it measures only.
// Launches may come from several host threads (the Lanczos chains): the
// shared tables are guarded by a mutex, and each thread keeps its own stack
// of open scopes.
namespace {
struct StatEntry {
    std::string name;
    int64_t launches = 0;
    double flops = 0, ms = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
bool g_stats_on = false;
std::string g_filter;  // empty: every family
std::vector<StatEntry> g_stats;
std::vector<hipEvent_t> g_event_pool;
thread_local std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> g_open;
std::mutex g_stats_mu;
hipEvent_t get_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    // timing only: no system-scope fence when the event completes, so a
    // record between two launches costs no cache write-back
    hipEvent_t e;
    (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    return e;
}
void resolve(StatEntry &st) {
    for (auto &pr : st.pending) {
        (void)hipEventSynchronize(pr.second);
        float ms = 0;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) st.ms += ms;
        g_event_pool.push_back(pr.first);
        g_event_pool.push_back(pr.second);
    }
    st.pending.clear();
}
int find_entry(const char *family) {
    for (size_t i = 0; i < g_stats.size(); i++)
        if (g_stats[i].name == family) return (int)i;
    return -1;
}
bool selected(const char *family) {
    if (g_filter.empty()) return true;
    const size_t n = std::strlen(family);
    for (size_t p = 0; p <= g_filter.size();) {
        size_t q = g_filter.find(',', p);
        if (q == std::string::npos) q = g_filter.size();
        if (q - p == n && g_filter.compare(p, n, family) == 0) return true;
        p = q + 1;
    }
    return false;
}
}  // namespace

void stats_enable(bool on) { g_stats_on = on; }
void stats_filter(const char *families) { g_filter = families ? families : ""; }
void stats_reset() {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    for (auto &st : g_stats) resolve(st);
    g_stats.clear();
}
bool stats_get(const char *family, int64_t *launches, double *ms, double *ddflops) {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    int i = find_entry(family);
    if (i < 0) {
        *launches = 0;
        *ms = 0;
        *ddflops = 0;
        return false;
    }
    resolve(g_stats[i]);
    *launches = g_stats[i].launches;
    *ms = g_stats[i].ms;
    *ddflops = g_stats[i].flops;
    return true;
}
StatScope::StatScope(const char *family, hipStream_t stream, double ddflops) {
    if (!g_stats_on || !selected(family)) return;
    s = stream;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    int i = find_entry(family);
    if (i < 0) {
        g_stats.push_back(StatEntry());
        g_stats.back().name = family;
        i = (int)g_stats.size() - 1;
    }
    g_stats[i].launches++;
    g_stats[i].flops += ddflops;
    hipEvent_t a = get_event(), b = get_event();
    (void)hipEventRecord(a, s);
    g_open.push_back({i, {a, b}});
    slot = (int)g_open.size() - 1;
}
StatScope::~StatScope() {
    if (slot < 0) return;
    auto pr = g_open[slot];
    (void)hipEventRecord(pr.second.second, s);
    std::lock_guard<std::mutex> lk(g_stats_mu);
    g_stats[pr.first].pending.push_back(pr.second);
    g_open.pop_back();
}

static constexpr int TB = 64;   // GEMM tile (rows and columns)
static constexpr int TK = 16;   // GEMM k-chunk staged in LDS
static constexpr int NT = 256;  // threads per GEMM workgroup

static inline unsigned cdiv(idx_t a, idx_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------ GEMM --
// alpha_kind: 0 general, 1 alpha == 1 (alpha*b == b exactly), 2 alpha == -1.
// beta_kind: 0 beta == 0, 1 beta == 1, 2 general.
// Tile (16*MT) x (16*MT) per 256-thread workgroup, MT x MT outputs per
// thread strided by 16, k staged through LDS in chunks of TK.
template <bool TRANSB, bool PTRS, int MT>
__global__ __launch_bounds__(NT) void k_gemm(idx_t m, idx_t n, idx_t k, dd alpha, int alpha_kind, const dd *__restrict__ A,
                                             idx_t lda, const dd *__restrict__ B, idx_t ldb, dd beta, int beta_kind,
                                             dd *__restrict__ C, idx_t ldc, idx_t sA, idx_t sB, idx_t sC, int lower_only,
                                             idx_t diag_offset, const dd *const *Ap, const dd *const *Bp, dd *const *Cp) {
    constexpr int T = 16 * MT;
    __shared__ dd As[TK][T];
    __shared__ dd Bs[TK][T];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const idx_t row0 = (idx_t)blockIdx.x * T, col0 = (idx_t)blockIdx.y * T;
    if (PTRS) {
        A = Ap[blockIdx.z];
        B = Bp[blockIdx.z];
        C = Cp[blockIdx.z];
    } else {
        A += (idx_t)blockIdx.z * sA;
        B += (idx_t)blockIdx.z * sB;
        C += (idx_t)blockIdx.z * sC;
    }
    // element (i,j) is "lower" when i + diag_offset >= j
    if (lower_only && row0 + T - 1 + diag_offset < col0) return;

    dd acc[MT][MT];
#pragma unroll
    for (int r = 0; r < MT; r++) {
#pragma unroll
        for (int c = 0; c < MT; c++) {
            idx_t i = row0 + tx + 16 * r, j = col0 + ty + 16 * c;
            dd v = mk(0.0);
            if (i < m && j < n && beta_kind != 0) {
                v = C[i + j * ldc];
                if (beta_kind == 2) v = dd_mul(beta, v);
            }
            acc[r][c] = v;
        }
    }
    for (idx_t k0 = 0; k0 < k; k0 += TK) {
        const int kc = (int)((k - k0) < TK ? (k - k0) : TK);
#pragma unroll
        for (int q = 0; q < MT; q++) {
            int e = tid + NT * q;
            int row = e % T, kk = e / T;
            idx_t i = row0 + row;
            As[kk][row] = (i < m && kk < kc) ? A[i + (k0 + kk) * lda] : mk(0.0);
        }
#pragma unroll
        for (int q = 0; q < MT; q++) {
            int e = tid + NT * q;
            int kk, col;
            if (TRANSB) {
                col = e % T;
                kk = e / T;
            } else {
                kk = e & 15;
                col = e >> 4;
            }
            idx_t j = col0 + col;
            dd b = mk(0.0);
            if (j < n && kk < kc) {
                b = TRANSB ? B[j + (k0 + kk) * ldb] : B[(k0 + kk) + j * ldb];
                if (alpha_kind == 0)
                    b = dd_mul(alpha, b);
                else if (alpha_kind == 2)
                    b = dd_neg(b);
            }
            Bs[kk][col] = b;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk++) {
            dd a[MT], b[MT];
#pragma unroll
            for (int r = 0; r < MT; r++) a[r] = As[kk][tx + 16 * r];
#pragma unroll
            for (int c = 0; c < MT; c++) b[c] = Bs[kk][ty + 16 * c];
#pragma unroll
            for (int r = 0; r < MT; r++) {
#pragma unroll
                for (int c = 0; c < MT; c++) acc[r][c] = dd_add_eq(acc[r][c], dd_mul(b[c], a[r]));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < MT; r++) {
#pragma unroll
        for (int c = 0; c < MT; c++) {
            idx_t i = row0 + tx + 16 * r, j = col0 + ty + 16 * c;
            if (i < m && j < n && (!lower_only || i + diag_offset >= j)) C[i + j * ldc] = acc[r][c];
        }
    }
}

// 32x32 tiles when 64x64 tiles would leave most of the 256 CUs idle.
int g_tile_policy = 0;
static bool use_small_tiles(idx_t m, idx_t n, int batch) {
    if (g_tile_policy == 1) return true;
    if (g_tile_policy == 2) return false;
    return (idx_t)cdiv(m, TB) * cdiv(n, TB) * batch < 1024;
}

static int kind_alpha(dd a) {
    if (a.hi == 1.0 && a.lo == 0.0) return 1;
    if (a.hi == -1.0 && a.lo == 0.0) return 2;
    return 0;
}
static int kind_beta(dd b) {
    if (b.hi == 0.0 && b.lo == 0.0) return 0;
    if (b.hi == 1.0 && b.lo == 0.0) return 1;
    return 2;
}

__global__ void k_gemm_scale_only(idx_t m, idx_t n, dd beta, int beta_kind, dd *C, idx_t ldc, idx_t sC) {
    C += (idx_t)blockIdx.z * sC;
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m || j >= n) return;
    dd &c = C[i + j * ldc];
    c = (beta_kind == 0) ? mk(0.0) : dd_mul(beta, c);
}

void gemm(hipStream_t s, char tb, idx_t m, idx_t n, idx_t k, dd alpha, const dd *A, idx_t lda, const dd *B, idx_t ldb,
          dd beta, dd *C, idx_t ldc, int batch, idx_t sA, idx_t sB, idx_t sC, bool lower_only, idx_t diag_offset) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    int ak = kind_alpha(alpha), bk = kind_beta(beta);
    bool alpha_zero = (alpha.hi == 0.0 && alpha.lo == 0.0);
    // Rgemm_omp.cpp:87-106: quick returns and the alpha == 0 path.
    if ((alpha_zero || k == 0) && bk == 1) return;
    if (alpha_zero) {
        dim3 g(cdiv(m, 256), (unsigned)n, (unsigned)batch);
        k_gemm_scale_only<<<g, 256, 0, s>>>(m, n, beta, bk, C, ldc, sC);
        return;
    }
    StatScope st("gemm", s, 2.0 * (double)m * n * k * batch * (lower_only ? 0.5 : 1.0));
    const int lo = lower_only ? 1 : 0;
    const bool tr = (tb == 'T' || tb == 't');
    if (use_small_tiles(m, n, batch)) {
        dim3 grid(cdiv(m, 32), cdiv(n, 32), (unsigned)batch);
        if (tr)
            k_gemm<true, false, 2><<<grid, NT, 0, s>>>(m, n, k, alpha, ak, A, lda, B, ldb, beta, bk, C, ldc, sA, sB, sC, lo,
                                                        diag_offset, nullptr, nullptr, nullptr);
        else
            k_gemm<false, false, 2><<<grid, NT, 0, s>>>(m, n, k, alpha, ak, A, lda, B, ldb, beta, bk, C, ldc, sA, sB, sC, lo,
                                                         diag_offset, nullptr, nullptr, nullptr);
        return;
    }
    dim3 grid(cdiv(m, TB), cdiv(n, TB), (unsigned)batch);
    if (tr)
        k_gemm<true, false, 4><<<grid, NT, 0, s>>>(m, n, k, alpha, ak, A, lda, B, ldb, beta, bk, C, ldc, sA, sB, sC, lo,
                                                    diag_offset, nullptr, nullptr, nullptr);
    else
        k_gemm<false, false, 4><<<grid, NT, 0, s>>>(m, n, k, alpha, ak, A, lda, B, ldb, beta, bk, C, ldc, sA, sB, sC, lo,
                                                     diag_offset, nullptr, nullptr, nullptr);
}

// C = A B in GEMM_KS parts over consecutive k ranges (each part summed in
// ascending k as gemm sums it), the parts then added in order: the Ozaki
// form's n x n products, whose per-element chains are GEMM_KS times shorter
// than Rgemm's.  The last part also takes the remainder of k.
__global__ void k_sum_parts(idx_t m, idx_t n, int parts, const dd *W, dd *C, idx_t ldc) {
    const idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    dd v = W[e];
    for (int p = 1; p < parts; p++) v = dd_add(v, W[(idx_t)p * m * n + e]);
    C[(e % m) + (e / m) * ldc] = v;
}
void gemm_split(hipStream_t s, idx_t m, idx_t n, idx_t k, const dd *A, idx_t lda, const dd *B, idx_t ldb, dd *C,
                idx_t ldc, dd *work) {
    if (m <= 0 || n <= 0) return;
    const idx_t q = k / GEMM_KS, r = k - q * GEMM_KS;
    if (q == 0) {
        gemm(s, 'N', m, n, k, mk(1.0), A, lda, B, ldb, mk(0.0), C, ldc);
        return;
    }
    // parts p = A(:, p q : (p+1) q) B(p q : (p+1) q, :) as one batch
    gemm(s, 'N', m, n, q, mk(1.0), A, lda, B, ldb, mk(0.0), work, m, GEMM_KS, q * lda, q, m * n);
    if (r > 0)
        gemm(s, 'N', m, n, r, mk(1.0), A + GEMM_KS * q * lda, lda, B + GEMM_KS * q, ldb, mk(1.0),
             work + (GEMM_KS - 1) * m * n, m);
    k_sum_parts<<<cdiv(m * n, 256), 256, 0, s>>>(m, n, GEMM_KS, work, C, ldc);
}

void gemm_ptrs(hipStream_t s, idx_t m, idx_t n, idx_t k, const dd *const *A, idx_t lda, const dd *const *B, idx_t ldb,
               dd *const *C, idx_t ldc, int batch) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    StatScope st("gemm", s, 2.0 * (double)m * n * k * batch);
    if (use_small_tiles(m, n, batch)) {
        dim3 grid(cdiv(m, 32), cdiv(n, 32), (unsigned)batch);
        k_gemm<false, true, 2><<<grid, NT, 0, s>>>(m, n, k, mk(1.0), 1, nullptr, lda, nullptr, ldb, mk(0.0), 0, nullptr, ldc,
                                                    0, 0, 0, 0, 0, A, B, C);
        return;
    }
    dim3 grid(cdiv(m, TB), cdiv(n, TB), (unsigned)batch);
    k_gemm<false, true, 4><<<grid, NT, 0, s>>>(m, n, k, mk(1.0), 1, nullptr, lda, nullptr, ldb, mk(0.0), 0, nullptr, ldc, 0, 0,
                                                0, 0, 0, A, B, C);
}

// --------------------------------------------- values as their own messages --
// Values passed between the workgroups of one launch (the fused Cholesky
// step, the block-inverse solve): each double is an agent-scope atomic,
// written once per launch (relaxed) into a slot that holds BI_EMPTY until then, and read (relaxed) until both words
// differ from it.  The value is its own message: no flag, no fence.
// BI_EMPTY is a signalling-NaN pattern, which no arithmetic result takes.
static constexpr unsigned long long BI_EMPTY = 0x7FF4A5A5A5A5A5A5ull;
__device__ __forceinline__ int relaxed_flag(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void put_dd(dd *p, dd v) {
    __hip_atomic_store((unsigned long long *)&p->hi, (unsigned long long)__double_as_longlong(v.hi), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store((unsigned long long *)&p->lo, (unsigned long long)__double_as_longlong(v.lo), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// The read is issued now and checked later (raw_ok), so its latency
// overlaps other work: the check of a value needs the load to be done.
struct RawDD {
    unsigned long long hi, lo;
};
__device__ __forceinline__ RawDD get_raw(const dd *p) {
    RawDD r;
    r.hi = __hip_atomic_load((unsigned long long *)&p->hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r.lo = __hip_atomic_load((unsigned long long *)&p->lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return r;
}
__device__ __forceinline__ bool raw_ok(RawDD r) { return r.hi != BI_EMPTY && r.lo != BI_EMPTY; }
__device__ __forceinline__ dd raw_dd(RawDD r) {
    return dd{__longlong_as_double((long long)r.hi), __longlong_as_double((long long)r.lo)};
}
__device__ __forceinline__ RawDD dd_raw(dd v) {
    return RawDD{(unsigned long long)__double_as_longlong(v.hi), (unsigned long long)__double_as_longlong(v.lo)};
}

// -------------------------------------------------------------- Cholesky --
static constexpr int PB = 64;  // Cholesky panel width
static constexpr int PD_NT = 1024;                                 // threads of the diagonal-block kernel
static constexpr int PD_UPD = PD_NT - 64;                          // update threads (waves 1..)
static constexpr int PD_OWN = (PB * (PB + 1) / 2 + PD_UPD - 1) / PD_UPD;  // elements per update thread

// Lane broadcast of a dd from a wave-uniform lane index.
__device__ __forceinline__ double readlane_d(double v, int k) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)b, k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ dd readlane_dd(dd v, int k) {
    dd r;
    r.hi = readlane_d(v.hi, k);
    r.lo = readlane_d(v.lo, k);
    return r;
}
// Lane shuffle of a dd from any lane (ds_bpermute).
__device__ __forceinline__ dd shfl_dd(dd v, int src) {
    dd r;
    r.hi = __shfl(v.hi, src, 64);
    r.lo = __shfl(v.lo, src, 64);
    return r;
}
// Packed column-major lower triangle of a PB x PB block: (j, k), j >= k.
__device__ __forceinline__ int lpack(int j, int k) { return k * PB - (k * (k - 1)) / 2 + (j - k); }
// Row r and column c (c <= r) of element e of the lower triangle of an
// nb x nb block, column by column.
__device__ __forceinline__ void tri_cm(int e, int nb, int &r, int &c) {
    c = 0;
    while (e >= nb - c) {
        e -= nb - c;
        c++;
    }
    r = c + e;
}

// Diagonal block of the blocked Cholesky (Rpotrf2 order per element:
// a_ij -= l_jk * l_ik for k ascending, skipping l_jk == 0; l_kk =
// sqrt(a_kk); l_ik = (1/l_kk) * a_ik).  One workgroup, one barrier per
// column.  In step k, wave 0 finishes column k+1 on its own: lane i applies
// column k to a_{i,k+1} (the l_ik it formed in step k-1 is still in its
// register; l_{k+1,k} comes by readlane), lane k+1 forms the pivot, and
// every lane scales its element by the broadcast reciprocal -- no LDS round
// trip on the pivot chain.  Meanwhile waves 1-15 apply the scaled column k
// (kept in LDS by wave 0 a step earlier) to columns k+2 and beyond, one
// product and one subtraction per element.  Each update thread owns a fixed
// set of elements of columns >= 2, enumerated column by column, so the
// finished columns are whole waves that skip their slot; it keeps them in
// registers and stores each to LDS once, when wave 0 is about to finish
// its column.
// FAST (the Ozaki form's large factorisations only): the pivot forms
// 1/l_kk directly from the double reciprocal square root of a_kk and one
// Newton step in dd, r = y + y (1 - a y^2) / 2 (about 2^-104 relative, the
// accuracy of QD's square root), and l_kk = a_kk r off the pivot chain;
// half the latency of a dd square root followed by a dd division.
// PUB: every reciprocal and scaled column is also published (put_dd) for
// the panel workgroups of the fused launch.
struct DiagLds {
    dd S[PB][PB + 1];
    dd Lb[2][PB];  // scaled column k at parity k & 1
    dd rl[PB];
    int failed;
};
static constexpr int PD_OWN2 = ((PB - 2) * (PB - 1) / 2 + PD_UPD - 1) / PD_UPD;  // update elements per thread
template <bool FAST, bool PUB>
__device__ __forceinline__ void diag_block(DiagLds &D, dd *A, idx_t lda, idx_t K, int nb, int64_t *info, dd *rinv,
                                           dd *pub) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    {
        // all global loads issued before the LDS stores
        constexpr int NL = PB * PB / PD_NT;
        dd v[NL];
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + PD_NT * q, i = e % PB, j = e / PB;
            v[q] = (i >= j && i < nb) ? A[(K + i) + (K + j) * lda] : mk(0.0);
        }
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + PD_NT * q, i = e % PB, j = e / PB;
            if (i >= j && i < nb) D.S[i][j] = v[q];
        }
    }
    // waves 1..: elements (i, j), 2 <= j <= i < nb, round-robin over the
    // column-major enumeration; own[q] packs i | j << 8, 0 for none (j = 0
    // never takes an update)
    int own[PD_OWN2];
    {
        const int t = tid - 64, m2 = nb - 2;
#pragma unroll
        for (int q = 0; q < PD_OWN2; q++) {
            own[q] = 0;
            const int e = t + PD_UPD * q;
            if (t >= 0 && m2 > 0 && e < m2 * (m2 + 1) / 2) {
                int ri, ci;
                tri_cm(e, m2, ri, ci);
                own[q] = (ri + 2) | ((ci + 2) << 8);
            }
        }
    }
    if (tid == 0) D.failed = 0;
    __syncthreads();
    // the owned elements, in registers from here on
    dd sv[PD_OWN2];
#pragma unroll
    for (int q = 0; q < PD_OWN2; q++) sv[q] = own[q] ? D.S[own[q] & 0xff][own[q] >> 8] : mk(0.0);
    dd *pubr = pub + PB * PB;
    // wave 0: finish column c (apply column c-1 unless c == 0, factor the
    // pivot, scale); li carries l_{lane,c-1} in and l_{lane,c} out
    dd li = mk(0.0);
    bool failed0 = false;  // wave 0: a pivot failed (D.failed is set too)
    auto finish_col = [&](int c) {
        const bool mine = lane >= c && lane < nb;
        dd a = mine ? D.S[lane][c] : mk(0.0);
        if (c > 0) {
            const dd lj = readlane_dd(li, c);
            if (mine && !dd_is_zero(lj)) a = dd_sub(a, dd_mul(lj, li));
        }
        int bad = 0;
        dd r = mk(0.0);
        if (lane == c) {
            if (dd_le(a, mk(0.0)) || dd_isnan(a)) {
                bad = 1;
            } else if (FAST) {
                const double y = rsqrt(a.hi);
                const dd e = d_sub_dd(1.0, dd_mul(a, dd_sqr_d(y)));
                double lo;
                const double hi = two_sum(y, e.hi * (0.5 * y), lo);
                r = dd{hi, lo};
                A[(K + c) + (K + c) * lda] = dd_mul(a, r);
            } else {
                const dd l = dd_sqrt(a);
                A[(K + c) + (K + c) * lda] = l;
                r = dd_div(mk(1.0), l);
            }
        }
        if (__builtin_amdgcn_readlane(bad, c)) {
            failed0 = true;
            if (lane == 0) {
                D.failed = 1;
                __hip_atomic_store(info, (int64_t)(K + c + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
        r = readlane_dd(r, c);
        if (lane == c) {
            D.rl[c] = r;
            if (PUB) put_dd(pubr + c, r);
        }
        li = mine && lane > c ? dd_mul(r, a) : mk(0.0);
        if (mine && lane > c) {
            A[(K + lane) + (K + c) * lda] = li;
            D.Lb[c & 1][lane] = li;
            if (PUB) put_dd(pub + c * PB + lane, li);
        }
    };
    if (wave == 0) finish_col(0);
    __syncthreads();
    // The largest column among a wave's q-th elements (wave-uniform): the
    // wave skips its q-th updates once that column is finished.  The q = 0
    // and q = 1 updates run branch-free in one block while both are live,
    // so their dependent chains interleave; q = 2 holds only a few elements.
    static_assert(PD_OWN2 == 3, "update element layout");
    int jmax[PD_OWN2];
#pragma unroll
    for (int q = 0; q < PD_OWN2; q++) {
        int m = own[q] >> 8;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
        jmax[q] = __builtin_amdgcn_readfirstlane(m);
    }
    // the loop always runs to the end, so every wave passes the same
    // barriers; after a failed pivot wave 0 does nothing more and the
    // update waves' further results are never stored
    for (int k = 0; k + 1 < nb; k++) {
        if (wave == 0) {
            if (!failed0) finish_col(k + 1);
        } else {
            // column k (scaled) to the owned elements of columns >= k + 2,
            // held in registers; an element goes to LDS once, when its
            // column is the one wave 0 finishes next (j = k + 2)
            const dd *lk = D.Lb[k & 1];
            auto upd = [&](int q) {
                const int i = own[q] & 0xff, j = own[q] >> 8;
                const dd lj = lk[j], lv = lk[i];
                const dd u = dd_sub(sv[q], dd_mul(lj, lv));
                const bool on = j >= k + 2 && !dd_is_zero(lj);
                sv[q].hi = on ? u.hi : sv[q].hi;
                sv[q].lo = on ? u.lo : sv[q].lo;
            };
            if (k + 2 <= jmax[0]) {
                upd(0);
                upd(1);
            } else if (k + 2 <= jmax[1]) {
                upd(1);
            }
            if (k + 2 <= jmax[2]) upd(2);
#pragma unroll
            for (int q = 0; q < PD_OWN2; q++) {
                const int i = own[q] & 0xff, j = own[q] >> 8;
                if (own[q] && j == k + 2) D.S[i][j] = sv[q];
            }
        }
        __syncthreads();
    }
    if (D.failed) return;
    for (int k = tid; k < nb; k += PD_NT) rinv[k] = D.rl[k];
}

template <bool FAST>
__global__ __launch_bounds__(PD_NT) void k_potrf_diag(dd *A, idx_t lda, idx_t K, int nb, int64_t *info, dd *rinv) {
    __shared__ DiagLds D;
    if (*info != 0) return;
    diag_block<FAST, false>(D, A, lda, K, nb, info, rinv, nullptr);
}

// B := B * L^{-T} on an nb-column panel (Rtrsm 'R','L','T','N' order with
// reciprocal scaling: b_rk *= rl_k, then b_rj -= l_jk * b_rk for j > k,
// skipping l_jk == 0), rows [r0, r0 + nrows).  L is the nb x nb lower block
// at Lb.  PR rows per workgroup are staged through LDS.  A wave holds four
// rows, 16 lanes each: lane c of a row holds columns c, 16 + c, 32 + c and
// 48 + c in registers, so every lane does useful work at every step (a
// step costs a wave one scaling, a shuffle and up to four updates for four
// rows), and the solved column is broadcast within each 16-lane group.
// Each element still sees the column loop's operation sequence.
static constexpr int PW_ROWS = 4;        // rows per wave
static constexpr int PR = 4 * PW_ROWS;   // rows per 256-thread workgroup
__global__ __launch_bounds__(256) void k_panel_rltn(dd *B, idx_t ldb, idx_t r0, idx_t nrows, const dd *Lb, idx_t lda, int nb,
                                                    const dd *rinv, const int64_t *info) {
    __shared__ dd Lc[PB * (PB + 1) / 2];
    __shared__ dd rl[PB];
    __shared__ dd Tt[PR][PB + 1];
    if (info && *info != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, c = lane & 15;
    const idx_t rb = r0 + (idx_t)blockIdx.x * PR;
    const int nr = (int)((r0 + nrows - rb) < PR ? (r0 + nrows - rb) : PR);
    {
        // every global load of the staging is issued before the LDS stores,
        // so their latencies overlap instead of adding up
        constexpr int NL = PB * PB / 256, NT = PR * PB / 256;
        dd lv[NL], tv[NT];
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + 256 * q, j = e % PB, k = e / PB;
            lv[q] = (j >= k && j < nb) ? Lb[j + (idx_t)k * lda] : mk(0.0);
        }
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const int e = tid + 256 * q, r = e % PR, j = e / PR;
            tv[q] = (r < nr && j < nb) ? B[(rb + r) + (idx_t)j * ldb] : mk(0.0);
        }
        const dd rv = tid < nb ? rinv[tid] : mk(0.0);
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int e = tid + 256 * q, j = e % PB, k = e / PB;
            if (j >= k && j < nb) Lc[lpack(j, k)] = lv[q];
        }
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const int e = tid + 256 * q, r = e % PR, j = e / PR;
            Tt[r][j] = tv[q];
        }
        if (tid < PB) rl[tid] = rv;
    }
    __syncthreads();
    const int row = PW_ROWS * wave + g;
    dd x[4];
#pragma unroll
    for (int b = 0; b < 4; b++) x[b] = (16 * b + c < nb) ? Tt[row][16 * b + c] : mk(0.0);
#pragma unroll
    for (int b = 0; b < 4; b++) {
        if (16 * b >= nb) break;
        const int kend = (nb - 16 * b) < 16 ? (nb - 16 * b) : 16;
        for (int kk = 0; kk < kend; kk++) {
            const int k = 16 * b + kk;
            dd l[4];
#pragma unroll
            for (int b2 = b; b2 < 4; b2++) {
                const int j = 16 * b2 + c;
                l[b2] = (j > k && j < nb) ? Lc[lpack(j, k)] : mk(0.0);
            }
            if (c == kk) x[b] = dd_mul(rl[k], x[b]);
            const dd xk = shfl_dd(x[b], (lane & 48) | kk);
#pragma unroll
            for (int b2 = b; b2 < 4; b2++)
                if (!dd_is_zero(l[b2])) x[b2] = dd_sub(x[b2], dd_mul(l[b2], xk));
        }
    }
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (16 * b + c < nb) Tt[row][16 * b + c] = x[b];
    __syncthreads();
    for (int e = tid; e < PR * nb; e += 256) {
        int r = e % PR, j = e / PR;
        if (r < nr) B[(rb + r) + (idx_t)j * ldb] = Tt[r][j];
    }
}

__global__ void k_zero_i64(int64_t *p) { *p = 0; }

// ------------------------------------ diagonal block and panel in one launch --
// One launch factors a diagonal block and solves the panel rows below it.
// The first workgroup to start (a ticket from a counter, so it is resident
// by construction) runs k_potrf_diag's schedule and publishes, as it goes,
// each pivot reciprocal r_k and each scaled column l_jk (j > k) into a
// buffer whose slots hold BI_EMPTY until written (every value is its own
// message: no flag, no fence).  Every other workgroup takes FR panel rows:
// its last wave receives the published columns into LDS in order, a few
// columns per round trip, and advances an LDS counter; the other waves hold
// two rows each (32 lanes per row, columns c and 32 + c) and apply column k
// as soon as the counter passes it.  A panel workgroup waits only for the
// diagonal workgroup, which waits for nothing, so the grid always drains;
// a wait that still runs out of its bound (or sees a failed pivot) leaves
// without storing and, unless a pivot failed, sets info to
// POTRF_WAIT_FAILED.  Each element sees the operation sequence of
// k_potrf_diag followed by k_panel_rltn: the results are identical.
// The publication buffers alternate between launches; the diagonal
// workgroup refills the other one with BI_EMPTY for the next launch.
static constexpr int FP_WAVES = PD_NT / 64 - 1;  // row waves of a panel workgroup
static constexpr int FR = 2 * FP_WAVES;          // panel rows per workgroup
static constexpr int FP_AHEAD = 4;               // columns read per round trip
static constexpr int PUB_DD = PB * PB + PB;      // one publication buffer: columns, then reciprocals
static_assert(POTRF_SCRATCH_DD >= PB + 2 * PUB_DD + 1, "Cholesky scratch layout");

// bound on a panel workgroup's consecutive empty polls (0: 2^22, seconds)
unsigned g_panel_spin = 0;
static unsigned panel_spin() { return g_panel_spin ? g_panel_spin : (1u << 22); }

struct PotrfPub {
    dd *rinv, *buf[2];
    unsigned long long *ctr;
    explicit PotrfPub(dd *scratch)
        : rinv(scratch), buf{scratch + PB, scratch + PB + PUB_DD},
          ctr(reinterpret_cast<unsigned long long *>(scratch + PB + 2 * PUB_DD)) {}
};

// info = 0, ticket counter = 0, both publication buffers empty.
__global__ void k_potrf_init(int64_t *info, dd *pub, unsigned long long *ctr) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const dd empty = {__longlong_as_double((long long)BI_EMPTY), __longlong_as_double((long long)BI_EMPTY)};
    if (e < 2 * PUB_DD) pub[e] = empty;
    if (e == 0) {
        *info = 0;
        *ctr = 0;
    }
}

// Digit planes of the panel rows for the Ozaki update that follows (the
// Ozaki form only): the 16 planes and tagged row exponents of rows
// [0, below) of the panel, as oz::gemm's slicing would write them.
struct PanelPlanes {
    int8_t *planes = nullptr;  // plane s, chunk kb of row r: planes[((s nkb + kb) rows_pad + r) 64 + k]
    unsigned *exps = nullptr;
    idx_t rows_pad = 0;
    unsigned tag = 0;
    // nkb = 1: the block's own 64 columns.  nkb > 1 (the last block of a
    // panel): the whole panel's 64 nkb columns, starting at column pcol,
    // the earlier blocks' read back from A (final by then)
    int nkb = 1;
    idx_t pcol = 0;
};

template <bool FAST, bool SL>
__global__ __launch_bounds__(PD_NT) void k_potrf_fused(dd *A, idx_t lda, idx_t K, int nb, idx_t below, int64_t *info,
                                                       dd *rinv, dd *pub, dd *pub_next, unsigned long long *ctr,
                                                       unsigned long long base, unsigned spin, PanelPlanes pl) {
    struct PanelLds {
        dd Lc[PB * (PB + 1) / 2];
        dd rl[PB];
        dd Tt[FR][PB + 1];
        int ready, stop;
    };
    __shared__ union {
        DiagLds d;
        PanelLds p;
    } sm;
    __shared__ unsigned long long ticket;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) ticket = __hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - base;
    __syncthreads();
    const unsigned long long role = ticket;
    if (__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    dd *pubr = pub + PB * PB;
    if (role == 0) {
        // ---- the diagonal block, publishing as it goes ----
        diag_block<FAST, true>(sm.d, A, lda, K, nb, info, rinv, pub);
        if (sm.d.failed) return;
        const dd empty = {__longlong_as_double((long long)BI_EMPTY), __longlong_as_double((long long)BI_EMPTY)};
        for (int e = tid; e < PUB_DD; e += PD_NT) pub_next[e] = empty;
        return;
    }
    // ---- panel rows [rb, rb + nr) ----
    auto &P = sm.p;
    const idx_t r0 = K + nb, rb = r0 + (idx_t)(role - 1) * FR;
    const int nr = (int)((r0 + below - rb) < FR ? (r0 + below - rb) : FR);
    dd *B = A + K * lda;
    {
        constexpr int NT = (FR * PB + PD_NT - 1) / PD_NT;
        dd tv[NT];
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const int e = tid + PD_NT * q, r = e % FR, j = e / FR;
            tv[q] = (e < FR * PB && r < nr && j < nb) ? B[(rb + r) + (idx_t)j * lda] : mk(0.0);
        }
#pragma unroll
        for (int q = 0; q < NT; q++) {
            const int e = tid + PD_NT * q, r = e % FR, j = e / FR;
            if (e < FR * PB) P.Tt[r][j] = tv[q];
        }
        if (tid == 0) {
            P.ready = 0;
            P.stop = 0;
        }
    }
    __syncthreads();
    if (wave == FP_WAVES) {
        // receiver: columns k0 .. k0 + FP_AHEAD - 1 read together, the
        // valid ones taken in order
        int k0 = 0;
        unsigned empty_polls = 0;
        while (k0 < nb) {
            RawDD v[FP_AHEAD], rr[FP_AHEAD];
#pragma unroll
            for (int d = 0; d < FP_AHEAD; d++) {
                const int kc = k0 + d;
                v[d] = RawDD{0, 0};
                rr[d] = RawDD{0, 0};
                if (kc < nb) {
                    if (lane > kc && lane < nb) v[d] = get_raw(pub + kc * PB + lane);
                    if (lane == 0) rr[d] = get_raw(pubr + kc);
                }
            }
            int got = 0;
#pragma unroll
            for (int d = 0; d < FP_AHEAD; d++) {
                const int kc = k0 + d;
                if (kc >= nb || got < d) break;
                const bool ok = (lane <= kc || lane >= nb || raw_ok(v[d])) && (lane != 0 || raw_ok(rr[d]));
                if (!__all(ok)) break;
                if (lane > kc && lane < nb) P.Lc[lpack(lane, kc)] = raw_dd(v[d]);
                if (lane == 0) P.rl[kc] = raw_dd(rr[d]);
                got++;
            }
            if (got) {
                k0 += got;
                empty_polls = 0;
                if (lane == 0) __hip_atomic_store(&P.ready, k0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                continue;
            }
            if (++empty_polls >= spin || __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                if (lane == 0) {
                    unsigned long long zero = 0;
                    __hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long *>(info), &zero,
                                                         (unsigned long long)POTRF_WAIT_FAILED, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    P.stop = 1;
                    __hip_atomic_store(&P.ready, PB + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    } else {
        // two rows per wave; lane c of a row holds columns c and 32 + c
        const int g = lane >> 5, c = lane & 31, row = 2 * wave + g;
        dd x[2];
#pragma unroll
        for (int b = 0; b < 2; b++) x[b] = (32 * b + c < nb) ? P.Tt[row][32 * b + c] : mk(0.0);
        int seen = 0;  // columns known to be in LDS
        bool stopped = false;
#pragma unroll
        for (int b = 0; b < 2; b++) {
            if (32 * b >= nb || stopped) break;
            const int kend = (nb - 32 * b) < 32 ? (nb - 32 * b) : 32;
            for (int kk = 0; kk < kend; kk++) {
                const int k = 32 * b + kk;
                if (k >= seen) {
                    while ((seen = __hip_atomic_load(&P.ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) <= k)
                        __builtin_amdgcn_s_sleep(1);
                    if (P.stop) {
                        stopped = true;
                        break;
                    }
                }
                dd l[2];
#pragma unroll
                for (int b2 = b; b2 < 2; b2++) {
                    const int j = 32 * b2 + c;
                    l[b2] = (j > k && j < nb) ? P.Lc[lpack(j, k)] : mk(0.0);
                }
                if (c == kk) x[b] = dd_mul(P.rl[k], x[b]);
                const dd xk = shfl_dd(x[b], (lane & 32) | kk);
#pragma unroll
                for (int b2 = b; b2 < 2; b2++)
                    if (!dd_is_zero(l[b2])) x[b2] = dd_sub(x[b2], dd_mul(l[b2], xk));
            }
        }
#pragma unroll
        for (int b = 0; b < 2; b++)
            if (32 * b + c < nb) P.Tt[row][32 * b + c] = x[b];
    }
    __syncthreads();
    if (P.stop) return;
    for (int e = tid; e < FR * nb; e += PD_NT) {
        const int r = e % FR, j = e / FR;
        if (r < nr) B[(rb + r) + (idx_t)j * lda] = P.Tt[r][j];
    }
    if (SL && pl.nkb == 1) {
        // 16 lanes per row, 4 consecutive k each; the row's exponent is the
        // maximum over its 16 lanes (k_slice's value: frexp of max |hi|)
        const int r = tid >> 4, q = tid & 15;
        if (r < FR) {
            dd av[4];
            double mx = 0.0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int kk = 4 * q + u;
                av[u] = (r < nr && kk < nb) ? P.Tt[r][kk] : mk(0.0);
                mx = fmax(mx, fabs(av[u].hi));
            }
#pragma unroll
            for (int o = 8; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 16));
            int e = 0;
            if (mx > 0.0) (void)frexp(mx, &e);
            if (r < nr) {
                const idx_t lr = rb - r0 + r;
                if (q == 0) pl.exps[lr] = oz::exp_tagged(pl.tag, e);
                unsigned w4[16];
#pragma unroll
                for (int t = 0; t < 16; t++) w4[t] = 0u;
#pragma unroll
                for (int u = 0; u < 4; u++) oz::put_digits<16>(av[u], e, w4, u);
#pragma unroll
                for (int t = 0; t < 16; t++)
                    *(unsigned *)(pl.planes + (((idx_t)t * pl.rows_pad + lr) * 64 + 4 * q)) = w4[t];
            }
        }
    } else if (SL) {
        // the whole panel (nkb <= 4 chunks of 64 columns) as k_rowexp and
        // k_slice form it: thread (r, g) = (tid % 32, tid / 32) takes row r,
        // columns [8 g, 8 g + 8), so a column's rows are read by consecutive
        // threads (the earlier blocks' columns from A, this block's from
        // LDS); the row's maximum meets in LDS over the 32 groups
        const int nkb = pl.nkb, own0 = 64 * (nkb - 1), ncol = 64 * nkb;
        const int r = tid & 31, g = tid >> 5;
        double *part = reinterpret_cast<double *>(&P.Lc[0]);  // [32 groups][32 rows]
        static_assert(sizeof(P.Lc) >= 32 * 32 * sizeof(double), "row maxima fit over Lc");
        const bool live = r < nr && 8 * g < ncol;
        dd av[8];
        double mx = 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = 8 * g + u;
            dd v = mk(0.0);
            if (live) v = c < own0 ? A[(rb + r) + (pl.pcol + c) * lda] : P.Tt[r][c - own0];
            av[u] = v;
            mx = fmax(mx, fabs(v.hi));
        }
        __syncthreads();  // Lc's last readers are done
        part[g * 32 + r] = mx;
        __syncthreads();
        double rm = 0.0;
#pragma unroll 8
        for (int h = 0; h < 32; h++) rm = fmax(rm, part[h * 32 + r]);
        int e = 0;
        if (rm > 0.0) (void)frexp(rm, &e);
        if (live) {
            const idx_t lr = rb - r0 + r;
            if (g == 0) pl.exps[lr] = oz::exp_tagged(pl.tag, e);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int c = 8 * g + 4 * h, kb = c / 64;
                unsigned w4[16];
#pragma unroll
                for (int t = 0; t < 16; t++) w4[t] = 0u;
#pragma unroll
                for (int u = 0; u < 4; u++) oz::put_digits<16>(av[4 * h + u], e, w4, u);
#pragma unroll
                for (int t = 0; t < 16; t++)
                    *(unsigned *)(pl.planes + ((((idx_t)t * nkb + kb) * pl.rows_pad + lr) * 64 + (c % 64))) = w4[t];
            }
        }
    }
}

// Diagonal block K (nb columns) and the `below` rows under it: one fused
// launch when there are rows below, the diagonal kernel alone otherwise.
// `launches` counts the fused launches of this factorisation (buffer
// parity and ticket base).
template <bool FAST>
static void potrf_step(hipStream_t s, idx_t K, int nb, idx_t below, dd *A, idx_t lda, int64_t *d_info,
                       const PotrfPub &pp, unsigned long long &base, int &launches, const PanelPlanes *pl = nullptr) {
    StatScope st("potrf", s, (double)nb * nb * nb / 3.0 + (double)below * nb * nb);
    if (below <= 0) {
        k_potrf_diag<FAST><<<1, PD_NT, 0, s>>>(A, lda, K, nb, d_info, pp.rinv);
        return;
    }
    const unsigned grid = 1 + cdiv(below, FR);
    const int par = launches & 1;
    if (pl)
        k_potrf_fused<FAST, true><<<grid, PD_NT, 0, s>>>(A, lda, K, nb, below, d_info, pp.rinv, pp.buf[par],
                                                         pp.buf[par ^ 1], pp.ctr, base, panel_spin(), *pl);
    else
        k_potrf_fused<FAST, false><<<grid, PD_NT, 0, s>>>(A, lda, K, nb, below, d_info, pp.rinv, pp.buf[par],
                                                          pp.buf[par ^ 1], pp.ctr, base, panel_spin(), PanelPlanes());
    base += grid;
    launches++;
}

static void potrf_begin(hipStream_t s, int64_t *d_info, const PotrfPub &pp) {
    k_potrf_init<<<cdiv(2 * PUB_DD, 256), 256, 0, s>>>(d_info, pp.buf[0], pp.ctr);
}

void potrf_lower(hipStream_t s, idx_t n, dd *A, idx_t lda, int64_t *d_info, dd *scratch) {
    const PotrfPub pp(scratch);
    potrf_begin(s, d_info, pp);
    unsigned long long base = 0;
    int launches = 0;
    for (idx_t K = 0; K < n; K += PB) {
        int nb = (int)std::min<idx_t>(PB, n - K);
        idx_t below = n - K - nb;
        potrf_step<false>(s, K, nb, below, A, lda, d_info, pp, base, launches);
        if (below > 0) {
            // trailing lower triangle: A22 += (-L21) * L21^T, k ascending
            dd *A21 = A + (K + nb) + K * lda;
            dd *A22 = A + (K + nb) + (K + nb) * lda;
            gemm(s, 'T', below, below, nb, mk(-1.0), A21, lda, A21, lda, mk(1.0), A22, lda, 1, 0, 0, 0, true, 0);
        }
    }
}

// Cholesky with Ozaki trailing updates (the precision path): panels of
// 64 or 256 columns are factored 64 columns at a time with the dd kernels
// above (diagonal blocks, panel solves), the rest of a panel updated after
// each 64 columns with an Ozaki product (k = 64); the lower triangle right
// of a panel is updated with the INT8-MFMA product A22 -= L21 L21^T (k =
// panel width).  Not bit-identical to Rpotrf; its accuracy is that of the Ozaki
// product (normwise dd for 16 digit planes).
int potrf_lower_oz(hipStream_t s, idx_t n, dd *A, idx_t lda, int64_t *d_info, dd *scratch, oz::Work &w, int S,
                   const OzLookAhead *la) {
    // panel width: 256 (four k chunks per Ozaki update) for large matrices,
    // 64 below 1024 so that small ones still take most updates on MFMA
    const idx_t OZ_NB = n > 1024 ? 256 : 64;
    // the split costs a second slicing of the panel and two events: it pays
    // only where the trailing updates are long (measured: m = 4551 gains 5%,
    // m = 800 loses 14% of the Cholesky time)
    if (n <= 1024) la = nullptr;
    // the pivot chain (m square roots and reciprocals in sequence) is the
    // critical path of every factorisation here: all take the faster pivot
    const bool fast_pivot = true;
    const PotrfPub pp(scratch);
    potrf_begin(s, d_info, pp);
    unsigned long long base = 0;
    int launches = 0;
    bool helper_pending = false;  // an update on la->hs that later work on s must wait for
    // with look-ahead, the last block's launch of each full panel writes the
    // digit planes of the whole panel (its rows are the rows below the
    // panel), which the look-ahead and the helper's update then share: one
    // slicing launch pair less on each stream.  Panels alternate between
    // two plane buffers, so the helper can still read one panel's planes
    // while the next panel's are written.
    const bool panel_planes = la && la->w2 && la->wp && S == 16;
    int panel_no = 0;
    for (idx_t P = 0; P < n; P += OZ_NB, panel_no++) {
        const idx_t pw = std::min<idx_t>(OZ_NB, n - P);
        // a one-block panel without look-ahead: its trailing update (k = 64)
        // reads the same planes as an in-panel update would
        const bool trail64 = pw == PB && !la;
        unsigned trail_tag = 0;
        oz::Work *pbuf = panel_planes ? (panel_no % 2 ? la->wp : la->w2) : nullptr;
        unsigned panel_tag = 0;
        idx_t panel_rows_pad = 0;
        for (idx_t K = P; K < P + pw; K += PB) {
            const int nb = (int)std::min<idx_t>(PB, P + pw - K);
            const idx_t below = n - K - nb;
            const idx_t rem = P + pw - K - nb;  // panel columns right of this block
            // with 16 digit planes the panel workgroups also write the
            // planes of their rows for the update (in-panel or trailing)
            // that follows
            PanelPlanes pl;
            const bool fuse = (rem > 0 || trail64) && S == 16 && nb == PB && below > 0;
            const bool last_planes = panel_planes && rem == 0 && nb == PB && pw % PB == 0 && below > 0;
            if (fuse) {
                pl.tag = oz::prepare_planes(s, w, below);
                if (!pl.tag) return -2;
                pl.planes = w.sa;
                pl.exps = w.ea;
                pl.rows_pad = (below + 63) / 64 * 64;
                if (trail64) trail_tag = pl.tag;
            } else if (last_planes) {
                pl.nkb = (int)(pw / PB);
                pl.pcol = P;
                pl.tag = oz::prepare_panel_planes(s, *pbuf, below, pl.nkb);
                if (!pl.tag) return -2;
                pl.planes = pbuf->sa;
                pl.exps = pbuf->ea;
                pl.rows_pad = (below + 63) / 64 * 64;
                panel_tag = pl.tag;
                panel_rows_pad = pl.rows_pad;
            }
            if (fast_pivot)
                potrf_step<true>(s, K, nb, below, A, lda, d_info, pp, base, launches,
                                 fuse || last_planes ? &pl : nullptr);
            else
                potrf_step<false>(s, K, nb, below, A, lda, d_info, pp, base, launches,
                                  fuse || last_planes ? &pl : nullptr);
            if (rem > 0) {
                // the panel's remaining columns, every row below: an Ozaki
                // product with k = 64 (on the FP64 VALU this rank-64 update
                // was the panel's largest cost)
                const dd *A21 = A + (K + nb) + K * lda;
                dd *C = A + (K + nb) + (K + nb) * lda;
                int rc = fuse ? oz::gemm_planes(s, w, pl.tag, below, rem, nb, -1, 1, C, lda, true)
                              : oz::gemm(s, w, 'N', 'T', below, rem, nb, -1, A21, lda, A21, lda, 1, C, lda, true, S);
                if (rc != 0) return rc;
            }
        }
        const idx_t Q = P + pw, below = n - Q;
        if (below <= 0) break;
        const dd *L21 = A + Q + P * lda;
        if (!la) {
            int rc = trail_tag ? oz::gemm_planes(s, w, trail_tag, below, below, pw, -1, 1, A + Q * (1 + lda), lda, true)
                               : oz::gemm(s, w, 'N', 'T', below, below, pw, -1, L21, lda, L21, lda, 1, A + Q * (1 + lda),
                                          lda, true, S);
            if (rc != 0) return rc;
            continue;
        }
        // the previous panel's remaining update covered these columns too
        if (helper_pending) {
            (void)hipStreamWaitEvent(s, la->done, 0);
            helper_pending = false;
        }
        const idx_t qw = std::min<idx_t>(OZ_NB, below), R = Q + qw;
        // next panel's columns [Q, R), every row below Q, on s
        const int pnkb = (int)(pw / PB);
        int rc = panel_tag ? oz::gemm_panel_planes(s, *pbuf, panel_tag, panel_rows_pad, pnkb, 0, below, qw, -1, 1,
                                                   A + Q * (1 + lda), lda, true)
                           : oz::gemm(s, w, 'N', 'T', below, qw, pw, -1, L21, lda, L21, lda, 1, A + Q * (1 + lda), lda,
                                      true, S);
        if (rc != 0) return rc;
        if (R < n) {
            // the rest: rows and columns from R, on the helper stream
            (void)hipEventRecord(la->ready, s);
            (void)hipStreamWaitEvent(la->hs, la->ready, 0);
            const dd *L31 = A + R + P * lda;
            rc = panel_tag && qw % PB == 0
                     ? oz::gemm_panel_planes(la->hs, *pbuf, panel_tag, panel_rows_pad, pnkb, qw, n - R, n - R, -1, 1,
                                             A + R * (1 + lda), lda, true)
                     : oz::gemm(la->hs, *la->w2, 'N', 'T', n - R, n - R, pw, -1, L31, lda, L31, lda, 1,
                                A + R * (1 + lda), lda, true, S);
            if (rc != 0) return rc;
            (void)hipEventRecord(la->done, la->hs);
            helper_pending = true;
        }
    }
    if (helper_pending) (void)hipStreamWaitEvent(s, la->done, 0);
    return 0;
}

// ------------------------------------------------------ triangular solves --
// Triangular solves with one 512-thread workgroup, in blocks of SB rows
// with one block of look-ahead.  In round R, wave 0 applies block R-1's
// solved values to block R (one row per lane) and then solves block R's
// diagonal part, broadcasting each solved value with readlane, while waves
// 1-7 apply block R-1's values to every later row (two rows per thread).  Each element still sees
// exactly the operation sequence of the column loop of Rtrsv.
static constexpr int SB = 32;
static constexpr int TS_NT = 512;          // threads of the triangular-solve kernels
static constexpr int TS_UPD = TS_NT - 64;  // update threads (waves 1..)

// The recovery of a cross-workgroup solve (below) whose bounded wait ran out:
// with redo non-null the single-workgroup solves run only when *redo (that
// solve's error word) is set, and then start from the saved right-hand side
// x0.  The word was written by an earlier kernel on the stream, so a plain
// load sees it; the restore is made visible to the workgroup by the barrier.
__device__ __forceinline__ bool trsv_redo(const int *redo, const dd *x0, dd *x, idx_t n) {
    if (!redo) return true;
    if (*redo == 0) return false;
    for (idx_t i = threadIdx.x; i < n; i += blockDim.x) x[i] = x0[i];
    __syncthreads();
    return true;
}

// x := L^{-1} x ('L','N','N'): for j ascending, if x_j != 0 { x_j /= L_jj;
// x_i -= x_j * L_ij for i > j }.
__global__ __launch_bounds__(TS_NT) void k_trsv_ln(idx_t n, const dd *L, idx_t lda, dd *x, const int *redo,
                                                   const dd *x0) {
    __shared__ dd D[SB][SB + 1];
    __shared__ dd v[2][SB];
    __shared__ int nzv[2][SB];
    if (!trsv_redo(redo, x0, x, n)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int nblk = (int)((n + SB - 1) / SB);
    for (int R = 0; R < nblk; R++) {
        const idx_t b = (idx_t)R * SB, bp = b - SB;  // this block and the previous (full) one
        const int nb = (int)((n - b) < SB ? (n - b) : SB);
        const int cur = R & 1, prv = cur ^ 1;
        if (tid < 64) {
            if (nb > 0) {
                for (int e = lane; e < SB * SB; e += 64) {
                    int r = e % SB, c = e / SB;
                    if (r < nb && c < nb && r >= c) D[r][c] = L[(b + r) + (b + c) * lda];
                }
                dd xr = lane < nb ? x[b + lane] : mk(0.0);
                if (R > 0 && lane < nb) {
                    // block R-1 (full, SB columns) applied to this row
                    dd a[SB];
#pragma unroll
                    for (int jj = 0; jj < SB; jj++) a[jj] = L[(b + lane) + (bp + jj) * lda];
#pragma unroll
                    for (int jj = 0; jj < SB; jj++)
                        if (nzv[prv][jj]) xr = dd_sub(xr, dd_mul(v[prv][jj], a[jj]));
                }
                __builtin_amdgcn_wave_barrier();
                // this lane's row of the diagonal block in registers (entries
                // right of the diagonal are never used); the step loop is
                // unrolled so that register and lane indices are constants
                dd drow[SB];
                const int lr = lane < nb ? lane : 0;
#pragma unroll
                for (int jj = 0; jj < SB; jj++) drow[jj] = D[lr][jj <= lr ? jj : lr];
#pragma unroll
                for (int jj = 0; jj < SB; jj++) {
                    if (jj < nb) {
                        const int nz = __builtin_amdgcn_readlane(!dd_is_zero(xr), jj);
                        if (lane == jj && nz) xr = dd_div(xr, drow[jj]);
                        const dd t = readlane_dd(xr, jj);
                        if (nz && lane > jj && lane < nb) xr = dd_sub(xr, dd_mul(t, drow[jj]));
                        if (lane == jj) {
                            v[cur][jj] = xr;
                            nzv[cur][jj] = nz;
                        }
                    }
                }
                if (lane < nb) x[b + lane] = xr;
            }
        } else if (R > 0) {
            // rows after this block get block R-1's update; two rows at a
            // time per thread, matrix entries loaded 16 columns at a time
            for (idx_t i = b + nb + (tid - 64); i < n; i += 2 * TS_UPD) {
                const idx_t i2 = i + TS_UPD;
                const bool two = i2 < n;
                const idx_t i2c = two ? i2 : i;
                dd x1 = x[i], x2 = x[i2c];
#pragma unroll
                for (int h = 0; h < SB; h += 16) {
                    dd a1[16], a2[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) {
                        a1[q] = L[i + (bp + h + q) * lda];
                        a2[q] = L[i2c + (bp + h + q) * lda];
                    }
#pragma unroll
                    for (int q = 0; q < 16; q++)
                        if (nzv[prv][h + q]) {
                            const dd vq = v[prv][h + q];
                            x1 = dd_sub(x1, dd_mul(vq, a1[q]));
                            x2 = dd_sub(x2, dd_mul(vq, a2[q]));
                        }
                }
                x[i] = x1;
                if (two) x[i2] = x2;
            }
        }
        __syncthreads();
    }
}

// x := L^{-T} x ('L','T','N'): for j descending, x_j = (x_j - sum over
// i > j, descending, of L_ij x_i) / L_jj.  Blocks from the bottom.
__global__ __launch_bounds__(TS_NT) void k_trsv_lt(idx_t n, const dd *L, idx_t lda, dd *x, const int *redo,
                                                   const dd *x0) {
    __shared__ dd D[SB][SB + 1];
    __shared__ dd v[2][SB];
    if (!trsv_redo(redo, x0, x, n)) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int nblk = (int)((n + SB - 1) / SB);
    // block R covers rows [lo(R), hi(R)); R = 0 is the bottom block
    auto blo = [&](int R) -> idx_t { idx_t t = n - (idx_t)(R + 1) * SB; return t > 0 ? t : 0; };
    auto bhi = [&](int R) -> idx_t { return n - (idx_t)R * SB; };
    for (int R = 0; R < nblk; R++) {
        const idx_t b = blo(R);
        const int nb = (int)(bhi(R) - b);
        const idx_t bp = R > 0 ? blo(R - 1) : 0;  // previous (lower, full) block: rows [bp, bp + SB)
        const int cur = R & 1, prv = cur ^ 1;
        if (tid < 64) {
            if (nb > 0) {
                for (int e = lane; e < SB * SB; e += 64) {
                    int r = e % SB, c = e / SB;
                    if (r < nb && c < nb && r >= c) D[r][c] = L[(b + r) + (b + c) * lda];
                }
                dd xr = lane < nb ? x[b + lane] : mk(0.0);
                if (R > 0 && lane < nb) {
#pragma unroll
                    for (int h = SB - 16; h >= 0; h -= 16) {
                        dd a[16];
#pragma unroll
                        for (int q = 0; q < 16; q++) a[q] = L[(bp + h + q) + (b + lane) * lda];
#pragma unroll
                        for (int q = 15; q >= 0; q--) xr = dd_sub(xr, dd_mul(a[q], v[prv][h + q]));
                    }
                }
                __builtin_amdgcn_wave_barrier();
                // L_(jj, lane) from LDS (read ahead of the chain: the step loop
                // is unrolled, so the addresses are known)
                const int lc = lane < nb ? lane : 0;
#pragma unroll
                for (int jj = SB - 1; jj >= 0; jj--) {
                    if (jj < nb) {
                        const dd djj = D[jj >= lc ? jj : lc][lc];
                        if (lane == jj) xr = dd_div(xr, djj);
                        const dd t = readlane_dd(xr, jj);
                        if (lane < jj) xr = dd_sub(xr, dd_mul(djj, t));
                    }
                }
                if (lane < nb) {
                    x[b + lane] = xr;
                    v[cur][lane] = xr;
                }
            }
        } else if (R > 0) {
            // rows above this block get block R-1's update; two rows at a
            // time per thread, 8 entries of each column segment at a time
            for (idx_t k = tid - 64; k < b; k += 2 * TS_UPD) {
                const idx_t k2 = k + TS_UPD;
                const bool two = k2 < b;
                const idx_t k2c = two ? k2 : k;
                dd x1 = x[k], x2 = x[k2c];
#pragma unroll
                for (int h = SB - 8; h >= 0; h -= 8) {
                    dd a1[8], a2[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        a1[q] = L[(bp + h + q) + k * lda];
                        a2[q] = L[(bp + h + q) + k2c * lda];
                    }
#pragma unroll
                    for (int q = 7; q >= 0; q--) {
                        const dd vq = v[prv][h + q];
                        x1 = dd_sub(x1, dd_mul(a1[q], vq));
                        x2 = dd_sub(x2, dd_mul(a2[q], vq));
                    }
                }
                x[k] = x1;
                if (two) x[k2] = x2;
            }
        }
        __syncthreads();
    }
}

// ---- triangular solves across workgroups ----
// The single-workgroup solves above are bound by the update waves'
// memory latency for large n.  These cooperative forms spread the updates
// over TC_UPD workgroups that each own 256 rows (one row per thread) and run
// concurrently with the diagonal chain:
//   workgroup 0, round R: wait until the owners of block R's rows have
//     applied blocks <= R-2; apply block R-1 to block R (row per lane); solve
//     block R's diagonal part; publish its values and the counter R+1;
//   updater u, round J: wait for block J; apply it to its rows in blocks
//     >= J+2 (in the reference's column order per row); publish progress J+1.
// Every element sees the same operation sequence as the column loop of
// Rtrsv.  The waits assume every workgroup of the grid is resident at once.
// A grid no larger than the CU count is launched plainly (cooperative
// launches of these grids made a whole theta solve about 4% slower); a
// larger one cooperatively.  Should a workgroup still not
// be resident (CUs or LDS held by other streams or processes), a bounded
// spin runs out, sets the error word and every workgroup leaves, so the
// grid always drains; a single-workgroup solve enqueued behind it then
// restores the saved right-hand side and solves again, in the same
// operation order, so the result never depends on co-residency.
static constexpr int TC_ROWS = 256;           // rows per updater workgroup
static constexpr int BO_ROWS = 128;           // the same for the block-inverse solves (four lanes per row)
static constexpr unsigned TC_SPIN = 1u << 22;  // default spin bound per wait (about a second)
unsigned g_trsv_spin = 0;
static unsigned trsv_spin() { return g_trsv_spin ? g_trsv_spin : TC_SPIN; }
__device__ __forceinline__ int tc_load(const int *p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void tc_store(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }
// Spin until *p >= target; false when the error word is set or the bound
// runs out (then it sets the error word).  The spin
// reads with relaxed agent-scope loads and takes the acquire once, after
// the flag is seen: an acquire per read would invalidate the L2 of the
// waiting workgroup's XCD on every spin.
__device__ bool tc_wait(const int *p, int target, int *err, unsigned spin) {
    for (unsigned it = 0; it < spin; it++) {
        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return true;
        }
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    tc_store(err, 1);
    return false;
}

// Cooperative L^{-1} x in the Rtrsv('L','N','N') order.  Workgroup 0 solves
// the 32-row blocks in sequence: wave 0 runs the division chain, waves 1-3
// stage the next block's diagonal and off-diagonal tiles in LDS meanwhile.
// Block R takes block R-1's update from wave 0; updater workgroup u owns
// rows [256u, 256u + 256) and applies each block J, once published, to its
// rows in blocks >= J+2.  flags: [0] blocks solved, [1 + u] blocks applied
// by updater u, [1 + nupd] error word, all zero at launch.  vg/nzg: solved
// values of every block and their nonzero flags.
__global__ __launch_bounds__(256) void k_trsv_ln_coop(idx_t n, const dd *L, idx_t lda, dd *x, int *flags, int nupd, dd *vg,
                                                      int *nzg, unsigned spin) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (int)((n + SB - 1) / SB);
    int *err = flags + 1 + nupd;
    if (blockIdx.x == 0) {
        __shared__ dd Dg[2][SB][SB + 1];  // Dg[buf][r][c] = L[b + r, b + c], c <= r
        __shared__ dd Od[2][SB][SB + 1];  // Od[buf][r][c] = L[b + r, b - SB + c]
        __shared__ int stop;
        auto stage = [&](int R, int buf, int t0, int nt) {
            const idx_t b = (idx_t)R * SB;
            const int nb = (int)((n - b) < SB ? (n - b) : SB);
            for (int e = t0; e < SB * SB; e += nt) {
                const int r = e & (SB - 1), c = e >> 5;
                if (r < nb && c <= r) Dg[buf][r][c] = L[(b + r) + (b + c) * lda];
                if (R > 0 && r < nb) Od[buf][r][c] = L[(b + r) + (b - SB + c) * lda];
            }
        };
        if (tid == 0) stop = 0;
        stage(0, 0, tid, 256);
        __syncthreads();
        dd prev = mk(0.0);  // wave 0: lane j holds the solved x of block R-1, row j
        int pnz = 0;
        for (int R = 0; R < nblk; R++) {
            const int buf = R & 1;
            if (wave != 0) {
                if (R + 1 < nblk) stage(R + 1, buf ^ 1, tid - 64, 192);
            } else {
                const idx_t b = (idx_t)R * SB;
                const int nb = (int)((n - b) < SB ? (n - b) : SB);
                // owners of this block's rows must have applied blocks <= R-2
                int ok = 1;
                if (lane == 0 && R >= 2) {
                    const int u0 = (int)(b / TC_ROWS), u1 = (int)((b + nb - 1) / TC_ROWS);
                    for (int u = u0; u <= u1 && ok; u++) ok = tc_wait(flags + 1 + u, R - 1, err, spin);
                }
                ok = __builtin_amdgcn_readlane(ok, 0);
                if (!ok) {
                    if (lane == 0) stop = 1;
                } else {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    const int lr = lane < nb ? lane : 0;
                    dd xr = x[b + lr];
                    int nzr = 0;
                    if (R > 0) {
#pragma unroll
                        for (int q = 0; q < SB; q++) {
                            const dd a = Od[buf][lr][q];
                            const int nz = __builtin_amdgcn_readlane(pnz, q);
                            const dd t = readlane_dd(prev, q);
                            if (nz) xr = dd_sub(xr, dd_mul(t, a));
                        }
                    }
#pragma unroll
                    for (int jj = 0; jj < SB; jj++) {
                        if (jj < nb) {
                            const dd djj = Dg[buf][lr][jj <= lr ? jj : lr];
                            const int nz = __builtin_amdgcn_readlane(!dd_is_zero(xr), jj);
                            if (lane == jj) {
                                nzr = nz;
                                if (nz) xr = dd_div(xr, djj);
                            }
                            const dd t = readlane_dd(xr, jj);
                            if (nz && lane > jj && lane < nb) xr = dd_sub(xr, dd_mul(t, djj));
                        }
                    }
                    // rows skip the entries that were zero before their division
                    pnz = lane < nb ? nzr : 0;
                    if (lane < nb) {
                        x[b + lane] = xr;
                        vg[b + lane] = xr;
                        nzg[b + lane] = pnz;
                    }
                    prev = xr;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    if (lane == 0) tc_store(flags, R + 1);
                }
            }
            __syncthreads();
            if (stop) return;
        }
        return;
    }
    const int u = blockIdx.x - 1;
    const idx_t lo = (idx_t)u * TC_ROWS, hi = lo + TC_ROWS < n ? lo + TC_ROWS : n;
    const idx_t i = lo + tid;
    __shared__ int go;
    dd xi = i < hi ? x[i] : mk(0.0);
    for (int J = 0; J < nblk; J++) {
        const idx_t first = (idx_t)(J + 2) * SB;  // rows that take block J here
        if (first >= hi) break;
        const bool act = i < hi && i >= first;
        const idx_t bp = (idx_t)J * SB;
        dd a[SB];
        if (act) {
#pragma unroll
            for (int q = 0; q < SB; q++) a[q] = L[i + (bp + q) * lda];
        }
        if (tid == 0) go = tc_wait(flags, J + 1, err, spin);
        __syncthreads();
        if (!go) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (act) {
#pragma unroll
            for (int q = 0; q < SB; q++)
                if (nzg[bp + q]) xi = dd_sub(xi, dd_mul(vg[bp + q], a[q]));
            x[i] = xi;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) tc_store(flags + 1 + u, J + 1);
    }
    if (tid == 0) tc_store(flags + 1 + u, nblk);
}

// Cooperative L^{-T} x in the Rtrsv('L','T','N') order: blocks from the
// bottom (block R = rows [blo(R), n - 32R)), otherwise as k_trsv_ln_coop.
__global__ __launch_bounds__(256) void k_trsv_lt_coop(idx_t n, const dd *L, idx_t lda, dd *x, int *flags, int nupd,
                                                      dd *vg, unsigned spin) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (int)((n + SB - 1) / SB);
    int *err = flags + 1 + nupd;
    auto blo = [&](int R) -> idx_t { idx_t t = n - (idx_t)(R + 1) * SB; return t > 0 ? t : 0; };
    if (blockIdx.x == 0) {
        __shared__ dd Dg[2][SB][SB + 1];  // Dg[buf][r][c] = L[b + r, b + c], c <= r
        __shared__ dd Od[2][SB][SB + 1];  // Od[buf][q][c] = L[blo(R-1) + q, b + c]
        __shared__ int stop;
        auto stage = [&](int R, int buf, int t0, int nt) {
            const idx_t b = blo(R);
            const int nb = (int)(n - (idx_t)R * SB - b);
            const idx_t bp = R > 0 ? blo(R - 1) : 0;
            for (int e = t0; e < SB * SB; e += nt) {
                const int r = e & (SB - 1), c = e >> 5;
                if (r < nb && c <= r) Dg[buf][r][c] = L[(b + r) + (b + c) * lda];
                if (R > 0 && c < nb) Od[buf][r][c] = L[(bp + r) + (b + c) * lda];
            }
        };
        if (tid == 0) stop = 0;
        stage(0, 0, tid, 256);
        __syncthreads();
        dd prev = mk(0.0);  // wave 0: lane q holds the solved x of block R-1, row q
        for (int R = 0; R < nblk; R++) {
            const int buf = R & 1;
            if (wave != 0) {
                if (R + 1 < nblk) stage(R + 1, buf ^ 1, tid - 64, 192);
            } else {
                const idx_t b = blo(R);
                const int nb = (int)(n - (idx_t)R * SB - b);
                int ok = 1;
                if (lane == 0 && R >= 2) {
                    const int u0 = (int)(b / TC_ROWS), u1 = (int)((b + nb - 1) / TC_ROWS);
                    for (int u = u0; u <= u1 && ok; u++) ok = tc_wait(flags + 1 + u, R - 1, err, spin);
                }
                ok = __builtin_amdgcn_readlane(ok, 0);
                if (!ok) {
                    if (lane == 0) stop = 1;
                } else {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    const int lc = lane < nb ? lane : 0;
                    dd xr = x[b + lc];
                    if (R > 0) {
#pragma unroll
                        for (int q = SB - 1; q >= 0; q--) xr = dd_sub(xr, dd_mul(Od[buf][q][lc], readlane_dd(prev, q)));
                    }
#pragma unroll
                    for (int jj = SB - 1; jj >= 0; jj--) {
                        if (jj < nb) {
                            const dd djj = Dg[buf][jj >= lc ? jj : lc][lc];
                            if (lane == jj) xr = dd_div(xr, djj);
                            const dd t = readlane_dd(xr, jj);
                            if (lane < jj) xr = dd_sub(xr, dd_mul(djj, t));
                        }
                    }
                    if (lane < nb) {
                        x[b + lane] = xr;
                        vg[(idx_t)R * SB + lane] = xr;
                    }
                    prev = xr;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    if (lane == 0) tc_store(flags, R + 1);
                }
            }
            __syncthreads();
            if (stop) return;
        }
        return;
    }
    const int u = blockIdx.x - 1;
    const idx_t lo = (idx_t)u * TC_ROWS, hi = lo + TC_ROWS < n ? lo + TC_ROWS : n;
    const idx_t k = lo + tid;
    __shared__ int go;
    dd xk = k < hi ? x[k] : mk(0.0);
    for (int J = 0; J < nblk; J++) {
        const idx_t lim = blo(J + 1);  // rows below lim take block J here (blocks >= J+2)
        if (J + 1 >= nblk || lo >= lim) break;
        const bool act = k < hi && k < lim;
        const idx_t bp = blo(J);
        dd a[SB];
        if (act) {
#pragma unroll
            for (int q = 0; q < SB; q++) a[q] = L[(bp + q) + k * lda];
        }
        if (tid == 0) go = tc_wait(flags, J + 1, err, spin);
        __syncthreads();
        if (!go) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (act) {
#pragma unroll
            for (int q = SB - 1; q >= 0; q--) xk = dd_sub(xk, dd_mul(a[q], vg[(idx_t)J * SB + q]));
            x[k] = xk;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) tc_store(flags + 1 + u, J + 1);
    }
    if (tid == 0) tc_store(flags + 1 + u, nblk);
}

// Workspace of the cross-workgroup solves (both forms): the solved block
// values, the saved right-hand side, the block-inverse solve's partials,
// the nonzero flags, then the wait words.
struct CoopWork {
    dd *vg, *x0, *pg;
    int *nzg, *flags;
};
static CoopWork coop_work(void *work, idx_t n) {
    const idx_t nblk = (n + SB - 1) / SB;
    char *w = (char *)work;
    CoopWork cw;
    cw.vg = (dd *)(w + 256 - (((uintptr_t)w) & 255));
    cw.x0 = cw.vg + nblk * SB;
    cw.pg = cw.x0 + nblk * SB;
    cw.nzg = (int *)(cw.pg + nblk * SB);
    cw.flags = cw.nzg + nblk * SB;
    return cw;
}
size_t trsv_coop_bytes(idx_t n) {
    static_assert(BO_ROWS <= TC_ROWS, "the wait words are counted for the smaller updater");
    const idx_t nblk = (n + SB - 1) / SB, nupd = (n + BO_ROWS - 1) / BO_ROWS;
    return 256 + sizeof(dd) * 3 * nblk * SB + sizeof(int) * nblk * SB + sizeof(int) * (2 + nupd);
}
__global__ void k_zero_ints(int *p, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0;
}
void zero_ints(hipStream_t s, int *p, int cnt) {
    if (cnt > 0) k_zero_ints<<<cdiv(cnt, 256), 256, 0, s>>>(p, cnt);
}
// A cross-workgroup solve over `grid` workgroups of `nt` threads: a plain
// launch when the grid fits the CU count, a cooperative one otherwise.
static hipError_t launch_coop(const void *fn, dim3 grid, void **args, hipStream_t s, int nt = 256) {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if ((int)grid.x <= cus) return hipLaunchKernel(fn, grid, dim3(nt), args, 0, s);
    return hipLaunchCooperativeKernel(fn, grid, dim3(nt), args, 0, s);
}
// The recovery enqueued behind every cross-workgroup solve: the
// single-workgroup solve in Rtrsv's order, a no-op unless the error word
// `err` of that solve is set.
static void trsv_recover(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x, const int *err,
                         const dd *x0) {
    if (trans == 'N' || trans == 'n')
        k_trsv_ln<<<1, TS_NT, 0, s>>>(n, L, lda, x, err, x0);
    else
        k_trsv_lt<<<1, TS_NT, 0, s>>>(n, L, lda, x, err, x0);
}
// Stream-ordered; returns 0, or 1 when the launch fails (nothing ran; the
// caller then uses the single-workgroup kernel).
int trsv_lower_coop(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x, void *work) {
    if (n <= 0) return 0;
    const int nupd = (int)((n + TC_ROWS - 1) / TC_ROWS);
    const CoopWork cw = coop_work(work, n);
    dd *vg = cw.vg;
    int *nzg = cw.nzg, *flags = cw.flags;
    zero_ints(s, flags, 2 + nupd);
    StatScope st("trsv", s, (double)n * n);
    copy(s, n, x, cw.x0);
    const bool ln = (trans == 'N' || trans == 'n');
    unsigned spin = trsv_spin();
    hipError_t e;
    if (ln) {
        void *args[] = {&n, &L, &lda, &x, &flags, (void *)&nupd, &vg, &nzg, &spin};
        e = launch_coop((const void *)k_trsv_ln_coop, dim3(1 + nupd), args, s);
    } else {
        void *args[] = {&n, &L, &lda, &x, &flags, (void *)&nupd, &vg, &spin};
        e = launch_coop((const void *)k_trsv_lt_coop, dim3(1 + nupd), args, s);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    trsv_recover(s, trans, n, L, lda, x, flags + 1 + nupd, cw.x0);
    return 0;
}


// ------------------------------------------ block-inverse triangular solves --
// The Ozaki form's Schur solves (not Rtrsv's operation order; the Ozaki
// form refines the result against B).  Blocks of BB rows are solved in a
// processing sequence (L x = b: from the top; L^T x = b: from the bottom),
// block s as x_s = Dinv_s p_s - sum_{d=2..BI_D} N_{s,d} x_{s-d}
// - N_{s,1} x_{s-1}, where Dinv_s inverts the diagonal block, N_{s,d} =
// Dinv_s T_{s,s-d}, T_{s,t} is the coupling block (L_{s,t}, or L_{t,s}^T)
// and p_s is b_s after the updater workgroups applied blocks
// <= s - BI_D - 1.  The operators are formed once per factor.
// Workgroup 0 runs one round per block.  In round s, wave 0 forms x_s (the
// one dependent step: a BB-term product with x_{s-1}); waves 1..BI_D
// prepare block s+1's terms (wave 1: Dinv p, wave d: N_{s+1,d} x_{s+1-d});
// wave BI_D + 1 sends x_{s-1} to x and to the updaters (its write-through
// stores would hold up a product wave's operator loads, which wait on the
// same in-order counter).  Wave 1 issues the read of block s+2's partial
// after its product and checks it a round later.  A product wave takes each row in two
// lanes (even and odd terms, then one exchange), which forms the same
// pairwise tree as one lane summing all BB terms.
static constexpr int BB = 32;  // rows per block of the block-inverse solves
static constexpr int BI_D = 6;
static constexpr int BI_OPS = (BI_D + 1) * BB * BB;  // dd per block and direction
static constexpr int BI_NT = 64 * (BI_D + 2);        // threads per workgroup
static constexpr int BI_RING = 8;                    // solved blocks kept in LDS
static_assert(BI_RING >= BI_D + 1 && (BI_RING & (BI_RING - 1)) == 0, "ring of solved blocks");
static_assert(BO_ROWS * 4 == BI_NT, "an updater row takes four lanes");

// Pairwise sum of N dd terms (depth log2(N) instead of an N-step chain).
template <int N>
__device__ __forceinline__ dd tree_sum(dd *t) {
#pragma unroll
    for (int w = N / 2; w >= 1; w >>= 1) {
#pragma unroll
        for (int q = 0; q < w; q++) t[q] = dd_add(t[q], t[q + w]);
    }
    return t[0];
}

__host__ __device__ __forceinline__ void seq_block(idx_t n, int trans, int s, idx_t &lo, int &nb) {
    if (!trans) {
        lo = (idx_t)s * BB;
        nb = (int)((n - lo) < BB ? (n - lo) : BB);
    } else {
        const idx_t hi = n - (idx_t)s * BB;
        lo = hi - BB > 0 ? hi - BB : 0;
        nb = (int)(hi - lo);
    }
}

// inv[s] = (diagonal block s)^{-1}, row-major BB x BB, zero outside; one
// wave per block, lane c forms column c by forward substitution.
__global__ __launch_bounds__(64) void k_blk_inv(idx_t n, int trans, const dd *L, idx_t lda, dd *inv) {
    __shared__ dd T[BB][BB + 1];
    const int s = blockIdx.x, lane = threadIdx.x;
    idx_t lo;
    int nb;
    seq_block(n, trans, s, lo, nb);
    for (int e = lane; e < BB * BB; e += 64) {
        const int i = e % BB, k = e / BB;
        T[i][k] = (i < nb && k <= i) ? L[(lo + i) + (lo + k) * lda] : mk(0.0);
    }
    __syncthreads();
    const int c = lane;
    dd y[BB];
#pragma unroll
    for (int i = 0; i < BB; i++) {
        dd acc = mk(i == c ? 1.0 : 0.0);
#pragma unroll
        for (int k = 0; k < i; k++)
            if (k >= c) acc = dd_sub(acc, dd_mul(T[i][k], y[k]));
        y[i] = (c < nb && i < nb && i >= c) ? dd_div(acc, T[i][i]) : mk(0.0);
    }
    if (c < BB)
#pragma unroll
        for (int i = 0; i < BB; i++) inv[((idx_t)s * BB + i) * BB + c] = y[i];
}

// ops[s][0] = Dinv_s (the inverse, transposed for L^T x = b);
// ops[s][d] = Dinv_s T_{s,s-d} for d = 1..BI_D (zero when s - d < 0).
// Each BB x BB operator is stored in the order the solve's lanes read it:
// element (i, q) at op_pos(i, q), so that the loads of term pair k (lane
// h BB + i takes (i, 2k + h)) cover 1 KB of consecutive memory.
__host__ __device__ __forceinline__ int op_pos(int i, int q) { return (q >> 1) * 2 * BB + (q & 1) * BB + i; }
__global__ __launch_bounds__(256) void k_blk_ops(idx_t n, int trans, const dd *L, idx_t lda, const dd *inv, dd *ops) {
    __shared__ dd Iv[BB][BB + 1];
    __shared__ dd T[BB][BB + 1];
    const int s = blockIdx.x, d = blockIdx.y, tid = threadIdx.x;
    idx_t lo;
    int nb;
    seq_block(n, trans, s, lo, nb);
    dd *out = ops + (idx_t)s * BI_OPS + (idx_t)d * BB * BB;
    for (int e = tid; e < BB * BB; e += 256) {
        const int i = e % BB, k = e / BB;
        Iv[i][k] = trans ? inv[((idx_t)s * BB + k) * BB + i] : inv[((idx_t)s * BB + i) * BB + k];
    }
    if (d == 0) {
        __syncthreads();
        for (int e = tid; e < BB * BB; e += 256) out[op_pos(e / BB, e % BB)] = Iv[e / BB][e % BB];
        return;
    }
    if (s - d < 0) {
        for (int e = tid; e < BB * BB; e += 256) out[e] = mk(0.0);
        return;
    }
    idx_t lo2;
    int nb2;
    seq_block(n, trans, s - d, lo2, nb2);
    for (int e = tid; e < BB * BB; e += 256) {
        const int k = e % BB, q = e / BB;
        dd v = mk(0.0);
        if (k < nb && q < nb2) v = trans ? L[(lo2 + q) + (lo + k) * lda] : L[(lo + k) + (lo2 + q) * lda];
        T[k][q] = v;
    }
    __syncthreads();
    for (int e = tid; e < BB * BB; e += 256) {
        const int i = e / BB, q = e % BB;
        dd acc = mk(0.0);
        for (int k = 0; k < nb; k++) acc = dd_add(acc, dd_mul(Iv[i][k], T[k][q]));
        out[op_pos(i, q)] = acc;
    }
}

// Row sums of the product waves: lane l < 32 holds the even terms of row
// l, lane l + 32 its odd terms; tree_sum<BB> over all BB terms ends with
// (even sum) + (odd sum), so each half sums its own and the halves swap.
__device__ __forceinline__ dd shfl_xor_dd(dd v, int m) {
    dd r;
    r.hi = __shfl_xor(v.hi, m, 64);
    r.lo = __shfl_xor(v.lo, m, 64);
    return r;
}
__device__ __forceinline__ dd pair_tree(dd *t, int h) {
    const dd mine = tree_sum<BB / 2>(t);
    const dd other = shfl_xor_dd(mine, 32);
    return h == 0 ? dd_add(mine, other) : dd_add(other, mine);
}
__global__ void k_binv_reset(idx_t len, dd *vg, dd *pg, int *flags, int nflags) {
    const idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    const dd empty = {__longlong_as_double((long long)BI_EMPTY), __longlong_as_double((long long)BI_EMPTY)};
    if (e < len) {
        vg[e] = empty;
        pg[e] = empty;
    }
    if (e < nflags) flags[e] = 0;
}

// vg: the solved blocks (x_J at J BB), written by workgroup 0;
// pg: each block's partial p_t (at t BB), written by the owner of its rows
// once blocks <= t - BI_D - 1 are applied.  Owner workgroup u (blockIdx
// u + 1) holds rows [BO_ROWS u, BO_ROWS u + BO_ROWS), four lanes per row,
// and applies block J to its rows in blocks >= J + BI_D + 1.  flags[1 +
// nupd] is the error word: a wait that runs out of its bound sets it and
// every wave leaves (the recovery behind the launch then solves again).
__global__ __launch_bounds__(BI_NT) void k_trsv_binv(idx_t n, int trans, const dd *L, idx_t lda, dd *x, const dd *ops,
                                                     int *flags, int nupd, dd *vg, dd *pg, const dd *LT, unsigned spin) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (int)((n + BB - 1) / BB);
    int *err = flags + 1 + nupd;
    if (blockIdx.x == 0) {
        __shared__ dd hp[2][BI_D][BB];  // prepared terms of a block, by block parity
        __shared__ dd xs[BI_RING][BB];  // solved blocks, ring by s mod BI_RING
        __shared__ dd pw[BB];           // wave 1's partial of the block it prepares
        __shared__ int stop;
        constexpr int PUB = BI_D + 1;   // the publishing wave
        const int r = lane & (BB - 1), h = lane >> 5;
        const int wop = wave == 1 ? 0 : wave;
        // row r (terms 2k + h) of the operator the wave applies next,
        // loaded a round ahead
        dd opr[BB / 2];
        auto load_row = [&](int t, int d) {
            if (t >= nblk) return;
            const dd *o = ops + (idx_t)t * BI_OPS + (idx_t)d * BB * BB + lane;
#pragma unroll
            for (int k = 0; k < BB / 2; k++) opr[k] = o[k * 2 * BB];
        };
        // wave 1: block t's partial (row lane), read during round t - 2;
        // blocks <= BI_D take no update and come from x
        RawDD pnext = {0, 0};
        auto read_partial = [&](int t) {
            idx_t lo;
            int nb;
            seq_block(n, trans, t, lo, nb);
            if (lane < nb) pnext = t >= BI_D + 1 ? get_raw(pg + (idx_t)t * BB + lane) : dd_raw(x[lo + lane]);
        };
        // waves 1..BI_D: terms of block t (< nblk) into hp[t & 1]
        auto prepare = [&](int t) {
            dd a = mk(0.0);
            if (wave == 1) {
                idx_t lo;
                int nb;
                seq_block(n, trans, t, lo, nb);
                bool have = lane >= nb || t < BI_D + 1 || raw_ok(pnext);
                for (unsigned it = 0; !__all(have); it++) {
                    if (it >= spin || relaxed_flag(err) != 0) {
                        if (lane == 0) {
                            tc_store(err, 1);
                            stop = 1;
                        }
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    if (!have) {
                        pnext = get_raw(pg + (idx_t)t * BB + lane);
                        have = raw_ok(pnext);
                    }
                }
                // broadcast through LDS (in order within the wave)
                if (lane < BB) pw[lane] = lane < nb ? raw_dd(pnext) : mk(0.0);
#pragma unroll
                for (int k = 0; k < BB / 2; k++) opr[k] = dd_mul(opr[k], pw[2 * k + h]);
                a = pair_tree(opr, h);
            } else if (t - wave >= 0) {
                const dd *xv = xs[(t - wave) & (BI_RING - 1)];
#pragma unroll
                for (int k = 0; k < BB / 2; k++) opr[k] = dd_mul(opr[k], xv[2 * k + h]);
                a = pair_tree(opr, h);
            }
            if (h == 0) hp[t & 1][wave - 1][r] = a;
            load_row(t + 1, wop);
            // last: the compiler waits for every load in flight before
            // issuing the operator loads above
            if (wave == 1 && t + 1 < nblk) read_partial(t + 1);
        };
        // wave PUB: block b to x, and to the updaters while they need it
        auto publish = [&](int b) {
            idx_t lo;
            int nb;
            seq_block(n, trans, b, lo, nb);
            if (lane < nb) {
                const dd v = xs[b & (BI_RING - 1)][lane];
                if (b + BI_D + 1 < nblk) put_dd(vg + (idx_t)b * BB + lane, v);
                x[lo + lane] = v;
            }
        };
        if (tid == 0) stop = 0;
        if (wave == 0) {
            load_row(1, 1);
        } else if (wave < PUB) {
            load_row(0, wop);
            if (wave == 1) read_partial(0);
        }
        __syncthreads();
        if (wave >= 1 && wave < PUB) prepare(0);
        __syncthreads();
        if (stop) return;
        for (int s = 0; s < nblk; s++) {
            if (wave == 0) {
                idx_t lo;
                int nb;
                seq_block(n, trans, s, lo, nb);
                dd xr = hp[s & 1][0][r];
#pragma unroll
                for (int d = 1; d < BI_D; d++) xr = dd_sub(xr, hp[s & 1][d][r]);
                if (s >= 1) {
                    // products in place: opr is reloaded for the next block below
                    const dd *xv = xs[(s - 1) & (BI_RING - 1)];
#pragma unroll
                    for (int k = 0; k < BB / 2; k++) opr[k] = dd_mul(opr[k], xv[2 * k + h]);
                    xr = dd_sub(xr, pair_tree(opr, h));
                }
                if (r >= nb) xr = mk(0.0);
                if (h == 0) xs[s & (BI_RING - 1)][r] = xr;
                load_row(s + 1, 1);
            } else if (wave < PUB) {
                if (s + 1 < nblk) prepare(s + 1);
            } else if (s >= 1) {
                publish(s - 1);
            }
            __syncthreads();
            if (stop) return;
        }
        if (wave == PUB) publish(nblk - 1);
        return;
    }
    const int u = blockIdx.x - 1;
    const int hq = tid & 3;  // terms 4k + hq of the row
    const idx_t rlo = (idx_t)u * BO_ROWS, rhi = rlo + BO_ROWS < n ? rlo + BO_ROWS : n;
    const idx_t i = rlo + (tid >> 2);
    constexpr int Q = BB / 4;
    dd xi = i < rhi ? x[i] : mk(0.0);
    dd a[Q];
    RawDD vr[Q];  // block J's values, read an iteration ahead
    auto load_l = [&](int J) {
        idx_t lo_n, bp;
        int nb_n, nbj;
        seq_block(n, trans, J + BI_D, lo_n, nb_n);
        seq_block(n, trans, J, bp, nbj);
        const bool act = i < rhi && (trans ? i < lo_n : i >= lo_n + nb_n);
#pragma unroll
        for (int k = 0; k < Q; k++) {
            const idx_t c = bp + 4 * k + hq;
            a[k] = !act ? mk(0.0) : !trans ? L[i + c * lda] : LT ? LT[i + c * n] : L[c + i * lda];
        }
    };
    auto read_v = [&](int J) {
#pragma unroll
        for (int k = 0; k < Q; k++) vr[k] = get_raw(vg + (idx_t)J * BB + 4 * k + hq);
    };
    auto all_ok = [&]() {
        bool ok = true;
#pragma unroll
        for (int k = 0; k < Q; k++) ok = ok && raw_ok(vr[k]);
        return ok;
    };
    if (BI_D + 1 < nblk) {
        load_l(0);
        read_v(0);
    }
    for (int J = 0; J + BI_D + 1 < nblk; J++) {
        // rows of sequence blocks >= J + BI_D + 1 take block J here
        idx_t lo_n, lo_f;
        int nb_n, nb_f;
        seq_block(n, trans, J + BI_D, lo_n, nb_n);
        if (trans ? rlo >= lo_n : rhi <= lo_n + nb_n) break;
        const bool act = i < rhi && (trans ? i < lo_n : i >= lo_n + nb_n);
        for (unsigned it = 0; !__all(all_ok()); it++) {
            if (it >= spin) {
                if (lane == 0) tc_store(err, 1);
                return;
            }
            if (relaxed_flag(err) != 0) return;
            __builtin_amdgcn_s_sleep(1);
            read_v(J);
        }
        dd pr[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) pr[k] = dd_mul(a[k], raw_dd(vr[k]));
        dd t = tree_sum<Q>(pr);
        dd o = shfl_xor_dd(t, 2);
        t = (hq & 2) == 0 ? dd_add(t, o) : dd_add(o, t);
        o = shfl_xor_dd(t, 1);
        t = (hq & 1) == 0 ? dd_add(t, o) : dd_add(o, t);
        if (act) {
            xi = dd_sub(xi, t);
            // rows of block J + BI_D + 1 are final now
            seq_block(n, trans, J + BI_D + 1, lo_f, nb_f);
            if (hq == 0 && i >= lo_f && i < lo_f + nb_f) put_dd(pg + (idx_t)(J + BI_D + 1) * BB + (i - lo_f), xi);
        }
        // block J + 1: its operands and values, checked next iteration;
        // issued after the partial above, which they would otherwise hold
        // back in the memory pipeline
        if (J + BI_D + 2 < nblk) {
            load_l(J + 1);
            read_v(J + 1);
        }
    }
}

// lt[c + r n] = L[r + c lda] on the 32 x 32 tiles that meet the lower
// triangle (tile row >= tile column).
__global__ __launch_bounds__(256) void k_transpose_lower(idx_t n, const dd *L, idx_t lda, dd *lt) {
    __shared__ dd T[32][33];
    const int bi = blockIdx.x, bj = blockIdx.y, tid = threadIdx.x;
    if (bi < bj) return;
    const idx_t r0 = (idx_t)bi * 32, c0 = (idx_t)bj * 32;
    for (int e = tid; e < 32 * 32; e += 256) {
        const int r = e % 32, c = e / 32;
        if (r0 + r < n && c0 + c < n) T[r][c] = L[(r0 + r) + (c0 + c) * lda];
    }
    __syncthreads();
    for (int e = tid; e < 32 * 32; e += 256) {
        const int c = e % 32, r = e / 32;
        if (r0 + r < n && c0 + c < n) lt[(c0 + c) + (r0 + r) * n] = T[r][c];
    }
}

size_t binv_ops_elems(idx_t n) { return (size_t)((n + BB - 1) / BB) * BI_OPS; }
void trsv_binv_prepare(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *ops_n, dd *ops_t, dd *inv, dd *lt) {
    if (n <= 0) return;
    const int nblk = (int)((n + BB - 1) / BB);
    if (lt) {
        const unsigned nt = (unsigned)((n + 31) / 32);
        k_transpose_lower<<<dim3(nt, nt), 256, 0, s>>>(n, L, lda, lt);
    }
    StatScope st("trsv_prep", s, 2.0 * nblk * (BB * BB * BB / 3.0 + BI_D * BB * BB * BB));
    k_blk_inv<<<nblk, 64, 0, s>>>(n, 0, L, lda, inv);
    k_blk_ops<<<dim3(nblk, BI_D + 1), 256, 0, s>>>(n, 0, L, lda, inv, ops_n);
    k_blk_inv<<<nblk, 64, 0, s>>>(n, 1, L, lda, inv);
    k_blk_ops<<<dim3(nblk, BI_D + 1), 256, 0, s>>>(n, 1, L, lda, inv, ops_t);
}
int trsv_binv(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x, const dd *ops, void *work,
              const dd *lt) {
    static_assert(BB == SB, "the block-inverse solves share the cross-workgroup workspace layout");
    if (n <= 0) return 0;
    const int nupd = (int)((n + BO_ROWS - 1) / BO_ROWS);
    const CoopWork cw = coop_work(work, n);
    dd *vg = cw.vg, *pg = cw.pg;
    int *flags = cw.flags;
    const idx_t len = ((n + BB - 1) / BB) * BB;
    k_binv_reset<<<cdiv(std::max<idx_t>(len, 2 + nupd), 256), 256, 0, s>>>(len, vg, pg, flags, 2 + nupd);
    StatScope st("trsv", s, (double)n * n);
    copy(s, n, x, cw.x0);
    const int tr = (trans == 'T' || trans == 't') ? 1 : 0;
    unsigned spin = trsv_spin();
    void *args[] = {&n, (void *)&tr, &L, &lda, &x, &ops, &flags, (void *)&nupd, &vg, &pg, (void *)&lt, &spin};
    const hipError_t e = launch_coop((const void *)k_trsv_binv, dim3(1 + nupd), args, s, BI_NT);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    // the recovery solves in Rtrsv's order: the Ozaki form refines the
    // result against B, and the two orders agree to dd rounding
    trsv_recover(s, tr ? 'T' : 'N', n, L, lda, x, flags + 1 + nupd, cw.x0);
    return 0;
}

void trsv_lower(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, dd *x) {
    if (n <= 0) return;
    StatScope st("trsv", s, (double)n * n);
    if (trans == 'N' || trans == 'n')
        k_trsv_ln<<<1, TS_NT, 0, s>>>(n, L, lda, x, nullptr, nullptr);
    else
        k_trsv_lt<<<1, TS_NT, 0, s>>>(n, L, lda, x, nullptr, nullptr);
}

// ---------------------------------------------- inverse of a lower triangle --
__global__ void k_set_identity(idx_t n, dd v, dd *A, idx_t lda) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) A[i + j * lda] = (i == j) ? v : mk(0.0);
}

// Rows [rb, rb+nb) of B := L^{-1} B for columns [0, ncols), with the
// updates from rows above already applied (Rtrsm 'L','L','N','N' order: if
// b_kj != 0 { b_kj = b_kj / a_kk; b_ij -= b_kj * a_ik for i > k }).  A wave
// owns a column, lane i its row i; the diagonal block is staged in LDS.
// FAST (the Ozaki form): b_k times the reciprocal 1/l_kk (formed once per
// launch, 64 divisions in parallel) in place of the division on the chain.
template <bool FAST>
__global__ __launch_bounds__(256) void k_trsm_llnn_diag(idx_t rb, int nb, idx_t ncols, const dd *L, idx_t lda, dd *B,
                                                        idx_t ldb) {
    __shared__ dd Lc[PB * (PB + 1) / 2];
    __shared__ dd rc[FAST ? PB : 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int e = tid; e < nb * nb; e += 256) {
        int j = e % nb, k = e / nb;
        if (j >= k) Lc[lpack(j, k)] = L[(rb + j) + (rb + k) * lda];
    }
    if (FAST && tid < nb) rc[tid] = dd_div(mk(1.0), L[(rb + tid) + (rb + tid) * lda]);
    __syncthreads();
    for (idx_t c = (idx_t)blockIdx.x * 4 + wave; c < ncols; c += (idx_t)gridDim.x * 4) {
        dd* col = B + rb + c * ldb;
        dd b = lane < nb ? col[lane] : mk(0.0);
        // step k's operands are read during step k-1, off the division chain
        // (lanes at or above nb read unused entries inside Lc)
        dd lnext = Lc[lpack(lane, 0)], dnext = FAST ? rc[0] : Lc[lpack(0, 0)];
        for (int k = 0; k < nb; k++) {
            const dd lk = lnext, dk = dnext;
            if (k + 1 < nb) {
                lnext = Lc[lpack(lane > k + 1 ? lane : k + 1, k + 1)];
                dnext = FAST ? rc[k + 1] : Lc[lpack(k + 1, k + 1)];
            }
            const int nz = __builtin_amdgcn_readlane(!dd_is_zero(b), k);
            if (!nz) continue;
            if (lane == k) b = FAST ? dd_mul(b, dk) : dd_div(b, dk);
            const dd bk = readlane_dd(b, k);
            if (lane > k && lane < nb) b = dd_sub(b, dd_mul(bk, lk));
        }
        if (lane < nb) col[lane] = b;
    }
}

// General B := L^{-1} B (m x n B); for the inverse, B starts as identity and
// only columns < rb+nb can be non-zero in the rows being solved.
static void trsm_llnn_impl(hipStream_t s, idx_t m, idx_t n, const dd *L, idx_t lda, dd *B, idx_t ldb, bool identity_rhs,
                           bool fast = false) {
    for (idx_t rb = 0; rb < m; rb += PB) {
        int nb = (int)std::min<idx_t>(PB, m - rb);
        idx_t ncols = identity_rhs ? std::min<idx_t>(n, rb + nb) : n;
        {
            StatScope st("trsm", s, (double)ncols * nb * nb);
            if (fast)
                k_trsm_llnn_diag<true><<<std::min<idx_t>(cdiv(ncols, 4), 1024), 256, 0, s>>>(rb, nb, ncols, L, lda, B,
                                                                                            ldb);
            else
                k_trsm_llnn_diag<false><<<std::min<idx_t>(cdiv(ncols, 4), 1024), 256, 0, s>>>(rb, nb, ncols, L, lda, B,
                                                                                             ldb);
        }
        idx_t below = m - rb - nb;
        if (below > 0) {
            // B[below, cols] += (-B[rb block, cols]) * L[below, rb block]: b_ij - b_kj*a_ik, k ascending
            gemm(s, 'N', below, ncols, nb, mk(-1.0), L + (rb + nb) + rb * lda, lda, B + rb, ldb, mk(1.0), B + (rb + nb), ldb);
        }
    }
}

void inv_lower(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *B, idx_t ldb, bool fast) {
    if (n <= 0) return;
    k_set_identity<<<dim3(cdiv(n, 256), (unsigned)n), 256, 0, s>>>(n, mk(1.0), B, ldb);
    trsm_llnn_impl(s, n, n, L, lda, B, ldb, true, fast);
}

void trsm_llnn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, dd *B, idx_t ldb) {
    if (m <= 0 || n <= 0) return;
    trsm_llnn_impl(s, m, n, A, lda, B, ldb, false);
}

// The Ozaki form's L^{-1}: blocked forward substitution on the identity in
// panels of 256 rows (64-row diagonal solves and dd updates inside a panel),
// the rows below a panel updated with the Ozaki product.
int inv_lower_oz(hipStream_t s, idx_t n, const dd *L, idx_t lda, dd *B, idx_t ldb, oz::Work &w, int S) {
    if (n <= 0) return 0;
    k_set_identity<<<dim3(cdiv(n, 256), (unsigned)n), 256, 0, s>>>(n, mk(1.0), B, ldb);
    const idx_t OZ_NB = 256;
    for (idx_t P = 0; P < n; P += OZ_NB) {
        const idx_t pw = std::min<idx_t>(OZ_NB, n - P);
        for (idx_t K = P; K < P + pw; K += PB) {
            const int nb = (int)std::min<idx_t>(PB, P + pw - K);
            const idx_t ncols = std::min<idx_t>(n, K + nb);
            {
                StatScope st("trsm", s, (double)ncols * nb * nb);
                k_trsm_llnn_diag<true><<<std::min<idx_t>(cdiv(ncols, 4), 1024), 256, 0, s>>>(K, nb, ncols, L, lda, B, ldb);
            }
            const idx_t rem = P + pw - K - nb;
            if (rem > 0)
                gemm(s, 'N', rem, ncols, nb, mk(-1.0), L + (K + nb) + K * lda, lda, B + K, ldb, mk(1.0), B + (K + nb),
                     ldb);
        }
        const idx_t below = n - P - pw;
        if (below > 0) {
            const idx_t ncols = std::min<idx_t>(n, P + pw);
            int rc = oz::gemm(s, w, 'N', 'N', below, ncols, pw, -1, L + (P + pw) + P * lda, lda, B + P, ldb, 1,
                              B + (P + pw), ldb, false, S);
            if (rc != 0) return rc;
        }
    }
    return 0;
}

// B := B * inv(A^T) for lower A (Rtrsm 'R','L','T','N', alpha = 1): the same
// panel kernels as the Cholesky, with the reciprocals one/a_kk.
__global__ void k_recips(idx_t n, const dd *A, idx_t lda, dd *r) {
    idx_t k = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) r[k] = dd_div(mk(1.0), A[k + k * lda]);
}
void trsm_rltn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, dd *B, idx_t ldb) {
    if (m <= 0 || n <= 0) return;
    dd *rinv = nullptr;
    hipMallocAsync((void **)&rinv, sizeof(dd) * n, s);
    k_recips<<<cdiv(n, 256), 256, 0, s>>>(n, A, lda, rinv);
    for (idx_t K = 0; K < n; K += PB) {
        int nb = (int)std::min<idx_t>(PB, n - K);
        // columns left of the panel were finished earlier: B[:, panel] += (-B[:, done]) * A[panel, done]^T
        if (K > 0) gemm(s, 'T', m, nb, K, mk(-1.0), B, ldb, A + K, lda, mk(1.0), B + K * ldb, ldb);
        k_panel_rltn<<<cdiv(m, PR), 256, 0, s>>>(B + K * ldb, ldb, 0, m, A + K + K * lda, lda, nb, rinv + K, nullptr);
    }
    hipFreeAsync(rinv, s);
}

// ------------------------------------------------------------------ TRMM --
// C[i,j] = b_ij*a_ii + sum_{k>i} a_ki*b_kj (k ascending) -- Rtrmm 'L','L','T','N'.
// When B is lower triangular too, terms with k < j vanish and are skipped.
__global__ __launch_bounds__(NT) void k_trmm_lltn(idx_t m, idx_t n, const dd *__restrict__ A, idx_t lda,
                                                 const dd *__restrict__ B, idx_t ldb, dd *__restrict__ C, idx_t ldc,
                                                 int b_lower) {
    __shared__ dd As[TK][TB + 1];
    __shared__ dd Bs[TK][TB + 1];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const idx_t row0 = (idx_t)blockIdx.x * TB, col0 = (idx_t)blockIdx.y * TB;
    dd acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            idx_t i = row0 + tx + 16 * r, j = col0 + ty + 16 * c;
            acc[r][c] = (i < m && j < n) ? dd_mul(B[i + j * ldb], A[i + i * lda]) : mk(0.0);
        }
    idx_t kbeg = row0 + 1;
    if (b_lower && col0 > kbeg) kbeg = col0;
    for (idx_t k0 = kbeg; k0 < m; k0 += TK) {
        const int kc = (int)((m - k0) < TK ? (m - k0) : TK);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int e = tid + NT * q;
            int kk = e & 15, rr = e >> 4;
            idx_t i = row0 + rr, j = col0 + rr;
            As[kk][rr] = (i < m && kk < kc) ? A[(k0 + kk) + i * lda] : mk(0.0);
            Bs[kk][rr] = (j < n && kk < kc) ? B[(k0 + kk) + j * ldb] : mk(0.0);
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk++) {
            idx_t k = k0 + kk;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                idx_t i = row0 + tx + 16 * r;
                if (k <= i) continue;
                dd a = As[kk][tx + 16 * r];
#pragma unroll
                for (int c = 0; c < 4; c++) acc[r][c] = dd_add_eq(acc[r][c], dd_mul(a, Bs[kk][ty + 16 * c]));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            idx_t i = row0 + tx + 16 * r, j = col0 + ty + 16 * c;
            if (i < m && j < n) C[i + j * ldc] = acc[r][c];
        }
}

// The same products on 16 x 16 tiles, one output element per thread (for
// grids of 64 x 64 tiles too small to fill the chip: a 300 x 300 product
// has 25 such tiles and 361 of these).  Each element keeps its chain:
// b_ij a_ii, then + a_ki b_kj for k > i ascending.
static constexpr int T16 = 16;
__global__ __launch_bounds__(256) void k_trmm_lltn16(idx_t m, idx_t n, const dd *__restrict__ A, idx_t lda,
                                                     const dd *__restrict__ B, idx_t ldb, dd *__restrict__ C, idx_t ldc,
                                                     int b_lower) {
    __shared__ dd As[T16][T16 + 1];
    __shared__ dd Bs[T16][T16 + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const idx_t row0 = (idx_t)blockIdx.x * T16, col0 = (idx_t)blockIdx.y * T16;
    const idx_t i = row0 + tx, j = col0 + ty;
    dd acc = (i < m && j < n) ? dd_mul(B[i + j * ldb], A[i + i * lda]) : mk(0.0);
    idx_t kbeg = row0 + 1;
    if (b_lower && col0 > kbeg) kbeg = col0;
    for (idx_t k0 = kbeg; k0 < m; k0 += T16) {
        // thread (tx, ty) stages A[k0 + tx, row0 + ty] and B[k0 + tx, col0 + ty]
        const idx_t k = k0 + tx, ra = row0 + ty, cb = col0 + ty;
        As[tx][ty] = (k < m && ra < m) ? A[k + ra * lda] : mk(0.0);
        Bs[tx][ty] = (k < m && cb < n) ? B[k + cb * ldb] : mk(0.0);
        __syncthreads();
        const int kc = (int)((m - k0) < T16 ? (m - k0) : T16);
        for (int kk = 0; kk < kc; kk++)
            if (k0 + kk > i) acc = dd_add_eq(acc, dd_mul(As[kk][tx], Bs[kk][ty]));
        __syncthreads();
    }
    if (i < m && j < n) C[i + j * ldc] = acc;
}

static void trmm_launch(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, const dd *B, idx_t ldb, dd *C,
                        idx_t ldc, int b_lower) {
    if ((idx_t)cdiv(m, TB) * cdiv(n, TB) < 512)
        k_trmm_lltn16<<<dim3(cdiv(m, T16), cdiv(n, T16)), 256, 0, s>>>(m, n, A, lda, B, ldb, C, ldc, b_lower);
    else
        k_trmm_lltn<<<dim3(cdiv(m, TB), cdiv(n, TB)), NT, 0, s>>>(m, n, A, lda, B, ldb, C, ldc, b_lower);
}
void trmm_lltn_self(hipStream_t s, idx_t n, const dd *A, idx_t lda, dd *C, idx_t ldc) {
    if (n <= 0) return;
    StatScope st("trmm", s, (double)n * n * n / 3.0);
    trmm_launch(s, n, n, A, lda, A, lda, C, ldc, 1);
}
void trmm_lltn(hipStream_t s, idx_t m, idx_t n, const dd *A, idx_t lda, const dd *B, idx_t ldb, dd *C, idx_t ldc) {
    if (m <= 0 || n <= 0) return;
    trmm_launch(s, m, n, A, lda, B, ldb, C, ldc, 0);
}

// SYRK lower/no-transpose: C := alpha*A*A^T + beta*C on the lower triangle.
void syrk_ln(hipStream_t s, idx_t n, idx_t k, dd alpha, const dd *A, idx_t lda, dd beta, dd *C, idx_t ldc) {
    gemm(s, 'T', n, n, k, alpha, A, lda, A, lda, beta, C, ldc, 1, 0, 0, 0, true, 0);
}

// ---------------------------------------------------------------- mat-vec --
// Mat-vecs inside Lanczos.  Each output is one dependent chain of dd
// additions in the reference's order, so the kernels are latency-bound on
// that chain.  A workgroup serves 64 outputs with four waves: wave 0
// (consumer) adds the products of tile t into the 64 chains, one per lane,
// while waves 1-3 (producers) form the products of tile t+1 into an LDS
// ring, each taking every third term of the tile, with their matrix loads
// issued two tiles ahead.
static constexpr int MU = 32;   // terms per tile
static constexpr int MP = 3;    // producer waves
static constexpr int MQ = (MU + MP - 1) / MP;

// The three mat-vecs differ only in how term u of tile t maps to a matrix
// element, a vector entry and the product's operand order; the consumer
// adds term u when mask bit u is set and keep(t, u) holds for its lane.
struct MvGemvN {  // y_i += (alpha*x_j)*A_ij, j ascending (Rgemv 'N')
    idx_t m, n, lda;
    const dd *A, *x;
    dd alpha;
    int ak;
    __device__ idx_t row(int lane, idx_t base) const { idx_t i = base + lane; return i < m ? i : m - 1; }
    __device__ idx_t vidx(int t, int u) const { idx_t j = (idx_t)t * MU + u; return j < n ? j : n - 1; }
    __device__ dd mat(idx_t r, int t, int u) const { return A[r + vidx(t, u) * lda]; }
    __device__ dd vec(int t, int u) const {
        dd xv = x[vidx(t, u)];
        return (ak == 1) ? xv : (ak == 2 ? dd_neg(xv) : dd_mul(alpha, xv));
    }
    __device__ dd prod(dd a, dd v) const { return dd_mul(v, a); }
    __device__ bool term_ok(int t, int u, const dd &) const { return (idx_t)t * MU + u < n; }
    __device__ bool keep(idx_t, int, int) const { return true; }
    __device__ bool full(idx_t, int t) const { return (idx_t)(t + 1) * MU <= n; }
};
struct MvTrmvLT {  // w_j = x_j*L_jj, += L_ij*x_i for i > j ascending (Rtrmv 'L','T')
    idx_t n, lda, ibeg;
    const dd *L, *x;
    __device__ idx_t row(int lane, idx_t base) const { idx_t j = base + lane; return j < n ? j : n - 1; }
    __device__ idx_t vidx(int t, int u) const { idx_t i = ibeg + (idx_t)t * MU + u; return i < n ? i : n - 1; }
    __device__ dd mat(idx_t r, int t, int u) const { return L[vidx(t, u) + r * lda]; }
    __device__ dd vec(int t, int u) const { return x[vidx(t, u)]; }
    __device__ dd prod(dd a, dd v) const { return dd_mul(a, v); }
    __device__ bool term_ok(int t, int u, const dd &) const { return ibeg + (idx_t)t * MU + u < n; }
    __device__ bool keep(idx_t j, int t, int u) const { return ibeg + (idx_t)t * MU + u > j; }
    __device__ bool full(idx_t j, int t) const { return ibeg + (idx_t)t * MU > j && ibeg + (idx_t)(t + 1) * MU <= n; }
};
struct MvTrmvLN {  // w_i = x_i*L_ii, += x_j*L_ij for j = i-1 down to 0, x_j != 0 (Rtrmv 'L','N')
    idx_t n, lda, jtop;
    const dd *L, *x;
    __device__ idx_t row(int lane, idx_t base) const { idx_t i = base + lane; return i < n ? i : n - 1; }
    __device__ idx_t vidx(int t, int u) const { idx_t j = jtop - (idx_t)t * MU - u; return j >= 0 ? j : 0; }
    __device__ dd mat(idx_t r, int t, int u) const { return L[r + vidx(t, u) * lda]; }
    __device__ dd vec(int t, int u) const { return x[vidx(t, u)]; }
    __device__ dd prod(dd a, dd v) const { return dd_mul(v, a); }
    __device__ bool term_ok(int t, int u, const dd &v) const { return jtop - (idx_t)t * MU - u >= 0 && !dd_is_zero(v); }
    __device__ bool keep(idx_t i, int t, int u) const { return jtop - (idx_t)t * MU - u < i; }
    __device__ bool full(idx_t i, int t) const { return jtop - (idx_t)t * MU < i; }
};

template <class K>
__device__ void mv_run(const K &k, int nt, idx_t base, dd acc, dd *out, idx_t nout) {
    __shared__ dd P[2][MU][64];
    __shared__ unsigned M[2][MP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const idx_t r = k.row(lane, base);
    if (wave > 0) {
        const int pw = wave - 1;
        dd a[2][MQ];
        dd xv[2];
        auto fetch = [&](int t, int slot) {
#pragma unroll
            for (int q = 0; q < MQ; q++) {
                const int u = pw + MP * q;
                if (u < MU) a[slot][q] = k.mat(r, t, u);
            }
            xv[slot] = lane < MU ? k.vec(t, lane) : mk(0.0);
        };
        auto produce = [&](int t, int slot, int buf) {
            unsigned mask = 0;
#pragma unroll
            for (int q = 0; q < MQ; q++) {
                const int u = pw + MP * q;
                if (u < MU) {
                    const dd v = readlane_dd(xv[slot], u);
                    P[buf][u][lane] = k.prod(a[slot][q], v);
                    if (k.term_ok(t, u, v)) mask |= 1u << u;
                }
            }
            if (lane == 0) M[buf][pw] = mask;
        };
        // tile t lives in register slot t & 1 and LDS buffer t & 1; the loop
        // is unrolled by two so that slots are compile-time indices
        fetch(0, 0);
        fetch(1, 1);
        if (nt > 0) produce(0, 0, 0);
        fetch(2, 0);
        __syncthreads();
        for (int t = 0; t < nt; t += 2) {
            if (t + 1 < nt) {
                produce(t + 1, 1, 1);
                fetch(t + 3, 1);
            }
            __syncthreads();
            if (t + 1 < nt) {
                if (t + 2 < nt) {
                    produce(t + 2, 0, 0);
                    fetch(t + 4, 0);
                }
                __syncthreads();
            }
        }
    } else {
        __syncthreads();
        for (int t = 0; t < nt; t++) {
            const int buf = t & 1;
            const unsigned mask = M[buf][0] | M[buf][1] | M[buf][2];
            dd pv[MU];
#pragma unroll
            for (int u = 0; u < MU; u++) pv[u] = P[buf][u][lane];
            __builtin_amdgcn_sched_barrier(0);  // all LDS reads issued before the chain
            if (mask == 0xffffffffu && k.full(base + lane, t)) {
#pragma unroll
                for (int u = 0; u < MU; u++) acc = dd_add_eq(acc, pv[u]);
            } else {
#pragma unroll
                for (int u = 0; u < MU; u++)
                    if (((mask >> u) & 1u) && k.keep(base + lane, t, u)) acc = dd_add_eq(acc, pv[u]);
            }
            __syncthreads();
        }
        if (base + lane < nout) out[base + lane] = acc;
    }
}

__global__ __launch_bounds__(256) void k_trmv_lt(idx_t n, const dd *L, idx_t lda, const dd *x, dd *w) {
    const idx_t j0 = (idx_t)blockIdx.x * 64;
    MvTrmvLT k{n, lda, j0 + 1, L, x};
    const int nt = k.ibeg < n ? (int)((n - k.ibeg + MU - 1) / MU) : 0;
    const idx_t jc = k.row(threadIdx.x & 63, j0);
    mv_run(k, nt, j0, dd_mul(x[jc], L[jc + jc * lda]), w, n);
}
__global__ __launch_bounds__(256) void k_trmv_ln(idx_t n, const dd *L, idx_t lda, const dd *x, dd *w) {
    const idx_t i0 = (idx_t)blockIdx.x * 64;
    MvTrmvLN k{n, lda, (i0 + 63 < n ? i0 + 63 : n - 1) - 1, L, x};
    const int nt = k.jtop >= 0 ? (int)((k.jtop + MU) / MU) : 0;
    const idx_t ic = k.row(threadIdx.x & 63, i0);
    const dd xi = x[ic];
    mv_run(k, nt, i0, dd_is_zero(xi) ? xi : dd_mul(xi, L[ic + ic * lda]), w, n);
}
void trmv_lower(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, const dd *x, dd *out) {
    if (n <= 0) return;
    StatScope st("mv", s, (double)n * n);
    if (trans == 'T' || trans == 't')
        k_trmv_lt<<<cdiv(n, 64), 256, 0, s>>>(n, L, lda, x, out);
    else
        k_trmv_ln<<<cdiv(n, 64), 256, 0, s>>>(n, L, lda, x, out);
}
// y_i = (beta handling) then += (alpha*x_j)*A_ij, j ascending (Rgemv 'N').
__global__ __launch_bounds__(256) void k_gemv_n(idx_t m, idx_t n, dd alpha, int ak, const dd *A, idx_t lda, const dd *x, dd beta, int bk,
                                                dd *y) {
    const idx_t i0 = (idx_t)blockIdx.x * 64;
    MvGemvN k{m, n, lda, A, x, alpha, ak};
    const idx_t ic = k.row(threadIdx.x & 63, i0);
    const dd init = (bk == 0) ? mk(0.0) : (bk == 1 ? y[ic] : dd_mul(beta, y[ic]));
    mv_run(k, (int)((n + MU - 1) / MU), i0, init, y, m);
}
void gemv_n(hipStream_t s, idx_t m, idx_t n, dd alpha, const dd *A, idx_t lda, const dd *x, dd beta, dd *y) {
    if (m <= 0) return;
    int ak = kind_alpha(alpha), bk = kind_beta(beta);
    bool az = (alpha.hi == 0.0 && alpha.lo == 0.0);
    if (n == 0 || (az && bk == 1)) return;
    if (az) {
        k_gemm_scale_only<<<dim3(cdiv(m, 256), 1, 1), 256, 0, s>>>(m, 1, beta, bk, y, m, 0);
        return;
    }
    StatScope st("mv", s, 2.0 * (double)m * n);
    k_gemv_n<<<cdiv(m, 64), 256, 0, s>>>(m, n, alpha, ak, A, lda, x, beta, bk, y);
}


// ------------------------------------------------- Ozaki form: fast order --
// Mat-vecs and dots of the Ozaki form, which does not keep the reference's
// summation order: sums are split into independent parts and merged, so no
// output waits on an n-long chain of dd additions.
// 'N'-type (y_i = sum_j A_ij x_j; kind 2: j <= i, the lower triangle): a
// 1024-thread workgroup serves 64 rows, wave w summing the w-th sixteenth of
// the columns; wave 0 merges the parts in order.
// 'N'-type mat-vec in a fixed split order: 8 rows per 1024-thread
// workgroup; thread t takes row t % 8 over column group t / 8 (128 groups,
// each a contiguous column range summed in order), and the 128 group sums
// of a row meet in a fixed pairwise LDS tree.  A wave loads 8 rows of 8
// columns per step (eight 128-byte segments).  kind 0: y = alpha A x +
// beta y; kind 2: y = L x with L lower triangular (columns <= the row).
constexpr int MV_R = 8, MV_G = 1024 / MV_R;
__device__ __forceinline__ void mvf_rows(int kind, idx_t m, idx_t n, const dd *A, idx_t lda, const dd *x, dd alpha,
                                         int ak, dd beta, int bk, dd *y, dd (*part)[MV_R]) {
    const int t = threadIdx.x, rl = t % MV_R, g = t / MV_R;
    const idx_t r = (idx_t)blockIdx.x * MV_R + rl;
    idx_t c0 = n * g / MV_G, c1 = n * (g + 1) / MV_G;
    if (kind == 2 && c1 > r + 1) c1 = r + 1;
    dd acc = mk(0.0);
    if (r < m) {
#pragma unroll 4
        for (idx_t c = c0; c < c1; c++) acc = dd_add(acc, dd_mul(A[r + c * lda], x[c]));
    }
    part[g][rl] = acc;
    __syncthreads();
#pragma unroll
    for (int half = MV_G / 2; half >= 1; half >>= 1) {
        if (g < half) part[g][rl] = dd_add(part[g][rl], part[g + half][rl]);
        __syncthreads();
    }
    if (g == 0 && r < m) {
        dd v = part[0][rl];
        if (kind == 0) {
            v = (ak == 1) ? v : (ak == 2 ? dd_neg(v) : dd_mul(alpha, v));
            if (bk == 1)
                v = dd_add(y[r], v);
            else if (bk == 2)
                v = dd_add(dd_mul(beta, y[r]), v);
        }
        y[r] = v;
    }
}
__global__ __launch_bounds__(1024) void k_mvf_n(int kind, idx_t m, idx_t n, const dd *A, idx_t lda, const dd *x,
                                                dd alpha, int ak, dd beta, int bk, dd *y) {
    __shared__ dd part[MV_G][MV_R];
    mvf_rows(kind, m, n, A, lda, x, alpha, ak, beta, bk, y, part);
}
// 'T'-type (w_j = sum_{i >= j} L_ij x_i): one wave per output, lanes over i,
// then a shuffle tree.
__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dd_add(v, shfl_dd(v, (threadIdx.x & 63) ^ o));
    return v;
}
__global__ __launch_bounds__(256) void k_mvf_lt(idx_t n, const dd *L, idx_t lda, const dd *x, dd *w) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const idx_t j = (idx_t)blockIdx.x * 4 + wave;
    if (j >= n) return;
    dd acc = mk(0.0);
    for (idx_t i = j + lane; i < n; i += 64) acc = dd_add(acc, dd_mul(L[i + j * lda], x[i]));
    acc = wave_sum_dd(acc);
    if (lane == 0) w[j] = acc;
}
void trmv_lower_fast(hipStream_t s, char trans, idx_t n, const dd *L, idx_t lda, const dd *x, dd *out) {
    if (n <= 0) return;
    StatScope st("mv", s, (double)n * n);
    if (trans == 'T' || trans == 't')
        k_mvf_lt<<<cdiv(n, 4), 256, 0, s>>>(n, L, lda, x, out);
    else
        k_mvf_n<<<cdiv(n, MV_R), 1024, 0, s>>>(2, n, n, L, lda, x, mk(1.0), 1, mk(0.0), 0, out);
}
void gemv_n_fast(hipStream_t s, idx_t m, idx_t n, dd alpha, const dd *A, idx_t lda, const dd *x, dd beta, dd *y) {
    if (m <= 0) return;
    const int ak = kind_alpha(alpha), bk = kind_beta(beta);
    const bool az = (alpha.hi == 0.0 && alpha.lo == 0.0);
    if (n == 0 || (az && bk == 1)) return;
    if (az) {
        k_gemm_scale_only<<<dim3(cdiv(m, 256), 1, 1), 256, 0, s>>>(m, 1, beta, bk, y, m, 0);
        return;
    }
    StatScope st("mv", s, 2.0 * (double)m * n);
    k_mvf_n<<<cdiv(m, MV_R), 1024, 0, s>>>(0, m, n, A, lda, x, alpha, ak, beta, bk, y);
}
// Dot with a fixed tree: per-thread sums over a grid stride, a workgroup
// tree, then the last workgroup to finish merges the workgroup sums in
// order (the ticket only picks which workgroup does it).
__global__ __launch_bounds__(256) void k_dot_fast(idx_t n, const dd *x, const dd *y, dd *out, dd *partial,
                                                  unsigned *counter) {
    __shared__ dd sm[256];
    __shared__ unsigned ticket;
    dd acc = mk(0.0);
    for (idx_t i = (idx_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (idx_t)gridDim.x * 256)
        acc = dd_add(acc, dd_mul(x[i], y[i]));
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
    __threadfence();
    if (threadIdx.x == 0) ticket = atomicAdd(counter, 1u);
    __syncthreads();
    if (ticket != gridDim.x - 1) return;
    if (threadIdx.x == 0) {
        __threadfence();
        dd r = mk(0.0);
        for (unsigned b = 0; b < gridDim.x; b++) r = dd_add(r, partial[b]);
        *out = r;
        *counter = 0u;
    }
}
// Many dots in the fixed-tree order: grid (G, ntasks), workgroup g of task
// t sums a grid stride of its products into partial[t G + g]; a second
// kernel adds each task's G partials in order.
__global__ __launch_bounds__(256) void k_dot_tasks_fast(const DotTask *tasks, dd *partial) {
    __shared__ dd sm[256];
    const DotTask tk = tasks[blockIdx.y];
    dd acc = mk(0.0);
    for (idx_t i = (idx_t)blockIdx.x * 256 + threadIdx.x; i < tk.n; i += (idx_t)gridDim.x * 256)
        acc = dd_add(acc, dd_mul(tk.x[i], tk.y[i]));
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(idx_t)blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}
__global__ void k_dot_tasks_merge(const DotTask *tasks, int ntasks, int G, const dd *partial) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntasks) return;
    dd acc = mk(0.0);
    for (int g = 0; g < G; g++) acc = dd_add(acc, partial[(idx_t)q * G + g]);
    *tasks[q].out = acc;
}
void dot_tasks_fast(hipStream_t s, const DotTask *d_tasks, int ntasks, idx_t maxn, int G, dd *partial) {
    if (ntasks <= 0) return;
    unsigned g = cdiv(maxn, 256 * 16);
    if (g > (unsigned)G) g = (unsigned)G;
    if (g == 0) g = 1;
    StatScope st("dot", s, 2.0 * (double)maxn * ntasks);
    k_dot_tasks_fast<<<dim3(g, (unsigned)ntasks), 256, 0, s>>>(d_tasks, partial);
    k_dot_tasks_merge<<<cdiv(ntasks, 64), 64, 0, s>>>(d_tasks, ntasks, (int)g, partial);
}
void dot_fast(hipStream_t s, idx_t n, const dd *x, const dd *y, dd *out, dd *work, unsigned *counter) {
    unsigned g = cdiv(n, 256 * 16);
    if (g > 256) g = 256;
    if (g == 0) g = 1;
    StatScope st("dot", s, 2.0 * (double)n);
    k_dot_fast<<<g, 256, 0, s>>>(n, x, y, out, work, counter);
}

// ----------------------------------------------------------- elementwise --
__global__ void k_fill(idx_t len, dd v, dd *x) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) x[i] = v;
}
__global__ void k_copy(idx_t len, const dd *x, dd *y) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) y[i] = x[i];
}
__global__ void k_scal(idx_t len, dd a, const dd *pa, dd *x) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pa) a = *pa;
    if (i < len) x[i] = dd_mul(a, x[i]);
}
__global__ void k_axpy(idx_t len, dd a, const dd *pa, int negate, const dd *x, dd *y) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pa) a = negate ? dd_neg(*pa) : *pa;
    if (dd_is_zero(a)) return;  // Raxpy: da == 0 returns
    if (i < len) y[i] = dd_add_eq(y[i], dd_mul(a, x[i]));
}
__global__ void k_zero_upper(idx_t n, dd *A) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < j && i < n) A[i + j * n] = mk(0.0);
}
// lower(i,j) = 0.5*(lower + 1*upper); upper = lower  (Lal::getSymmetrize)
__global__ void k_symmetrize(idx_t n, dd *A) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n || i <= j) return;
    dd v = dd_add_eq(A[i + j * n], dd_mul(mk(1.0), A[j + i * n]));
    v = dd_mul(mk(0.5), v);
    A[i + j * n] = v;
    A[j + i * n] = v;
}

// The QD operator set elementwise (qd/include/qd/dd_inline.h:26-393,
// qd/src/dd_real.cpp:47-69): 0 a+b, 1 a-b, 2 a*b, 3 a/b, 4 sqrt(a), 5 a+=b,
// 6 a-=b, 7 a*=b, 8 a*hi(b), 9 a/hi(b), 10 a+hi(b), 11 hi(a)/b.
__global__ void k_dd_vec_op(int op, idx_t n, const dd *__restrict__ a, const dd *__restrict__ b, dd *__restrict__ out) {
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const dd x = a[i], y = b[i];
    dd r;
    switch (op) {
    case 0: r = dd_add(x, y); break;
    case 1: r = dd_sub(x, y); break;
    case 2: r = dd_mul(x, y); break;
    case 3: r = dd_div(x, y); break;
    case 4: r = dd_sqrt(x); break;
    case 5: r = dd_add_eq(x, y); break;
    case 6: r = dd_sub(x, y); break;
    case 7: r = dd_mul_eq(x, y); break;
    case 8: r = dd_mul_d(x, y.hi); break;
    case 9: r = dd_div_d(x, y.hi); break;
    case 10: r = dd_add_d(x, y.hi); break;
    default: r = dd_div(mk(x.hi), y); break;
    }
    out[i] = r;
}
// Level-1 BLAS with strides (Raxpy_omp.cpp:79-112, Rscal.cpp, Rcopy.cpp):
// element i is x[x0 + i incx], y[y0 + i incy], the start offsets placing a
// negative stride's first element at the far end as the reference does.
// Elements are independent (x and y must not overlap).
__global__ void k_axpy_strided(idx_t n, dd a, const dd *x, idx_t x0, idx_t incx, dd *y, idx_t y0, idx_t incy) {
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dd &yy = y[y0 + i * incy];
    yy = dd_add_eq(yy, dd_mul(a, x[x0 + i * incx]));  // dy += da * dx
}
__global__ void k_scal_strided(idx_t n, dd a, dd *x, idx_t incx) {
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dd &xx = x[i * incx];
    xx = dd_mul(a, xx);  // dx = da * dx
}
__global__ void k_copy_strided(idx_t n, const dd *x, idx_t x0, idx_t incx, dd *y, idx_t y0, idx_t incy) {
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[y0 + i * incy] = x[x0 + i * incx];
}
void axpy_strided(hipStream_t s, idx_t n, dd a, const dd *x, idx_t incx, dd *y, idx_t incy) {
    if (n <= 0 || dd_is_zero(a)) return;
    const idx_t x0 = incx < 0 ? (1 - n) * incx : 0, y0 = incy < 0 ? (1 - n) * incy : 0;
    k_axpy_strided<<<cdiv(n, 256), 256, 0, s>>>(n, a, x, x0, incx, y, y0, incy);
}
void scal_strided(hipStream_t s, idx_t n, dd a, dd *x, idx_t incx) {
    if (n <= 0 || incx <= 0) return;
    k_scal_strided<<<cdiv(n, 256), 256, 0, s>>>(n, a, x, incx);
}
void copy_strided(hipStream_t s, idx_t n, const dd *x, idx_t incx, dd *y, idx_t incy) {
    if (n <= 0) return;
    const idx_t x0 = incx < 0 ? (1 - n) * incx : 0, y0 = incy < 0 ? (1 - n) * incy : 0;
    k_copy_strided<<<cdiv(n, 256), 256, 0, s>>>(n, x, x0, incx, y, y0, incy);
}

void dd_vec_op(hipStream_t s, int op, idx_t n, const dd *a, const dd *b, dd *out) {
    if (n > 0) k_dd_vec_op<<<cdiv(n, 256), 256, 0, s>>>(op, n, a, b, out);
}

void fill(hipStream_t s, idx_t len, dd v, dd *x) {
    if (len > 0) k_fill<<<cdiv(len, 256), 256, 0, s>>>(len, v, x);
}
void copy(hipStream_t s, idx_t len, const dd *x, dd *y) {
    if (len > 0 && x != y) k_copy<<<cdiv(len, 256), 256, 0, s>>>(len, x, y);
}
void scal(hipStream_t s, idx_t len, dd a, dd *x) {
    if (len > 0) k_scal<<<cdiv(len, 256), 256, 0, s>>>(len, a, nullptr, x);
}
void scal_dev(hipStream_t s, idx_t len, const dd *a, dd *x) {
    if (len > 0) k_scal<<<cdiv(len, 256), 256, 0, s>>>(len, mk(0.0), a, x);
}
void axpy(hipStream_t s, idx_t len, dd a, const dd *x, dd *y) {
    if (len > 0 && !(a.hi == 0.0 && a.lo == 0.0)) k_axpy<<<cdiv(len, 256), 256, 0, s>>>(len, a, nullptr, 0, x, y);
}
void axpy_dev(hipStream_t s, idx_t len, const dd *a, bool negate, const dd *x, dd *y) {
    if (len > 0) k_axpy<<<cdiv(len, 256), 256, 0, s>>>(len, mk(0.0), a, negate ? 1 : 0, x, y);
}
void set_identity(hipStream_t s, idx_t n, dd v, dd *A) {
    if (n > 0) k_set_identity<<<dim3(cdiv(n, 256), (unsigned)n), 256, 0, s>>>(n, v, A, n);
}
void zero_upper(hipStream_t s, idx_t n, dd *A) {
    if (n > 0) k_zero_upper<<<dim3(cdiv(n, 256), (unsigned)n), 256, 0, s>>>(n, A);
}
void symmetrize(hipStream_t s, idx_t n, dd *A) {
    if (n > 1) k_symmetrize<<<dim3(cdiv(n, 256), (unsigned)n), 256, 0, s>>>(n, A);
}

// ------------------------------------------------------------------ dots --
// Chunk t of the static partition of [0,n) into T chunks.
__device__ __forceinline__ void chunk_range(idx_t n, int T, int t, idx_t &start, idx_t &cnt) {
    idx_t q = n / T, r = n % T;
    cnt = q + (t < r ? 1 : 0);
    start = (idx_t)t * q + (t < r ? t : r);
}
__device__ __forceinline__ dd chunk_sum(const dd *x, const dd *y, idx_t start, idx_t cnt) {
    dd s = mk(0.0);
    for (idx_t i = start; i < start + cnt; i++) s = dd_add_eq(s, dd_mul(x[i], y[i]));
    return s;
}

// Long dots: one workgroup per (chunk, task).  All lanes form the products
// x_i*y_i of a tile into LDS; lane 0 adds them in index order (the chunk's
// running sum), so the per-chunk result is exactly Rdot_omp's thread-local
// sum.  A second kernel adds the T partial sums of each task in chunk order.
static constexpr int DOT_TILE = 1024;
__global__ __launch_bounds__(256) void k_dot_chunks(const DotTask *tasks, int T, dd *partial) {
    __shared__ dd prod[DOT_TILE];
    const DotTask tk = tasks[blockIdx.y];
    const int chunk = blockIdx.x;
    idx_t st, cnt;
    chunk_range(tk.n, T, chunk, st, cnt);
    dd acc = mk(0.0);
    for (idx_t t0 = 0; t0 < cnt; t0 += DOT_TILE) {
        const int tl = (int)((cnt - t0) < DOT_TILE ? (cnt - t0) : DOT_TILE);
        for (int e = threadIdx.x; e < tl; e += 256) prod[e] = dd_mul(tk.x[st + t0 + e], tk.y[st + t0 + e]);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int e = 0; e < tl; e++) acc = dd_add_eq(acc, prod[e]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(idx_t)blockIdx.y * T + chunk] = acc;
}
__global__ void k_dot_combine(const DotTask *tasks, int ntasks, int T, const dd *partial) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntasks) return;
    dd acc = mk(0.0);
    for (int c = 0; c < T; c++) acc = dd_add_eq(acc, partial[(idx_t)q * T + c]);
    *tasks[q].out = acc;
}
void dot_tasks_long(hipStream_t s, const DotTask *d_tasks, int ntasks, int T, dd *partial) {
    if (ntasks <= 0) return;
    k_dot_chunks<<<dim3((unsigned)T, (unsigned)ntasks), 256, 0, s>>>(d_tasks, T, partial);
    k_dot_combine<<<(ntasks + 63) / 64, 64, 0, s>>>(d_tasks, ntasks, T, partial);
}
__global__ void k_set_task(DotTask *t, const dd *x, const dd *y, idx_t n, dd *out) {
    t->x = x;
    t->y = y;
    t->n = n;
    t->out = out;
}

void dot(hipStream_t s, idx_t n, const dd *x, const dd *y, dd *out, int T, dd *work, unsigned *counter) {
    StatScope st("dot", s, 2.0 * (double)n);
    // work holds the task record followed by T partial sums
    DotTask *task = reinterpret_cast<DotTask *>(work);
    dd *partial = work + 4;
    (void)counter;
    k_set_task<<<1, 1, 0, s>>>(task, x, y, n, out);
    if (n / T <= 64)
        dot_batch(s, task, 1, T);  // short chunks: one workgroup, a lane per chunk
    else
        dot_tasks_long(s, task, 1, T, partial);
}

// Many dots in one launch; maxn is the longest.  Short chunks: one
// workgroup per task; long ones: one workgroup per (chunk, task).
void dot_tasks(hipStream_t s, const DotTask *d_tasks, int ntasks, idx_t maxn, int T, dd *partial) {
    if (ntasks <= 0) return;
    StatScope st("dot", s, 2.0 * (double)maxn * ntasks);
    if (maxn / T <= 64)
        dot_batch(s, d_tasks, ntasks, T);
    else
        dot_tasks_long(s, d_tasks, ntasks, T, partial);
}

// Batched dots: one workgroup per task, partial sums in LDS (T <= 4096).
__global__ __launch_bounds__(256) void k_dot_batch(const DotTask *tasks, int T) {
    extern __shared__ dd part[];
    const DotTask tk = tasks[blockIdx.x];
    for (int t = threadIdx.x; t < T; t += 256) {
        idx_t st, cnt;
        chunk_range(tk.n, T, t, st, cnt);
        part[t] = chunk_sum(tk.x, tk.y, st, cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dd acc = mk(0.0);
        for (int t = 0; t < T; t++) acc = dd_add_eq(acc, part[t]);
        *tk.out = acc;
    }
}
void dot_batch(hipStream_t s, const DotTask *d_tasks, int ntasks, int T) {
    if (ntasks > 0) k_dot_batch<<<ntasks, 256, sizeof(dd) * T, s>>>(d_tasks, T);
}
__global__ __launch_bounds__(256) void k_dot_batch_off(const DotTaskOff *tasks, const dd *ybase, int T) {
    extern __shared__ dd part[];
    const DotTaskOff tk = tasks[blockIdx.x];
    const dd *y = ybase + tk.yoff;
    for (int t = threadIdx.x; t < T; t += 256) {
        idx_t st, cnt;
        chunk_range(tk.n, T, t, st, cnt);
        part[t] = chunk_sum(tk.x, y, st, cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dd acc = mk(0.0);
        for (int t = 0; t < T; t++) acc = dd_add_eq(acc, part[t]);
        *tk.out = acc;
    }
}
void dot_batch_off(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T) {
    if (ntasks > 0) k_dot_batch_off<<<ntasks, 256, sizeof(dd) * T, s>>>(d_tasks, ybase, T);
}
// Long offset tasks: one workgroup per (chunk, task) as k_dot_chunks, then
// the chunk sums of each task in chunk order.
__global__ __launch_bounds__(256) void k_dot_chunks_off(const DotTaskOff *tasks, const dd *ybase, int T, dd *partial) {
    __shared__ dd prod[DOT_TILE];
    const DotTaskOff tk = tasks[blockIdx.y];
    const dd *y = ybase + tk.yoff;
    idx_t st, cnt;
    chunk_range(tk.n, T, blockIdx.x, st, cnt);
    dd acc = mk(0.0);
    for (idx_t t0 = 0; t0 < cnt; t0 += DOT_TILE) {
        const int tl = (int)((cnt - t0) < DOT_TILE ? (cnt - t0) : DOT_TILE);
        for (int e = threadIdx.x; e < tl; e += 256) prod[e] = dd_mul(tk.x[st + t0 + e], y[st + t0 + e]);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int e = 0; e < tl; e++) acc = dd_add_eq(acc, prod[e]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(idx_t)blockIdx.y * T + blockIdx.x] = acc;
}
__global__ void k_dot_combine_off(const DotTaskOff *tasks, int ntasks, int T, const dd *partial) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntasks) return;
    dd acc = mk(0.0);
    for (int c = 0; c < T; c++) acc = dd_add_eq(acc, partial[(idx_t)q * T + c]);
    *tasks[q].out = acc;
}
void dot_tasks_off(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T, dd *partial) {
    if (ntasks <= 0) return;
    k_dot_chunks_off<<<dim3((unsigned)T, (unsigned)ntasks), 256, 0, s>>>(d_tasks, ybase, T, partial);
    k_dot_combine_off<<<cdiv(ntasks, 64), 64, 0, s>>>(d_tasks, ntasks, T, partial);
}
// The Ozaki form's order: each chunk as 256 strided partial sums merged by
// a pairwise LDS tree (no sequential chain over the chunk), the chunks then
// added in order as above.
__global__ __launch_bounds__(256) void k_dot_chunks_off_fast(const DotTaskOff *tasks, const dd *ybase, int T,
                                                             dd *partial) {
    __shared__ dd sm[256];
    const DotTaskOff tk = tasks[blockIdx.y];
    const dd *y = ybase + tk.yoff;
    idx_t st, cnt;
    chunk_range(tk.n, T, blockIdx.x, st, cnt);
    dd acc = mk(0.0);
    for (idx_t e = threadIdx.x; e < cnt; e += 256) acc = dd_add(acc, dd_mul(tk.x[st + e], y[st + e]));
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(idx_t)blockIdx.y * T + blockIdx.x] = sm[0];
}
void dot_tasks_off_fast(hipStream_t s, const DotTaskOff *d_tasks, int ntasks, const dd *ybase, int T, dd *partial) {
    if (ntasks <= 0) return;
    k_dot_chunks_off_fast<<<dim3((unsigned)T, (unsigned)ntasks), 256, 0, s>>>(d_tasks, ybase, T, partial);
    k_dot_combine_off<<<cdiv(ntasks, 64), 64, 0, s>>>(d_tasks, ntasks, T, partial);
}

// out_i = A_ii for an n x n column-major block.
__global__ void k_gather_diag(idx_t n, const dd *A, dd *out) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = A[i + i * n];
}
void gather_diag(hipStream_t s, idx_t n, const dd *A, dd *out) {
    if (n > 0) k_gather_diag<<<cdiv(n, 256), 256, 0, s>>>(n, A, out);
}

// Lanczos vector updates (Lal::getMinEigen, sdpa_linear.cpp:64-82), fused:
// qold = q; q = v*r (Rcopy, Rcopy, Rscal).
__global__ void k_lanczos_q(idx_t n, dd v, const dd *r, dd *q, dd *qold) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        qold[i] = q[i];
        q[i] = dd_mul(v, r[i]);
    }
}
void lanczos_q(hipStream_t s, idx_t n, dd v, const dd *r, dd *q, dd *qold) {
    if (n > 0) k_lanczos_q<<<cdiv(n, 256), 256, 0, s>>>(n, v, r, q, qold);
}
// The same with the scalar on the device: v = sc[3] (the reciprocal of the
// previous step's beta, formed by k_lanczos_tail_dev).
__global__ void k_lanczos_q_dev(idx_t n, const dd *sc, const dd *r, dd *q, dd *qold) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const dd v = sc[3];
        qold[i] = q[i];
        q[i] = dd_mul(v, r[i]);
    }
}
void lanczos_q_dev(hipStream_t s, idx_t n, const dd *sc, const dd *r, dd *q, dd *qold) {
    if (n > 0) k_lanczos_q_dev<<<cdiv(n, 256), 256, 0, s>>>(n, sc, r, q, qold);
}
// alpha = q.w; r = w; r += (-alpha) q; r += (-beta) qold (Raxpy skips a
// zero scalar); rr = r.r.  Dots in the T-chunk order; out = {alpha, rr}.
// One workgroup: thread t owns chunk t of both dots and the elements of
// that chunk in the update.  With sc (device scalars) beta is sc[2] and the
// kernel then leaves the next step's sc[2] = sqrt(rr), sc[3] = 1/sc[2] (the
// host's Lal::getMinEigen scalar steps, in the same dd operations).
__global__ __launch_bounds__(256) void k_lanczos_tail(idx_t n, int T, const dd *q, const dd *qold, const dd *w, dd *r, dd nbeta,
                                                      dd *out, dd *sc) {
    extern __shared__ dd part[];
    __shared__ dd s_alpha;
    if (sc) nbeta = dd_neg(sc[2]);
    for (int t = threadIdx.x; t < T; t += 256) {
        idx_t st, cnt;
        chunk_range(n, T, t, st, cnt);
        part[t] = chunk_sum(q, w, st, cnt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dd acc = mk(0.0);
        for (int t = 0; t < T; t++) acc = dd_add_eq(acc, part[t]);
        s_alpha = acc;
        out[0] = acc;
    }
    __syncthreads();
    const dd na = dd_neg(s_alpha);
    const bool za = dd_is_zero(na), zb = dd_is_zero(nbeta);
    __syncthreads();
    for (int t = threadIdx.x; t < T; t += 256) {
        idx_t st, cnt;
        chunk_range(n, T, t, st, cnt);
        dd acc = mk(0.0);
        for (idx_t i = st; i < st + cnt; i++) {
            dd ri = w[i];
            if (!za) ri = dd_add_eq(ri, dd_mul(na, q[i]));
            if (!zb) ri = dd_add_eq(ri, dd_mul(nbeta, qold[i]));
            r[i] = ri;
            acc = dd_add_eq(acc, dd_mul(ri, ri));
        }
        part[t] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dd acc = mk(0.0);
        for (int t = 0; t < T; t++) acc = dd_add_eq(acc, part[t]);
        out[1] = acc;
        if (sc) {
            const dd beta = dd_sqrt(acc);
            sc[2] = beta;
            sc[3] = dd_div(mk(1.0), beta);
        }
    }
}
void lanczos_tail(hipStream_t s, idx_t n, int T, const dd *q, const dd *qold, const dd *w, dd *r, dd nbeta, dd *out) {
    StatScope st("dot", s, 6.0 * (double)n);
    k_lanczos_tail<<<1, 256, sizeof(dd) * T, s>>>(n, T, q, qold, w, r, nbeta, out, nullptr);
}
void lanczos_tail_dev(hipStream_t s, idx_t n, int T, const dd *q, const dd *qold, const dd *w, dd *r, dd *sc, dd *out) {
    StatScope st("dot", s, 6.0 * (double)n);
    k_lanczos_tail<<<1, 256, sizeof(dd) * T, s>>>(n, T, q, qold, w, r, mk(0.0), out, sc);
}
// The same step in a fixed split order (the Ozaki form): 256 contiguous
// chunks, each summed in order, then a pairwise LDS tree for each of the
// two dots.
__device__ __forceinline__ dd block_tree_sum(dd v, dd *sm) {
    sm[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) sm[threadIdx.x] = dd_add(sm[threadIdx.x], sm[threadIdx.x + h]);
        __syncthreads();
    }
    dd r = sm[0];
    __syncthreads();
    return r;
}
// The tail's body on threads 0..255 of a workgroup of at least 256 threads
// (the others only take part in the barriers); sm: 256 dd of LDS.
__device__ __forceinline__ void lanczos_tail_fast_body(idx_t n, const dd *q, const dd *qold, const dd *w, dd *r, dd *out,
                                                       dd *sc, dd *sm) {
    const int t = threadIdx.x;
    const bool tl = t < 256;
    const dd nbeta = dd_neg(sc[2]);
    const idx_t st = tl ? n * t / 256 : 0, en = tl ? n * (t + 1) / 256 : 0;
    dd acc = mk(0.0);
    for (idx_t i = st; i < en; i++) acc = dd_add(acc, dd_mul(q[i], w[i]));
    if (tl) sm[t] = acc;
    __syncthreads();
#pragma unroll
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) sm[t] = dd_add(sm[t], sm[t + h]);
        __syncthreads();
    }
    const dd alpha = sm[0];
    __syncthreads();
    const dd na = dd_neg(alpha);
    acc = mk(0.0);
    for (idx_t i = st; i < en; i++) {
        dd ri = dd_add(dd_add(w[i], dd_mul(na, q[i])), dd_mul(nbeta, qold[i]));
        r[i] = ri;
        acc = dd_add(acc, dd_mul(ri, ri));
    }
    if (tl) sm[t] = acc;
    __syncthreads();
#pragma unroll
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) sm[t] = dd_add(sm[t], sm[t + h]);
        __syncthreads();
    }
    if (t == 0) {
        const dd rr = sm[0];
        out[0] = alpha;
        out[1] = rr;
        const dd beta = dd_sqrt(rr);
        sc[2] = beta;
        sc[3] = dd_div(mk(1.0), beta);
    }
}
__global__ __launch_bounds__(256) void k_lanczos_tail_fast(idx_t n, const dd *q, const dd *qold, const dd *w, dd *r,
                                                           dd *out, dd *sc) {
    __shared__ dd sm[256];
    lanczos_tail_fast_body(n, q, qold, w, r, out, sc, sm);
}
void lanczos_tail_fast(hipStream_t s, idx_t n, const dd *q, const dd *qold, const dd *w, dd *r, dd *sc, dd *out) {
    StatScope st("dot", s, 6.0 * (double)n);
    k_lanczos_tail_fast<<<1, 256, 0, s>>>(n, q, qold, w, r, out, sc);
}

// A Lanczos step of the Ozaki form in three launches, with the same
// operations as lanczos_q_dev, trmv_lower_fast('T'), gemv_n_fast,
// trmv_lower_fast('N') and lanczos_tail_fast in sequence:
// k_mvf_lt_q: workgroup b forms q_i = sc[3] r_i for i >= 4b in LDS (the
// value lanczos_q_dev stores) and moves q to qold and stores the new q for
// its own four indices, then computes w_j = sum_{i >= j} L_ij q_i as
// k_mvf_lt does.  n <= LZ_MAXN.
constexpr idx_t LZ_MAXN = 4096;
__global__ __launch_bounds__(256) void k_mvf_lt_q(idx_t n, const dd *L, idx_t lda, const dd *r, const dd *sc, dd *q,
                                                  dd *qold, dd *w) {
    extern __shared__ dd xq[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const idx_t j0 = (idx_t)blockIdx.x * 4;
    const dd v = sc[3];
    for (idx_t i = j0 + threadIdx.x; i < n; i += 256) xq[i - j0] = dd_mul(v, r[i]);
    if (threadIdx.x < 4 && j0 + threadIdx.x < n) {
        const idx_t i = j0 + threadIdx.x;
        qold[i] = q[i];
        q[i] = dd_mul(v, r[i]);
    }
    __syncthreads();
    const idx_t j = j0 + wave;
    if (j >= n) return;
    dd acc = mk(0.0);
    for (idx_t i = j + lane; i < n; i += 64) acc = dd_add(acc, dd_mul(L[i + j * lda], xq[i - j0]));
    acc = wave_sum_dd(acc);
    if (lane == 0) w[j] = acc;
}
// k_mvf_ln_tail: w = L x as k_mvf_n (kind 2); the last workgroup to finish
// (ticket on ctr, left 0) then runs lanczos_tail_fast's body.  No
// workgroup waits for another.
__global__ __launch_bounds__(1024) void k_mvf_ln_tail(idx_t n, const dd *L, idx_t lda, const dd *x, const dd *q,
                                                      const dd *qold, dd *w, dd *r, dd *out, dd *sc, unsigned *ctr) {
    __shared__ dd part[MV_G][MV_R];
    __shared__ unsigned ticket;
    mvf_rows(2, n, n, L, lda, x, mk(1.0), 1, mk(0.0), 0, w, part);
    // w's stores come from wave 0 only (mvf_rows): its thread 0 publishes
    // them with the ticket
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        ticket = atomicAdd(ctr, 1u);
    }
    __syncthreads();
    if (ticket != gridDim.x - 1) return;
    __threadfence();
    lanczos_tail_fast_body(n, q, qold, w, r, out, sc, &part[0][0]);
    if (threadIdx.x == 0) *ctr = 0u;
}
bool lanczos_step_fast(hipStream_t s, idx_t n, const dd *L, const dd *M, dd *r, dd *q, dd *qold, dd *w, dd *tmp,
                       dd *sc, dd *out, unsigned *ctr) {
    if (n <= 0 || n > LZ_MAXN) return false;
    {
        StatScope st("mv", s, (double)n * n);
        k_mvf_lt_q<<<cdiv(n, 4), 256, sizeof(dd) * n, s>>>(n, L, n, r, sc, q, qold, w);
    }
    gemv_n_fast(s, n, n, mk(1.0), M, n, w, mk(0.0), tmp);
    StatScope st("mv", s, (double)n * n + 6.0 * (double)n);
    k_mvf_ln_tail<<<cdiv(n, MV_R), 1024, 0, s>>>(n, L, n, tmp, q, qold, w, r, out, sc, ctr);
    return true;
}

// Max of |x_i|: per-thread running max, tree merge, last workgroup merges.
// max |x_i| (exact, so independent of the grid) and, with y non-null,
// y = x in the same pass; four independent loads per thread in flight
__global__ __launch_bounds__(256) void k_max_abs(idx_t len, const dd *x, dd *y, dd *out, dd *partial,
                                                 unsigned *counter) {
    __shared__ dd sm[256];
    dd best = mk(0.0);
    const idx_t stride = (idx_t)gridDim.x * 256;
    idx_t i = (idx_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < len; i += 4 * stride) {
        dd v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = x[i + q * stride];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (y) y[i + q * stride] = v[q];
            const dd t = dd_abs(v[q]);
            if (dd_gt(t, best)) best = t;
        }
    }
    for (; i < len; i += stride) {
        const dd v = x[i];
        if (y) y[i] = v;
        const dd t = dd_abs(v);
        if (dd_gt(t, best)) best = t;
    }
    sm[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w && dd_gt(sm[threadIdx.x + w], sm[threadIdx.x])) sm[threadIdx.x] = sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
    __threadfence();
    __shared__ unsigned ticket;
    if (threadIdx.x == 0) ticket = atomicAdd(counter, 1u);
    __syncthreads();
    if (ticket != gridDim.x - 1) return;
    // the last workgroup reduces the partials (exact, any order)
    __threadfence();
    dd r = mk(0.0);
    for (unsigned b = threadIdx.x; b < gridDim.x; b += 256)
        if (dd_gt(partial[b], r)) r = partial[b];
    sm[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w && dd_gt(sm[threadIdx.x + w], sm[threadIdx.x])) sm[threadIdx.x] = sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *out = sm[0];
        *counter = 0u;
    }
}
void max_abs(hipStream_t s, idx_t len, const dd *x, dd *out, dd *work, unsigned *counter, dd *copy_to) {
    unsigned g = std::min<unsigned>(cdiv(len, 1024), (unsigned)MAX_ABS_BLOCKS);
    if (g == 0) g = 1;
    k_max_abs<<<g, 256, 0, s>>>(len, x, copy_to, out, work, counter);
}

}  // namespace sdk

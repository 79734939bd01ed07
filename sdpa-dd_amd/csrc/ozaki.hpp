// Ozaki-scheme dd matrix products on INT8 MFMA (gfx950).
//
// A dd matrix is split row-wise (or column-wise for the right operand)
// into S signed 7-bit digit planes: with e_r the exponent of row r's
// largest |hi| (frexp), every a_rk = 2^e_r * sum_s d_s,rk 2^-7(s+1) plus a
// remainder below 2^(e_r - 7S).  The digit extraction is exact dd
// arithmetic.  Products of planes accumulate exactly in int32 per level
// d = s + t (levels d < S are kept); the levels are recombined into dd with
// a Horner sum and scaled by 2^(e_r + e_c).  The error of an entry is below
// about k * S * 2^(e_r + e_c - 7S - 7), i.e. normwise relative to the
// largest entries of its row of A and column of B.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dd.hpp"

namespace oz {

using idx_t = int64_t;

constexpr int KC = 64;        // k per MFMA chunk (v_mfma_i32_16x16x64_i8)
constexpr int MAX_S = 16;     // digit planes (7 bits each)
constexpr idx_t MAX_K = 8192;  // int32 level sums stay exact up to this k

// Workspace for one product: digit planes and exponents of both operands.
// Row exponents are stored tagged, (tag << 12) | (e + 2048), with a tag
// that grows with every operand sliced: a row without a nonzero keeps an
// older (smaller) tag and reads as exponent 0, so the arrays need no
// clearing between products, only when they are new or the tag wraps.
struct Work {
    int8_t *sa = nullptr, *sb = nullptr;
    unsigned *ea = nullptr, *eb = nullptr;  // tagged row exponents
    size_t cap_a = 0, cap_b = 0, cap_ea = 0, cap_eb = 0;
    unsigned tag = 0;
    bool fresh = true;  // ea/eb hold no tagged values yet
    bool reserve(idx_t mpad, idx_t npad, idx_t kpad, int S, int esize = 1);
    unsigned next_tag(hipStream_t s);
    void release();
};

// C = beta*C + alpha*(op(A) op(B)) with alpha in {1, -1}, beta in {0, 1}:
//   op(A) = A (transa 'N', m x k) or A^T ('T', A is k x m);
//   op(B) = B (transb 'N', B is k x n) or B^T ('T', B is n x k).
// lower: only entries with i >= j are written.  k <= MAX_K per call (the
// caller splits longer products).  S in {8, 12, 14, 16}: INT8 planes of 7
// bits on INT8 MFMA; S in {-4 .. -7}: the FP64-slice variant, |S| digits of
// 18 bits held in doubles and accumulated by FP64 MFMA (exact while
// k <= MAX_K), error about k 2^(10 - 18|S|) max|A_i.| max|B_.j|.  Returns 0,
// -1 for bad arguments, -2 when the workspace cannot grow.
int gemm(hipStream_t s, Work &w, char transa, char transb, idx_t m, idx_t n, idx_t k, int alpha, const dd *A, idx_t lda,
         const dd *B, idx_t ldb, int beta, dd *C, idx_t ldc, bool lower, int S);

// C = beta*C + alpha*(M M(0:n, :)^T) for an m x k operand M (k <= KC, S =
// 16) whose digit planes (rows padded to roundup(m, 64)) and exponents,
// tagged `tag`, are already in w.sa / w.ea -- written by the kernel that
// produced M.  The caller reserves the workspace (prepare_planes) before
// that kernel runs.  Same results as gemm('N', 'T', ...) on M.
int gemm_planes(hipStream_t s, Work &w, unsigned tag, idx_t m, idx_t n, idx_t k, int alpha, int beta, dd *C, idx_t ldc,
                bool lower);
// Panel planes: an m-row operand with k = 64 nkb (S = 16), its planes laid
// out as k_slice lays them (plane s, chunk kb at rows (s nkb + kb) stride,
// stride = roundup(m, 64)) and written by the kernel that produced it.
// prepare_panel_planes reserves them and returns their tag (0 when the
// workspace cannot grow); gemm_panel_planes forms C = beta*C + alpha*(M'
// M'(0:n, :)^T) for M' = rows [row0, row0 + m) (row0 a multiple of 64,
// row0 + roundup(m, 64) <= stride).  Same results as gemm('N', 'T', ...)
// on M'.
unsigned prepare_panel_planes(hipStream_t s, Work &w, idx_t m, int nkb);
int gemm_panel_planes(hipStream_t s, Work &w, unsigned tag, idx_t stride, int nkb, idx_t row0, idx_t m, idx_t n,
                      int alpha, int beta, dd *C, idx_t ldc, bool lower);
// Reserves planes for an m-row operand with k <= KC and S = 16 and returns
// the tag its producer writes (0 when the workspace cannot grow).
unsigned prepare_planes(hipStream_t s, Work &w, idx_t m);

// ---- digit splitting, shared by the slicing kernels and the producers ----
// Row exponents are stored as (tag << 12) | (e + EXP_BIAS).
constexpr int EXP_BIAS = 2048;
__host__ __device__ __forceinline__ unsigned exp_tagged(unsigned tag, int e) { return (tag << 12) | (unsigned)(e + EXP_BIAS); }
__device__ __forceinline__ int exp_of(unsigned b, unsigned tag) {
    return (b >> 12) == tag ? (int)(b & 4095u) - EXP_BIAS : 0;
}
// The S signed 7-bit digits of a 2^-e (|a| 2^-e < 1), most significant
// first, packed as bytes into word[s] at byte position `pos` (0..3): digit
// d = trunc(128 x) in [-127, 127], x <- 128 x - d, the remainder carried
// exactly by a TwoSum.
template <int S>
__device__ __forceinline__ void put_digits(dd a, int e, unsigned *word, int pos) {
    double xh = ldexp(a.hi, -e), xl = ldexp(a.lo, -e);
#pragma unroll
    for (int s = 0; s < S; s++) {
        const double th = xh * 128.0, tl = xl * 128.0;
        const double d = trunc(th);
        const double rh = th - d;
        const double sm = rh + tl, bb = sm - rh;
        xl = (rh - (sm - bb)) + (tl - bb);
        xh = sm;
        word[s] |= ((unsigned)((int)d & 0xff)) << (8 * pos);
    }
}

}  // namespace oz

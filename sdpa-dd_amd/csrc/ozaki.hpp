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

}  // namespace oz

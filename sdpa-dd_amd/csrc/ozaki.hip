// Ozaki-scheme dd products on INT8 MFMA: digit-plane slicing, the
// level-accumulating MFMA kernel and the dd recombination.
#include "ozaki.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace oz {

typedef int v4i __attribute__((ext_vector_type(4)));

static inline unsigned cdiv(idx_t a, idx_t b) { return (unsigned)((a + b - 1) / b); }
static inline idx_t roundup(idx_t a, idx_t b) { return (a + b - 1) / b * b; }

bool Work::reserve(idx_t mpad, idx_t npad, idx_t kpad, int S, int esize) {
    auto grow = [](void **p, size_t &cap, size_t want) {
        if (want <= cap) return true;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        cap = 0;
        if (hipMalloc(p, want) != hipSuccess) return false;
        cap = want;
        return true;
    };
    const size_t old_ea = cap_ea, old_eb = cap_eb;
    bool ok = grow((void **)&sa, cap_a, (size_t)esize * S * mpad * kpad) &&
              grow((void **)&sb, cap_b, (size_t)esize * S * npad * kpad) &&
              grow((void **)&ea, cap_ea, sizeof(unsigned) * mpad) && grow((void **)&eb, cap_eb, sizeof(unsigned) * npad);
    if (cap_ea != old_ea || cap_eb != old_eb) fresh = true;
    return ok;
}
// Clears both exponent arrays (a kernel, not a memset: the stream's work may
// be captured into a graph by a caller).
__global__ void k_exp_clear(idx_t rows, unsigned *ex);
unsigned Work::next_tag(hipStream_t s) {
    if (fresh || tag >= (1u << 20) - 1) {
        const idx_t na = (idx_t)(cap_ea / sizeof(unsigned)), nb = (idx_t)(cap_eb / sizeof(unsigned));
        if (na > 0) k_exp_clear<<<cdiv(na, 256), 256, 0, s>>>(na, ea);
        if (nb > 0) k_exp_clear<<<cdiv(nb, 256), 256, 0, s>>>(nb, eb);
        tag = 0;
        fresh = false;
    }
    return ++tag;
}
void Work::release() {
    if (sa) (void)hipFree(sa);
    if (sb) (void)hipFree(sb);
    if (ea) (void)hipFree(ea);
    if (eb) (void)hipFree(eb);
    *this = Work();
}

// Row exponents: e_r with max_k |M(r,k).hi| = f 2^e, f in [0.5, 1) (0 for
// a zero row); element (r, k) is M[r*rs + k*cs].  Then |M(r,k)| 2^-e < 1.
// The grid matches k_slice: a thread takes 8 k of one row, the eight
// threads of a row reduce in LDS, and chunks combine with an integer
// atomicMax on the tagged value (tag << 12) | (e + EXP_BIAS) (order-free).
// Row exponents and slicing: a workgroup covers SL_R rows of one 64-wide k
// chunk with 8 threads per row (8 k values each).  Small row blocks keep
// the grid wide for the k = 64 Cholesky updates (one chunk).
constexpr int SL_R = 16, SL_T = SL_R * 8;
__global__ void k_exp_clear(idx_t rows, unsigned *ex) {
    idx_t r = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) ex[r] = 0u;
}
__global__ __launch_bounds__(SL_T) void k_rowexp(idx_t rows, idx_t k, const dd *M, idx_t rs, idx_t cs, unsigned *ex,
                                                 unsigned tag) {
    __shared__ double part[8][SL_R];
    const int t = threadIdx.x, kq = t / SL_R, rl = t % SL_R;
    const idx_t r = (idx_t)blockIdx.x * SL_R + rl;
    double mx = 0.0;
    if (r < rows) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const idx_t kk = (idx_t)blockIdx.y * KC + kq * 8 + q;
            if (kk < k) mx = fmax(mx, fabs(M[r * rs + kk * cs].hi));
        }
    }
    part[kq][rl] = mx;
    __syncthreads();
    if (kq == 0 && r < rows) {
#pragma unroll
        for (int w = 1; w < 8; w++) mx = fmax(mx, part[w][rl]);
        if (mx > 0.0) {
            int e = 0;
            (void)frexp(mx, &e);
            atomicMax(ex + r, exp_tagged(tag, e));
        }
    }
}

// Digit planes of rows [SL_R bx, SL_R bx + SL_R) over k chunk by: plane s of row r,
// chunk kb lives at out[((s nkb + kb) rows_pad + r) 64 + (k mod 64)].
// x = a 2^-e (|x| < 1); digit d = trunc(128 x) in [-127, 127], x <- 128 x - d
// (exact: the remainder is formed with a TwoSum).  Padded rows and k are 0.
typedef int v2i __attribute__((ext_vector_type(2)));
// ONE (k <= 64, a single chunk): the workgroup holds whole rows, so it also
// forms their exponents (as k_rowexp: the same maximum, written tagged even
// for a zero row, which decodes to 0 as an untouched entry does).
template <int S, bool ONE>
__global__ __launch_bounds__(SL_T) void k_slice(idx_t rows, idx_t k, const dd *M, idx_t rs, idx_t cs, unsigned *exps,
                                                unsigned tag, int8_t *out, idx_t rows_pad, int nkb) {
    __shared__ double part[8][SL_R];
    const int t = threadIdx.x, kq = t / SL_R, kb = blockIdx.y;  // thread: row t % SL_R, k values 8 kq .. 8 kq + 7
    const int rl = t % SL_R;
    const idx_t r = (idx_t)blockIdx.x * SL_R + rl;
    unsigned wd[S][2];
    dd av[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const idx_t kk = (idx_t)kb * KC + kq * 8 + q;
        av[q] = (r < rows && kk < k) ? M[r * rs + kk * cs] : mk(0.0);
    }
    int e = 0;
    if (ONE) {
        double mx = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) mx = fmax(mx, fabs(av[q].hi));
        part[kq][rl] = mx;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < 8; w++) mx = fmax(mx, part[w][rl]);
        if (mx > 0.0) (void)frexp(mx, &e);
        if (kq == 0 && r < rows) exps[r] = exp_tagged(tag, e);
    } else {
        e = r < rows ? exp_of(exps[r], tag) : 0;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        unsigned w4[S];
#pragma unroll
        for (int s = 0; s < S; s++) w4[s] = 0u;
#pragma unroll
        for (int q = 0; q < 4; q++) put_digits<S>(av[4 * h + q], e, w4, q);
#pragma unroll
        for (int s = 0; s < S; s++) wd[s][h] = w4[s];
    }
    if (r >= rows_pad) return;
#pragma unroll
    for (int s = 0; s < S; s++) {
        v2i v = {(int)wd[s][0], (int)wd[s][1]};
        *(v2i *)(out + ((((idx_t)s * nkb + kb) * rows_pad + r) * KC + kq * 8)) = v;
    }
}

// 64 x 64 output tile per 512-thread workgroup; wave w computes rows
// 16 (w % 4) .. +16 and columns 32 (w / 4) .. +32 (two 16 x 16 MFMA tiles).
// Per 64-wide k chunk the planes of both operands are staged in LDS (16-byte
// chunks XOR-swizzled by row so that a fragment read is conflict-free);
// level d = s + t of the product accumulates in acc[d] (int32, exact).
// An output element from its level sums: level d (int32, exact) carries
// weight 2^-7(d+2).  Levels are combined four at a time in int64, exactly:
// a group a0 2^21 + a1 2^14 + a2 2^7 + a3 with |a_d| < 2^31 stays below
// 2^53, so it is an exact double; the groups (unit 2^-7(4q+5)) are then
// summed in dd from the smallest, three dd additions where a Horner sum
// over the levels takes fifteen.  The result is in units of 2^-35.
static constexpr int RECOMB_UNIT = 35;
template <int S, class L>
__device__ __forceinline__ dd recombine(L lev) {
    constexpr int NG = (S + 3) / 4;
    double g[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) {
        long long G = 0;
#pragma unroll
        for (int d = 4 * q; d < 4 * q + 4; d++) G = G * 128 + (d < S ? (long long)lev(d) : 0ll);
        g[q] = (double)G;
    }
    dd h = mk(g[NG - 1]);
#pragma unroll
    for (int q = NG - 2; q >= 0; q--) {
        h.hi *= 0x1p-28;
        h.lo *= 0x1p-28;
        h = dd_add_d(h, g[q]);
    }
    return h;
}

// Tile order of the 64 x 64 kernel.  Workgroups are dealt to the eight
// XCDs in turn (workgroup L runs on XCD L mod 8), and each XCD has its own
// L2.  The tiles are grouped in super-blocks of XB_R x XB_C = 32 tiles,
// listed column-major (for a lower-triangular product only those that reach
// the lower triangle: super-column cs starts at super-row cs / 2), and the
// list is dealt to the XCDs in turn: XCD x takes super-blocks x, x + 8,
// x + 16, ..., workgroup L tile (L / 8) mod 32 of list entry
// x + 8 ((L / 8) / 32).  The 32 CUs of an XCD then work on one super-block
// at a time, whose XB_R row tiles and XB_C column tiles of digit planes come
// from HBM or the last-level cache once and from L2 after that.  Tiles
// outside the grid (partial super-blocks) leave at once.
static constexpr int XB_R = 8, XB_C = 4, XB = XB_R * XB_C, NXCD = 8;
__host__ __device__ __forceinline__ idx_t xcd_blocks(idx_t nsr, idx_t nsc, bool lower) {
    if (!lower) return nsr * nsc;
    idx_t t = 0;
    for (idx_t cs = 0; cs < nsc; cs++) t += nsr - cs / 2 > 0 ? nsr - cs / 2 : 0;
    return t;
}
static idx_t xcd_grid(idx_t mt, idx_t nt, bool lower) {
    const idx_t nsb = xcd_blocks((mt + XB_R - 1) / XB_R, (nt + XB_C - 1) / XB_C, lower);
    return NXCD * XB * ((nsb + NXCD - 1) / NXCD);
}
__device__ __forceinline__ bool xcd_tile(idx_t mt, idx_t nt, bool lower, idx_t &tr, idx_t &tc) {
    const idx_t L = blockIdx.x, j = L / NXCD, t = j % XB;
    const idx_t nsr = (mt + XB_R - 1) / XB_R, nsc = (nt + XB_C - 1) / XB_C;
    const idx_t sb = L % NXCD + NXCD * (j / XB);
    idx_t cs, rs;
    if (!lower) {
        if (sb >= nsr * nsc) return false;
        cs = sb / nsr;
        rs = sb % nsr;
    } else {
        idx_t base = 0;
        for (cs = 0;; cs++) {
            if (cs >= nsc) return false;
            const idx_t cnt = nsr - cs / 2 > 0 ? nsr - cs / 2 : 0;
            if (sb < base + cnt) break;
            base += cnt;
        }
        rs = cs / 2 + (sb - base);
    }
    tr = rs * XB_R + t % XB_R;
    tc = cs * XB_C + t / XB_R;
    return tr < mt && tc < nt;
}

template <int S>
__global__ __launch_bounds__(512) void k_gemm(idx_t m, idx_t n, int nkb, const int8_t *SA, idx_t mpad, const unsigned *EA,
                                              unsigned ta, const int8_t *SB, idx_t npad, const unsigned *EB, unsigned tb,
                                              dd *C, idx_t ldc, int alpha,
                                              int beta, int lower) {
    __shared__ v4i LA[S][64][4];
    __shared__ v4i LB[S][64][4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    idx_t tr, tc;
    // mpad and npad are the planes' row strides; the tile grid follows m, n
    if (!xcd_tile((m + 63) / 64, (n + 63) / 64, lower != 0, tr, tc)) return;
    const idx_t r0 = tr * 64, c0 = tc * 64;
    if (lower && r0 + 63 < c0) return;
    const int wr = wave & 3, wc = wave >> 2;
    const int fr = lane & 15, fq = lane >> 4;
    v4i acc[S][2];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d][0] = acc[d][1] = v4i{0, 0, 0, 0};
    for (int kb = 0; kb < nkb; kb++) {
        // every global load of the chunk is issued before the LDS stores,
        // so the chunk pays one memory latency instead of one per load
        constexpr int NQ = S * 256 / 512;
        v4i va[NQ], vb[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = tid + 512 * q, sl = c >> 8, row = (c >> 2) & 63, kq = c & 3;
            va[q] = *(const v4i *)(SA + ((((idx_t)sl * nkb + kb) * mpad + r0 + row) * KC + kq * 16));
            vb[q] = *(const v4i *)(SB + ((((idx_t)sl * nkb + kb) * npad + c0 + row) * KC + kq * 16));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = tid + 512 * q, sl = c >> 8, row = (c >> 2) & 63, kq = c & 3;
            const int sw = kq ^ ((row >> 2) & 3);
            LA[sl][row][sw] = va[q];
            LB[sl][row][sw] = vb[q];
        }
        __syncthreads();
        const int ar = wr * 16 + fr;
        v4i af[S];
#pragma unroll
        for (int s = 0; s < S; s++) af[s] = LA[s][ar][fq ^ ((ar >> 2) & 3)];
        const int br0 = wc * 32 + fr, br1 = br0 + 16;
#pragma unroll
        for (int t = 0; t < S; t++) {
            const v4i b0 = LB[t][br0][fq ^ ((br0 >> 2) & 3)];
            const v4i b1 = LB[t][br1][fq ^ ((br1 >> 2) & 3)];
#pragma unroll
            for (int s = 0; s + t < S; s++) {
                acc[s + t][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[s], b0, acc[s + t][0], 0, 0, 0);
                acc[s + t][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[s], b1, acc[s + t][1], 0, 0, 0);
            }
        }
    }
    // recombine: sum_d acc[d] 2^-7(d+2) (recombine<S>)
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {
        const idx_t j = c0 + wc * 32 + ct * 16 + fr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const idx_t i = r0 + wr * 16 + fq * 4 + r;
            if (i >= m || j >= n || (lower && i < j)) continue;
            dd h = recombine<S>([&](int d) { return acc[d][ct][r]; });
            const int ex = exp_of(EA[i], ta) + exp_of(EB[j], tb) - RECOMB_UNIT;
            h.hi = ldexp(h.hi, ex);
            h.lo = ldexp(h.lo, ex);
            if (alpha < 0) h = dd_neg(h);
            dd *cp = C + i + j * ldc;
            *cp = beta ? dd_add(*cp, h) : h;
        }
    }
}

// 32 x 32 output tile per 256-thread workgroup (wave w: the 16 x 16 tile at
// rows 16 (w % 2), columns 16 (w / 2)), for products whose 64 x 64 grid
// would leave most CUs idle: a quarter of the LDS, so two workgroups share
// a CU.  Same staging and recombination per element as k_gemm (the level
// sums are exact integers, so the results are identical).
template <int S>
__global__ __launch_bounds__(256) void k_gemm32(idx_t m, idx_t n, int nkb, const int8_t *SA, idx_t mpad,
                                                const unsigned *EA, unsigned ta, const int8_t *SB, idx_t npad,
                                                const unsigned *EB, unsigned tb, dd *C, idx_t ldc, int alpha, int beta,
                                                int lower) {
    __shared__ v4i LA[S][32][4];
    __shared__ v4i LB[S][32][4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const idx_t r0 = (idx_t)blockIdx.x * 32, c0 = (idx_t)blockIdx.y * 32;
    if (lower && r0 + 31 < c0) return;
    const int wr = wave & 1, wc = wave >> 1;
    const int fr = lane & 15, fq = lane >> 4;
    v4i acc[S];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d] = v4i{0, 0, 0, 0};
    for (int kb = 0; kb < nkb; kb++) {
        constexpr int NQ = S * 128 / 256;
        v4i va[NQ], vb[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = tid + 256 * q, sl = c >> 7, row = (c >> 2) & 31, kq = c & 3;
            va[q] = *(const v4i *)(SA + ((((idx_t)sl * nkb + kb) * mpad + r0 + row) * KC + kq * 16));
            vb[q] = *(const v4i *)(SB + ((((idx_t)sl * nkb + kb) * npad + c0 + row) * KC + kq * 16));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int c = tid + 256 * q, sl = c >> 7, row = (c >> 2) & 31, kq = c & 3;
            const int sw = kq ^ ((row >> 2) & 3);
            LA[sl][row][sw] = va[q];
            LB[sl][row][sw] = vb[q];
        }
        __syncthreads();
        const int ar = wr * 16 + fr, br = wc * 16 + fr;
        v4i af[S];
#pragma unroll
        for (int s = 0; s < S; s++) af[s] = LA[s][ar][fq ^ ((ar >> 2) & 3)];
#pragma unroll
        for (int t = 0; t < S; t++) {
            const v4i b = LB[t][br][fq ^ ((br >> 2) & 3)];
#pragma unroll
            for (int s = 0; s + t < S; s++) acc[s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[s], b, acc[s + t], 0, 0, 0);
        }
    }
    const idx_t j = c0 + wc * 16 + fr;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const idx_t i = r0 + wr * 16 + fq * 4 + r;
        if (i >= m || j >= n || (lower && i < j)) continue;
        dd h = recombine<S>([&](int d) { return acc[d][r]; });
        const int ex = exp_of(EA[i], ta) + exp_of(EB[j], tb) - RECOMB_UNIT;
        h.hi = ldexp(h.hi, ex);
        h.lo = ldexp(h.lo, ex);
        if (alpha < 0) h = dd_neg(h);
        dd *cp = C + i + j * ldc;
        *cp = beta ? dd_add(*cp, h) : h;
    }
}

template <int S>
static void product(hipStream_t s, Work &w, idx_t m, idx_t n, idx_t k, int nkb, unsigned ta, unsigned tb, bool same,
                    idx_t mpad, idx_t npad, int alpha, int beta, dd *C, idx_t ldc, bool lower);

template <int S>
static void launch(hipStream_t s, Work &w, idx_t m, idx_t n, idx_t k, const dd *A, idx_t ars, idx_t acs, const dd *B,
                   idx_t brs, idx_t bcs, int alpha, int beta, dd *C, idx_t ldc, bool lower) {
    const int nkb = (int)cdiv(k, KC);
    const idx_t mpad = roundup(m, 64), npad = roundup(n, 64);
    const unsigned ta = w.next_tag(s);
    if (nkb == 1) {
        k_slice<S, true><<<dim3(cdiv(mpad, SL_R), 1), SL_T, 0, s>>>(m, k, A, ars, acs, w.ea, ta, w.sa, mpad, 1);
    } else {
        k_rowexp<<<dim3(cdiv(m, SL_R), nkb), SL_T, 0, s>>>(m, k, A, ars, acs, w.ea, ta);
        k_slice<S, false><<<dim3(cdiv(mpad, SL_R), nkb), SL_T, 0, s>>>(m, k, A, ars, acs, w.ea, ta, w.sa, mpad, nkb);
    }
    // a product of a matrix with (a leading block of rows of) its own
    // transpose -- L L^T, and the Cholesky panel's A21 A21(0:n, :)^T -- has
    // B's rows among A's: one set of planes serves both, B's rows read with
    // A's padding (every B row the tiles touch, < roundup(n, 64) <= mpad,
    // lies in A's planes; columns beyond n are never stored)
    const bool same = (A == B && ars == brs && acs == bcs && n <= m);
    const unsigned tb = same ? ta : w.next_tag(s);
    if (!same) {
        if (nkb == 1) {
            k_slice<S, true><<<dim3(cdiv(npad, SL_R), 1), SL_T, 0, s>>>(n, k, B, brs, bcs, w.eb, tb, w.sb, npad, 1);
        } else {
            k_rowexp<<<dim3(cdiv(n, SL_R), nkb), SL_T, 0, s>>>(n, k, B, brs, bcs, w.eb, tb);
            k_slice<S, false><<<dim3(cdiv(npad, SL_R), nkb), SL_T, 0, s>>>(n, k, B, brs, bcs, w.eb, tb, w.sb, npad, nkb);
        }
    }
    product<S>(s, w, m, n, k, nkb, ta, same ? ta : tb, same, mpad, npad, alpha, beta, C, ldc, lower);
}

template <int S>
static void product_at(hipStream_t s, const int8_t *sa, const unsigned *ea, idx_t astride, unsigned ta, const int8_t *sb,
                       const unsigned *eb, idx_t bstride, unsigned tb, idx_t m, idx_t n, idx_t k, int nkb, int alpha,
                       int beta, dd *C, idx_t ldc, bool lower);
// The tile kernel of a product whose planes are in w (B's in w.sa when
// `same`).
template <int S>
static void product(hipStream_t s, Work &w, idx_t m, idx_t n, idx_t k, int nkb, unsigned ta, unsigned tb, bool same,
                    idx_t mpad, idx_t npad, int alpha, int beta, dd *C, idx_t ldc, bool lower) {
    product_at<S>(s, w.sa, w.ea, mpad, ta, same ? w.sa : w.sb, same ? w.ea : w.eb, same ? mpad : npad, tb, m, n, k, nkb,
                  alpha, beta, C, ldc, lower);
}
// The tile kernel on planes at sa / sb (row strides astride, bstride: the
// planes of chunk kb of plane s start at row (s nkb + kb) stride), timed as
// family "ozaki" with the product's algorithmic 2mnk dd-flops (the MFMA
// work is S(S+1)/2 times that in int8 multiply-adds).
template <int S>
static void product_at(hipStream_t s, const int8_t *sa, const unsigned *ea, idx_t astride, unsigned ta, const int8_t *sb,
                       const unsigned *eb, idx_t bstride, unsigned tb, idx_t m, idx_t n, idx_t k, int nkb, int alpha,
                       int beta, dd *C, idx_t ldc, bool lower) {
    sdk::StatScope st("ozaki", s, 2.0 * (double)m * n * k * (lower ? 0.5 : 1.0));
    const idx_t mpad = roundup(m, 64), npad = roundup(n, 64);
    // 64 x 64 tiles hold a CU each (128 KB of LDS at S = 16): below about
    // two per CU, the 32 x 32 kernel fills the chip better
    const idx_t t64 = cdiv(mpad, 64) * cdiv(npad, 64) / (lower ? 2 : 1);
    const bool small = sdk::g_tile_policy == 1 || (sdk::g_tile_policy == 0 && t64 < 512);
    if (small)
        k_gemm32<S><<<dim3(cdiv(mpad, 32), cdiv(npad, 32)), 256, 0, s>>>(m, n, nkb, sa, astride, ea, ta, sb, bstride, eb,
                                                                          tb, C, ldc, alpha, beta, lower ? 1 : 0);
    else
        k_gemm<S><<<dim3((unsigned)xcd_grid(mpad / 64, npad / 64, lower)), 512, 0, s>>>(
            m, n, nkb, sa, astride, ea, ta, sb, bstride, eb, tb, C, ldc, alpha, beta, lower ? 1 : 0);
}

int gemm_panel_planes(hipStream_t s, Work &w, unsigned tag, idx_t stride, int nkb, idx_t row0, idx_t m, idx_t n,
                      int alpha, int beta, dd *C, idx_t ldc, bool lower) {
    if (m < 0 || n < 0 || n > m || nkb < 1 || (idx_t)nkb * KC > MAX_K || (alpha != 1 && alpha != -1) ||
        (beta != 0 && beta != 1) || !tag || row0 < 0 || row0 % 64 != 0 || row0 + roundup(m, 64) > stride)
        return -1;
    if (m == 0 || n == 0) return 0;
    const int8_t *pa = w.sa + row0 * KC;
    const unsigned *pe = w.ea + row0;
    product_at<16>(s, pa, pe, stride, tag, pa, pe, stride, tag, m, n, (idx_t)nkb * KC, nkb, alpha, beta, C, ldc, lower);
    return 0;
}
unsigned prepare_panel_planes(hipStream_t s, Work &w, idx_t m, int nkb) {
    if (!w.reserve(roundup(m, 64), roundup(m, 64), (idx_t)KC * nkb, 16, 1)) return 0;
    return w.next_tag(s);
}

unsigned prepare_planes(hipStream_t s, Work &w, idx_t m) {
    if (!w.reserve(roundup(m, 64), roundup(m, 64), KC, 16, 1)) return 0;
    return w.next_tag(s);
}

int gemm_planes(hipStream_t s, Work &w, unsigned tag, idx_t m, idx_t n, idx_t k, int alpha, int beta, dd *C, idx_t ldc,
                bool lower) {
    if (m < 0 || n < 0 || n > m || k < 1 || k > KC || (alpha != 1 && alpha != -1) || (beta != 0 && beta != 1) || !tag)
        return -1;
    if (m == 0 || n == 0) return 0;
    const idx_t mpad = roundup(m, 64);
    product<16>(s, w, m, n, k, 1, tag, tag, true, mpad, roundup(n, 64), alpha, beta, C, ldc, lower);
    return 0;
}


// ------------------------------------------------------ FP64-slice variant --
// The same splitting with FB-bit digits held in doubles: d = trunc(2^FB x).
// Plane products accumulate per level in FP64 MFMA (v_mfma_f64_16x16x4f64);
// |d| < 2^18, so a level sum over k <= 8192 with at most 7 products per k
// stays below 2^52 and is exact.  Recombination: sum_d acc[d] 2^-FB(d+2).
constexpr int FB = 18;
constexpr int KC2 = 8;  // k per LDS chunk (two MFMA k-steps of 4)
typedef double v4d __attribute__((ext_vector_type(4)));

// plane s of row r, chunk kb: out[((s nkb + kb) rows_pad + r) KC2 + (k mod KC2)]
template <int S>
__global__ __launch_bounds__(512) void k_slice_f64(idx_t rows, idx_t k, const dd *M, idx_t rs, idx_t cs,
                                                   const unsigned *exps, unsigned tag, double *out, idx_t rows_pad,
                                                   int nkb) {
    const int t = threadIdx.x, kq = t >> 6, kb = blockIdx.y;
    const idx_t r = (idx_t)blockIdx.x * 64 + (t & 63);
    if (r >= rows_pad) return;
    const idx_t kk = (idx_t)kb * KC2 + kq;
    const int e = r < rows ? exp_of(exps[r], tag) : 0;
    const dd a = (r < rows && kk < k) ? M[r * rs + kk * cs] : mk(0.0);
    double xh = ldexp(a.hi, -e), xl = ldexp(a.lo, -e);
    constexpr double SC = (double)(1 << FB);
#pragma unroll
    for (int s = 0; s < S; s++) {
        const double th = xh * SC, tl = xl * SC;
        const double d = trunc(th);
        const double rh = th - d;
        const double sm = rh + tl, bb = sm - rh;
        xl = (rh - (sm - bb)) + (tl - bb);
        xh = sm;
        out[(((idx_t)s * nkb + kb) * rows_pad + r) * KC2 + kq] = d;
    }
}

// 64 x 64 output tile per 512-thread workgroup, wave layout as k_gemm.
template <int S>
__global__ __launch_bounds__(512) void k_gemm_f64(idx_t m, idx_t n, int nkb, const double *SA, idx_t mpad,
                                                  const unsigned *EA, unsigned ta, const double *SB, idx_t npad,
                                                  const unsigned *EB, unsigned tb, dd *C, idx_t ldc, int alpha, int beta,
                                                  int lower) {
    __shared__ double LA[S][64][KC2 + 1];
    __shared__ double LB[S][64][KC2 + 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const idx_t r0 = (idx_t)blockIdx.x * 64, c0 = (idx_t)blockIdx.y * 64;
    if (lower && r0 + 63 < c0) return;
    const int wr = wave & 3, wc = wave >> 2;
    const int fr = lane & 15, fq = lane >> 4;
    v4d acc[S][2];
#pragma unroll
    for (int d = 0; d < S; d++) acc[d][0] = acc[d][1] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int kb = 0; kb < nkb; kb++) {
        __syncthreads();
        for (int c = tid; c < S * 64 * KC2; c += 512) {
            const int sl = c / (64 * KC2), row = (c / KC2) & 63, kq = c % KC2;
            LA[sl][row][kq] = SA[(((idx_t)sl * nkb + kb) * mpad + r0 + row) * KC2 + kq];
            LB[sl][row][kq] = SB[(((idx_t)sl * nkb + kb) * npad + c0 + row) * KC2 + kq];
        }
        __syncthreads();
#pragma unroll
        for (int k4 = 0; k4 < KC2; k4 += 4) {
            double af[S];
#pragma unroll
            for (int s = 0; s < S; s++) af[s] = LA[s][wr * 16 + fr][k4 + fq];
#pragma unroll
            for (int t = 0; t < S; t++) {
                const double b0 = LB[t][wc * 32 + fr][k4 + fq];
                const double b1 = LB[t][wc * 32 + 16 + fr][k4 + fq];
#pragma unroll
                for (int s = 0; s + t < S; s++) {
                    acc[s + t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], b0, acc[s + t][0], 0, 0, 0);
                    acc[s + t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], b1, acc[s + t][1], 0, 0, 0);
                }
            }
        }
    }
    constexpr double STEP = 1.0 / (double)(1 << FB);
#pragma unroll
    for (int ct = 0; ct < 2; ct++) {
        const idx_t j = c0 + wc * 32 + ct * 16 + fr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const idx_t i = r0 + wr * 16 + fq + 4 * r;  // v_mfma_f64_16x16x4f64: row g + 4r in lane j + 16g
            if (i >= m || j >= n || (lower && i < j)) continue;
            dd h = mk(acc[S - 1][ct][r]);
#pragma unroll
            for (int d = S - 2; d >= 0; d--) {
                h.hi *= STEP;
                h.lo *= STEP;
                h = dd_add_d(h, acc[d][ct][r]);
            }
            const int ex = exp_of(EA[i], ta) + exp_of(EB[j], tb) - 2 * FB;
            h.hi = ldexp(h.hi, ex);
            h.lo = ldexp(h.lo, ex);
            if (alpha < 0) h = dd_neg(h);
            dd *cp = C + i + j * ldc;
            *cp = beta ? dd_add(*cp, h) : h;
        }
    }
}

template <int S>
static void launch_f64(hipStream_t s, Work &w, idx_t m, idx_t n, idx_t k, const dd *A, idx_t ars, idx_t acs,
                       const dd *B, idx_t brs, idx_t bcs, int alpha, int beta, dd *C, idx_t ldc, bool lower) {
    const int nkb = (int)cdiv(k, KC2), nkb64 = (int)cdiv(k, KC);
    const idx_t mpad = roundup(m, 64), npad = roundup(n, 64);
    double *sa = (double *)w.sa, *sb = (double *)w.sb;
    const unsigned ta = w.next_tag(s);
    k_rowexp<<<dim3(cdiv(m, SL_R), nkb64), SL_T, 0, s>>>(m, k, A, ars, acs, w.ea, ta);
    k_slice_f64<S><<<dim3(cdiv(mpad, 64), nkb), 512, 0, s>>>(m, k, A, ars, acs, w.ea, ta, sa, mpad, nkb);
    const bool same = (A == B && ars == brs && acs == bcs && m == n);
    const unsigned tb = same ? ta : w.next_tag(s);
    if (!same) {
        k_rowexp<<<dim3(cdiv(n, SL_R), nkb64), SL_T, 0, s>>>(n, k, B, brs, bcs, w.eb, tb);
        k_slice_f64<S><<<dim3(cdiv(npad, 64), nkb), 512, 0, s>>>(n, k, B, brs, bcs, w.eb, tb, sb, npad, nkb);
    }
    sdk::StatScope st("ozaki", s, 2.0 * (double)m * n * k * (lower ? 0.5 : 1.0));
    k_gemm_f64<S><<<dim3(cdiv(mpad, 64), cdiv(npad, 64)), 512, 0, s>>>(m, n, nkb, sa, mpad, w.ea, ta, same ? sa : sb,
                                                                      npad, same ? w.ea : w.eb, tb, C, ldc, alpha, beta,
                                                                      lower ? 1 : 0);
}

int gemm(hipStream_t s, Work &w, char transa, char transb, idx_t m, idx_t n, idx_t k, int alpha, const dd *A, idx_t lda,
         const dd *B, idx_t ldb, int beta, dd *C, idx_t ldc, bool lower, int S) {
    if (m < 0 || n < 0 || k < 0 || k > MAX_K || (alpha != 1 && alpha != -1) || (beta != 0 && beta != 1)) return -1;
    if (m == 0 || n == 0) return 0;
    if (k == 0) return beta == 1 ? 0 : -1;  // an empty product only leaves C as it is
    const bool nt = (transb == 'T' || transb == 't'), ta = (transa == 'T' || transa == 't');
    if (!nt && transb != 'N' && transb != 'n') return -1;
    if (!ta && transa != 'N' && transa != 'n') return -1;
    const idx_t kpad = roundup(k, KC);
    const bool f64 = S < 0;
    if (f64 && (S < -7 || S > -4)) return -1;
    if (!w.reserve(roundup(m, 64), roundup(n, 64), kpad, f64 ? -S : S, f64 ? 8 : 1)) return -2;
    // element (i, kk) of op(A): A[i + kk lda] for 'N', A[kk + i lda] for 'T';
    // element (j, kk) of op(B)^T: B[kk + j ldb] for 'N', B[j + kk ldb] for 'T'
    const idx_t ars = ta ? lda : 1, acs = ta ? 1 : lda;
    const idx_t brs = nt ? 1 : ldb, bcs = nt ? ldb : 1;
    switch (S) {
    case -7: launch_f64<7>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case -6: launch_f64<6>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case -5: launch_f64<5>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case -4: launch_f64<4>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case 16: launch<16>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case 14: launch<14>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case 12: launch<12>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    case 8: launch<8>(s, w, m, n, k, A, ars, acs, B, brs, bcs, alpha, beta, C, ldc, lower); break;
    default: return -1;
    }
    return 0;
}

}  // namespace oz

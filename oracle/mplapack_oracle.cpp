// ORACLE — test infrastructure only.
// CPU restatement of the MPLAPACK 2.0.1 dd routines on the SDPA-DD Newton
// step.  Each function cites the reference routine it restates; loop orders
// and the form of every dd update (`c += t*a`, `b = b - t*c`, `t = one/a`
// then `t*b`, ...) follow the reference because the rounding depends on them.
#include "mplapack_oracle.hpp"

#include <algorithm>

namespace orc {

int g_dot_threads = 1;

// mplapack/Rgemm_omp.cpp:41-127, Rgemm_NN_omp.cpp:36-65, Rgemm_NT_omp.cpp:31-57.
// Only the NN and NT forms are reached on the path.
void Rgemm(char ta, char tb, mint m, mint n, mint k, dd_real alpha, const dd_real *A, mint lda, const dd_real *B,
           mint ldb, dd_real beta, dd_real *C, mint ldc) {
    const dd_real zero = 0.0, one = 1.0;
    if (m == 0 || n == 0 || (((alpha == zero) || (k == 0)) && (beta == one))) return;
    if (alpha == zero) {
        for (mint j = 0; j < n; j++)
            for (mint i = 0; i < m; i++) C[i + j * ldc] = (beta == zero) ? zero : beta * C[i + j * ldc];
        return;
    }
    if (ta != 'N' && ta != 'n') std::abort();  // transposed A is never reached on this path
    bool notb = (tb == 'N' || tb == 'n');
    for (mint j = 0; j < n; j++) {
        if (beta == 0.0) {
            for (mint i = 0; i < m; i++) C[i + j * ldc] = 0.0;
        } else if (beta != 1.0) {
            for (mint i = 0; i < m; i++) C[i + j * ldc] = beta * C[i + j * ldc];
        }
    }
    // Columns are independent (Rgemm_NN_omp.cpp:54 parallelises over them too).
#pragma omp parallel for schedule(static)
    for (mint j = 0; j < n; j++) {
        for (mint l = 0; l < k; l++) {
            dd_real temp = alpha * (notb ? B[l + j * ldb] : B[j + l * ldb]);
            for (mint i = 0; i < m; i++) C[i + j * ldc] += temp * A[i + l * lda];
        }
    }
}

// mplapack/Rsyrk.cpp:158-176 (Lower, No transpose).
void Rsyrk_LN(mint n, mint k, dd_real alpha, const dd_real *A, mint lda, dd_real beta, dd_real *C, mint ldc) {
    const dd_real zero = 0.0, one = 1.0;
    if (n == 0 || (((alpha == zero) || (k == 0)) && (beta == one))) return;
    if (alpha == zero) {
        for (mint j = 0; j < n; j++)
            for (mint i = j; i < n; i++) C[i + j * ldc] = (beta == zero) ? zero : beta * C[i + j * ldc];
        return;
    }
    // Columns are independent: each C element keeps the serial loop's order.
#pragma omp parallel for schedule(dynamic, 4) if (n * (k + 1) > 8192)
    for (mint j = 0; j < n; j++) {
        if (beta == zero) {
            for (mint i = j; i < n; i++) C[i + j * ldc] = zero;
        } else if (beta != one) {
            for (mint i = j; i < n; i++) C[i + j * ldc] = beta * C[i + j * ldc];
        }
        for (mint l = 0; l < k; l++) {
            if (A[j + l * lda] != zero) {
                dd_real temp = alpha * A[j + l * lda];
                for (mint i = j; i < n; i++) C[i + j * ldc] += temp * A[i + l * lda];
            }
        }
    }
}

// mplapack/Rtrsm.cpp:256-283 (Right, Lower, Transpose, Non-unit):
// B := alpha * B * inv(A**T).
void Rtrsm_RLTN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb) {
    const dd_real zero = 0.0, one = 1.0;
    if (m == 0 || n == 0) return;
    if (alpha == zero) {
        for (mint j = 0; j < n; j++)
            for (mint i = 0; i < m; i++) B[i + j * ldb] = zero;
        return;
    }
    // Rows of B are independent: each thread runs the loop nest on its own
    // range of rows, so every element sees the serial operation sequence.
    const mint RB = 16;
#pragma omp parallel for schedule(dynamic, 1) if (m * n > 8192)
    for (mint i0 = 0; i0 < m; i0 += RB) {
        const mint i1 = std::min(m, i0 + RB);
        for (mint k = 0; k < n; k++) {
            dd_real temp = one / A[k + k * lda];
            for (mint i = i0; i < i1; i++) B[i + k * ldb] = temp * B[i + k * ldb];
            for (mint j = k + 1; j < n; j++) {
                if (A[j + k * lda] != zero) {
                    temp = A[j + k * lda];
                    for (mint i = i0; i < i1; i++) B[i + j * ldb] = B[i + j * ldb] - temp * B[i + k * ldb];
                }
            }
            if (alpha != one)
                for (mint i = i0; i < i1; i++) B[i + k * ldb] = alpha * B[i + k * ldb];
        }
    }
}

// mplapack/Rtrsm.cpp:128-142 (Left, Lower, No transpose, Non-unit):
// B := alpha * inv(A) * B.
void Rtrsm_LLNN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb) {
    const dd_real zero = 0.0, one = 1.0;
    if (m == 0 || n == 0) return;
    if (alpha == zero) {
        for (mint j = 0; j < n; j++)
            for (mint i = 0; i < m; i++) B[i + j * ldb] = zero;
        return;
    }
    // Columns of B are independent.
#pragma omp parallel for schedule(dynamic, 1) if (m * n > 8192)
    for (mint j = 0; j < n; j++) {
        if (alpha != one)
            for (mint i = 0; i < m; i++) B[i + j * ldb] = alpha * B[i + j * ldb];
        for (mint k = 0; k < m; k++) {
            if (B[k + j * ldb] != zero) {
                B[k + j * ldb] = B[k + j * ldb] / A[k + k * lda];
                for (mint i = k + 1; i < m; i++) B[i + j * ldb] = B[i + j * ldb] - B[k + j * ldb] * A[i + k * lda];
            }
        }
    }
}

// mplapack/Rtrmm.cpp:157-169 (Left, Lower, Transpose, Non-unit):
// B := alpha * A**T * B.
void Rtrmm_LLTN(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, dd_real *B, mint ldb) {
    const dd_real zero = 0.0;
    if (m == 0 || n == 0) return;
    if (alpha == zero) {
        for (mint j = 0; j < n; j++)
            for (mint i = 0; i < m; i++) B[i + j * ldb] = zero;
        return;
    }
    // Columns of B are independent.
#pragma omp parallel for schedule(dynamic, 1) if (m * n > 8192)
    for (mint j = 0; j < n; j++) {
        for (mint i = 0; i < m; i++) {
            dd_real temp = B[i + j * ldb];
            temp = temp * A[i + i * lda];
            for (mint k = i + 1; k < m; k++) temp += A[k + i * lda] * B[k + j * ldb];
            B[i + j * ldb] = alpha * temp;
        }
    }
}

// mplapack/Rtrsv.cpp:134-145 (Lower, No transpose, Non-unit, incx=1).
void Rtrsv_LN(mint n, const dd_real *A, mint lda, dd_real *x) {
    const dd_real zero = 0.0;
    for (mint j = 0; j < n; j++) {
        if (x[j] != zero) {
            x[j] = x[j] / A[j + j * lda];
            dd_real temp = x[j];
            for (mint i = j + 1; i < n; i++) x[i] = x[i] - temp * A[i + j * lda];
        }
    }
}

// mplapack/Rtrsv.cpp:200-211 (Lower, Transpose, Non-unit, incx=1).
void Rtrsv_LT(mint n, const dd_real *A, mint lda, dd_real *x) {
    for (mint j = n - 1; j >= 0; j--) {
        dd_real temp = x[j];
        for (mint i = n - 1; i >= j + 1; i--) temp = temp - A[i + j * lda] * x[i];
        temp = temp / A[j + j * lda];
        x[j] = temp;
    }
}

// mplapack/Rtrmv.cpp:134-145 (Lower, No transpose, Non-unit, incx=1).
void Rtrmv_LN(mint n, const dd_real *A, mint lda, dd_real *x) {
    const dd_real zero = 0.0;
    for (mint j = n - 1; j >= 0; j--) {
        if (x[j] != zero) {
            dd_real temp = x[j];
            for (mint i = n - 1; i >= j + 1; i--) x[i] += temp * A[i + j * lda];
            x[j] = x[j] * A[j + j * lda];
        }
    }
}

// mplapack/Rtrmv.cpp:203-213 (Lower, Transpose, Non-unit, incx=1).
void Rtrmv_LT(mint n, const dd_real *A, mint lda, dd_real *x) {
    for (mint j = 0; j < n; j++) {
        dd_real temp = x[j];
        temp = temp * A[j + j * lda];
        for (mint i = j + 1; i < n; i++) temp += A[i + j * lda] * x[i];
        x[j] = temp;
    }
}

// mplapack/Rgemv.cpp:118-171 (No transpose, incx=incy=1).
void Rgemv_N(mint m, mint n, dd_real alpha, const dd_real *A, mint lda, const dd_real *x, dd_real beta, dd_real *y) {
    const dd_real zero = 0.0, one = 1.0;
    if (m == 0 || n == 0 || ((alpha == zero) && (beta == one))) return;
    if (beta != one) {
        for (mint i = 0; i < m; i++) y[i] = (beta == zero) ? zero : beta * y[i];
    }
    if (alpha == zero) return;
    for (mint j = 0; j < n; j++) {
        dd_real temp = alpha * x[j];
        for (mint i = 0; i < m; i++) y[i] += temp * A[i + j * lda];
    }
}

// mplapack/Rdot_omp.cpp:80-120.  Unit strides: an OpenMP static partition
// of [0,n) over g_dot_threads chunks (q = n/T, the first n%T chunks one
// longer), each summed in order from zero, the partial sums then added in
// chunk order.  Other strides: one sequential sum.
dd_real Rdot(mint n, const dd_real *x, mint incx, const dd_real *y, mint incy) {
    dd_real temp = 0.0;
    if (incx == 1 && incy == 1) {
        mint T = g_dot_threads;
        mint q = n / T, r = n % T;
        for (mint t = 0; t < T; t++) {
            mint cnt = q + (t < r ? 1 : 0);
            mint start = t * q + (t < r ? t : r);
            dd_real templ = 0.0;
            for (mint i = start; i < start + cnt; i++) templ += x[i] * y[i];
            temp += templ;
        }
        return temp;
    }
    mint ix = 0, iy = 0;
    if (incx < 0) ix = (-n + 1) * incx;
    if (incy < 0) iy = (-n + 1) * incy;
    for (mint i = 0; i < n; i++) {
        temp += x[ix] * y[iy];
        ix += incx;
        iy += incy;
    }
    return temp;
}

// mplapack/Raxpy.cpp:38-53 + Raxpy_omp.cpp:79-112.
void Raxpy(mint n, dd_real a, const dd_real *x, mint incx, dd_real *y, mint incy) {
    const dd_real zero = 0.0;
    if (n <= 0) return;
    if (a == zero) return;
    if (incx == 1 && incy == 1) {  // elementwise (Raxpy_omp.cpp:96-101 runs this in parallel)
#pragma omp parallel for schedule(static) if (n > 100000)
        for (mint i = 0; i < n; i++) y[i] += a * x[i];
        return;
    }
    mint ix = 0, iy = 0;
    if (incx < 0) ix = (-n + 1) * incx;
    if (incy < 0) iy = (-n + 1) * incy;
    for (mint i = 0; i < n; i++) {
        y[iy] += a * x[ix];
        ix += incx;
        iy += incy;
    }
}

// mplapack/Rscal.cpp: dx = da * dx elementwise.
void Rscal(mint n, dd_real a, dd_real *x, mint incx) {
    if (n <= 0 || incx <= 0) return;
    for (mint i = 0; i < n; i++) x[i * incx] = a * x[i * incx];
}

// mplapack/Rcopy_omp.cpp.
void Rcopy(mint n, const dd_real *x, mint incx, dd_real *y, mint incy) {
    mint ix = 0, iy = 0;
    if (incx < 0) ix = (-n + 1) * incx;
    if (incy < 0) iy = (-n + 1) * incy;
    for (mint i = 0; i < n; i++) {
        y[iy] = x[ix];
        ix += incx;
        iy += incy;
    }
}

// mplapack/Rswap.cpp.
void Rswap(mint n, dd_real *x, mint incx, dd_real *y, mint incy) {
    mint ix = 0, iy = 0;
    if (incx < 0) ix = (-n + 1) * incx;
    if (incy < 0) iy = (-n + 1) * incy;
    for (mint i = 0; i < n; i++) {
        dd_real t = x[ix];
        x[ix] = y[iy];
        y[iy] = t;
        ix += incx;
        iy += incy;
    }
}

// mplapack/Rpotrf2.cpp:94-146 (Lower): recursive halving, sqrt pivots.
void Rpotrf2_L(mint n, dd_real *a, mint lda, mint &info) {
    info = 0;
    if (n == 0) return;
    const dd_real zero = 0.0, one = 1.0;
    if (n == 1) {
        if (a[0] <= zero || isnan_dd(a[0])) {
            info = 1;
            return;
        }
        a[0] = sqrt(a[0]);
        return;
    }
    mint n1 = n / 2, n2 = n - n1, iinfo = 0;
    Rpotrf2_L(n1, a, lda, iinfo);
    if (iinfo != 0) {
        info = iinfo;
        return;
    }
    Rtrsm_RLTN(n2, n1, one, a, lda, &a[n1], lda);
    Rsyrk_LN(n2, n1, -one, &a[n1], lda, one, &a[n1 + n1 * lda], lda);
    Rpotrf2_L(n2, &a[n1 + n1 * lda], lda, iinfo);
    if (iinfo != 0) info = iinfo + n1;
}

// mplapack/Rpotrf.cpp:75-151 (Lower), block size 64 (iMlaenv.cpp:222-229).
void Rpotrf_L(mint n, dd_real *a, mint lda, mint &info) {
    info = 0;
    if (n == 0) return;
    const mint nb = 64;
    const dd_real one = 1.0;
    if (nb <= 1 || nb >= n) {
        Rpotrf2_L(n, a, lda, info);
        return;
    }
    for (mint j = 1; j <= n; j += nb) {
        mint jb = std::min(nb, n - j + 1);
        Rsyrk_LN(jb, j - 1, -one, &a[j - 1], lda, one, &a[(j - 1) + (j - 1) * lda], lda);
        Rpotrf2_L(jb, &a[(j - 1) + (j - 1) * lda], lda, info);
        if (info != 0) {
            info += j - 1;
            return;
        }
        if (j + jb <= n) {
            Rgemm('N', 'T', n - j - jb + 1, jb, j - 1, -one, &a[(j + jb) - 1], lda, &a[j - 1], lda, one,
                  &a[((j + jb) - 1) + (j - 1) * lda], lda);
            Rtrsm_RLTN(n - j - jb + 1, jb, one, &a[(j - 1) + (j - 1) * lda], lda, &a[((j + jb) - 1) + (j - 1) * lda], lda);
        }
    }
}

// ---- tridiagonal eigen helpers (mplapack/Rlapy2.cpp, Rlartg.cpp, Rlae2.cpp,
// Rlaev2.cpp, Rlasr.cpp, Rlanst.cpp, Rlascl.cpp, Rlaset.cpp) ----------------
static inline dd_real pow2(const dd_real &a) { return a * a; }  // mplapack_utils_dd.h: pow2
static inline dd_real sign(const dd_real &a, const dd_real &b) {
    dd_real t = abs(a);
    if (b < 0.0) t = -t;
    return t;
}
static inline dd_real lamchE() { return dd_real(dd_eps); }
static inline dd_real lamchS() { return dd_real(dd_min_normalized); }

dd_real Rlapy2(dd_real x, dd_real y) {
    bool xn = x != x, yn = y != y;
    dd_real rv = 0.0;
    if (xn) rv = x;
    if (yn) rv = y;
    if (!(xn || yn)) {
        dd_real xa = abs(x), ya = abs(y);
        dd_real w = max(xa, ya), z = min(xa, ya);
        if (z == dd_real(0.0))
            rv = w;
        else
            rv = w * sqrt(dd_real(1.0) + pow2(z / w));
    }
    return rv;
}

// Rlartg computes its scaling threshold as
// 2**castINTEGER(log(safmin/eps)/log(2)/2) with safmin = 2^-969, eps = 2^-104:
// the dd quotient is -432.5 to ~30 digits, its high word is exactly -432.5 and
// the cast truncates to -432, so the threshold is 2^-432 exactly.
void Rlartg(dd_real f, dd_real g, dd_real &cs, dd_real &sn, dd_real &r) {
    const dd_real zero = 0.0, one = 1.0;
    const dd_real safmn2 = npwr(dd_real(2.0), -432);
    const dd_real safmx2 = one / safmn2;
    if (g == zero) {
        cs = one;
        sn = zero;
        r = f;
        return;
    }
    if (f == zero) {
        cs = zero;
        sn = one;
        r = g;
        return;
    }
    dd_real f1 = f, g1 = g;
    dd_real scale = max(abs(f1), abs(g1));
    if (scale >= safmx2) {
        mint count = 0;
        do {
            count++;
            f1 = f1 * safmn2;
            g1 = g1 * safmn2;
            scale = max(abs(f1), abs(g1));
        } while (scale >= safmx2 && count < 20);
        r = sqrt(pow2(f1) + pow2(g1));
        cs = f1 / r;
        sn = g1 / r;
        for (mint i = 1; i <= count; i++) r = r * safmx2;
    } else if (scale <= safmn2) {
        mint count = 0;
        do {
            count++;
            f1 = f1 * safmx2;
            g1 = g1 * safmx2;
            scale = max(abs(f1), abs(g1));
        } while (scale <= safmn2);
        r = sqrt(pow2(f1) + pow2(g1));
        cs = f1 / r;
        sn = g1 / r;
        for (mint i = 1; i <= count; i++) r = r * safmn2;
    } else {
        r = sqrt(pow2(f1) + pow2(g1));
        cs = f1 / r;
        sn = g1 / r;
    }
    if (abs(f) > abs(g) && cs < zero) {
        cs = -cs;
        sn = -sn;
        r = -r;
    }
}

static void lae2_core(const dd_real &a, const dd_real &b, const dd_real &c, dd_real &rt1, dd_real &rt2, dd_real &df,
                      dd_real &rt, dd_real &tb, dd_real &ab, int &sgn1) {
    const dd_real one = 1.0, two = 2.0, zero = 0.0, half = 0.5;
    dd_real sm = a + c;
    df = a - c;
    dd_real adf = abs(df);
    tb = b + b;
    ab = abs(tb);
    dd_real acmx, acmn;
    if (abs(a) > abs(c)) {
        acmx = a;
        acmn = c;
    } else {
        acmx = c;
        acmn = a;
    }
    if (adf > ab)
        rt = adf * sqrt(one + pow2(ab / adf));
    else if (adf < ab)
        rt = ab * sqrt(one + pow2(adf / ab));
    else
        rt = ab * sqrt(two);
    if (sm < zero) {
        rt1 = half * (sm - rt);
        sgn1 = -1;
        rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    } else if (sm > zero) {
        rt1 = half * (sm + rt);
        sgn1 = 1;
        rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    } else {
        rt1 = half * rt;
        rt2 = -half * rt;
        sgn1 = 1;
    }
}

void Rlae2(dd_real a, dd_real b, dd_real c, dd_real &rt1, dd_real &rt2) {
    dd_real df, rt, tb, ab;
    int sgn1;
    lae2_core(a, b, c, rt1, rt2, df, rt, tb, ab, sgn1);
}

void Rlaev2(dd_real a, dd_real b, dd_real c, dd_real &rt1, dd_real &rt2, dd_real &cs1, dd_real &sn1) {
    const dd_real one = 1.0, zero = 0.0;
    dd_real df, rt, tb, ab;
    int sgn1;
    lae2_core(a, b, c, rt1, rt2, df, rt, tb, ab, sgn1);
    int sgn2;
    dd_real cs;
    if (df >= zero) {
        cs = df + rt;
        sgn2 = 1;
    } else {
        cs = df - rt;
        sgn2 = -1;
    }
    dd_real acs = abs(cs);
    if (acs > ab) {
        dd_real ct = -tb / cs;
        sn1 = one / sqrt(one + ct * ct);
        cs1 = ct * sn1;
    } else {
        if (ab == zero) {
            cs1 = one;
            sn1 = zero;
        } else {
            dd_real tn = -cs / tb;
            cs1 = one / sqrt(one + tn * tn);
            sn1 = tn * cs1;
        }
    }
    if (sgn1 == sgn2) {
        dd_real tn = cs1;
        cs1 = -sn1;
        sn1 = tn;
    }
}

// Rlasr with side='R', pivot='V': plane rotations applied to columns j, j+1
// of the m x n matrix a, forward (F) or backward (B) over j.
static void Rlasr_RV(bool forward, mint m, mint n, const dd_real *c, const dd_real *s, dd_real *a, mint lda) {
    const dd_real one = 1.0, zero = 0.0;
    if (m == 0 || n == 0) return;
    for (mint jj = 0; jj < n - 1; jj++) {
        mint j = forward ? jj : (n - 2 - jj);
        dd_real ct = c[j], st = s[j];
        if ((ct != one) || (st != zero)) {
            for (mint i = 0; i < m; i++) {
                dd_real temp = a[i + (j + 1) * lda];
                a[i + (j + 1) * lda] = ct * temp - st * a[i + j * lda];
                a[i + j * lda] = st * temp + ct * a[i + j * lda];
            }
        }
    }
}

static dd_real Rlanst_M(mint n, const dd_real *d, const dd_real *e) {
    if (n <= 0) return dd_real(0.0);
    dd_real anorm = abs(d[n - 1]);
    for (mint i = 0; i < n - 1; i++) {
        dd_real sum = abs(d[i]);
        if (anorm < sum || isnan_dd(sum)) anorm = sum;
        sum = abs(e[i]);
        if (anorm < sum || isnan_dd(sum)) anorm = sum;
    }
    return anorm;
}

// Rlascl type 'G' on an m x 1 column.
static void Rlascl_G(dd_real cfrom, dd_real cto, mint m, dd_real *a) {
    const dd_real zero = 0.0, one = 1.0;
    if (m == 0) return;
    dd_real smlnum = lamchS();
    dd_real bignum = one / smlnum;
    dd_real cfromc = cfrom, ctoc = cto, mul;
    bool done = false;
    while (!done) {
        dd_real cfrom1 = cfromc * smlnum;
        if (cfrom1 == cfromc) {
            mul = ctoc / cfromc;
            done = true;
        } else {
            dd_real cto1 = ctoc / bignum;
            if (cto1 == ctoc) {
                mul = ctoc;
                done = true;
                cfromc = one;
            } else if (abs(cfrom1) > abs(ctoc) && ctoc != zero) {
                mul = smlnum;
                done = false;
                cfromc = cfrom1;
            } else if (abs(cto1) > abs(cfromc)) {
                mul = bignum;
                done = false;
                ctoc = cto1;
            } else {
                mul = ctoc / cfromc;
                done = true;
            }
        }
        for (mint i = 0; i < m; i++) a[i] = a[i] * mul;
    }
}

// mplapack/Rsteqr.cpp:32-503, compz = 'I' or 'N'.  Indices below are the
// reference's 1-based ones; d(i) is d[i-1].
void Rsteqr(char compz, mint n, dd_real *d, dd_real *e, dd_real *z, mint ldz, dd_real *work, mint &info) {
    const dd_real one = 1.0, zero = 0.0, two = 2.0, three = 3.0;
    const mint maxit = 30;
    int icompz = (compz == 'I' || compz == 'i') ? 2 : 0;
    info = 0;
    if (n == 0) return;
    if (n == 1) {
        if (icompz == 2) z[0] = one;
        return;
    }
#define D(i) d[(i) - 1]
#define E(i) e[(i) - 1]
    dd_real eps = lamchE();
    dd_real eps2 = pow2(eps);
    dd_real safmin = lamchS();
    dd_real safmax = one / safmin;
    dd_real ssfmax = sqrt(safmax) / three;
    dd_real ssfmin = sqrt(safmin) / eps2;
    if (icompz == 2) {
        for (mint j = 0; j < n; j++)
            for (mint i = 0; i < n; i++) z[i + j * ldz] = zero;
        for (mint i = 0; i < n; i++) z[i + i * ldz] = one;
    }
    mint nmaxit = n * maxit, jtot = 0, l1 = 1, nm1 = n - 1;
    mint m = 0, l = 0, lsv = 0, lend = 0, lendsv = 0, iscale = 0;
    dd_real anorm, p, rt1, rt2, c, s, g, r, f, b, tst;
    for (;;) {  // statement 10
        if (l1 > n) goto sorting;
        if (l1 > 1) E(l1 - 1) = zero;
        m = n;
        if (l1 <= nm1) {
            for (mint mm = l1; mm <= nm1; mm++) {
                tst = abs(E(mm));
                if (tst == zero) {
                    m = mm;
                    break;
                }
                if (tst <= (sqrt(abs(D(mm))) * sqrt(abs(D(mm + 1)))) * eps) {
                    E(mm) = zero;
                    m = mm;
                    break;
                }
            }
        }
        l = l1;
        lsv = l;
        lend = m;
        lendsv = lend;
        l1 = m + 1;
        if (lend == l) continue;
        anorm = Rlanst_M(lend - l + 1, &D(l), &E(l));
        iscale = 0;
        if (anorm == zero) continue;
        if (anorm > ssfmax) {
            iscale = 1;
            Rlascl_G(anorm, ssfmax, lend - l + 1, &D(l));
            Rlascl_G(anorm, ssfmax, lend - l, &E(l));
        } else if (anorm < ssfmin) {
            iscale = 2;
            Rlascl_G(anorm, ssfmin, lend - l + 1, &D(l));
            Rlascl_G(anorm, ssfmin, lend - l, &E(l));
        }
        if (abs(D(lend)) < abs(D(l))) {
            lend = lsv;
            l = lendsv;
        }
        if (lend > l) {
            // QL iteration: look for a small subdiagonal element (statement 40).
            for (;;) {
                m = lend;
                if (l != lend) {
                    for (mint mm = l; mm <= lend - 1; mm++) {
                        tst = pow2(abs(E(mm)));
                        if (tst <= (eps2 * abs(D(mm))) * abs(D(mm + 1)) + safmin) {
                            m = mm;
                            break;
                        }
                    }
                }
                if (m < lend) E(m) = zero;
                p = D(l);
                if (m == l) {  // statement 80
                    D(l) = p;
                    l++;
                    if (l <= lend) continue;
                    break;
                }
                if (m == l + 1) {
                    if (icompz > 0) {
                        Rlaev2(D(l), E(l), D(l + 1), rt1, rt2, c, s);
                        work[l - 1] = c;
                        work[n - 1 + l - 1] = s;
                        Rlasr_RV(false, n, 2, &work[l - 1], &work[n - 1 + l - 1], &z[(l - 1) * ldz], ldz);
                    } else {
                        Rlae2(D(l), E(l), D(l + 1), rt1, rt2);
                    }
                    D(l) = rt1;
                    D(l + 1) = rt2;
                    E(l) = zero;
                    l += 2;
                    if (l <= lend) continue;
                    break;
                }
                if (jtot == nmaxit) break;
                jtot++;
                g = (D(l + 1) - p) / (two * E(l));
                r = Rlapy2(g, one);
                g = D(m) - p + (E(l) / (g + sign(r, g)));
                s = one;
                c = one;
                p = zero;
                for (mint i = m - 1; i >= l; i--) {
                    f = s * E(i);
                    b = c * E(i);
                    Rlartg(g, f, c, s, r);
                    if (i != m - 1) E(i + 1) = r;
                    g = D(i + 1) - p;
                    r = (D(i) - g) * s + two * c * b;
                    p = s * r;
                    D(i + 1) = g + p;
                    g = c * r - b;
                    if (icompz > 0) {
                        work[i - 1] = c;
                        work[n - 1 + i - 1] = -s;
                    }
                }
                if (icompz > 0) {
                    mint mm = m - l + 1;
                    Rlasr_RV(false, n, mm, &work[l - 1], &work[n - 1 + l - 1], &z[(l - 1) * ldz], ldz);
                }
                D(l) = D(l) - p;
                E(l) = g;
            }
        } else {
            // QR iteration (statement 90).
            for (;;) {
                m = lend;
                if (l != lend) {
                    for (mint mm = l; mm >= lend + 1; mm--) {
                        tst = pow2(abs(E(mm - 1)));
                        if (tst <= (eps2 * abs(D(mm))) * abs(D(mm - 1)) + safmin) {
                            m = mm;
                            break;
                        }
                    }
                }
                if (m > lend) E(m - 1) = zero;
                p = D(l);
                if (m == l) {  // statement 130
                    D(l) = p;
                    l--;
                    if (l >= lend) continue;
                    break;
                }
                if (m == l - 1) {
                    if (icompz > 0) {
                        Rlaev2(D(l - 1), E(l - 1), D(l), rt1, rt2, c, s);
                        work[m - 1] = c;
                        work[n - 1 + m - 1] = s;
                        Rlasr_RV(true, n, 2, &work[m - 1], &work[n - 1 + m - 1], &z[(l - 2) * ldz], ldz);
                    } else {
                        Rlae2(D(l - 1), E(l - 1), D(l), rt1, rt2);
                    }
                    D(l - 1) = rt1;
                    D(l) = rt2;
                    E(l - 1) = zero;
                    l -= 2;
                    if (l >= lend) continue;
                    break;
                }
                if (jtot == nmaxit) break;
                jtot++;
                g = (D(l - 1) - p) / (two * E(l - 1));
                r = Rlapy2(g, one);
                g = D(m) - p + (E(l - 1) / (g + sign(r, g)));
                s = one;
                c = one;
                p = zero;
                mint lm1 = l - 1;
                for (mint i = m; i <= lm1; i++) {
                    f = s * E(i);
                    b = c * E(i);
                    Rlartg(g, f, c, s, r);
                    if (i != m) E(i - 1) = r;
                    g = D(i) - p;
                    r = (D(i + 1) - g) * s + two * c * b;
                    p = s * r;
                    D(i) = g + p;
                    g = c * r - b;
                    if (icompz > 0) {
                        work[i - 1] = c;
                        work[n - 1 + i - 1] = s;
                    }
                }
                if (icompz > 0) {
                    mint mm = l - m + 1;
                    Rlasr_RV(true, n, mm, &work[m - 1], &work[n - 1 + m - 1], &z[(m - 1) * ldz], ldz);
                }
                D(l) = D(l) - p;
                E(lm1) = g;
            }
        }
        // statement 140: undo scaling.
        if (iscale == 1) {
            Rlascl_G(ssfmax, anorm, lendsv - lsv + 1, &D(lsv));
            Rlascl_G(ssfmax, anorm, lendsv - lsv, &E(lsv));
        } else if (iscale == 2) {
            Rlascl_G(ssfmin, anorm, lendsv - lsv + 1, &D(lsv));
            Rlascl_G(ssfmin, anorm, lendsv - lsv, &E(lsv));
        }
        if (jtot < nmaxit) continue;
        for (mint i = 1; i <= n - 1; i++)
            if (E(i) != zero) info++;
        return;
    }
sorting:
    // statement 160: selection sort of eigenvalues, swapping vectors.
    if (icompz == 0) {
        std::abort();  // not reached: the path always asks for vectors
    }
    for (mint ii = 2; ii <= n; ii++) {
        mint i = ii - 1, k = i;
        p = D(i);
        for (mint j = ii; j <= n; j++) {
            if (D(j) < p) {
                k = j;
                p = D(j);
            }
        }
        if (k != i) {
            D(k) = D(i);
            D(i) = p;
            Rswap(n, &z[(i - 1) * ldz], 1, &z[(k - 1) * ldz], 1);
        }
    }
#undef D
#undef E
}

// ---- symmetric eigenvalues (mplapack/Rsyev.cpp with jobz 'N', uplo 'L') ----
// Rlascl's multiplier sequence (mplapack/Rlascl.cpp:100-185): apply(mul) is
// called once per pass.
template <class F>
static void lascl_passes(dd_real cfrom, dd_real cto, F apply) {
    const dd_real zero = 0.0, one = 1.0;
    dd_real smlnum = lamchS();
    dd_real bignum = one / smlnum;
    dd_real cfromc = cfrom, ctoc = cto, mul;
    bool done = false;
    while (!done) {
        dd_real cfrom1 = cfromc * smlnum;
        if (cfrom1 == cfromc) {
            mul = ctoc / cfromc;
            done = true;
        } else {
            dd_real cto1 = ctoc / bignum;
            if (cto1 == ctoc) {
                mul = ctoc;
                done = true;
                cfromc = one;
            } else if (abs(cfrom1) > abs(ctoc) && ctoc != zero) {
                mul = smlnum;
                done = false;
                cfromc = cfrom1;
            } else if (abs(cto1) > abs(cfromc)) {
                mul = bignum;
                done = false;
                ctoc = cto1;
            } else {
                mul = ctoc / cfromc;
                done = true;
            }
        }
        apply(mul);
    }
}

// mplapack/Rnrm2.cpp:32-60 (scaled sum of squares, unit stride)
static dd_real Rnrm2(mint n, const dd_real *x) {
    const dd_real zero = 0.0, one = 1.0;
    if (n < 1) return zero;
    if (n == 1) return abs(x[0]);
    dd_real scale = zero, ssq = one;
    for (mint i = 0; i < n; i++) {
        if (x[i] != zero) {
            dd_real absxi = abs(x[i]);
            if (scale < absxi) {
                ssq = one + ssq * pow2(scale / absxi);
                scale = absxi;
            } else {
                ssq += pow2(absxi / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

// mplapack/Rlarfg.cpp:32-80 (unit stride): Householder reflector for (alpha, x)
static void Rlarfg(mint n, dd_real &alpha, dd_real *x, dd_real &tau) {
    const dd_real zero = 0.0, one = 1.0;
    if (n <= 1) {
        tau = zero;
        return;
    }
    dd_real xnorm = Rnrm2(n - 1, x);
    if (xnorm == zero) {
        tau = zero;
        return;
    }
    dd_real beta = -sign(Rlapy2(alpha, xnorm), alpha);
    dd_real safmin = lamchS() / lamchE();
    mint knt = 0;
    if (abs(beta) < safmin) {
        dd_real rsafmn = one / safmin;
        do {
            knt++;
            Rscal(n - 1, rsafmn, x, 1);
            beta = beta * rsafmn;
            alpha = alpha * rsafmn;
        } while (abs(beta) < safmin && knt < 20);
        xnorm = Rnrm2(n - 1, x);
        beta = -sign(Rlapy2(alpha, xnorm), alpha);
    }
    tau = (beta - alpha) / beta;
    Rscal(n - 1, one / (alpha - beta), x, 1);
    for (mint j = 1; j <= knt; j++) beta = beta * safmin;
    alpha = beta;
}

// mplapack/Rsymv.cpp, uplo 'L', beta 0, unit strides: y = alpha A x
static void Rsymv_L(mint n, dd_real alpha, const dd_real *a, mint lda, const dd_real *x, dd_real *y) {
    const dd_real zero = 0.0;
    for (mint i = 0; i < n; i++) y[i] = zero;
    if (alpha == zero) return;
    for (mint j = 0; j < n; j++) {
        dd_real temp1 = alpha * x[j], temp2 = zero;
        y[j] += temp1 * a[j + j * lda];
        for (mint i = j + 1; i < n; i++) {
            y[i] += temp1 * a[i + j * lda];
            temp2 += a[i + j * lda] * x[i];
        }
        y[j] += alpha * temp2;
    }
}

// mplapack/Rsyr2.cpp, uplo 'L', unit strides: A += alpha (x y^T + y x^T)
static void Rsyr2_L(mint n, dd_real alpha, const dd_real *x, const dd_real *y, dd_real *a, mint lda) {
    const dd_real zero = 0.0;
    if (n == 0 || alpha == zero) return;
    for (mint j = 0; j < n; j++) {
        if (x[j] != zero || y[j] != zero) {
            dd_real temp1 = alpha * y[j], temp2 = alpha * x[j];
            for (mint i = j; i < n; i++) a[i + j * lda] += x[i] * temp1 + y[i] * temp2;
        }
    }
}

// mplapack/Rsytd2.cpp, uplo 'L' (Rsytrd runs it unblocked for n <= 32)
static void Rsytd2_L(mint n, dd_real *a, mint lda, dd_real *d, dd_real *e, dd_real *tau) {
    const dd_real zero = 0.0, one = 1.0, half = 1.0 / 2.0;
    if (n <= 0) return;
    for (mint i = 1; i <= n - 1; i++) {
        dd_real taui;
        dd_real *col = &a[i + (i - 1) * lda];  // A(i+1, i)
        Rlarfg(n - i, col[0], &a[std::min(i + 2, n) - 1 + (i - 1) * lda], taui);
        e[i - 1] = col[0];
        if (taui != zero) {
            col[0] = one;
            dd_real *sub = &a[i + i * lda];  // A(i+1, i+1)
            Rsymv_L(n - i, taui, sub, lda, col, &tau[i - 1]);
            dd_real alpha = -half * taui * Rdot(n - i, &tau[i - 1], 1, col, 1);
            Raxpy(n - i, alpha, col, 1, &tau[i - 1], 1);
            Rsyr2_L(n - i, -one, col, &tau[i - 1], sub, lda);
            col[0] = e[i - 1];
        }
        d[i - 1] = a[(i - 1) + (i - 1) * lda];
        tau[i - 1] = taui;
    }
    d[n - 1] = a[(n - 1) + (n - 1) * lda];
}

// mplapack/Rsterf.cpp: eigenvalues of a symmetric tridiagonal by the
// Pal-Walker-Kahan root-free QL/QR.  Indices are the reference's 1-based
// ones; d(i) is d[i-1].  The final ascending sort (Rlasrt 'I') is done with
// std::sort: it only permutes values.
void Rsterf(mint n, dd_real *dv, dd_real *ev, mint &info) {
#define D(i) dv[(i) - 1]
#define E(i) ev[(i) - 1]
    const dd_real zero = 0.0, one = 1.0, two = 2.0, three = 3.0;
    const mint maxit = 30;
    info = 0;
    if (n <= 1) return;
    dd_real eps = lamchE();
    dd_real eps2 = pow2(eps);
    dd_real safmin = lamchS();
    dd_real safmax = one / safmin;
    dd_real ssfmax = sqrt(safmax) / three;
    dd_real ssfmin = sqrt(safmin) / eps2;
    mint nmaxit = n * maxit;
    dd_real sigma = zero;
    mint jtot = 0;
    mint l1 = 1;
    mint m = 0, l = 0, lsv = 0, lend = 0, lendsv = 0, iscale = 0;
    dd_real anorm, p, rte, rt1, rt2, r, c, s, gamma, bb, oldc, oldgam, alpha;
    for (;;) {
        // statement 10
        if (l1 > n) {
            std::sort(dv, dv + n);
            return;
        }
        if (l1 > 1) E(l1 - 1) = zero;
        for (m = l1; m <= n - 1; m++) {
            if (abs(E(m)) <= (sqrt(abs(D(m))) * sqrt(abs(D(m + 1)))) * eps) {
                E(m) = zero;
                goto found30;
            }
        }
        m = n;
    found30:
        l = l1;
        lsv = l;
        lend = m;
        lendsv = lend;
        l1 = m + 1;
        if (lend == l) continue;
        anorm = Rlanst_M(lend - l + 1, &D(l), &E(l));
        iscale = 0;
        if (anorm == zero) continue;
        if (anorm > ssfmax) {
            iscale = 1;
            Rlascl_G(anorm, ssfmax, lend - l + 1, &D(l));
            Rlascl_G(anorm, ssfmax, lend - l, &E(l));
        } else if (anorm < ssfmin) {
            iscale = 2;
            Rlascl_G(anorm, ssfmin, lend - l + 1, &D(l));
            Rlascl_G(anorm, ssfmin, lend - l, &E(l));
        }
        for (mint i = l; i <= lend - 1; i++) E(i) = pow2(E(i));
        if (abs(D(lend)) < abs(D(l))) {
            lend = lsv;
            l = lendsv;
        }
        if (lend >= l) {
            // QL iteration
            for (;;) {
                if (l != lend) {
                    for (m = l; m <= lend - 1; m++)
                        if (abs(E(m)) <= eps2 * abs(D(m) * D(m + 1))) goto found70;
                }
                m = lend;
            found70:
                if (m < lend) E(m) = zero;
                p = D(l);
                if (m == l) {
                    D(l) = p;
                    l++;
                    if (l <= lend) continue;
                    break;
                }
                if (m == l + 1) {
                    rte = sqrt(E(l));
                    Rlae2(D(l), rte, D(l + 1), rt1, rt2);
                    D(l) = rt1;
                    D(l + 1) = rt2;
                    E(l) = zero;
                    l += 2;
                    if (l <= lend) continue;
                    break;
                }
                if (jtot == nmaxit) break;
                jtot++;
                rte = sqrt(E(l));
                sigma = (D(l + 1) - p) / (two * rte);
                r = Rlapy2(sigma, one);
                sigma = p - (rte / (sigma + sign(r, sigma)));
                c = one;
                s = zero;
                gamma = D(m) - sigma;
                p = gamma * gamma;
                for (mint i = m - 1; i >= l; i--) {
                    bb = E(i);
                    r = p + bb;
                    if (i != m - 1) E(i + 1) = s * r;
                    oldc = c;
                    c = p / r;
                    s = bb / r;
                    oldgam = gamma;
                    alpha = D(i);
                    gamma = c * (alpha - sigma) - s * oldgam;
                    D(i + 1) = oldgam + (alpha - gamma);
                    if (c != zero)
                        p = (gamma * gamma) / c;
                    else
                        p = oldc * bb;
                }
                E(l) = s * p;
                D(l) = sigma + gamma;
            }
        } else {
            // QR iteration
            for (;;) {
                for (m = l; m >= lend + 1; m--)
                    if (abs(E(m - 1)) <= eps2 * abs(D(m) * D(m - 1))) goto found120;
                m = lend;
            found120:
                if (m > lend) E(m - 1) = zero;
                p = D(l);
                if (m == l) {
                    D(l) = p;
                    l--;
                    if (l >= lend) continue;
                    break;
                }
                if (m == l - 1) {
                    rte = sqrt(E(l - 1));
                    Rlae2(D(l), rte, D(l - 1), rt1, rt2);
                    D(l) = rt1;
                    D(l - 1) = rt2;
                    E(l - 1) = zero;
                    l -= 2;
                    if (l >= lend) continue;
                    break;
                }
                if (jtot == nmaxit) break;
                jtot++;
                rte = sqrt(E(l - 1));
                sigma = (D(l - 1) - p) / (two * rte);
                r = Rlapy2(sigma, one);
                sigma = p - (rte / (sigma + sign(r, sigma)));
                c = one;
                s = zero;
                gamma = D(m) - sigma;
                p = gamma * gamma;
                for (mint i = m; i <= l - 1; i++) {
                    bb = E(i);
                    r = p + bb;
                    if (i != m) E(i - 1) = s * r;
                    oldc = c;
                    c = p / r;
                    s = bb / r;
                    oldgam = gamma;
                    alpha = D(i + 1);
                    gamma = c * (alpha - sigma) - s * oldgam;
                    D(i) = oldgam + (alpha - gamma);
                    if (c != zero)
                        p = (gamma * gamma) / c;
                    else
                        p = oldc * bb;
                }
                E(l - 1) = s * p;
                D(l) = sigma + gamma;
            }
        }
        // statement 150
        if (iscale == 1) Rlascl_G(ssfmax, anorm, lendsv - lsv + 1, &D(lsv));
        if (iscale == 2) Rlascl_G(ssfmin, anorm, lendsv - lsv + 1, &D(lsv));
        if (jtot >= nmaxit) {
            for (mint i = 1; i <= n - 1; i++)
                if (E(i) != zero) info++;
            return;
        }
    }
#undef D
#undef E
}

// mplapack/Rsyev.cpp with jobz 'N', uplo 'L': eigenvalues of the symmetric
// matrix whose lower triangle is in a (overwritten) into w, ascending.
void Rsyev_NL(mint n, dd_real *a, mint lda, dd_real *w, mint &info) {
    const dd_real zero = 0.0, one = 1.0;
    info = 0;
    if (n == 0) return;
    if (n == 1) {
        w[0] = a[0];
        return;
    }
    dd_real safmin = lamchS();
    dd_real eps = lamchE();
    dd_real smlnum = safmin / eps;
    dd_real bignum = one / smlnum;
    dd_real rmin = sqrt(smlnum);
    dd_real rmax = sqrt(bignum);
    // Rlansy 'M' on the lower triangle
    dd_real anrm = zero;
    for (mint j = 0; j < n; j++)
        for (mint i = j; i < n; i++) {
            dd_real sum = abs(a[i + j * lda]);
            if (anrm < sum || isnan_dd(sum)) anrm = sum;
        }
    bool iscale = false;
    dd_real sigma = zero;
    if (anrm > zero && anrm < rmin) {
        iscale = true;
        sigma = rmin / anrm;
    } else if (anrm > rmax) {
        iscale = true;
        sigma = rmax / anrm;
    }
    if (iscale)  // Rlascl type 'L' from one to sigma
        lascl_passes(one, sigma, [&](const dd_real &mul) {
            for (mint j = 0; j < n; j++)
                for (mint i = j; i < n; i++) a[i + j * lda] = a[i + j * lda] * mul;
        });
    std::vector<dd_real> e(n), tau(n);
    Rsytd2_L(n, a, lda, w, e.data(), tau.data());
    Rsterf(n, w, e.data(), info);
    if (iscale) {
        mint imax = (info == 0) ? n : info - 1;
        Rscal(imax, one / sigma, w, 1);
    }
}

}  // namespace orc

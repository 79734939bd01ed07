// Small-block eigenvalues and LP-entry kernels; see smallblk.hpp.
#include "smallblk.hpp"

#include "ddr.hpp"

namespace sdk {

namespace {

constexpr int NMAX = SMALL_EIG_MAXN;
// Rlamch('E') = 2^-104 and Rlamch('S') = 2^-969 of the dd build
// (mplapack/Rlamch.cpp:36-47)
constexpr double kEps = 4.93038065763132e-32;
constexpr double kSafmin = 2.0041683600089728e-292;

__device__ ddr pow2(const ddr &a) { return a * a; }  // mplapack_utils_dd.h: pow2
__device__ ddr sign(const ddr &a, const ddr &b) {
    ddr t = abs(a);
    if (b < 0.0) t = -t;
    return t;
}
__device__ bool isnan_r(const ddr &a) { return a.v.hi != a.v.hi || a.v.lo != a.v.lo; }

// mplapack/Rlapy2.cpp
__device__ ddr lapy2(const ddr &x, const ddr &y) {
    bool xn = isnan_r(x), yn = isnan_r(y);
    ddr rv = 0.0;
    if (xn) rv = x;
    if (yn) rv = y;
    if (!(xn || yn)) {
        ddr xa = abs(x), ya = abs(y);
        ddr w = max(xa, ya), z = min(xa, ya);
        rv = (z == ddr(0.0)) ? w : w * sqrt(ddr(1.0) + pow2(z / w));
    }
    return rv;
}

// mplapack/Rlae2.cpp: eigenvalues rt1 (larger |.|) and rt2 of [[a, b], [b, c]]
__device__ void lae2(const ddr &a, const ddr &b, const ddr &c, ddr &rt1, ddr &rt2) {
    const ddr one = 1.0, two = 2.0, zero = 0.0, half = 0.5;
    ddr sm = a + c;
    ddr df = a - c;
    ddr adf = abs(df);
    ddr tb = b + b;
    ddr ab = abs(tb);
    ddr acmx, acmn;
    if (abs(a) > abs(c)) {
        acmx = a;
        acmn = c;
    } else {
        acmx = c;
        acmn = a;
    }
    ddr rt;
    if (adf > ab)
        rt = adf * sqrt(one + pow2(ab / adf));
    else if (adf < ab)
        rt = ab * sqrt(one + pow2(adf / ab));
    else
        rt = ab * sqrt(two);
    if (sm < zero) {
        rt1 = half * (sm - rt);
        rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    } else if (sm > zero) {
        rt1 = half * (sm + rt);
        rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
    } else {
        rt1 = half * rt;
        rt2 = -half * rt;
    }
}

// mplapack/Rlanst.cpp, norm 'M'
__device__ ddr lanst_m(int n, const ddr *d, const ddr *e) {
    if (n <= 0) return ddr(0.0);
    ddr anorm = abs(d[n - 1]);
    for (int i = 0; i < n - 1; i++) {
        ddr sum = abs(d[i]);
        if (anorm < sum || isnan_r(sum)) anorm = sum;
        sum = abs(e[i]);
        if (anorm < sum || isnan_r(sum)) anorm = sum;
    }
    return anorm;
}

// mplapack/Rlascl.cpp's multiplier passes from cfrom to cto, applied to the
// n-vector a (type 'G', one column).
__device__ void lascl_g(ddr cfrom, ddr cto, int n, ddr *a) {
    const ddr zero = 0.0, one = 1.0;
    ddr smlnum = kSafmin;
    ddr bignum = one / smlnum;
    ddr cfromc = cfrom, ctoc = cto, mul;
    bool done = false;
    while (!done) {
        ddr cfrom1 = cfromc * smlnum;
        if (cfrom1 == cfromc) {
            mul = ctoc / cfromc;
            done = true;
        } else {
            ddr cto1 = ctoc / bignum;
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
        for (int i = 0; i < n; i++) a[i] = a[i] * mul;
    }
}

// mplapack/Rsterf.cpp (1-based indices as there; d(i) is dv[i-1]).  Returns
// info; on success the minimum eigenvalue is the minimum of dv (the
// reference's final Rlasrt only orders the values).
__device__ int sterf(int n, ddr *dv, ddr *ev) {
#define D(i) dv[(i) - 1]
#define E(i) ev[(i) - 1]
    const ddr zero = 0.0, one = 1.0, two = 2.0, three = 3.0;
    const int maxit = 30;
    if (n <= 1) return 0;
    ddr eps = kEps;
    ddr eps2 = pow2(eps);
    ddr safmin = kSafmin;
    ddr safmax = one / safmin;
    ddr ssfmax = sqrt(safmax) / three;
    ddr ssfmin = sqrt(safmin) / eps2;
    const int nmaxit = n * maxit;
    ddr sigma = zero;
    int jtot = 0, l1 = 1, m = 0, l = 0, lsv = 0, lend = 0, lendsv = 0, iscale = 0;
    ddr anorm, p, rte, rt1, rt2, r, c, s, gamma, bb, oldc, oldgam, alpha;
    for (;;) {
        if (l1 > n) return 0;
        if (l1 > 1) E(l1 - 1) = zero;
        for (m = l1; m <= n - 1; m++)
            if (abs(E(m)) <= (sqrt(abs(D(m))) * sqrt(abs(D(m + 1)))) * eps) {
                E(m) = zero;
                goto split_found;
            }
        m = n;
    split_found:
        l = l1;
        lsv = l;
        lend = m;
        lendsv = lend;
        l1 = m + 1;
        if (lend == l) continue;
        anorm = lanst_m(lend - l + 1, &D(l), &E(l));
        iscale = 0;
        if (anorm == zero) continue;
        if (anorm > ssfmax) {
            iscale = 1;
            lascl_g(anorm, ssfmax, lend - l + 1, &D(l));
            lascl_g(anorm, ssfmax, lend - l, &E(l));
        } else if (anorm < ssfmin) {
            iscale = 2;
            lascl_g(anorm, ssfmin, lend - l + 1, &D(l));
            lascl_g(anorm, ssfmin, lend - l, &E(l));
        }
        for (int i = l; i <= lend - 1; i++) E(i) = pow2(E(i));
        if (abs(D(lend)) < abs(D(l))) {
            lend = lsv;
            l = lendsv;
        }
        if (lend >= l) {
            for (;;) {  // QL
                if (l != lend) {
                    for (m = l; m <= lend - 1; m++)
                        if (abs(E(m)) <= eps2 * abs(D(m) * D(m + 1))) goto ql_found;
                }
                m = lend;
            ql_found:
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
                    lae2(D(l), rte, D(l + 1), rt1, rt2);
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
                r = lapy2(sigma, one);
                sigma = p - (rte / (sigma + sign(r, sigma)));
                c = one;
                s = zero;
                gamma = D(m) - sigma;
                p = gamma * gamma;
                for (int i = m - 1; i >= l; i--) {
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
            for (;;) {  // QR
                for (m = l; m >= lend + 1; m--)
                    if (abs(E(m - 1)) <= eps2 * abs(D(m) * D(m - 1))) goto qr_found;
                m = lend;
            qr_found:
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
                    lae2(D(l), rte, D(l - 1), rt1, rt2);
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
                r = lapy2(sigma, one);
                sigma = p - (rte / (sigma + sign(r, sigma)));
                c = one;
                s = zero;
                gamma = D(m) - sigma;
                p = gamma * gamma;
                for (int i = m; i <= l - 1; i++) {
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
        if (iscale == 1) lascl_g(ssfmax, anorm, lendsv - lsv + 1, &D(lsv));
        if (iscale == 2) lascl_g(ssfmin, anorm, lendsv - lsv + 1, &D(lsv));
        if (jtot >= nmaxit) {
            int info = 0;
            for (int i = 1; i <= n - 1; i++)
                if (E(i) != zero) info++;
            return info;
        }
    }
#undef D
#undef E
}

// mplapack/Rnrm2.cpp (unit stride)
__device__ ddr nrm2(int n, const ddr *x) {
    const ddr zero = 0.0, one = 1.0;
    if (n < 1) return zero;
    if (n == 1) return abs(x[0]);
    ddr scale = zero, ssq = one;
    for (int i = 0; i < n; i++) {
        if (x[i] != zero) {
            ddr absxi = abs(x[i]);
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

// mplapack/Rlarfg.cpp (unit stride): reflector for (alpha, x[0 .. n-2])
__device__ void larfg(int n, ddr &alpha, ddr *x, ddr &tau) {
    const ddr zero = 0.0, one = 1.0;
    if (n <= 1) {
        tau = zero;
        return;
    }
    ddr xnorm = nrm2(n - 1, x);
    if (xnorm == zero) {
        tau = zero;
        return;
    }
    ddr beta = -sign(lapy2(alpha, xnorm), alpha);
    ddr safmin = ddr(kSafmin) / ddr(kEps);
    int knt = 0;
    if (abs(beta) < safmin) {
        ddr rsafmn = one / safmin;
        do {
            knt++;
            for (int i = 0; i < n - 1; i++) x[i] = rsafmn * x[i];
            beta = beta * rsafmn;
            alpha = alpha * rsafmn;
        } while (abs(beta) < safmin && knt < 20);
        xnorm = nrm2(n - 1, x);
        beta = -sign(lapy2(alpha, xnorm), alpha);
    }
    tau = (beta - alpha) / beta;
    ddr sc = one / (alpha - beta);
    for (int i = 0; i < n - 1; i++) x[i] = sc * x[i];
    for (int j = 1; j <= knt; j++) beta = beta * safmin;
    alpha = beta;
}

// Rdot with the T-way partition (mplapack/Rdot_omp.cpp:80-120), unit strides
__device__ ddr dot_part(int n, const ddr *x, const ddr *y, int T) {
    ddr temp = 0.0;
    int q = n / T, r = n % T;
    for (int t = 0; t < T; t++) {
        int cnt = q + (t < r ? 1 : 0);
        int start = t * q + (t < r ? t : r);
        ddr templ = 0.0;
        for (int i = start; i < start + cnt; i++) templ += x[i] * y[i];
        temp += templ;
    }
    return temp;
}

// One wave per task.  a: the n x n work matrix in LDS (column-major, lda n).
__global__ __launch_bounds__(64) void k_small_min_eig(const SmallEigTask *tasks, int T, dd *out, int *info_out) {
    // LDS holds plain dd storage (a __shared__ object cannot have a
    // constructor); ddr is a layout-identical view of it.
    __shared__ dd lds[4 * NMAX * NMAX + 3 * NMAX + 3];
    __shared__ int s_scale;
    ddr *Lm = reinterpret_cast<ddr *>(lds), *Mm = Lm + NMAX * NMAX, *Tm = Mm + NMAX * NMAX, *a = Tm + NMAX * NMAX;
    ddr *dv = a + NMAX * NMAX, *ev = dv + NMAX, *tau = ev + NMAX;
    ddr &s_alpha = tau[NMAX], &s_taui = tau[NMAX + 1], &s_sigma = tau[NMAX + 2];
    const SmallEigTask tk = tasks[blockIdx.x];
    const int n = tk.n, lane = threadIdx.x, nsq = n * n;
    const ddr zero = 0.0, one = 1.0;
    for (int q = lane; q < nsq; q += 64) {
        Lm[q] = ddr(tk.L[q]);
        Mm[q] = ddr(tk.M[q]);
    }
    __syncthreads();
    // DLS2 = M L^T: Rgemm 'N','T' with alpha 1, beta 0 (Rgemm_NT_omp order:
    // column j, then l, then c_ij += (alpha l_jl) m_il)
    for (int q = lane; q < nsq; q += 64) {
        const int i = q % n, j = q / n;
        ddr c = zero;
        for (int l = 0; l < n; l++) {
            ddr temp = one * Lm[j + l * n];
            c += temp * Mm[i + l * n];
        }
        Tm[q] = c;
    }
    __syncthreads();
    // DLS1 = L DLS2: Rgemm 'N','N' (c_ij += (alpha t_lj) l_il)
    for (int q = lane; q < nsq; q += 64) {
        const int i = q % n, j = q / n;
        ddr c = zero;
        for (int l = 0; l < n; l++) {
            ddr temp = one * Tm[l + j * n];
            c += temp * Lm[i + l * n];
        }
        a[q] = c;
    }
    __syncthreads();
    if (n == 1) {
        if (lane == 0) {
            out[blockIdx.x] = a[0].v;
            info_out[blockIdx.x] = 0;
        }
        return;
    }
    // Rsyev: scale to [rmin, rmax] when outside (Rlansy 'M' on the lower
    // triangle, Rlascl type 'L')
    if (lane == 0) {
        ddr safmin = kSafmin, eps = kEps;
        ddr smlnum = safmin / eps;
        ddr bignum = one / smlnum;
        ddr rmin = sqrt(smlnum), rmax = sqrt(bignum);
        ddr anrm = zero;
        for (int j = 0; j < n; j++)
            for (int i = j; i < n; i++) {
                ddr sum = abs(a[i + j * n]);
                if (anrm < sum || isnan_r(sum)) anrm = sum;
            }
        s_scale = 0;
        s_sigma = zero;
        if (anrm > zero && anrm < rmin) {
            s_scale = 1;
            s_sigma = rmin / anrm;
        } else if (anrm > rmax) {
            s_scale = 1;
            s_sigma = rmax / anrm;
        }
        if (s_scale) {
            // Rlascl passes from one to sigma over the lower triangle
            ddr cfromc = one, ctoc = s_sigma, mul;
            ddr bn = one / ddr(kSafmin), sm = kSafmin;
            bool done = false;
            while (!done) {
                ddr cfrom1 = cfromc * sm;
                if (cfrom1 == cfromc) {
                    mul = ctoc / cfromc;
                    done = true;
                } else {
                    ddr cto1 = ctoc / bn;
                    if (cto1 == ctoc) {
                        mul = ctoc;
                        done = true;
                        cfromc = one;
                    } else if (abs(cfrom1) > abs(ctoc) && ctoc != zero) {
                        mul = sm;
                        done = false;
                        cfromc = cfrom1;
                    } else if (abs(cto1) > abs(cfromc)) {
                        mul = bn;
                        done = false;
                        ctoc = cto1;
                    } else {
                        mul = ctoc / cfromc;
                        done = true;
                    }
                }
                for (int j = 0; j < n; j++)
                    for (int i = j; i < n; i++) a[i + j * n] = a[i + j * n] * mul;
            }
        }
    }
    __syncthreads();
    // Rsytd2, uplo 'L' (1-based i as in the reference)
    const ddr half = 1.0 / 2.0;
    for (int i = 1; i <= n - 1; i++) {
        const int len = n - i;               // length of the reflector
        ddr *col = &a[i + (i - 1) * n];      // A(i+1, i): v, with v[0] the pivot
        ddr *sub = &a[i + i * n];            // A(i+1, i+1), lda n
        ddr *w = &tau[i - 1];                // Rsymv's y
        if (lane == 0) {
            ddr taui;
            larfg(len, col[0], &a[(i + 2 <= n ? i + 2 : n) - 1 + (i - 1) * n], taui);
            ev[i - 1] = col[0];
            s_taui = taui;
            if (taui != zero) col[0] = one;
        }
        __syncthreads();
        const ddr taui = s_taui;
        if (taui != zero) {
            // Rsymv 'L' (beta 0): lane r forms y(r) in the reference's order
            if (lane < len) {
                const int r = lane;
                ddr y = zero;
                for (int j = 0; j < r; j++) {
                    ddr temp1 = taui * col[j];
                    y += temp1 * sub[r + j * n];
                }
                ddr temp1 = taui * col[r];
                y += temp1 * sub[r + r * n];
                ddr temp2 = zero;
                for (int k = r + 1; k < len; k++) temp2 += sub[k + r * n] * col[k];
                y += taui * temp2;
                w[r] = y;
            }
            __syncthreads();
            if (lane == 0) s_alpha = -half * taui * dot_part(len, w, col, T);
            __syncthreads();
            const ddr alpha = s_alpha;
            if (lane < len) w[lane] += alpha * col[lane];  // Raxpy
            __syncthreads();
            // Rsyr2 'L' with alpha -1: lane r updates row r of the trailing block
            if (lane < len) {
                const int r = lane;
                const ddr mone = -one;
                for (int j = 0; j <= r; j++) {
                    if (col[j] != zero || w[j] != zero) {
                        ddr temp1 = mone * w[j], temp2 = mone * col[j];
                        sub[r + j * n] += col[r] * temp1 + w[r] * temp2;
                    }
                }
            }
            __syncthreads();
            if (lane == 0) col[0] = ev[i - 1];
        }
        if (lane == 0) {
            dv[i - 1] = a[(i - 1) + (i - 1) * n];
            tau[i - 1] = taui;
        }
        __syncthreads();
    }
    if (lane == 0) {
        dv[n - 1] = a[(n - 1) + (n - 1) * n];
        int info = sterf(n, dv, ev);
        ddr mn = dv[0];
        if (info == 0) {
            for (int i = 1; i < n; i++)
                if (dv[i] < mn) mn = dv[i];
            if (s_scale) mn = (one / s_sigma) * mn;  // Rscal by 1/sigma
        }
        out[blockIdx.x] = mn.v;
        info_out[blockIdx.x] = info;
    }
}

__global__ void k_lp_plus(idx_t n, const dd *a, const dd *b, dd sc, dd *ret) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ret[i] = dd_add(a[i], dd_mul(b[i], sc));
}
__global__ void k_lp_triple(idx_t n, const dd *a, const dd *b, const dd *c, dd *ret) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ret[i] = dd_mul(dd_mul(a[i], b[i]), c[i]);
}
__global__ void k_lp_prod(idx_t n, const dd *a, const dd *b, dd *out) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dd_mul(a[i], b[i]);
}
__global__ void k_lp_eig(idx_t n, const dd *m, const dd *l, dd *out) {
    idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dd_mul(dd_mul(m[i], l[i]), l[i]);
}
// One workgroup: a flag per lane, OR-reduced into *info.
__global__ __launch_bounds__(256) void k_lp_inv_chol(idx_t n, const dd *a, dd *invChol, dd *inv, int64_t *info) {
    int bad = 0;
    for (idx_t i = threadIdx.x; i < n; i += 256) {
        const dd x = a[i];
        if (dd_lt(x, mk(0.0))) bad = 1;
        invChol[i] = dd_div(mk(1.0), dd_sqrt(x));
        if (inv) inv[i] = dd_div(mk(1.0), x);
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) *info = bad;
}

inline unsigned blocks(idx_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void small_min_eig(hipStream_t s, const SmallEigTask *d_tasks, int ntasks, int T, dd *out, int *info) {
    if (ntasks > 0) k_small_min_eig<<<ntasks, 64, 0, s>>>(d_tasks, T, out, info);
}
void lp_plus(hipStream_t s, idx_t n, const dd *a, const dd *b, dd sc, dd *ret) {
    if (n > 0) k_lp_plus<<<blocks(n), 256, 0, s>>>(n, a, b, sc, ret);
}
void lp_triple(hipStream_t s, idx_t n, const dd *a, const dd *b, const dd *c, dd *ret) {
    if (n > 0) k_lp_triple<<<blocks(n), 256, 0, s>>>(n, a, b, c, ret);
}
void lp_prod(hipStream_t s, idx_t n, const dd *a, const dd *b, dd *out) {
    if (n > 0) k_lp_prod<<<blocks(n), 256, 0, s>>>(n, a, b, out);
}
void lp_eig(hipStream_t s, idx_t n, const dd *m, const dd *l, dd *out) {
    if (n > 0) k_lp_eig<<<blocks(n), 256, 0, s>>>(n, m, l, out);
}
void lp_inv_chol(hipStream_t s, idx_t n, const dd *a, dd *invChol, dd *inv, int64_t *info) {
    k_lp_inv_chol<<<1, 256, 0, s>>>(n, a, invChol, inv, info);
}

}  // namespace sdk

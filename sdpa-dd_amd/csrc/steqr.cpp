// Tridiagonal QL/QR eigen-solver in dd; see steqr.hpp.
// Follows mplapack/Rsteqr.cpp:32-503 and its helpers Rlaev2, Rlae2, Rlapy2,
// Rlartg, Rlasr ('R','V',*), Rlanst('M'), Rlascl('G') operation for
// operation, with 1-based indices where the reference uses them.
#include "steqr.hpp"

#include <cmath>

namespace sdp {

namespace {
const double kEps = 4.93038065763132e-32;             // 2^-104, Rlamch("E")
const double kSafmin = 2.0041683600089728e-292;       // 2^-969, Rlamch("S")

ddr pow2(const ddr &a) { return a * a; }
ddr sign(const ddr &a, const ddr &b) {
    ddr t = abs(a);
    if (b < 0.0) t = -t;
    return t;
}
bool isnan_d(const ddr &a) { return a.v.hi != a.v.hi || a.v.lo != a.v.lo; }

ddr lapy2(const ddr &x, const ddr &y) {
    bool xn = isnan_d(x), yn = isnan_d(y);
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

// Rlartg: the scaling threshold 2^castINTEGER(log(safmin/eps)/log(2)/2)
// equals 2^-432 (the dd quotient's high word is -432.5, truncated).
void lartg(const ddr &f, const ddr &g, ddr &cs, ddr &sn, ddr &r) {
    const ddr zero = 0.0, one = 1.0;
    const ddr safmn2 = ddr(std::ldexp(1.0, -432));
    const ddr safmx2 = one / safmn2;
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
    ddr f1 = f, g1 = g;
    ddr scale = max(abs(f1), abs(g1));
    if (scale >= safmx2) {
        int count = 0;
        do {
            count++;
            f1 = f1 * safmn2;
            g1 = g1 * safmn2;
            scale = max(abs(f1), abs(g1));
        } while (scale >= safmx2 && count < 20);
        r = sqrt(pow2(f1) + pow2(g1));
        cs = f1 / r;
        sn = g1 / r;
        for (int i = 1; i <= count; i++) r = r * safmx2;
    } else if (scale <= safmn2) {
        int count = 0;
        do {
            count++;
            f1 = f1 * safmx2;
            g1 = g1 * safmx2;
            scale = max(abs(f1), abs(g1));
        } while (scale <= safmn2);
        r = sqrt(pow2(f1) + pow2(g1));
        cs = f1 / r;
        sn = g1 / r;
        for (int i = 1; i <= count; i++) r = r * safmn2;
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

// Shared part of Rlae2/Rlaev2: eigenvalues of [[a,b],[b,c]].
void eig2(const ddr &a, const ddr &b, const ddr &c, ddr &rt1, ddr &rt2, ddr &df, ddr &rt, ddr &tb, ddr &ab, int &sgn1) {
    const ddr one = 1.0, two = 2.0, zero = 0.0, half = 0.5;
    ddr sm = a + c;
    df = a - c;
    ddr adf = abs(df);
    tb = b + b;
    ab = abs(tb);
    ddr acmx, acmn;
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
void laev2(const ddr &a, const ddr &b, const ddr &c, ddr &rt1, ddr &rt2, ddr &cs1, ddr &sn1) {
    const ddr one = 1.0, zero = 0.0;
    ddr df, rt, tb, ab;
    int sgn1, sgn2;
    eig2(a, b, c, rt1, rt2, df, rt, tb, ab, sgn1);
    ddr cs;
    if (df >= zero) {
        cs = df + rt;
        sgn2 = 1;
    } else {
        cs = df - rt;
        sgn2 = -1;
    }
    if (abs(cs) > ab) {
        ddr ct = -tb / cs;
        sn1 = one / sqrt(one + ct * ct);
        cs1 = ct * sn1;
    } else if (ab == zero) {
        cs1 = one;
        sn1 = zero;
    } else {
        ddr tn = -cs / tb;
        cs1 = one / sqrt(one + tn * tn);
        sn1 = tn * cs1;
    }
    if (sgn1 == sgn2) {
        ddr tn = cs1;
        cs1 = -sn1;
        sn1 = tn;
    }
}
// Rlasr('R','V', forward or backward) on columns of z, rows [r0, r1).
void lasr_rv(bool forward, int64_t ncols, const ddr *c, const ddr *s, ddr *a, int64_t lda, int64_t r0, int64_t r1) {
    const ddr one = 1.0, zero = 0.0;
    for (int64_t jj = 0; jj < ncols - 1; jj++) {
        int64_t j = forward ? jj : (ncols - 2 - jj);
        ddr ct = c[j], st = s[j];
        if ((ct != one) || (st != zero)) {
            for (int64_t i = r0; i < r1; i++) {
                ddr temp = a[i + (j + 1) * lda];
                a[i + (j + 1) * lda] = ct * temp - st * a[i + j * lda];
                a[i + j * lda] = st * temp + ct * a[i + j * lda];
            }
        }
    }
}
ddr lanst_m(int64_t n, const ddr *d, const ddr *e) {
    if (n <= 0) return ddr(0.0);
    ddr anorm = abs(d[n - 1]);
    for (int64_t i = 0; i < n - 1; i++) {
        ddr sum = abs(d[i]);
        if (anorm < sum || isnan_d(sum)) anorm = sum;
        sum = abs(e[i]);
        if (anorm < sum || isnan_d(sum)) anorm = sum;
    }
    return anorm;
}
void lascl_g(const ddr &cfrom, const ddr &cto, int64_t m, ddr *a) {
    const ddr zero = 0.0, one = 1.0;
    if (m == 0) return;
    ddr smlnum = kSafmin, bignum = one / smlnum;
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
                cfromc = cfrom1;
            } else if (abs(cto1) > abs(cfromc)) {
                mul = bignum;
                ctoc = cto1;
            } else {
                mul = ctoc / cfromc;
                done = true;
            }
        }
        for (int64_t i = 0; i < m; i++) a[i] = a[i] * mul;
    }
}
}  // namespace

int64_t steqr_rows(int64_t n, ddr *d, ddr *e, ddr *z, int64_t ldz, ddr *work, int64_t r0, int64_t r1) {
    const ddr one = 1.0, zero = 0.0, two = 2.0, three = 3.0;
    int64_t info = 0;
    if (n == 0) return 0;
    if (r1 > n) r1 = n;
    if (n == 1) {
        if (r0 <= 0 && r1 >= 1) z[0] = one;
        return 0;
    }
#define D(i) d[(i) - 1]
#define E(i) e[(i) - 1]
    ddr eps = kEps;
    ddr eps2 = pow2(eps);
    ddr safmin = kSafmin;
    ddr safmax = one / safmin;
    ddr ssfmax = sqrt(safmax) / three;
    ddr ssfmin = sqrt(safmin) / eps2;
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = r0; i < r1; i++) z[i + j * ldz] = (i == j) ? one : zero;
    const int64_t nmaxit = n * 30, nm1 = n - 1;
    int64_t jtot = 0, l1 = 1, m = 0, l = 0, lsv = 0, lend = 0, lendsv = 0, iscale = 0;
    ddr anorm, p, rt1, rt2, c, s, g, r, f, b, tst;
    for (;;) {
        if (l1 > n) break;  // to the final sort
        if (l1 > 1) E(l1 - 1) = zero;
        m = n;
        if (l1 <= nm1) {
            for (int64_t mm = l1; mm <= nm1; mm++) {
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
        if (abs(D(lend)) < abs(D(l))) {
            lend = lsv;
            l = lendsv;
        }
        if (lend > l) {  // QL
            for (;;) {
                m = lend;
                if (l != lend)
                    for (int64_t mm = l; mm <= lend - 1; mm++) {
                        tst = pow2(abs(E(mm)));
                        if (tst <= (eps2 * abs(D(mm))) * abs(D(mm + 1)) + safmin) {
                            m = mm;
                            break;
                        }
                    }
                if (m < lend) E(m) = zero;
                p = D(l);
                if (m == l) {
                    D(l) = p;
                    l++;
                    if (l <= lend) continue;
                    break;
                }
                if (m == l + 1) {
                    laev2(D(l), E(l), D(l + 1), rt1, rt2, c, s);
                    work[l - 1] = c;
                    work[n - 1 + l - 1] = s;
                    lasr_rv(false, 2, &work[l - 1], &work[n - 1 + l - 1], &z[(l - 1) * ldz], ldz, r0, r1);
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
                r = lapy2(g, one);
                g = D(m) - p + (E(l) / (g + sign(r, g)));
                s = one;
                c = one;
                p = zero;
                for (int64_t i = m - 1; i >= l; i--) {
                    f = s * E(i);
                    b = c * E(i);
                    lartg(g, f, c, s, r);
                    if (i != m - 1) E(i + 1) = r;
                    g = D(i + 1) - p;
                    r = (D(i) - g) * s + two * c * b;
                    p = s * r;
                    D(i + 1) = g + p;
                    g = c * r - b;
                    work[i - 1] = c;
                    work[n - 1 + i - 1] = -s;
                }
                lasr_rv(false, m - l + 1, &work[l - 1], &work[n - 1 + l - 1], &z[(l - 1) * ldz], ldz, r0, r1);
                D(l) = D(l) - p;
                E(l) = g;
            }
        } else {  // QR
            for (;;) {
                m = lend;
                if (l != lend)
                    for (int64_t mm = l; mm >= lend + 1; mm--) {
                        tst = pow2(abs(E(mm - 1)));
                        if (tst <= (eps2 * abs(D(mm))) * abs(D(mm - 1)) + safmin) {
                            m = mm;
                            break;
                        }
                    }
                if (m > lend) E(m - 1) = zero;
                p = D(l);
                if (m == l) {
                    D(l) = p;
                    l--;
                    if (l >= lend) continue;
                    break;
                }
                if (m == l - 1) {
                    laev2(D(l - 1), E(l - 1), D(l), rt1, rt2, c, s);
                    work[m - 1] = c;
                    work[n - 1 + m - 1] = s;
                    lasr_rv(true, 2, &work[m - 1], &work[n - 1 + m - 1], &z[(l - 2) * ldz], ldz, r0, r1);
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
                r = lapy2(g, one);
                g = D(m) - p + (E(l - 1) / (g + sign(r, g)));
                s = one;
                c = one;
                p = zero;
                const int64_t lm1 = l - 1;
                for (int64_t i = m; i <= lm1; i++) {
                    f = s * E(i);
                    b = c * E(i);
                    lartg(g, f, c, s, r);
                    if (i != m) E(i - 1) = r;
                    g = D(i) - p;
                    r = (D(i + 1) - g) * s + two * c * b;
                    p = s * r;
                    D(i) = g + p;
                    g = c * r - b;
                    work[i - 1] = c;
                    work[n - 1 + i - 1] = s;
                }
                lasr_rv(true, l - m + 1, &work[m - 1], &work[n - 1 + m - 1], &z[(m - 1) * ldz], ldz, r0, r1);
                D(l) = D(l) - p;
                E(lm1) = g;
            }
        }
        if (iscale == 1) {
            lascl_g(ssfmax, anorm, lendsv - lsv + 1, &D(lsv));
            lascl_g(ssfmax, anorm, lendsv - lsv, &E(lsv));
        } else if (iscale == 2) {
            lascl_g(ssfmin, anorm, lendsv - lsv + 1, &D(lsv));
            lascl_g(ssfmin, anorm, lendsv - lsv, &E(lsv));
        }
        if (jtot < nmaxit) continue;
        for (int64_t i = 1; i <= n - 1; i++)
            if (E(i) != zero) info++;
        return info;
    }
    // selection sort, swapping the formed rows of the vectors
    for (int64_t ii = 2; ii <= n; ii++) {
        int64_t i = ii - 1, k = i;
        p = D(i);
        for (int64_t j = ii; j <= n; j++)
            if (D(j) < p) {
                k = j;
                p = D(j);
            }
        if (k != i) {
            D(k) = D(i);
            D(i) = p;
            for (int64_t q = r0; q < r1; q++) {
                ddr t = z[q + (i - 1) * ldz];
                z[q + (i - 1) * ldz] = z[q + (k - 1) * ldz];
                z[q + (k - 1) * ldz] = t;
            }
        }
    }
#undef D
#undef E
    return info;
}

}  // namespace sdp

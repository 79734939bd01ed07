// ORACLE — test infrastructure only. Nothing in the product links this file.
//
// A CPU restatement of the double-double ("dd") arithmetic that SDPA-DD uses:
// the QD 2.3.23 dd_real type configured with its defaults (sloppy add, sloppy
// divide, no FMA; qd/include/qd/dd_inline.h:26-393, qd/include/qd/inline.h:36-113,
// qd/src/dd_real.cpp:47-69 inside external/qd/download/qd-2.3.23.tar.gz).
//
// The operator set mirrors QD's overloads so that restated reference
// expressions resolve to the same rounding sequence under C++ overload
// resolution (e.g. `dd / int` picks dd/double, `double * dd` picks dd*double).
// Operands are combined in the order the published error-free transformations
// prescribe (Knuth TwoSum, Dekker QuickTwoSum, Dekker/Veltkamp TwoProduct with a
// 2^27+1 splitter): the reference's results depend on that order bit for bit.
//
// TwoProduct here uses Dekker's splitting (as the reference build does on
// x86-64 without FMA); the product code uses an FMA.  Both return the exact
// error term, so they agree bitwise; tests check that.
//
// Must be compiled with -ffp-contract=off.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

namespace orc {

namespace eft {
// s = fl(a+b), err = a+b-s exactly, assuming |a| >= |b|.
inline double quick_two_sum(double a, double b, double &err) {
    double s = a + b;
    err = b - (s - a);
    return s;
}
// s = fl(a+b), err exact, no magnitude assumption (Knuth).
inline double two_sum(double a, double b, double &err) {
    double s = a + b;
    double v = s - a;
    err = (a - (s - v)) + (b - v);
    return s;
}
inline double two_diff(double a, double b, double &err) {
    double s = a - b;
    double v = s - a;
    err = (a - (s - v)) - (b + v);
    return s;
}
// Veltkamp split of a into 26+26 significant bits; large inputs are scaled
// by 2^-28 first to avoid overflow in the splitter product.
inline void split(double a, double &hi, double &lo) {
    const double splitter = 134217729.0;          // 2^27 + 1
    const double thresh = 6.69692879491417e+299;  // 2^996
    if (a > thresh || a < -thresh) {
        a *= 3.7252902984619140625e-09;  // 2^-28
        double t = splitter * a;
        hi = t - (t - a);
        lo = a - hi;
        hi *= 268435456.0;  // 2^28
        lo *= 268435456.0;
    } else {
        double t = splitter * a;
        hi = t - (t - a);
        lo = a - hi;
    }
}
// p = fl(a*b), err = a*b-p exactly (Dekker).
inline double two_prod(double a, double b, double &err) {
    double p = a * b;
    double ah, al, bh, bl;
    split(a, ah, al);
    split(b, bh, bl);
    err = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
    return p;
}
inline double two_sqr(double a, double &err) {
    double q = a * a;
    double h, l;
    split(a, h, l);
    err = ((h * h - q) + 2.0 * h * l) + l * l;
    return q;
}
}  // namespace eft

struct dd_real {
    double x[2];
    dd_real() { x[0] = 0.0; x[1] = 0.0; }
    dd_real(double h) { x[0] = h; x[1] = 0.0; }
    dd_real(int h) { x[0] = static_cast<double>(h); x[1] = 0.0; }
    dd_real(double h, double l) { x[0] = h; x[1] = l; }

    static dd_real add(double a, double b) {
        double e;
        double s = eft::two_sum(a, b, e);
        return dd_real(s, e);
    }
    static dd_real sqr(double a) {
        double e;
        double p = eft::two_sqr(a, e);
        return dd_real(p, e);
    }
    static dd_real sloppy_add(const dd_real &a, const dd_real &b) {
        double e;
        double s = eft::two_sum(a.x[0], b.x[0], e);
        e += (a.x[1] + b.x[1]);
        s = eft::quick_two_sum(s, e, e);
        return dd_real(s, e);
    }
    static dd_real sloppy_div(const dd_real &a, const dd_real &b);

    dd_real &operator+=(double a) {
        double s2;
        double s1 = eft::two_sum(x[0], a, s2);
        s2 += x[1];
        x[0] = eft::quick_two_sum(s1, s2, x[1]);
        return *this;
    }
    dd_real &operator+=(const dd_real &a) {
        double e;
        double s = eft::two_sum(x[0], a.x[0], e);
        e += x[1];
        e += a.x[1];
        x[0] = eft::quick_two_sum(s, e, x[1]);
        return *this;
    }
    dd_real &operator-=(double a) {
        double s2;
        double s1 = eft::two_diff(x[0], a, s2);
        s2 += x[1];
        x[0] = eft::quick_two_sum(s1, s2, x[1]);
        return *this;
    }
    dd_real &operator-=(const dd_real &a) {
        double e;
        double s = eft::two_diff(x[0], a.x[0], e);
        e += x[1];
        e -= a.x[1];
        x[0] = eft::quick_two_sum(s, e, x[1]);
        return *this;
    }
    dd_real operator-() const { return dd_real(-x[0], -x[1]); }
    dd_real &operator*=(double a) {
        double p2;
        double p1 = eft::two_prod(x[0], a, p2);
        p2 += x[1] * a;
        x[0] = eft::quick_two_sum(p1, p2, x[1]);
        return *this;
    }
    dd_real &operator*=(const dd_real &a) {
        double p2;
        double p1 = eft::two_prod(x[0], a.x[0], p2);
        p2 += a.x[1] * x[0];
        p2 += a.x[0] * x[1];
        x[0] = eft::quick_two_sum(p1, p2, x[1]);
        return *this;
    }
    dd_real &operator/=(double a);
    dd_real &operator/=(const dd_real &a);
    dd_real &operator=(double a) {
        x[0] = a;
        x[1] = 0.0;
        return *this;
    }
    bool is_zero() const { return x[0] == 0.0; }
    bool is_negative() const { return x[0] < 0.0; }
};

// ---- additions / subtractions ------------------------------------------
inline dd_real operator+(const dd_real &a, double b) {
    double s2;
    double s1 = eft::two_sum(a.x[0], b, s2);
    s2 += a.x[1];
    s1 = eft::quick_two_sum(s1, s2, s2);
    return dd_real(s1, s2);
}
inline dd_real operator+(const dd_real &a, const dd_real &b) { return dd_real::sloppy_add(a, b); }
inline dd_real operator+(double a, const dd_real &b) { return b + a; }
inline dd_real operator-(const dd_real &a, double b) {
    double s2;
    double s1 = eft::two_diff(a.x[0], b, s2);
    s2 += a.x[1];
    s1 = eft::quick_two_sum(s1, s2, s2);
    return dd_real(s1, s2);
}
inline dd_real operator-(const dd_real &a, const dd_real &b) {
    double e;
    double s = eft::two_diff(a.x[0], b.x[0], e);
    e += a.x[1];
    e -= b.x[1];
    s = eft::quick_two_sum(s, e, e);
    return dd_real(s, e);
}
inline dd_real operator-(double a, const dd_real &b) {
    double s2;
    double s1 = eft::two_diff(a, b.x[0], s2);
    s2 -= b.x[1];
    s1 = eft::quick_two_sum(s1, s2, s2);
    return dd_real(s1, s2);
}

// ---- multiplications ----------------------------------------------------
inline dd_real operator*(const dd_real &a, double b) {
    double p2;
    double p1 = eft::two_prod(a.x[0], b, p2);
    p2 += (a.x[1] * b);
    p1 = eft::quick_two_sum(p1, p2, p2);
    return dd_real(p1, p2);
}
inline dd_real operator*(const dd_real &a, const dd_real &b) {
    double p2;
    double p1 = eft::two_prod(a.x[0], b.x[0], p2);
    p2 += (a.x[0] * b.x[1] + a.x[1] * b.x[0]);
    p1 = eft::quick_two_sum(p1, p2, p2);
    return dd_real(p1, p2);
}
inline dd_real operator*(double a, const dd_real &b) { return b * a; }

// ---- divisions ------------------------------------------------------------
inline dd_real operator/(const dd_real &a, double b) {
    double q1 = a.x[0] / b;
    double p2;
    double p1 = eft::two_prod(q1, b, p2);
    double e;
    double s = eft::two_diff(a.x[0], p1, e);
    e += a.x[1];
    e -= p2;
    double q2 = (s + e) / b;
    dd_real r;
    r.x[0] = eft::quick_two_sum(q1, q2, r.x[1]);
    return r;
}
inline dd_real dd_real::sloppy_div(const dd_real &a, const dd_real &b) {
    double q1 = a.x[0] / b.x[0];
    dd_real r = b * q1;
    double s2;
    double s1 = eft::two_diff(a.x[0], r.x[0], s2);
    s2 -= r.x[1];
    s2 += a.x[1];
    double q2 = (s1 + s2) / b.x[0];
    r.x[0] = eft::quick_two_sum(q1, q2, r.x[1]);
    return r;
}
inline dd_real operator/(const dd_real &a, const dd_real &b) { return dd_real::sloppy_div(a, b); }
inline dd_real operator/(double a, const dd_real &b) { return dd_real(a) / b; }
inline dd_real &dd_real::operator/=(double a) {
    *this = *this / a;
    return *this;
}
inline dd_real &dd_real::operator/=(const dd_real &a) {
    *this = *this / a;
    return *this;
}

// ---- comparisons ------------------------------------------------------------
inline bool operator==(const dd_real &a, double b) { return a.x[0] == b && a.x[1] == 0.0; }
inline bool operator==(const dd_real &a, const dd_real &b) { return a.x[0] == b.x[0] && a.x[1] == b.x[1]; }
inline bool operator==(double a, const dd_real &b) { return a == b.x[0] && b.x[1] == 0.0; }
inline bool operator!=(const dd_real &a, double b) { return a.x[0] != b || a.x[1] != 0.0; }
inline bool operator!=(const dd_real &a, const dd_real &b) { return a.x[0] != b.x[0] || a.x[1] != b.x[1]; }
inline bool operator!=(double a, const dd_real &b) { return a != b.x[0] || b.x[1] != 0.0; }
inline bool operator>(const dd_real &a, double b) { return a.x[0] > b || (a.x[0] == b && a.x[1] > 0.0); }
inline bool operator>(const dd_real &a, const dd_real &b) {
    return a.x[0] > b.x[0] || (a.x[0] == b.x[0] && a.x[1] > b.x[1]);
}
inline bool operator>(double a, const dd_real &b) { return a > b.x[0] || (a == b.x[0] && b.x[1] < 0.0); }
inline bool operator<(const dd_real &a, double b) { return a.x[0] < b || (a.x[0] == b && a.x[1] < 0.0); }
inline bool operator<(const dd_real &a, const dd_real &b) {
    return a.x[0] < b.x[0] || (a.x[0] == b.x[0] && a.x[1] < b.x[1]);
}
inline bool operator<(double a, const dd_real &b) { return a < b.x[0] || (a == b.x[0] && b.x[1] > 0.0); }
inline bool operator>=(const dd_real &a, double b) { return a.x[0] > b || (a.x[0] == b && a.x[1] >= 0.0); }
inline bool operator>=(const dd_real &a, const dd_real &b) {
    return a.x[0] > b.x[0] || (a.x[0] == b.x[0] && a.x[1] >= b.x[1]);
}
inline bool operator<=(const dd_real &a, double b) { return a.x[0] < b || (a.x[0] == b && a.x[1] <= 0.0); }
inline bool operator<=(const dd_real &a, const dd_real &b) {
    return a.x[0] < b.x[0] || (a.x[0] == b.x[0] && a.x[1] <= b.x[1]);
}
inline bool operator>=(double a, const dd_real &b) { return b <= a; }
inline bool operator<=(double a, const dd_real &b) { return b >= a; }

// ---- functions ------------------------------------------------------------
inline dd_real abs(const dd_real &a) { return (a.x[0] < 0.0) ? -a : a; }
inline dd_real sqr(const dd_real &a) {
    double p2;
    double p1 = eft::two_sqr(a.x[0], p2);
    p2 += 2.0 * a.x[0] * a.x[1];
    p2 += a.x[1] * a.x[1];
    double s2;
    double s1 = eft::quick_two_sum(p1, p2, s2);
    return dd_real(s1, s2);
}
// Karp's square root: x ~ 1/sqrt(a.hi) in double, one correction in dd.
inline dd_real sqrt(const dd_real &a) {
    if (a.is_zero()) return dd_real(0.0);
    if (a.is_negative()) return dd_real(std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::quiet_NaN());
    double x = 1.0 / std::sqrt(a.x[0]);
    double ax = a.x[0] * x;
    return dd_real::add(ax, (a - dd_real::sqr(ax)).x[0] * (x * 0.5));
}
// std::max / std::min semantics used by the reference through `using namespace std`.
inline const dd_real &max(const dd_real &a, const dd_real &b) { return (a < b) ? b : a; }
inline const dd_real &min(const dd_real &a, const dd_real &b) { return (b < a) ? b : a; }
inline bool isnan_dd(const dd_real &a) { return std::isnan(a.x[0]) || std::isnan(a.x[1]); }

// Binary powering with dd squaring; reciprocal for negative exponents.
inline dd_real npwr(const dd_real &a, int n) {
    if (n == 0) return dd_real(1.0);
    dd_real r = a;
    dd_real s = 1.0;
    int N = n < 0 ? -n : n;
    if (N > 1) {
        while (N > 0) {
            if (N % 2 == 1) s *= r;
            N /= 2;
            if (N > 0) r = sqr(r);
        }
    } else {
        s = r;
    }
    if (n < 0) return 1.0 / s;
    return s;
}

// Decimal string -> dd the way QD's dd_real::read does it (qd/src/dd_real.cpp
// read()): digits accumulated as r = r*10 + d in dd, then one multiplication by
// 10^e computed with npwr.  Returns 0 on success, -1 on a malformed string.
inline int read_dd(const char *s, dd_real &a) {
    const char *p = s;
    int sign = 0, point = -1, nd = 0, e = 0;
    bool done = false;
    dd_real r = 0.0;
    while (*p == ' ') p++;
    char ch;
    while (!done && (ch = *p) != '\0') {
        if (ch >= '0' && ch <= '9') {
            r *= 10.0;
            r += static_cast<double>(ch - '0');
            nd++;
        } else if (ch == '.') {
            if (point >= 0) return -1;
            point = nd;
        } else if (ch == '-' || ch == '+') {
            if (sign != 0 || nd > 0) return -1;
            sign = (ch == '-') ? -1 : 1;
        } else if (ch == 'E' || ch == 'e') {
            if (std::sscanf(p + 1, "%d", &e) != 1) return -1;
            done = true;
        } else {
            return -1;
        }
        p++;
    }
    if (point >= 0) e -= (nd - point);
    if (e != 0) r *= npwr(dd_real(10.0), e);
    a = (sign == -1) ? -r : r;
    return 0;
}

// QD constants used by the LAPACK helpers (qd/src/dd_const.cpp:33-34).
static const double dd_eps = 4.93038065763132e-32;             // 2^-104
static const double dd_min_normalized = 2.0041683600089728e-292;  // 2^-968

}  // namespace orc

// Double-double ("dd") arithmetic for CDNA4 FP64 lanes and the host.
//
// Semantics follow the dd_real type of QD 2.3.23 as SDPA-DD builds it
// (defaults: sloppy add, sloppy divide): the same error-free transformations
// combined in the same order, so results are bit-identical to the CPU
// reference.  TwoProduct uses a fused multiply-add, which yields the exact
// product error that the reference's Dekker split also yields.
//
// Every translation unit including this header must be compiled with
// -ffp-contract=off: the rounding of each step is part of the contract.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define DD_HD __host__ __device__ __forceinline__

struct dd {
    double hi, lo;
};

DD_HD dd mk(double h) { return dd{h, 0.0}; }
DD_HD dd mk(double h, double l) { return dd{h, l}; }

// ---- error-free transformations -------------------------------------------
DD_HD double quick_two_sum(double a, double b, double &err) {
    double s = a + b;
    err = b - (s - a);
    return s;
}
DD_HD double two_sum(double a, double b, double &err) {
    double s = a + b;
    double v = s - a;
    err = (a - (s - v)) + (b - v);
    return s;
}
DD_HD double two_diff(double a, double b, double &err) {
    double s = a - b;
    double v = s - a;
    err = (a - (s - v)) - (b + v);
    return s;
}
DD_HD double two_prod(double a, double b, double &err) {
    double p = a * b;
    err = fma(a, b, -p);
    return p;
}

// ---- QD operator set ----------------------------------------------------------
// a + b (sloppy_add)
DD_HD dd dd_add(dd a, dd b) {
    double e;
    double s = two_sum(a.hi, b.hi, e);
    e += (a.lo + b.lo);
    s = quick_two_sum(s, e, e);
    return dd{s, e};
}
// a += b
DD_HD dd dd_add_eq(dd a, dd b) {
    double e;
    double s = two_sum(a.hi, b.hi, e);
    e += a.lo;
    e += b.lo;
    double l;
    double h = quick_two_sum(s, e, l);
    return dd{h, l};
}
// a - b  (and a -= b, which rounds identically)
DD_HD dd dd_sub(dd a, dd b) {
    double e;
    double s = two_diff(a.hi, b.hi, e);
    e += a.lo;
    e -= b.lo;
    s = quick_two_sum(s, e, e);
    return dd{s, e};
}
// a + double
DD_HD dd dd_add_d(dd a, double b) {
    double s2;
    double s1 = two_sum(a.hi, b, s2);
    s2 += a.lo;
    s1 = quick_two_sum(s1, s2, s2);
    return dd{s1, s2};
}
// a - double
DD_HD dd dd_sub_d(dd a, double b) {
    double s2;
    double s1 = two_diff(a.hi, b, s2);
    s2 += a.lo;
    s1 = quick_two_sum(s1, s2, s2);
    return dd{s1, s2};
}
// double - b
DD_HD dd d_sub_dd(double a, dd b) {
    double s2;
    double s1 = two_diff(a, b.hi, s2);
    s2 -= b.lo;
    s1 = quick_two_sum(s1, s2, s2);
    return dd{s1, s2};
}
// a * b
DD_HD dd dd_mul(dd a, dd b) {
    double p2;
    double p1 = two_prod(a.hi, b.hi, p2);
    p2 += (a.hi * b.lo + a.lo * b.hi);
    p1 = quick_two_sum(p1, p2, p2);
    return dd{p1, p2};
}
// a *= b
DD_HD dd dd_mul_eq(dd a, dd b) {
    double p2;
    double p1 = two_prod(a.hi, b.hi, p2);
    p2 += b.lo * a.hi;
    p2 += b.hi * a.lo;
    double l;
    double h = quick_two_sum(p1, p2, l);
    return dd{h, l};
}
// a * double (and double * a)
DD_HD dd dd_mul_d(dd a, double b) {
    double p2;
    double p1 = two_prod(a.hi, b, p2);
    p2 += (a.lo * b);
    p1 = quick_two_sum(p1, p2, p2);
    return dd{p1, p2};
}
// a / double
DD_HD dd dd_div_d(dd a, double b) {
    double q1 = a.hi / b;
    double p2;
    double p1 = two_prod(q1, b, p2);
    double e;
    double s = two_diff(a.hi, p1, e);
    e += a.lo;
    e -= p2;
    double q2 = (s + e) / b;
    double l;
    double h = quick_two_sum(q1, q2, l);
    return dd{h, l};
}
// a / b (sloppy_div)
DD_HD dd dd_div(dd a, dd b) {
    double q1 = a.hi / b.hi;
    dd r = dd_mul_d(b, q1);
    double s2;
    double s1 = two_diff(a.hi, r.hi, s2);
    s2 -= r.lo;
    s2 += a.lo;
    double q2 = (s1 + s2) / b.hi;
    double l;
    double h = quick_two_sum(q1, q2, l);
    return dd{h, l};
}
DD_HD dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }
DD_HD dd dd_abs(dd a) { return (a.hi < 0.0) ? dd_neg(a) : a; }
// dd square of a double (exact)
DD_HD dd dd_sqr_d(double a) {
    double e;
    double p = two_prod(a, a, e);
    return dd{p, e};
}
// dd square of a dd
DD_HD dd dd_sqr(dd a) {
    double p2;
    double p1 = two_prod(a.hi, a.hi, p2);
    p2 += 2.0 * a.hi * a.lo;
    p2 += a.lo * a.lo;
    double s2;
    double s1 = quick_two_sum(p1, p2, s2);
    return dd{s1, s2};
}
// Karp's square root: one Newton correction of 1/sqrt(hi) in dd.
DD_HD dd dd_sqrt(dd a) {
    if (a.hi == 0.0) return dd{0.0, 0.0};
    if (a.hi < 0.0) return dd{NAN, NAN};
    double x = 1.0 / sqrt(a.hi);
    double ax = a.hi * x;
    double c = dd_sub(a, dd_sqr_d(ax)).hi * (x * 0.5);
    double e;
    double s = two_sum(ax, c, e);
    return dd{s, e};
}

// ---- comparisons (QD semantics) ----------------------------------------------
DD_HD bool dd_eq(dd a, dd b) { return a.hi == b.hi && a.lo == b.lo; }
DD_HD bool dd_ne(dd a, dd b) { return a.hi != b.hi || a.lo != b.lo; }
DD_HD bool dd_lt(dd a, dd b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
DD_HD bool dd_gt(dd a, dd b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }
DD_HD bool dd_le(dd a, dd b) { return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo); }
DD_HD bool dd_ge(dd a, dd b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }
DD_HD bool dd_isnan(dd a) { return a.hi != a.hi || a.lo != a.lo; }
DD_HD bool dd_is_zero(dd a) { return a.hi == 0.0 && a.lo == 0.0; }

// Fused "c += x*y" and "c -= x*y" in the operator forms the BLAS loops use.
DD_HD dd dd_fma_acc(dd c, dd x, dd y) { return dd_add_eq(c, dd_mul(x, y)); }
DD_HD dd dd_fms(dd c, dd x, dd y) { return dd_sub(c, dd_mul(x, y)); }

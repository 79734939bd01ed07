// dd scalar with QD's operator overload set, for the solver's scalar
// control logic on the host and for scalar device code.  Overload resolution reproduces the reference's
// mixed dd/double/int expressions (e.g. `dd / int` resolves to dd/double,
// `double * dd` to dd*double), and every operator delegates to the shared
// rounding sequences in dd.hpp.
#pragma once
#include "dd.hpp"

struct ddr {
    dd v;
    DD_HD ddr() : v{0.0, 0.0} {}
    DD_HD ddr(double h) : v{h, 0.0} {}
    DD_HD ddr(int h) : v{(double)h, 0.0} {}
    DD_HD ddr(dd x) : v(x) {}
    DD_HD ddr(double h, double l) : v{h, l} {}
    DD_HD double hi() const { return v.hi; }
    DD_HD ddr &operator+=(const ddr &a) { v = dd_add_eq(v, a.v); return *this; }
    DD_HD ddr &operator+=(double a) { v = dd_add_d(v, a); return *this; }
    DD_HD ddr &operator-=(const ddr &a) { v = dd_sub(v, a.v); return *this; }
    DD_HD ddr &operator*=(const ddr &a) { v = dd_mul_eq(v, a.v); return *this; }
    DD_HD ddr &operator*=(double a) { v = dd_mul_d(v, a); return *this; }
    DD_HD ddr &operator/=(double a) { v = dd_div_d(v, a); return *this; }
    DD_HD ddr &operator/=(const ddr &a) { v = dd_div(v, a.v); return *this; }
    DD_HD ddr operator-() const { return ddr(dd_neg(v)); }
};

DD_HD ddr operator+(const ddr &a, const ddr &b) { return ddr(dd_add(a.v, b.v)); }
DD_HD ddr operator+(const ddr &a, double b) { return ddr(dd_add_d(a.v, b)); }
DD_HD ddr operator+(double a, const ddr &b) { return ddr(dd_add_d(b.v, a)); }
DD_HD ddr operator-(const ddr &a, const ddr &b) { return ddr(dd_sub(a.v, b.v)); }
DD_HD ddr operator-(const ddr &a, double b) { return ddr(dd_sub_d(a.v, b)); }
DD_HD ddr operator-(double a, const ddr &b) { return ddr(d_sub_dd(a, b.v)); }
DD_HD ddr operator*(const ddr &a, const ddr &b) { return ddr(dd_mul(a.v, b.v)); }
DD_HD ddr operator*(const ddr &a, double b) { return ddr(dd_mul_d(a.v, b)); }
DD_HD ddr operator*(double a, const ddr &b) { return ddr(dd_mul_d(b.v, a)); }
DD_HD ddr operator/(const ddr &a, const ddr &b) { return ddr(dd_div(a.v, b.v)); }
DD_HD ddr operator/(const ddr &a, double b) { return ddr(dd_div_d(a.v, b)); }
DD_HD ddr operator/(double a, const ddr &b) { return ddr(dd_div(mk(a), b.v)); }

DD_HD bool operator<(const ddr &a, const ddr &b) { return dd_lt(a.v, b.v); }
DD_HD bool operator>(const ddr &a, const ddr &b) { return dd_gt(a.v, b.v); }
DD_HD bool operator<=(const ddr &a, const ddr &b) { return dd_le(a.v, b.v); }
DD_HD bool operator>=(const ddr &a, const ddr &b) { return dd_ge(a.v, b.v); }
DD_HD bool operator==(const ddr &a, const ddr &b) { return dd_eq(a.v, b.v); }
DD_HD bool operator!=(const ddr &a, const ddr &b) { return dd_ne(a.v, b.v); }
DD_HD bool operator<(const ddr &a, double b) { return a.v.hi < b || (a.v.hi == b && a.v.lo < 0.0); }
DD_HD bool operator>(const ddr &a, double b) { return a.v.hi > b || (a.v.hi == b && a.v.lo > 0.0); }
DD_HD bool operator<=(const ddr &a, double b) { return a.v.hi < b || (a.v.hi == b && a.v.lo <= 0.0); }
DD_HD bool operator>=(const ddr &a, double b) { return a.v.hi > b || (a.v.hi == b && a.v.lo >= 0.0); }
DD_HD bool operator<(double a, const ddr &b) { return a < b.v.hi || (a == b.v.hi && b.v.lo > 0.0); }
DD_HD bool operator>(double a, const ddr &b) { return a > b.v.hi || (a == b.v.hi && b.v.lo < 0.0); }
DD_HD bool operator<=(double a, const ddr &b) { return b >= a; }
DD_HD bool operator>=(double a, const ddr &b) { return b <= a; }

DD_HD ddr abs(const ddr &a) { return ddr(dd_abs(a.v)); }
DD_HD ddr sqrt(const ddr &a) { return ddr(dd_sqrt(a.v)); }
// std::max / std::min as the reference reaches them through `using namespace std`.
DD_HD ddr max(const ddr &a, const ddr &b) { return (a < b) ? b : a; }
DD_HD ddr min(const ddr &a, const ddr &b) { return (b < a) ? b : a; }

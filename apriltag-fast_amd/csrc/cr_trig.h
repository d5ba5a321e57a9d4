// Correctly rounded float atan2, cos and sin on the device: the libm calls
// of refine_edges (fast/decode.c:740-744), whose results the reference's
// goldens hold correctly rounded.
//
// Ziv's strategy: the double-precision function (the device library's,
// within a few ulps of double) rounded to float is the correctly rounded
// float unless the double result lies within that error of a float
// rounding boundary (a midpoint between two floats). Then, about once in
// 2^26 calls, the value is recomputed in double-double arithmetic
// (about 100 bits) and that is rounded: sin and cos by a Taylor series after
// reduction by pi/2 held in three doubles, atan2 by one Newton step from
// the double result (theta + atan(delta), delta from the double-double sin
// and cos of theta).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace atg {
namespace cr {

struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
    const double s = __dadd_rn(a, b), bb = __dsub_rn(s, a);
    return {s, __dadd_rn(__dsub_rn(a, __dsub_rn(s, bb)), __dsub_rn(b, bb))};
}
__device__ __forceinline__ dd fast_two_sum(double a, double b) {  // |a| >= |b|
    const double s = __dadd_rn(a, b);
    return {s, __dsub_rn(b, __dsub_rn(s, a))};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
    const double p = __dmul_rn(a, b);
    return {p, __fma_rn(a, b, -p)};
}
__device__ __forceinline__ dd add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo = __dadd_rn(s.lo, t.hi);
    s = fast_two_sum(s.hi, s.lo);
    s.lo = __dadd_rn(s.lo, t.lo);
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd neg(dd a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ dd mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo = __dadd_rn(p.lo, __dadd_rn(__dmul_rn(a.hi, b.lo), __dmul_rn(a.lo, b.hi)));
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    p.lo = __fma_rn(a.lo, b, p.lo);
    return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd div_d(dd a, double b) {
    const double q1 = __ddiv_rn(a.hi, b);
    const dd r = add(a, neg(two_prod(q1, b)));
    const double q2 = __ddiv_rn(r.hi, b);
    return fast_two_sum(q1, q2);
}

// pi/2 = P1 + P2 + P3 to about 160 bits
constexpr double kP1 = 0x1.921fb54442d18p+0, kP2 = 0x1.1a62633145c07p-54, kP3 = -0x1.f1976b7ed8fbcp-110;

// sin and cos of x (|x| <= 4) in double-double: x - k pi/2 = r with
// |r| <= pi/4 (k P1 and k P2 exact as double-doubles), the series of sin r
// and cos r to r^27, then the quadrant.
__device__ void sincos_dd(double x, dd *s, dd *c) {
    const double k = rint(__dmul_rn(x, 0.63661977236758138));  // 2/pi
    dd r = add({x, 0.0}, neg(two_prod(k, kP1)));
    r = add(r, neg(two_prod(k, kP2)));
    r = add(r, {-__dmul_rn(k, kP3), 0.0});
    const dd r2 = mul(r, r);
    dd ts = r, tc = {1.0, 0.0};
    dd sn = r, cs = {1.0, 0.0};
    for (int n = 1; n <= 13; n++) {
        ts = div_d(mul(ts, r2), -(double)((2 * n) * (2 * n + 1)));
        tc = div_d(mul(tc, r2), -(double)((2 * n - 1) * (2 * n)));
        sn = add(sn, ts);
        cs = add(cs, tc);
    }
    const int q = ((int)k) & 3;
    if (q == 0) {
        *s = sn;
        *c = cs;
    } else if (q == 1) {
        *s = cs;
        *c = neg(sn);
    } else if (q == 2) {
        *s = neg(sn);
        *c = neg(cs);
    } else {
        *s = neg(cs);
        *c = sn;
    }
}

// The float nearest to hi + lo (|lo| <= ulp(hi) / 2): float(hi) unless hi
// is exactly halfway between two floats, when lo's sign decides.
__device__ __forceinline__ float round_dd(dd v) {
    const float f = (float)v.hi;
    const double fd = (double)f;
    if (fd == v.hi || v.lo == 0.0) return f;
    const float g = nextafterf(f, v.hi > fd ? INFINITY : -INFINITY);
    const double mid = 0.5 * (fd + (double)g);  // exact
    if (v.hi != mid) return f;
    // hi is the midpoint: the side of lo
    return (v.lo > 0.0) == ((double)g > fd) ? g : f;
}

// Whether r (a double within eb of the exact value) rounds to float
// safely: no float rounding boundary lies within eb of it.
__device__ __forceinline__ bool rounds_safely(double r, double eb) {
    const float f = (float)r;
    const double fd = (double)f;
    const double up = (double)nextafterf(f, INFINITY);
    const double dn = (double)nextafterf(f, -INFINITY);
    const double m_hi = 0.5 * (fd + up), m_lo = 0.5 * (fd + dn);
    return fabs(r - m_hi) > eb && fabs(r - m_lo) > eb;
}

// 2^-50: a bound, with margin, on the relative error of the device's
// double atan2, sin and cos
constexpr double kRelErr = 0x1p-50;

// slow: take the double-double path whenever it applies (a check of it)
__device__ __forceinline__ float atan2f_cr(float yf, float xf, bool slow = false) {
    const double y = (double)yf, x = (double)xf;
    const double t = atan2(y, x);
    if ((!slow && rounds_safely(t, fabs(t) * kRelErr)) || !(t != 0.0) || !isfinite(y) || !isfinite(x) || y == 0.0 ||
        x == 0.0)
        return (float)t;
    // one Newton step: atan2(y, x) = t + atan(delta),
    // delta = (y cos t - x sin t) / (x cos t + y sin t)
    dd s, c;
    sincos_dd(t, &s, &c);
    const dd num = add(mul_d(c, y), neg(mul_d(s, x)));
    const double den = __dadd_rn(__dmul_rn(x, c.hi), __dmul_rn(y, s.hi));
    const double delta = __ddiv_rn(num.hi, den);
    return round_dd(two_sum(t, delta));
}

// cos and sin: correctly rounded for |x| <= 4 (the refinement's angles
// are within pi/2)
__device__ __forceinline__ float cosf_cr(float xf, bool slow = false) {
    const double x = (double)xf;
    const double r = cos(x);
    if ((!slow && rounds_safely(r, fabs(r) * kRelErr)) || !(fabs(x) <= 4.0)) return (float)r;
    dd s, c;
    sincos_dd(x, &s, &c);
    return round_dd(c);
}

__device__ __forceinline__ float sinf_cr(float xf, bool slow = false) {
    const double x = (double)xf;
    const double r = sin(x);
    if ((!slow && rounds_safely(r, fabs(r) * kRelErr)) || !(fabs(x) <= 4.0) || x == 0.0) return (float)r;
    dd s, c;
    sincos_dd(x, &s, &c);
    return round_dd(s);
}

}  // namespace cr
}  // namespace atg

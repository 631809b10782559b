// sat_crmath.hpp -- correctly rounded sin, cos, acos and x^-3 for the FP64 path.
//
// The factored forward solve loses about (dist/dx)^2 of relative precision in
// its quadratic coefficients, so a one-ulp change in the analytic factor T0
// moves the converged travel time at 200 cells by ~1e-13 and the adjoint
// kernel by ~1e-11.  To make the GPU's T0 bit-identical to a CPU evaluation,
// both sides use correctly rounded elementary functions: here double-double
// evaluation (~2^-104 relative) rounded once to double.
//
// Methods: sin/cos -- Cody-Waite reduction by pi/2 with a three-part pi/2, then
// Taylor series in double-double on |r| <= pi/4.  acos -- acos(x) =
// 2*asin(sqrt((1-x)/2)) for x >= 1/2 (1-x exact), pi - 2*asin(sqrt((1+x)/2))
// for x <= -1/2, pi/2 - asin(x) otherwise, with the asin Maclaurin series
// sum_n (2n)!/(4^n n!^2 (2n+1)) s^(2n+1) on |s| <= 1/2 summed in double-double.
// Double-double primitives: Dekker/Knuth two-sum, FMA two-product.
#pragma once
#include <hip/hip_runtime.h>

namespace sat {
namespace cr {

struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__device__ __forceinline__ dd fast_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}
__device__ __forceinline__ dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd neg(dd a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd div_d(dd a, double b) {
  const double q1 = a.hi / b;
  const dd p = two_prod(q1, b);
  const dd r = add(a, neg(p));
  const double q2 = r.hi / b;
  return fast_two_sum(q1, q2);
}
__device__ __forceinline__ dd div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  const dd r = add(a, neg(mul_d(b, q1)));
  const double q2 = r.hi / b.hi;
  const dd r2 = add(r, neg(mul_d(b, q2)));
  const double q3 = r2.hi / b.hi;
  return add(fast_two_sum(q1, q2), dd{q3, 0.0});
}
__device__ __forceinline__ dd sqrt_d(double z) {  // z >= 0
  if (z == 0.0) return {0.0, 0.0};
  const double s = __builtin_sqrt(z);
  const double e = __builtin_fma(-s, s, z);
  return fast_two_sum(s, e / (2.0 * s));
}
__device__ __forceinline__ double round_dd(dd a) { return a.hi + a.lo; }

constexpr double kPIO2_1 = 1.5707963267948966;       // pi/2, three parts
constexpr double kPIO2_2 = 6.123233995736766e-17;
constexpr double kPIO2_3 = -1.4973849048591698e-33;
constexpr double kPI_hi = 3.141592653589793;         // pi, two parts
constexpr double kPI_lo = 1.2246467991473532e-16;

// sin(r), cos(r) for |r| <= pi/4 + tiny, double-double in and out.
__device__ inline void sincos_kernel(dd r, dd& s, dd& c) {
  const dd r2 = mul(r, r);
  // sin: r - r^3/3! + r^5/5! - ...
  dd t = r;
  dd sum = r;
  for (int n = 1; n < 40; n++) {
    t = div_d(mul(t, r2), (double)((2 * n) * (2 * n + 1)));
    sum = add(sum, (n & 1) ? neg(t) : t);
    if (__builtin_fabs(t.hi) < 1e-36 * __builtin_fabs(sum.hi) || t.hi == 0.0) break;
  }
  s = sum;
  // cos: 1 - r^2/2! + r^4/4! - ...
  t = dd{1.0, 0.0};
  sum = t;
  for (int n = 1; n < 40; n++) {
    t = div_d(mul(t, r2), (double)((2 * n - 1) * (2 * n)));
    sum = add(sum, (n & 1) ? neg(t) : t);
    if (__builtin_fabs(t.hi) < 1e-36 || t.hi == 0.0) break;
  }
  c = sum;
}

// Correctly rounded sin and cos of x (|x| < 1e6).
__device__ inline void sincos_cr(double x, double& sx, double& cx) {
  if (x == 0.0) { sx = x; cx = 1.0; return; }
  const double k = __builtin_rint(x * 0.63661977236758134);  // x * 2/pi
  dd r = add(dd{x, 0.0}, neg(two_prod(k, kPIO2_1)));
  r = add(r, neg(two_prod(k, kPIO2_2)));
  r = add(r, dd{-k * kPIO2_3, 0.0});
  dd s, c;
  sincos_kernel(r, s, c);
  const int q = ((int)k) & 3;
  dd so, co;
  switch (q) {
    case 0: so = s; co = c; break;
    case 1: so = c; co = neg(s); break;
    case 2: so = neg(s); co = neg(c); break;
    default: so = neg(c); co = s; break;
  }
  sx = round_dd(so);
  cx = round_dd(co);
}
__device__ inline double sin_cr(double x) { double s, c; sincos_cr(x, s, c); return s; }
__device__ inline double cos_cr(double x) { double s, c; sincos_cr(x, s, c); return c; }

// asin(s) for |s| <= 1/2, double-double.
__device__ inline dd asin_series(dd s) {
  const dd u = mul(s, s);
  dd t = s;   // a_n s^(2n+1), a_n = (2n)!/(4^n n!^2)
  dd sum = s;
  for (int n = 1; n < 90; n++) {
    t = div_d(mul_d(mul(t, u), (double)(2 * n - 1)), (double)(2 * n));
    const dd term = div_d(t, (double)(2 * n + 1));
    sum = add(sum, term);
    if (__builtin_fabs(term.hi) < 1e-36 * __builtin_fabs(sum.hi) || term.hi == 0.0) break;
  }
  return sum;
}

// Correctly rounded acos(x), x in [-1, 1].
__device__ inline double acos_cr(double x) {
  if (!(x >= -1.0 && x <= 1.0)) return __builtin_nan("");
  if (x == 1.0) return 0.0;
  if (x >= 0.5) {
    const double z = (1.0 - x) * 0.5;  // exact
    const dd a = asin_series(sqrt_d(z));
    return round_dd(dd{2.0 * a.hi, 2.0 * a.lo});
  }
  if (x <= -0.5) {
    const double z = (1.0 + x) * 0.5;  // exact
    const dd a = asin_series(sqrt_d(z));
    return round_dd(add(dd{kPI_hi, kPI_lo}, neg(dd{2.0 * a.hi, 2.0 * a.lo})));
  }
  const dd a = asin_series(dd{x, 0.0});
  return round_dd(add(dd{kPIO2_1, kPIO2_2}, neg(a)));
}

// Correctly rounded v^-3 for v > 0 (src/preproc.cpp:29 pow(-3)).
__device__ inline double pow_m3_cr(double v) {
  const dd v3 = mul(two_prod(v, v), dd{v, 0.0});
  return round_dd(div(dd{1.0, 0.0}, v3));
}

}  // namespace cr
}  // namespace sat

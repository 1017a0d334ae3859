// devmath.hpp -- FP64 reciprocal square root and sine/cosine for the pair kernels.
//
// rsqrt: the hardware estimate v_rsq_f64 (measured relative error 5.2e-8 on
//   log-uniform inputs in [1e-8, 1e8]) refined by one third-order step (rsqrt_nr,
//   exact to rounding) or one Newton step (rsqrt_fast, relative error below 4.1e-15,
//   always from below; used per quadrature-point pair of regular cell pairs).
// sincos: argument reduction x = n pi/2 + r with |r| <= pi/4 by the three-term
//   Cody-Waite scheme (W. J. Cody, W. Waite, Software Manual for the Elementary
//   Functions, 1980) evaluated with fused multiply-adds against pi/2 split into
//   P1 + P2 + P3 (each the double nearest to the remaining tail of pi/2); then the
//   Taylor series of sin (to r^15) and cos (to r^16), whose truncation errors on
//   [-pi/4, pi/4] are below 7e-17 and 3e-18 relative; the quadrant n mod 4 selects
//   and negates.
#pragma once
#include <hip/hip_runtime.h>

namespace bempp_amd {
namespace dev {

// One third-order step from the hardware estimate y (relative error below 6e-8):
// with e = 1 - x y^2, 1/sqrt(x) = y (1 - e)^{-1/2} = y (1 + e/2 + 3e^2/8 + O(e^3)),
// so the result is exact to rounding (e^3 ~ 1e-22).
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  const double t = fma(0.375, e, 0.5);
  return fma(y * e, t, y);
}

// One Newton step: y (1 + e/2) = 1/sqrt(x) (1 - 3e^2/8 + O(e^3)), |e| <= 1.1e-7.
__device__ __forceinline__ double rsqrt_fast(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, 0.5, y);
}

// Valid for |x| <= 1e8 (n = rint(2x/pi) fits an int and every reduction step is
// exact or singly rounded); the host rejects wavenumbers with |k| diam > 1e8.
__device__ __forceinline__ void sincos_cw(double x, double* s_out, double* c_out) {
  constexpr double kTwoOverPi = 0.6366197723675814;        // 2/pi
  constexpr double kP1 = 1.5707963267948966;               // pi/2 = P1 + P2 + P3
  constexpr double kP2 = 6.123233995736766e-17;
  constexpr double kP3 = -1.4973849048591698e-33;
  const double n = rint(x * kTwoOverPi);
  double r = fma(-n, kP1, x);
  r = fma(-n, kP2, r);
  r = fma(-n, kP3, r);
  const double r2 = r * r;
  // sin r = r + r^3 (s3 + r^2 (s5 + ... + r^2 s15))
  double ps = -7.647163731819816e-13;
  ps = fma(ps, r2, 1.6059043836821613e-10);
  ps = fma(ps, r2, -2.505210838544172e-08);
  ps = fma(ps, r2, 2.7557319223985893e-06);
  ps = fma(ps, r2, -0.0001984126984126984);
  ps = fma(ps, r2, 0.008333333333333333);
  ps = fma(ps, r2, -0.16666666666666666);
  const double s = fma(ps * r2, r, r);
  // cos r = 1 + r^2 (c2 + r^2 (c4 + ... + r^2 c16))
  double pc = 4.779477332387385e-14;
  pc = fma(pc, r2, -1.1470745597729725e-11);
  pc = fma(pc, r2, 2.08767569878681e-09);
  pc = fma(pc, r2, -2.755731922398589e-07);
  pc = fma(pc, r2, 2.48015873015873e-05);
  pc = fma(pc, r2, -0.001388888888888889);
  pc = fma(pc, r2, 0.041666666666666664);
  pc = fma(pc, r2, -0.5);
  const double c = fma(pc, r2, 1.0);
  // quadrant q = n mod 4: sin x = (s, c, -s, -c)[q], cos x = (c, -s, -c, s)[q]
  const int q = (int)n;
  const bool swap = q & 1;
  const double ss = swap ? c : s;
  const double cc = swap ? s : c;
  // negate by flipping the sign bit: bit 1 of q for sin, bit 1 of q+1 for cos
  const long long sgn_s = (long long)(q & 2) << 62;
  const long long sgn_c = (long long)((q + 1) & 2) << 62;
  *s_out = __longlong_as_double(__double_as_longlong(ss) ^ sgn_s);
  *c_out = __longlong_as_double(__double_as_longlong(cc) ^ sgn_c);
}

// sin and cos of a small argument without reduction, for the phase k (r - R) of a
// quadrature-point pair relative to its cell pair's centroid distance R (the factor
// e^{ikR} is applied once per cell pair). TIER 1: |x| <= 0.5, TIER 2: |x| <= 0.25.
//   sin x = x + x^3 P(x^2),  cos x = 1 + x^2 C(x^2)
// with P, C the weighted minimax polynomials (absolute error of sin/cos) of degrees
// 4/5 (tier 1) and 3/4 (tier 2), computed by Lawson's iteration at 50 digits; the
// truncation errors are below 1e-17 and the float64 Horner evaluation errors below
// 6e-17 (sin) and 1.2e-16 (cos) on the whole interval.
// SQRT2 = true: the argument arrives multiplied by sqrt 2, x = sqrt(2) a, and the results are
// cos a and sqrt(2) sin a: the same polynomials in u = x^2 = 2 a^2 with the coefficient of u^n
// divided by 2^(n+1) (exact), i.e. sin a / a = 1 + (u/2) P(u/2), cos a = 1 + (u/2) C(u/2).
template <int TIER = 1, bool SQRT2 = false>
__device__ __forceinline__ void sincos_small(double x, double* s_out, double* c_out) {
  static_assert(TIER == 1 || TIER == 2, "phase tiers: 1 (|x| <= 0.5), 2 (|x| <= 0.25)");
  constexpr double f1 = SQRT2 ? 0.5 : 1.0, f2 = f1 * f1, f3 = f2 * f1, f4 = f2 * f2, f5 = f4 * f1, f6 = f3 * f3;
  const double u = x * x;
  double p, c;
  if constexpr (TIER == 1) {
    p = fma(u, -2.4927410039626402e-08 * f5, 2.7556952897882737e-06 * f4);
    p = fma(p, u, -0.00019841269341158048 * f3);
    p = fma(p, u, 0.008333333333023885 * f2);
    p = fma(p, u, -0.16666666666666 * f1);
    c = fma(u, 2.077788258235537e-09 * f6, -2.7556985499996286e-07 * f5);
    c = fma(c, u, 2.480158674730518e-05 * f4);
    c = fma(c, u, -0.00138888888884234 * f3);
    c = fma(c, u, 0.04166666666666487 * f2);
    c = fma(c, u, -0.5 * f1);
  } else {
    p = fma(u, 2.751683000751954e-06 * f4, -0.0001984124679978689 * f3);
    p = fma(p, u, 0.00833333332798428 * f2);
    p = fma(p, u, -0.1666666666666255 * f1);
    c = fma(u, -2.751887867547197e-07 * f5, 2.4801560959230762e-05 * f4);
    c = fma(c, u, -0.0013888888880693657 * f3);
    c = fma(c, u, 0.04166666666665554 * f2);
    c = fma(c, u, -0.49999999999999994 * f1);
  }
  *s_out = fma(x * u, p, x);
  *c_out = fma(u, c, 1.0);
}

}  // namespace dev
}  // namespace bempp_amd

// devmath.hpp -- FP64 reciprocal square root and sine/cosine for the pair kernels.
//
// rsqrt: the hardware estimate v_rsq_f64 (measured relative error 5.2e-8 on
//   log-uniform inputs in [1e-8, 1e8]) refined by one third-order step.
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

// sin and cos of a small argument |x| <= 0.5 without reduction: Taylor series to
// x^13 (sin) and x^14 (cos); truncation below 5e-17 and 1e-18 relative on [-0.5, 0.5].
// Used for the phase k (r - R) of a quadrature-point pair relative to its cell pair's
// centroid distance R; the factor e^{ikR} is applied once per cell pair.
__device__ __forceinline__ void sincos_small(double x, double* s_out, double* c_out) {
  const double x2 = x * x;
  double ps = 1.6059043836821613e-10;
  ps = fma(ps, x2, -2.505210838544172e-08);
  ps = fma(ps, x2, 2.7557319223985893e-06);
  ps = fma(ps, x2, -0.0001984126984126984);
  ps = fma(ps, x2, 0.008333333333333333);
  ps = fma(ps, x2, -0.16666666666666666);
  *s_out = fma(ps * x2, x, x);
  double pc = -1.1470745597729725e-11;
  pc = fma(pc, x2, 2.08767569878681e-09);
  pc = fma(pc, x2, -2.755731922398589e-07);
  pc = fma(pc, x2, 2.48015873015873e-05);
  pc = fma(pc, x2, -0.001388888888888889);
  pc = fma(pc, x2, 0.041666666666666664);
  pc = fma(pc, x2, -0.5);
  *c_out = fma(pc, x2, 1.0);
}

}  // namespace dev
}  // namespace bempp_amd

// Sine/cosine and reciprocal for the hot kernels.
//
// sincos_d: reduction by pi/2 with k = rint(2x/pi) in the FMA form
//   r = fma(-k, P1, x); r = fma(-k, P2, r)
// where P1 = pi/2 rounded to double and P2 = (pi/2 - P1) rounded.  The first FMA is exact:
// x and k*P1 both lie on the 2^-52 grid once |x| >= 1 and their difference is below 1, so
// no argument size needs a fallback (the error of r is ~ulp(r) + |k| 1.5e-33).  Then
// Taylor polynomials on [-pi/4, pi/4] (coefficients 1/n!, truncation < 1e-19) and the
// quadrant k mod 4.
// sincos_small_d: |x| <= 4 without reduction: the same polynomials at x/2, then the
// double-angle formulas (error < 1e-13 at |x| = 4, < 1e-19 for |x| <= 2).
// rcp_d: hardware reciprocal estimate refined by two Newton steps (correctly rounded to
// about one ulp).
#pragma once
#include <hip/hip_runtime.h>

namespace pes {

constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPio2Hi = 1.5707963267948966;     // pi/2 rounded to double
constexpr double kPio2Lo = 6.123233995736766e-17;  // pi/2 - kPio2Hi, rounded
constexpr double kInv2Pi = 0.15915494309189535;    // 1/(2 pi)

// sin/cos on |r| <= 1 (Taylor to r^19 / r^20)
__device__ inline void sincos_poly_d(double r, double* s, double* c) {
  const double r2 = r * r;
  double ps = 8.22063524662433e-18;
  ps = fma(ps, r2, -2.8114572543455206e-15);
  ps = fma(ps, r2, 7.647163731819816e-13);
  ps = fma(ps, r2, -1.6059043836821613e-10);
  ps = fma(ps, r2, 2.505210838544172e-08);
  ps = fma(ps, r2, -2.7557319223985893e-06);
  ps = fma(ps, r2, 0.0001984126984126984);
  ps = fma(ps, r2, -0.008333333333333333);
  ps = fma(ps, r2, 0.16666666666666666);
  *s = fma(-r * r2, ps, r);
  double pc = 1.5619206968586225e-16;
  pc = fma(pc, r2, -4.779477332387385e-14);
  pc = fma(pc, r2, 1.1470745597729725e-11);
  pc = fma(pc, r2, -2.08767569878681e-09);
  pc = fma(pc, r2, 2.755731922398589e-07);
  pc = fma(pc, r2, -2.48015873015873e-05);
  pc = fma(pc, r2, 0.001388888888888889);
  pc = fma(pc, r2, -0.041666666666666664);
  pc = fma(pc, r2, 0.5);
  *c = fma(-r2, pc, 1.0);
}

// FP64 sin/cos of any finite argument
__device__ inline void sincos_d(double x, double* s, double* c) {
  const double k = rint(x * kTwoOverPi);
  double r = fma(-k, kPio2Hi, x);
  r = fma(-k, kPio2Lo, r);
  double sr, cr;
  sincos_poly_d(r, &sr, &cr);
  const int q = (int)fma(-4.0, floor(k * 0.25), k);  // k mod 4, exact for |k| < 2^53
  const double ss = (q & 1) ? cr : sr;
  const double cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// FP64 sin/cos for |x| <= 4
__device__ inline void sincos_small_d(double x, double* s, double* c) {
  double sh, ch;
  sincos_poly_d(0.5 * x, &sh, &ch);
  *s = 2.0 * sh * ch;
  *c = (ch - sh) * (ch + sh);
}

__device__ inline double rcp_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// FP64 argument reduced to revolutions in FP64, then the hardware FP32 sine/cosine
// (v_sin_f32 / v_cos_f32 take revolutions); absolute error ~1e-7
__device__ inline void sincos_rev_f(double x, double* s, double* c) {
  double r = x * kInv2Pi;
  r -= rint(r);
  const float rf = (float)r;
  *s = (double)__builtin_amdgcn_sinf(rf);
  *c = (double)__builtin_amdgcn_cosf(rf);
}

// the same for |x| <= 4 (|x / 2 pi| < 0.64 revolutions: no reduction needed)
__device__ inline void sincos_rev_small_f(double x, double* s, double* c) {
  const float rf = (float)(x * kInv2Pi);
  *s = (double)__builtin_amdgcn_sinf(rf);
  *c = (double)__builtin_amdgcn_cosf(rf);
}

__device__ inline double rcp_f(double x) { return (double)__builtin_amdgcn_rcpf((float)x); }

}  // namespace pes

// Range-reduced sine/cosine for the hot kernels.
//
// Cody-Waite reduction by pi/2 (k = rint(2x/pi), three-part split of pi/2 whose
// first two parts have 29 significant bits so k*C1, k*C2 are exact for |k| < 2^24)
// followed by Taylor polynomials on [-pi/4, pi/4] (coefficients 1/n!), quadrant by
// k mod 4.  The split constants are pi/2 rounded to 29 + 29 + 53 bits, computed with
// mpmath.  Arguments beyond the reduction's exact range fall back to the library.
#pragma once
#include <hip/hip_runtime.h>

namespace pes {

constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPio2_1 = 1.570796325802803;        // 29 bits
constexpr double kPio2_2 = 9.920935774287987e-10;    // next 29 bits
constexpr double kPio2_3 = 2.2517417741562176e-18;   // remainder
constexpr double kPio2_f1 = 1.5707963109016418;      // 26 bits (float-output path)
constexpr double kPio2_f2 = 1.5893254773528196e-08;  // remainder

// sin/cos on |r| <= pi/4, FP64 (truncation < 1e-19)
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

__device__ inline void quadrant(int q, double sr, double cr, double* s, double* c) {
  double ss = (q & 1) ? cr : sr;
  double cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// FP64 sin/cos, |x| < 2.6e7 on the fast path
__device__ inline void sincos_d(double x, double* s, double* c) {
  if (fabs(x) > 2.5e7) {
    sincos(x, s, c);
    return;
  }
  double k = rint(x * kTwoOverPi);
  double r = fma(-k, kPio2_1, x);
  r = fma(-k, kPio2_2, r);
  r = fma(-k, kPio2_3, r);
  double sr, cr;
  sincos_poly_d(r, &sr, &cr);
  quadrant(((int)k) & 3, sr, cr, s, c);
}

// FP32 sin/cos on |r| <= pi/4 (truncation < 2e-9)
__device__ inline void sincos_poly_f(float r, float* s, float* c) {
  const float r2 = r * r;
  float ps = 2.7557319223985893e-06f;
  ps = fmaf(ps, r2, -0.0001984126984126984f);
  ps = fmaf(ps, r2, 0.008333333333333333f);
  ps = fmaf(ps, r2, -0.16666666666666666f);
  *s = fmaf(r * r2, ps, r);
  float pc = 2.08767569878681e-09f;
  pc = fmaf(pc, r2, -2.755731922398589e-07f);
  pc = fmaf(pc, r2, 2.48015873015873e-05f);
  pc = fmaf(pc, r2, -0.001388888888888889f);
  pc = fmaf(pc, r2, 0.041666666666666664f);
  pc = fmaf(pc, r2, -0.5f);
  *c = fmaf(r2, pc, 1.0f);
}

// FP32 result for an FP64 argument: the reduction is done in FP64
__device__ inline void sincos_df(double x, float* s, float* c) {
  double k = rint(x * kTwoOverPi);
  double r = fma(-k, kPio2_f1, x);
  r = fma(-k, kPio2_f2, r);
  float sr, cr;
  sincos_poly_f((float)r, &sr, &cr);
  int q = ((int)k) & 3;
  float ss = (q & 1) ? cr : sr;
  float cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// FP32 argument of modest size (|x| < ~1e3): reduce in FP32 with a two-part pi/2
__device__ inline void sincos_f(float x, float* s, float* c) {
  float k = rintf(x * 0.63661977236758134f);
  float r = fmaf(-k, 1.5707963705062866f, x);
  r = fmaf(-k, -4.3711388286737929e-08f, r);
  float sr, cr;
  sincos_poly_f(r, &sr, &cr);
  int q = ((int)k) & 3;
  float ss = (q & 1) ? cr : sr;
  float cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// precision-generic front ends
__device__ inline void sc_small(double x, double* s, double* c) { sincos_d(x, s, c); }
__device__ inline void sc_small(float x, float* s, float* c) { sincos_f(x, s, c); }

}  // namespace pes

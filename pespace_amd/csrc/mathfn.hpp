// Sine/cosine and reciprocal for the hot kernels.
//
// sincos_d: reduction by pi/2 with k = rint(2x/pi) in the FMA form
//   r = fma(-k, P1, x); r = fma(-k, P2, r)
// where P1 = pi/2 rounded to double and P2 = (pi/2 - P1) rounded.  The first FMA is exact:
// x and k*P1 both lie on the 2^-52 grid once |x| >= 1 and their difference is below 1, so
// no argument size needs a fallback (the error of r is ~ulp(r) + |k| 1.5e-33).  Then
// Taylor polynomials on [-pi/4, pi/4] (coefficients 1/n!, truncation < 1e-19) and the
// quadrant k mod 4.
// Table sincos, the hot-loop variants: reduction by h = 2 pi / 256 in the same FMA form (the
// first FMA is again exact; k is rounded with the 1.5 * 2^52 trick, whose encoding also gives
// k mod 256 without a conversion), sin and cos of the remainder |r| <= pi/256 from short Taylor
// polynomials (truncation < 5e-18), and the angle-addition formulas with sin, cos of k h read
// from a 256-entry table of correctly rounded values held in LDS (4 KiB): about 16 VALU
// instructions and one LDS read instead of ~37, error within ~2 ulp.  sincos_tab_s (arguments
// of a few radians) takes one reduction step; sincos_tab_sum takes a + x with a already reduced
// (tab_reduce, both steps) and reduces x with both steps; sincos_tab_gmem reads the table entry
// from constant memory instead of LDS.
// rcp_d: hardware reciprocal estimate refined by two Newton steps (correctly rounded to
// about one ulp).
#pragma once
#include <hip/hip_runtime.h>

#include "trig_table.inc"

namespace pes {

// sin, cos of 2 pi i / kTrigN in LDS; every kernel that uses the table sincos fills it first with
// trig_table_load()
__shared__ double2 g_trig[kTrigN];

__device__ inline void trig_table_load() {
  for (int i = threadIdx.x; i < kTrigN; i += blockDim.x) g_trig[i] = make_double2(kTrigTab[i][0], kTrigTab[i][1]);
  __syncthreads();
}

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

constexpr double kTabScale = 40.74366543152521;      // 256 / (2 pi)
constexpr double kTabStepHi = 0.02454369260617026;   // 2 pi / 256 rounded to double
constexpr double kTabStepLo = 9.567553118338697e-19; // 2 pi / 256 - kTabStepHi, rounded

// 1.5 * 2^52: for |y| < 2^51, y + kRoundMagic rounds y to the nearest integer n and the low
// 32 bits of the result's encoding hold n mod 2^32 (two's complement)
constexpr double kRoundMagic = 6755399441055744.0;

// table part of the table sincos: km = k + kRoundMagic, r = x - k 2 pi / 256 already formed
__device__ inline void sincos_tab_kr(double km, double r, double* s, double* c) {
  const int idx = __double2loint(km) & (kTrigN - 1);  // k mod 256
  const double2 t = g_trig[idx];
  const double r2 = r * r;
  const double sr = fma(r * r2, fma(r2, 1.0 / 120.0, -1.0 / 6.0), r);
  const double cr = fma(r2, fma(r2, fma(r2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
  *s = fma(t.x, cr, t.y * sr);
  *c = fma(t.y, cr, -t.x * sr);
}

// one reduction step: the omitted k (2 pi / 256 - kTabStepHi) is at most 3.9e-17 |x|
// absolute, for arguments of a few radians below the table's own error
__device__ inline void sincos_tab_s(double x, double* s, double* c) {
  const double km = fma(x, kTabScale, kRoundMagic);
  const double k = km - kRoundMagic;
  sincos_tab_kr(km, fma(-k, kTabStepHi, x), s, c);
}

// x = k 2 pi / kTrigN + r with k the nearest integer to x kTrigN / (2 pi) (|x| < 5e13): r and
// the low 32 bits of k
__device__ inline void tab_reduce(double x, double& r, int& k_lo) {
  const double km = fma(x, kTabScale, kRoundMagic);
  const double k = km - kRoundMagic;
  r = fma(-k, kTabStepLo, fma(-k, kTabStepHi, x));
  k_lo = __double2loint(km);
}

// sin, cos of a + x with a given reduced (a = ka 2 pi / kTrigN + ra, ka's low bits): x is
// reduced on its own and the two remainders added, so a + x is never rounded; the remainder
// is then up to 2 pi / kTrigN, so the sine series takes one more term
__device__ inline void sincos_tab_sum(double ra, int ka_lo, double x, double* s, double* c) {
  double rx;
  int kx_lo;
  tab_reduce(x, rx, kx_lo);
  const double r = ra + rx;
  const double2 t = g_trig[(ka_lo + kx_lo) & (kTrigN - 1)];
  const double r2 = r * r;
  const double sr = fma(r * r2, fma(r2, fma(r2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0), r);
  const double cr = fma(r2, fma(r2, fma(r2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
  *s = fma(t.x, cr, t.y * sr);
  *c = fma(t.y, cr, -t.x * sr);
}

// sin, cos of x (|x| < 5e13) with the table entry read from global (constant) memory instead of
// LDS, for kernels that keep no LDS copy of the table: both reduction steps, the polynomials of
// sincos_tab_kr (|r| <= pi / kTrigN)
__device__ inline void sincos_tab_gmem(double x, double* s, double* c) {
  double r;
  int k_lo;
  tab_reduce(x, r, k_lo);
  const int idx = k_lo & (kTrigN - 1);
  const double ts = kTrigTab[idx][0], tc = kTrigTab[idx][1];
  const double r2 = r * r;
  const double sr = fma(r * r2, fma(r2, 1.0 / 120.0, -1.0 / 6.0), r);
  const double cr = fma(r2, fma(r2, fma(r2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
  *s = fma(ts, cr, tc * sr);
  *c = fma(tc, cr, -ts * sr);
}

__device__ inline double rcp_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Hardware FP32 sine/cosine of a small FP64 argument (|x| well below 256 revolutions; v_sin_f32
// / v_cos_f32 take revolutions), widened to FP64: absolute error of a few 1e-7.
__device__ inline void sincos_f32(double x, double* s, double* c) {
  const float r = (float)(x * kInv2Pi);
  *s = (double)__builtin_amdgcn_sinf(r);
  *c = (double)__builtin_amdgcn_cosf(r);
}

// Hardware FP32 reciprocal widened to FP64 (relative error ~1 FP32 ulp).
__device__ inline double rcp_f32(double x) { return (double)__builtin_amdgcn_rcpf((float)x); }

// FP64 argument reduced to revolutions in FP64, then the hardware FP32 sine/cosine
// (v_sin_f32 / v_cos_f32 take revolutions), with the rounding of the argument to FP32 undone
// to first order and the pair brought back to unit modulus by one Newton step: what remains
// is the FP32 units' own phase error (a few 1e-8 rad)
__device__ inline void sincos_rev_fc(double x, double* s, double* c) {
  double r = x * kInv2Pi;
  r -= rint(r);
  const float rf = (float)r;
  const double dr = (r - (double)rf) * (2.0 * 3.14159265358979323846);
  const double sf = (double)__builtin_amdgcn_sinf(rf), cf = (double)__builtin_amdgcn_cosf(rf);
  const double s1 = fma(cf, dr, sf), c1 = fma(-sf, dr, cf);
  const double nrm = fma(-0.5, fma(s1, s1, c1 * c1), 1.5);
  *s = s1 * nrm;
  *c = c1 * nrm;
}

}  // namespace pes

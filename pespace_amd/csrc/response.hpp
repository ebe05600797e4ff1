// Detector side of the hot path on the device: analytic constellation orbits,
// Marsat-Baker FD single links and FD Michelson TDI.
//
// orbit:  pespace/detector/orbit.py:189-228, 381-419 (heliocentric),
//         orbit.py:128-139, 278-322 (geocentric)
// links:  pespace/detector/antenna.py:429-460, 496-543
// TDI:    pespace/detector/tdi.py:1683-1703, 1725-1793
#pragma once
#include "phenom.hpp"

namespace pes {

constexpr double AU_SEC = AU_SI / C_SI;
constexpr double SQRT3 = 1.7320508075688772935274463415058723669428;

// Per-detector constants (uploaded once per context).
struct DetectorDev {
  int kind;          // 0 heliocentric, 1 geocentric
  int tdi_gen;       // 15 or 20
  int orthogonal;    // 1: A,E,T  0: X,Y,Z
  int n_bins;        // real bins
  int bin_offset;    // offset of this detector's first bin in the concatenated arrays
  int n_tiles;       // padded bins / TILE
  double L;          // arm length [s]
  double rev0, rot0;
  double rp;         // helio r' = L/(2 sqrt3); geo r = L/sqrt3
  double ck[3], sk[3];                  // helio node phases kappa_i
  double slam, clam, sbet, cbet, omega_rot;  // geo reference source / rotation
  double invL;                          // 1 / L
  // constellation offsets xd_i in the fused kernel, products precomputed per detector:
  // heliocentric  xd_i = (sc hsk_i - (1+sa^2) hck_i, sc hck_i - (1+ca^2) hsk_i,
  //                      -sqrt3 (ca hck_i + sa hsk_i)) with sa, ca of the revolution angle, sc = sa ca;
  // geocentric    xd_i = (gxs sk_i + gxc ck_i, gys sk_i + gyc ck_i, gzs sk_i) with the rotation
  //               angle of node i
  double hck[3], hsk[3];
  double gxs, gxc, gys, gyc, gzs;
  // link projections n_a.v = (sqrt3/L) (al_a v.V - be_a v.U) with xd_i = al_i U + be_i V:
  // c3L = sqrt3/L; heliocentric hsa = c3L hsk, hca = c3L hck; geocentric U, V scaled by c3L
  double c3L;
  double hsa[3], hca[3];
  double sgxs, sgxc, sgys, sgyc, sgzs;
  // observation duty cycle (duty = 0: always observing).  The detector observes for duty_on
  // seconds of every duty_period, the windows starting at duty_t0 + n duty_period, with smooth
  // ramps of duty_taper seconds inside each window; the response of a mode at a bin is weighted
  // by the window at its time t = t_f + t_c
  int duty;
  double duty_period, duty_on, duty_t0, duty_taper, duty_inv_period, duty_inv_taper;
};

// The duty window w(t) and dw/dt: s = (t - t0) mod period; w = S(s / taper) on the rising
// ramp, 1 on the plateau, S((on - s) / taper) on the falling ramp, 0 while off, with the C1
// smoothstep S(x) = x^2 (3 - 2 x) (taper 0: a hard window, dw/dt = 0)
PES_HD void duty_window(const DetectorDev& d, double t, double& w, double& dw) {
  const double u = (t - d.duty_t0) * d.duty_inv_period;
  const double s = (u - floor(u)) * d.duty_period;
  w = 0.0;
  dw = 0.0;
  if (s >= d.duty_on) return;
  const double up = s, dn = d.duty_on - s;  // distances into the window from its two ends
  const double x = d.duty_inv_taper > 0.0 ? fmin(up, dn) * d.duty_inv_taper : 1.0;  // 0: no ramps
  if (x >= 1.0) {
    w = 1.0;
    return;
  }
  w = x * x * (3.0 - 2.0 * x);
  const double g = 6.0 * x * (1.0 - x) * d.duty_inv_taper;
  dw = up <= dn ? g : -g;
}

struct ModeList {  // the modes of a context (ids into kModeM etc.), passed to kernels by value
  int n;
  int id[6];
};

// Host side: the orbit constants of a detector (heliocentric orbit.py:189-228; geocentric
// orbit.py:112-139 with the given rotation rate and reference source); arguments validated by
// the caller.
inline void orbit_constants(DetectorDev& d, int orbit_kind, double arm_length_m, double rotation_initial,
                            double revolution_initial, double omega_rotation, double lambda_ref, double beta_ref) {
  d.kind = orbit_kind;
  d.L = arm_length_m / C_SI;
  d.rot0 = rotation_initial;
  d.rev0 = revolution_initial;
  d.invL = 1.0 / d.L;
  if (orbit_kind == 0) {  // orbit.py:189-228
    d.rp = d.L / (2.0 * std::sqrt(3.0));
    double k1 = rotation_initial, k2 = k1 + 2.0 * PI / 3.0, k3 = k2 + 2.0 * PI / 3.0;
    d.ck[0] = std::cos(k1); d.sk[0] = std::sin(k1);
    d.ck[1] = std::cos(k2); d.sk[1] = std::sin(k2);
    d.ck[2] = std::cos(k3); d.sk[2] = std::sin(k3);
    for (int i = 0; i < 3; ++i) {
      d.hck[i] = d.rp * d.ck[i];
      d.hsk[i] = d.rp * d.sk[i];
    }
    d.c3L = std::sqrt(3.0) * d.invL;
    for (int i = 0; i < 3; ++i) {
      d.hca[i] = d.c3L * d.hck[i];
      d.hsa[i] = d.c3L * d.hsk[i];
    }
  } else {  // orbit.py:77-139
    d.rp = d.L / std::sqrt(3.0);
    d.slam = std::sin(lambda_ref); d.clam = std::cos(lambda_ref);
    d.sbet = std::sin(beta_ref); d.cbet = std::cos(beta_ref);
    d.omega_rot = omega_rotation;
    d.gxs = d.rp * d.sbet * d.clam;
    d.gxc = d.rp * d.slam;
    d.gys = d.rp * d.sbet * d.slam;
    d.gyc = -d.rp * d.clam;
    d.gzs = -d.rp * d.cbet;
    d.c3L = std::sqrt(3.0) * d.invL;
    d.sgxs = d.c3L * d.gxs;
    d.sgxc = d.c3L * d.gxc;
    d.sgys = d.c3L * d.gys;
    d.sgyc = d.c3L * d.gyc;
    d.sgzs = d.c3L * d.gzs;
  }
}

struct Vec3 {
  double x, y, z;
};
PES_HD double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// node positions x_i (SSB, seconds) and link unit vectors n_a at time t
__device__ inline void constellation(const DetectorDev& d, double t, Vec3 n[3], Vec3 x[3]) {
  const double omega = 2.0 * PI / YEAR_SI;
  double alpha = omega * t + d.rev0;
  double sa, ca;
  sincos(alpha, &sa, &ca);
  Vec3 xd[3];
  if (d.kind == 0) {
    double sc = sa * ca, sa2 = 1.0 + sa * sa, ca2 = 1.0 + ca * ca;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xd[i].x = d.rp * (sc * d.sk[i] - sa2 * d.ck[i]);
      xd[i].y = d.rp * (sc * d.ck[i] - ca2 * d.sk[i]);
      xd[i].z = d.rp * (-SQRT3 * (ca * d.ck[i] + sa * d.sk[i]));
    }
  } else {
    double k1 = d.omega_rot * t + d.rot0;
    double s1, c1;
    sincos(k1, &s1, &c1);
    // kappa_2,3 = kappa_1 + 2pi/3, + 4pi/3 by rotation
    const double cr = -0.5, sr = 0.5 * SQRT3;
    double sk[3] = {s1, s1 * cr + c1 * sr, 0.0};
    double ck[3] = {c1, c1 * cr - s1 * sr, 0.0};
    sk[2] = sk[1] * cr + ck[1] * sr;
    ck[2] = ck[1] * cr - sk[1] * sr;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xd[i].x = d.rp * (d.sbet * d.clam * sk[i] + d.slam * ck[i]);
      xd[i].y = d.rp * (d.sbet * d.slam * sk[i] - d.clam * ck[i]);
      xd[i].z = d.rp * (-d.cbet * sk[i]);
    }
  }
  const double invL = d.invL;  // 1 / L, rounded on the host exactly as a device division would
  n[0] = {(xd[2].x - xd[1].x) * invL, (xd[2].y - xd[1].y) * invL, (xd[2].z - xd[1].z) * invL};
  n[1] = {(xd[0].x - xd[2].x) * invL, (xd[0].y - xd[2].y) * invL, (xd[0].z - xd[2].z) * invL};
  n[2] = {(xd[1].x - xd[0].x) * invL, (xd[1].y - xd[0].y) * invL, (xd[1].z - xd[0].z) * invL};
  double x0x = AU_SEC * ca, x0y = AU_SEC * sa;
#pragma unroll
  for (int i = 0; i < 3; ++i) x[i] = {xd[i].x + x0x, xd[i].y + x0y, xd[i].z};
}

struct Cplx {
  double re, im;
};
PES_HD Cplx cmul(Cplx a, Cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
PES_HD Cplx cadd(Cplx a, Cplx b) { return {a.re + b.re, a.im + b.im}; }
PES_HD Cplx csub(Cplx a, Cplx b) { return {a.re - b.re, a.im - b.im}; }
PES_HD Cplx cscale(Cplx a, double s) { return {a.re * s, a.im * s}; }

__device__ inline double sinc(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

// six links 12,21,23,32,31,13
struct Links {
  Cplx l12, l21, l23, l32, l31, l13;
};

// Add one mode's single-link response at frequency f (antenna.py:496-543).
// hp, hc already include the t_c shift; t = tf + t_c.
__device__ inline void add_mode_links(const DetectorDev& d, double f, double t, Cplx hp, Cplx hc, const Vec3& p,
                                      const Vec3& q, const Vec3& k, Links& L) {
  if (d.duty) {  // the detector's duty window at the mode's time t = t_f + t_c
    double w, dw;
    duty_window(d, t, w, dw);
    hp = {hp.re * w, hp.im * w};
    hc = {hc.re * w, hc.im * w};
  }
  Vec3 n[3], x[3];
  constellation(d, t, n, x);
  Cplx nhn[3];
  double kn[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double np = dot(n[a], p), nq = dot(n[a], q);
    double Pp = np * np - nq * nq;  // n.e+.n
    double Pc = 2.0 * np * nq;      // n.ex.n
    nhn[a] = {Pp * hp.re + Pc * hc.re, Pp * hp.im + Pc * hc.im};
    kn[a] = dot(k, n[a]);
  }
  double kx12 = k.x * (x[0].x + x[1].x) + k.y * (x[0].y + x[1].y) + k.z * (x[0].z + x[1].z);
  double kx23 = k.x * (x[1].x + x[2].x) + k.y * (x[1].y + x[2].y) + k.z * (x[1].z + x[2].z);
  double kx31 = k.x * (x[2].x + x[0].x) + k.y * (x[2].y + x[0].y) + k.z * (x[2].z + x[0].z);
  double xL = PI * f * d.L;
  double s32 = sinc(xL * (1.0 - kn[0])), s23 = sinc(xL * (1.0 + kn[0]));
  double s13 = sinc(xL * (1.0 - kn[1])), s31 = sinc(xL * (1.0 + kn[1]));
  double s21 = sinc(xL * (1.0 - kn[2])), s12 = sinc(xL * (1.0 + kn[2]));
  double pf = -PI * f;
  Cplx e12, e23, e31;
  sincos(pf * (d.L + kx12), &e12.im, &e12.re);
  sincos(pf * (d.L + kx23), &e23.im, &e23.re);
  sincos(pf * (d.L + kx31), &e31.im, &e31.re);
  // pre = -i xL:  (-i xL) * z = (xL z.im, -xL z.re)
  Cplx c3 = cmul({xL * nhn[2].im, -xL * nhn[2].re}, e12);
  Cplx c1 = cmul({xL * nhn[0].im, -xL * nhn[0].re}, e23);
  Cplx c2 = cmul({xL * nhn[1].im, -xL * nhn[1].re}, e31);
  L.l12 = cadd(L.l12, cscale(c3, s12));
  L.l21 = cadd(L.l21, cscale(c3, s21));
  L.l23 = cadd(L.l23, cscale(c1, s23));
  L.l32 = cadd(L.l32, cscale(c1, s32));
  L.l31 = cadd(L.l31, cscale(c2, s31));
  L.l13 = cadd(L.l13, cscale(c2, s13));
}

// TDI (tdi.py:1683-1793): returns the three output channels (A,E,T or X,Y,Z).
__device__ inline void tdi_combine(const DetectorDev& d, double f, const Links& l, Cplx out[3]) {
  Cplx D;
  sincos(-2.0 * PI * f * d.L, &D.im, &D.re);
  Cplx D2 = cmul(D, D);
  Cplx pf = {1.0 - D2.re, -D2.im};
  if (d.tdi_gen == 20) {
    Cplx D4 = cmul(D2, D2);
    Cplx D6 = cmul(D4, D2);
    pf = {1.0 - D2.re - D4.re + D6.re, -D2.im - D4.im + D6.im};
  }
  Cplx X = cmul(pf, cadd(csub(l.l13, l.l12), cmul(D, csub(l.l31, l.l21))));
  Cplx Y = cmul(pf, cadd(csub(l.l21, l.l23), cmul(D, csub(l.l12, l.l32))));
  Cplx Z = cmul(pf, cadd(csub(l.l32, l.l31), cmul(D, csub(l.l23, l.l13))));
  if (d.orthogonal) {
    const double r2 = 1.0 / sqrt(2.0), r6 = 1.0 / sqrt(6.0), r3 = 1.0 / sqrt(3.0);
    out[0] = cscale(csub(Z, X), r2);
    out[1] = cscale(cadd(csub(X, cscale(Y, 2.0)), Z), r6);
    out[2] = cscale(cadd(cadd(X, Y), Z), r3);
  } else {
    out[0] = X;
    out[1] = Y;
    out[2] = Z;
  }
}

}  // namespace pes

// FP64 C++/OpenMP restatement of the reference's per-bin likelihood, the CPU baseline of the
// benchmark.  TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and tests/ call it; the
// product never does.
//
// What it computes, per parameter row and detector bin (the algorithm of the reference's taichi
// backend, on the same interpolated waveform as the GPU path):
//   * waveform of every mode from node tables on a log grid: cubic Hermite in u = ln f of
//     A, Psi (and t_f = dPsi/df / 2 pi), h+ = K+ A e^{-i Psi} / d_L, hx = Kx A e^{-i Psi} / d_L
//     (interface_ripple.py:42-51 container contract), cut above the mode's cut frequency;
//   * t_c shift h e^{-2 pi i f t_c}, t = t_f + t_c (antenna.py:434-449);
//   * constellation at t: heliocentric orbit.py:363-419, geocentric orbit.py:260-322;
//   * polarization basis p, q, k (antenna.py:70-83, 112-114), n.e+.n = (n.p)^2 - (n.q)^2,
//     n.ex.n = 2 (n.p)(n.q);
//   * six Marsat-Baker links summed over modes (antenna.py:496-543);
//   * Michelson TDI 1.5 / 2.0, A, E, T = (X+Y+Z)/sqrt3 or X, Y, Z (tdi.py:1683-1793);
//   * Whittle sum -sum_c w_c |d_c - h_c|^2 with w_c = 2 df / S_c (likelihood_ti.py:58-67).
// The gradient comes from forward-mode dual numbers over the 11 row parameters (T = Dual<11>);
// the intrinsic parameters (chirp mass, mass ratio, spins) enter through node-table derivatives
// supplied by the caller, the others through the closed-form dependence above.
//
// Build: g++ -O3 -march=x86-64-v3 -ffast-math -fopenmp -shared -fPIC (oracle/build_cpu.py; fast
// math as the taichi backend compiles by default).
#include <omp.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double C_SI = 299792458.0;
constexpr double AU_SEC = 149597870700.0 / C_SI;
constexpr double YEAR_SI = 31557600.0;
constexpr double PC_SI = 3.085677581491367e16;
constexpr int NP = 11;

// ---- scalar types: double, and forward-mode duals over the 11 parameters -------------------
template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() : v(0.0) { std::memset(d, 0, sizeof d); }
  Dual(double x) : v(x) { std::memset(d, 0, sizeof d); }  // NOLINT: implicit lift of constants
};
template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, double s) {
  Dual<N> r(a.v * s);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
template <int N>
inline Dual<N> operator*(double s, const Dual<N>& a) { return a * s; }
template <int N>
inline Dual<N> operator+(const Dual<N>& a, double s) {
  Dual<N> r = a;
  r.v += s;
  return r;
}
template <int N>
inline Dual<N> operator+(double s, const Dual<N>& a) { return a + s; }
template <int N>
inline Dual<N> operator-(double s, const Dual<N>& a) { return (-a) + s; }
template <int N>
inline Dual<N> operator-(const Dual<N>& a, double s) { return a + (-s); }
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const double ib = 1.0 / b.v;
  Dual<N> r(a.v * ib);
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
inline void sincos_t(const Dual<N>& x, Dual<N>& s, Dual<N>& c) {
  const double sv = std::sin(x.v), cv = std::cos(x.v);
  s = Dual<N>(sv);
  c = Dual<N>(cv);
  for (int i = 0; i < N; ++i) {
    s.d[i] = cv * x.d[i];
    c.d[i] = -sv * x.d[i];
  }
}
inline void sincos_t(double x, double& s, double& c) {
  s = std::sin(x);
  c = std::cos(x);
}
inline double value(double x) { return x; }
template <int N>
inline double value(const Dual<N>& x) { return x.v; }

// sin(x)/x, 1 at 0 (utils.py:60-79)
inline double sinc_t(double x) { return x == 0.0 ? 1.0 : std::sin(x) / x; }
template <int N>
inline Dual<N> sinc_t(const Dual<N>& x) {
  double v, dv;
  if (std::fabs(x.v) < 1e-4) {
    const double x2 = x.v * x.v;
    v = 1.0 - x2 / 6.0;
    dv = -x.v / 3.0;
  } else {
    const double s = std::sin(x.v), c = std::cos(x.v);
    v = s / x.v;
    dv = (c - v) / x.v;
  }
  Dual<N> r(v);
  for (int i = 0; i < N; ++i) r.d[i] = dv * x.d[i];
  return r;
}

template <class T>
struct Cx {
  T re, im;
};
template <class T>
inline Cx<T> cadd(const Cx<T>& a, const Cx<T>& b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
inline Cx<T> csub(const Cx<T>& a, const Cx<T>& b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
inline Cx<T> cmul(const Cx<T>& a, const Cx<T>& b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <class T>
inline Cx<T> cscale(const Cx<T>& a, const T& s) { return {a.re * s, a.im * s}; }
template <class T>
inline Cx<T> cmuld(const Cx<T>& a, double br, double bi) { return {a.re * br - a.im * bi, a.re * bi + a.im * br}; }
template <class T>
inline Cx<T> cexpi(const T& phase) {  // exp(i phase)
  T s, c;
  sincos_t(phase, s, c);
  return {c, s};
}

template <class T>
struct V3 {
  T x, y, z;
};
template <class T, class U>
inline T dot(const V3<T>& a, const V3<U>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// ---- inputs -----------------------------------------------------------------------------
struct Det {
  int kind, gen, orth, n, off;  // off: first bin in the concatenated arrays
  double L, rot0, rev0, omega_rot, slam, clam, sbet, cbet;
  int duty;                                      // observation duty cycle (0: always observing)
  double duty_period, duty_on, duty_t0, duty_taper;
};

// The duty window at time t (oracle/detector.py duty_window): s = (t - t0) mod period, the
// smoothstep S(x) = x^2 (3 - 2x) of the distance into [0, on) over the taper, 0 outside; the
// window's position s carries t's derivatives (the floor does not)
template <class T>
T duty_w(const Det& d, const T& t) {
  const double u = (value(t) - d.duty_t0) / d.duty_period;
  const T s = t - (d.duty_t0 + std::floor(u) * d.duty_period);
  const double sv = value(s);
  if (sv >= d.duty_on) return T(0.0);
  if (d.duty_taper <= 0.0) return T(1.0);
  const T x = (sv <= d.duty_on - sv ? s : T(d.duty_on) - s) * (1.0 / d.duty_taper);
  if (value(x) >= 1.0) return T(1.0);
  return x * x * (T(3.0) - x * 2.0);
}

struct Model {
  int n_modes, n_nodes;
  double u0, du;
  const double* nodes;       // [n_modes][n_nodes][4]: A, dA/du, Psi, dPsi/du
  const double* nodes_grad;  // [n_modes][n_nodes][4][4]: d/d(Mc, q, chi1, chi2) of the same four, or NULL
  const double* kfac;        // [n_modes][4]: K+ (re, im), Kx (re, im), mode phase factor included
  const double* kfac_d;      // [n_modes][2][4]: d/d iota, d/d phi_ref of the same four
  const double* mode_cut;    // [n_modes] Hz
};

// the parameter row as scalars of type T (seeds for the duals)
template <class T>
inline T seed(const double* params, int i) { return T(params[i]); }
template <>
inline Dual<NP> seed<Dual<NP>>(const double* params, int i) {
  Dual<NP> r(params[i]);
  r.d[i] = 1.0;
  return r;
}

// one value of the node tables with its intrinsic-parameter derivatives (components 0..3)
template <class T>
inline T node_val(double v, const double* g, int k) { (void)g; (void)k; return T(v); }
template <>
inline Dual<NP> node_val<Dual<NP>>(double v, const double* g, int k) {
  Dual<NP> r(v);
  if (g)
    for (int i = 0; i < 4; ++i) r.d[i] = g[4 * i + k];
  return r;
}

template <class T>
void constellation(const Det& d, const T& t, V3<T> n[3], V3<T> x[3]) {
  T sa, ca;
  sincos_t((2.0 * PI / YEAR_SI) * t + d.rev0, sa, ca);
  V3<T> xd[3];
  if (d.kind == 0) {  // orbit.py:189-228, 381-419
    const double rp = d.L / (2.0 * std::sqrt(3.0));
    for (int i = 0; i < 3; ++i) {
      const double k = d.rot0 + i * 2.0 * PI / 3.0, ck = std::cos(k), sk = std::sin(k);
      xd[i] = {rp * (sa * ca * sk - (1.0 + sa * sa) * ck), rp * (sa * ca * ck - (1.0 + ca * ca) * sk),
               rp * (-std::sqrt(3.0)) * (ca * ck + sa * sk)};
    }
  } else {  // orbit.py:112-139, 278-322
    const double r = d.L / std::sqrt(3.0);
    for (int i = 0; i < 3; ++i) {
      T sk, ck;
      sincos_t(d.omega_rot * t + (d.rot0 + i * 2.0 * PI / 3.0), sk, ck);
      xd[i] = {r * (d.sbet * d.clam * sk + d.slam * ck), r * (d.sbet * d.slam * sk - d.clam * ck), r * (-d.cbet) * sk};
    }
  }
  const double iL = 1.0 / d.L;
  n[0] = {(xd[2].x - xd[1].x) * iL, (xd[2].y - xd[1].y) * iL, (xd[2].z - xd[1].z) * iL};
  n[1] = {(xd[0].x - xd[2].x) * iL, (xd[0].y - xd[2].y) * iL, (xd[0].z - xd[2].z) * iL};
  n[2] = {(xd[1].x - xd[0].x) * iL, (xd[1].y - xd[0].y) * iL, (xd[1].z - xd[0].z) * iL};
  for (int i = 0; i < 3; ++i) x[i] = {xd[i].x + AU_SEC * ca, xd[i].y + AU_SEC * sa, xd[i].z};
}

// lnL of one row on the bins [b0, b1) of every detector, stride s (weights times s)
template <class T>
T lnl_row(const std::vector<Det>& dets, const double* f_all, const double* w_all, const double* dre_all,
          const double* dim_all, int NB, const Model& M, const double* params, int stride, int n_threads) {
  const T dL = seed<T>(params, 4), phi = seed<T>(params, 5), iota = seed<T>(params, 6);
  const T lam = seed<T>(params, 7), bet = seed<T>(params, 8), psi = seed<T>(params, 9), tc = seed<T>(params, 10);
  (void)phi;
  (void)iota;
  T sl, cl, sb, cb, sp, cp;
  sincos_t(lam, sl, cl);
  sincos_t(bet, sb, cb);
  sincos_t(psi, sp, cp);
  const V3<T> p = {sl * cp - cl * sb * sp, T(0.0) - cl * cp - sl * sb * sp, cb * sp};
  const V3<T> q = {T(0.0) - sl * sp - cl * sb * cp, cl * sp - sl * sb * cp, cb * cp};
  const V3<T> k = {T(0.0) - cb * cl, T(0.0) - cb * sl, T(0.0) - sb};
  const T inv_dist = T(1.0) / (dL * (1.0e6 * PC_SI));
  // harmonic factors with their iota and phi_ref derivatives
  std::vector<Cx<T>> Kp(M.n_modes), Kc(M.n_modes);
  for (int m = 0; m < M.n_modes; ++m) {
    const double* kf = M.kfac + 4 * m;
    T v[4];
    for (int j = 0; j < 4; ++j) {
      v[j] = T(kf[j]);
      if constexpr (!std::is_same<T, double>::value) {
        v[j].d[6] = M.kfac_d[8 * m + j];
        v[j].d[5] = M.kfac_d[8 * m + 4 + j];
      }
    }
    Kp[m] = {v[0] * inv_dist, v[1] * inv_dist};
    Kc[m] = {v[2] * inv_dist, v[3] * inv_dist};
  }
  T total(0.0);
  const int NN = M.n_nodes;
#pragma omp parallel num_threads(n_threads)
  {
    T acc(0.0);
#pragma omp for schedule(dynamic, 64)
    for (int bi = 0; bi < NB; bi += stride) {
      int di = 0;
      while (bi >= dets[di].off + dets[di].n) ++di;
      const Det& d = dets[di];
      const double f = f_all[bi];
      const double x = (std::log(f) - M.u0) / M.du;
      int j = (int)std::floor(x);
      j = j < 0 ? 0 : (j > NN - 2 ? NN - 2 : j);
      const double s = x - j, s1 = 1.0 - s;
      const double h00 = (1.0 + 2.0 * s) * s1 * s1, h10 = s * s1 * s1, h01 = s * s * (3.0 - 2.0 * s),
                   h11 = -s * s * s1;
      const double d00 = 6.0 * s * (s - 1.0), d10 = s1 * (1.0 - 3.0 * s), d01 = -d00, d11 = s * (3.0 * s - 2.0);
      Cx<T> ex = cexpi(T(-2.0 * PI * f) * tc);  // the t_c shift
      Cx<T> lk[6];
      for (int i = 0; i < 6; ++i) lk[i] = {T(0.0), T(0.0)};
      for (int m = 0; m < M.n_modes; ++m) {
        if (f > M.mode_cut[m]) continue;
        const double* n0 = M.nodes + ((size_t)m * NN + j) * 4;
        const double* g0 = M.nodes_grad ? M.nodes_grad + ((size_t)m * NN + j) * 16 : nullptr;
        const double* g1 = g0 ? g0 + 16 : nullptr;
        const T A0 = node_val<T>(n0[0], g0, 0), dA0 = node_val<T>(n0[1], g0, 1);
        const T P0 = node_val<T>(n0[2], g0, 2), dP0 = node_val<T>(n0[3], g0, 3);
        const T A1 = node_val<T>(n0[4], g1, 0), dA1 = node_val<T>(n0[5], g1, 1);
        const T P1 = node_val<T>(n0[6], g1, 2), dP1 = node_val<T>(n0[7], g1, 3);
        const T A = h00 * A0 + (h10 * M.du) * dA0 + h01 * A1 + (h11 * M.du) * dA1;
        const T Psi = h00 * P0 + (h10 * M.du) * dP0 + h01 * P1 + (h11 * M.du) * dP1;
        const T dPds = d00 * P0 + (d10 * M.du) * dP0 + d01 * P1 + (d11 * M.du) * dP1;
        const T tf = dPds * (1.0 / (2.0 * PI * M.du * f));
        const Cx<T> h = cmul(cexpi(T(0.0) - Psi), ex);  // e^{-i Psi} e^{-2 pi i f t_c}
        const T t = tf + tc;
        const T Aw = d.duty ? A * duty_w(d, t) : A;
        const Cx<T> hA = {h.re * Aw, h.im * Aw};
        const Cx<T> hp = cmul(Kp[m], hA), hc = cmul(Kc[m], hA);
        V3<T> nv[3], xv[3];
        constellation(d, t, nv, xv);
        T nhn_re[3], nhn_im[3], kn[3];
        for (int a = 0; a < 3; ++a) {
          const T np = dot(nv[a], p), nq = dot(nv[a], q);
          const T Pp = np * np - nq * nq, Pc = 2.0 * np * nq;
          nhn_re[a] = Pp * hp.re + Pc * hc.re;
          nhn_im[a] = Pp * hp.im + Pc * hc.im;
          kn[a] = dot(k, nv[a]);
        }
        const T k12 = dot(k, V3<T>{xv[0].x + xv[1].x, xv[0].y + xv[1].y, xv[0].z + xv[1].z});
        const T k23 = dot(k, V3<T>{xv[1].x + xv[2].x, xv[1].y + xv[2].y, xv[1].z + xv[2].z});
        const T k31 = dot(k, V3<T>{xv[2].x + xv[0].x, xv[2].y + xv[0].y, xv[2].z + xv[0].z});
        const double xL = PI * f * d.L;
        // the -i xL prefactor and the delay phases, per pair
        const Cx<T> e12 = cexpi(T(-PI * f) * (k12 + d.L)), e23 = cexpi(T(-PI * f) * (k23 + d.L)),
                    e31 = cexpi(T(-PI * f) * (k31 + d.L));
        const Cx<T> c3 = cmul(Cx<T>{xL * nhn_im[2], T(0.0) - xL * nhn_re[2]}, e12);
        const Cx<T> c1 = cmul(Cx<T>{xL * nhn_im[0], T(0.0) - xL * nhn_re[0]}, e23);
        const Cx<T> c2 = cmul(Cx<T>{xL * nhn_im[1], T(0.0) - xL * nhn_re[1]}, e31);
        // links 12, 21, 23, 32, 31, 13
        lk[0] = cadd(lk[0], cscale(c3, sinc_t(xL * (1.0 + kn[2]))));
        lk[1] = cadd(lk[1], cscale(c3, sinc_t(xL * (1.0 - kn[2]))));
        lk[2] = cadd(lk[2], cscale(c1, sinc_t(xL * (1.0 + kn[0]))));
        lk[3] = cadd(lk[3], cscale(c1, sinc_t(xL * (1.0 - kn[0]))));
        lk[4] = cadd(lk[4], cscale(c2, sinc_t(xL * (1.0 + kn[1]))));
        lk[5] = cadd(lk[5], cscale(c2, sinc_t(xL * (1.0 - kn[1]))));
      }
      // TDI (tdi.py:1683-1793)
      const double Dr = std::cos(2.0 * PI * f * d.L), Di = -std::sin(2.0 * PI * f * d.L);
      const double D2r = Dr * Dr - Di * Di, D2i = 2.0 * Dr * Di;
      double pfr = 1.0 - D2r, pfi = -D2i;
      if (d.gen == 20) {
        const double D4r = D2r * D2r - D2i * D2i, D4i = 2.0 * D2r * D2i;
        const double D6r = D4r * D2r - D4i * D2i, D6i = D4r * D2i + D4i * D2r;
        pfr = 1.0 - D2r - D4r + D6r;
        pfi = -D2i - D4i + D6i;
      }
      auto comb = [&](const Cx<T>& a, const Cx<T>& b, const Cx<T>& c, const Cx<T>& e) {
        return cmuld(cadd(csub(a, b), cmuld(csub(c, e), Dr, Di)), pfr, pfi);
      };
      const Cx<T> X = comb(lk[5], lk[0], lk[4], lk[1]);
      const Cx<T> Y = comb(lk[1], lk[2], lk[0], lk[3]);
      const Cx<T> Z = comb(lk[3], lk[4], lk[2], lk[5]);
      Cx<T> out[3];
      if (d.orth) {
        const T r2(1.0 / std::sqrt(2.0)), r6(1.0 / std::sqrt(6.0)), r3(1.0 / std::sqrt(3.0));
        out[0] = cscale(csub(Z, X), r2);
        out[1] = cscale(cadd(csub(X, cscale(Y, T(2.0))), Z), r6);
        out[2] = cscale(cadd(cadd(X, Y), Z), r3);
      } else {
        out[0] = X;
        out[1] = Y;
        out[2] = Z;
      }
      for (int c = 0; c < 3; ++c) {
        const double w = w_all[(size_t)c * NB + bi] * stride;
        if (w == 0.0) continue;
        const T rr = dre_all[(size_t)c * NB + bi] - out[c].re, ri = dim_all[(size_t)c * NB + bi] - out[c].im;
        acc = acc - w * (rr * rr + ri * ri);
      }
    }
#pragma omp critical
    total = total + acc;
  }
  return total;
}

}  // namespace

extern "C" {

// One row: lnL (and, if grad != NULL, its gradient over the 11 parameters) on the network.
// Detector arrays are concatenated: f_all[NB], w_all / dre_all / dim_all [3][NB] (channel-major,
// w = 2 df / S, 0 for unused channels); detector d owns bins [det_off[d], det_off[d] + det_n[d]).
// det_i: [n_det][4] = kind, tdi generation (15, 20), orthogonal, duty cycle (0 / 1);
// det_d: [n_det][10] = L [s], rot0, rev0, omega_rotation, lambda_ref, beta_ref, and the duty
// cycle's period, on duration, start time, taper [s].
// stride: evaluate every stride-th bin with its weight multiplied by stride (a bounded sample).
int cpu_lnl_row(int n_det, const int* det_i, const double* det_d, const int* det_off, const int* det_n, int NB,
                const double* f_all, const double* w_all, const double* dre_all, const double* dim_all, int n_modes,
                int n_nodes, double u0, double du, const double* nodes, const double* nodes_grad, const double* kfac,
                const double* kfac_d, const double* mode_cut, const double* params, int stride, int n_threads,
                double* lnl, double* grad) {
  if (n_det < 1 || NB < 1 || n_modes < 1 || n_nodes < 2 || stride < 1 || n_threads < 1 || !lnl) return 1;
  std::vector<Det> dets(n_det);
  for (int i = 0; i < n_det; ++i) {
    Det& d = dets[i];
    d.kind = det_i[4 * i];
    d.gen = det_i[4 * i + 1];
    d.orth = det_i[4 * i + 2];
    d.off = det_off[i];
    d.n = det_n[i];
    const double* dd = det_d + 10 * i;
    d.L = dd[0];
    d.rot0 = dd[1];
    d.rev0 = dd[2];
    d.omega_rot = dd[3];
    d.slam = std::sin(dd[4]);
    d.clam = std::cos(dd[4]);
    d.sbet = std::sin(dd[5]);
    d.cbet = std::cos(dd[5]);
    d.duty = det_i[4 * i + 3];
    d.duty_period = dd[6];
    d.duty_on = dd[7];
    d.duty_t0 = dd[8];
    d.duty_taper = dd[9];
  }
  Model M{n_modes, n_nodes, u0, du, nodes, nodes_grad, kfac, kfac_d, mode_cut};
  if (grad) {
    if (!nodes_grad || !kfac_d) return 1;
    const Dual<NP> r = lnl_row<Dual<NP>>(dets, f_all, w_all, dre_all, dim_all, NB, M, params, stride, n_threads);
    *lnl = r.v;
    for (int i = 0; i < NP; ++i) grad[i] = r.d[i];
  } else {
    *lnl = lnl_row<double>(dets, f_all, w_all, dre_all, dim_all, NB, M, params, stride, n_threads);
  }
  return 0;
}

}  // extern "C"

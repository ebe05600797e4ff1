// Fused per-(bin, row) forward response and adjoint.
//
// The forward response, the residual and lnL are FP64 in every precision mode (PREC template
// parameter, PES_PREC_* values), so lnL does not depend on the mode.  The modes differ in the
// adjoint's recompute of each mode's response (the gradient only):
//   0 FP64: every sine and cosine from the FP64 table sincos, every reciprocal from rcp_d; the
//     phasor of the total phase Psi + 2 pi f (t_c + k.x0) (the waveform phase, the t_c shift and
//     the barycentric delay, antenna.py:435, 530-533, which the reference applies as three
//     factors) from one table sincos of 2 pi f t_c (reduced once per bin and row) plus
//     Psi + 2 pi f k.x0 (reduced per mode), the remainders added: the total is never rounded and
//     each term keeps the rounding it has as a separate factor.
//   1 MIXED: in the adjoint's recompute, the pair phasors come from the FP32 hardware
//     sine/cosine and the sinc denominators from the FP32 reciprocal; the orbit angle keeps the
//     FP64 table sincos (its error is multiplied by 2 pi f AU/c ~ 150 rad in the delay phase)
//     and all algebra stays FP64.  The gradient then carries FP32-level relative errors (~1e-7).
// In both modes the gradient kernel's adjoint takes each mode's total-phase phasor from the
// forward pass (kept in LDS, mode_geom UPH), so that phase is formed once per (bin, mode, row).
// (The response-only kernels in the mixed mode form it from the FP32 hardware sine/cosine of the
// phase summed and reduced in FP64.)
// Why FP32 goes no further: without packed FP32 (off by build rule) an FP32 FMA issues at the
// FP64 rate on gfx950, and the mode loops are ~90% FMA/add/mul, so FP32 pays only in the
// transcendental units; an FP32 forward would bias lnL coherently (per-row constants and the
// slowly varying transfer functions round the same way over many bins: |dlnL| ~ eps SNR^2,
// ~0.4 far from the peak), far outside the |dlnL| < 1e-3 bound.
//
// Algebraic rearrangements relative to antenna.py:496-543 (same values):
//   * the constellation offsets sum to zero, so k.(x_i + x_j) = 2 k.x0 - k.xd_a for link
//     pair a with nodes (i, j) and third node a: exp(-i pi f (L + k.(x_i + x_j))) =
//     exp(-i pi f L) exp(-2 pi i f k.x0) z_a with z_a = exp(+i pi f k.xd_a);
//   * z_0 z_1 z_2 = 1 (the offsets sum to zero), so z_2 = conj(z_0 z_1): two small sincos per
//     mode give all three pair phasors;
//   * sinc(xL(1 +- k.n_a)): sin(xL +- a) = sin xL cos a +- cos xL sin a with
//     a = xL k.n_a = pi f k.(xd_i - xd_j), so exp(i a) = z_i conj(z_j) is a product of the
//     pair phasors (no further sincos);
//   * the factor (-i xL) exp(-i xL) common to every link term of every mode is applied once
//     per bin, inside the TDI prefactor;
//   * n_a.v = (xd_i.v - xd_j.v)/L for v in {p, q, k} from six scalar products per mode (the
//     node offsets are two fixed vectors weighted by equally spaced phases, see ModeG);
//   * n.e+.n = (n.p)^2 - (n.q)^2, n.ex.n = 2 (n.p)(n.q) with e+ = pp - qq, ex = pq + qp (the
//     factor 2 is carried by the stored cross factor);
//   * T = (X + Y + Z)/sqrt3 from the differences of the two links of each pair, and its
//     adjoint folded into the link adjoints.
// This header is included inside the anonymous namespace of pespace_amd.hip.

struct Cx {
  double re, im;
};
__device__ inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline Cx sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline Cx scale(Cx a, double s) { return {a.re * s, a.im * s}; }
__device__ inline Cx conjx(Cx a) { return {a.re, -a.im}; }
__device__ inline Cx axpy(double s, Cx a, Cx y) { return {fma(s, a.re, y.re), fma(s, a.im, y.im)}; }

struct V3 {
  double x, y, z;
};
__device__ inline double dot3(const V3& a, const V3& b) { return fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)); }

template <int PREC>
struct Trig {
  // FP32 hardware sine/cosine for the pair phasors and the total phase, FP32 reciprocals for the
  // sinc denominators (the adjoint's recompute in the mixed mode; the forward is always PREC 0)
  static constexpr bool kFp32 = PREC >= 1;
};

// per-bin quantities shared by every row and mode of a tile lane (kept small: they stay in
// registers for all rows of the tile)
struct BinPre {
  double f, pif, xL, sxL, cxL;  // f, pi f, pi f L, sin and cos of pi f L
  // cubic-Hermite weights on [node j, node j+1] at this bin, u = ln f, s = (u - u_j)/du:
  //   value = v0 + hb1 (v1 - v0) + hb2 v0' + hb3 v1'      (v' = dv/du)
  //   t_f   = he1 (P1 - P0) + he2 P0' + he3 P1'            (dPsi/df / 2 pi)
  double hb1, hb2, hb3, he1, he2, he3;
  int j;
};

__device__ inline void bin_prepare(const DetectorDev& det, double f, double s, int j, double du, BinPre& b) {
  b.f = f;
  b.j = j;
  b.pif = PI * f;
  const double inv2pidf = 1.0 / (2.0 * PI * du * f);
  const double s1 = 1.0 - s;
  b.hb1 = s * s * (3.0 - 2.0 * s);
  b.hb2 = s * s1 * s1 * du;
  b.hb3 = -s * s * s1 * du;
  b.he1 = 6.0 * s * s1 * inv2pidf;
  b.he2 = s1 * (1.0 - 3.0 * s) * du * inv2pidf;
  b.he3 = s * (3.0 * s - 2.0) * du * inv2pidf;
  const double xL = PI * f * det.L;
  double sx, cx;
  sincos_d(xL, &sx, &cx);
  b.xL = xL;
  b.sxL = sx;
  b.cxL = cx;
}

// TDI constants of one bin, formed in FP64 from their product forms (1 - D^n cancels at
// low f); recomputed per row from sin, cos of xL rather than held in registers
struct TdiC {
  Cx D, pf;  // delay exp(-2 i xL) and generation prefactor times (-i xL) exp(-i xL)
  Cx pfT;    // T-channel factor pf (1 - D) / sqrt(3), with the same common factor
};
__device__ inline void tdi_consts(const DetectorDev& det, const BinPre& b, TdiC& t) {
  const double sx = b.sxL, cx = b.cxL;
  // D = exp(-2 i xL);  1 - D^n = 2 i sin(n xL) exp(-i n xL)
  const double s2 = 2.0 * sx * cx, c2 = (cx - sx) * (cx + sx);
  t.D = {c2, -s2};
  // 1 - D^2 = 2 i sin(2xL) exp(-2 i xL) = 2 sin(2xL) (sin(2xL) + i cos(2xL))
  double pfr = 2.0 * s2 * s2, pfi = 2.0 * s2 * c2;
  if (det.tdi_gen == 20) {  // (1 - D^2)(1 - D^4), 1 - D^4 = 2 sin(4xL) (sin(4xL) + i cos(4xL))
    const double s4 = 2.0 * s2 * c2, c4 = (c2 - s2) * (c2 + s2);
    const double b_re = 2.0 * s4 * s4, b_im = 2.0 * s4 * c4;
    const double r = pfr * b_re - pfi * b_im, i = pfr * b_im + pfi * b_re;
    pfr = r;
    pfi = i;
  }
  // common factor of every link term: (-i xL) exp(-i xL) = xL (-sin xL, -cos xL)
  {
    const double fr = -b.xL * sx, fi = -b.xL * cx;
    const double r = pfr * fr - pfi * fi, i = pfr * fi + pfi * fr;
    pfr = r;
    pfi = i;
  }
  t.pf = {pfr, pfi};
  // pf (1 - D) / sqrt(3),  1 - D = 2 i sin(xL) exp(-i xL)
  const double o_re = 2.0 * sx * sx, o_im = 2.0 * sx * cx;
  const double r3 = 0.57735026918962576451;
  t.pfT = {(pfr * o_re - pfi * o_im) * r3, (pfr * o_im + pfi * o_re) * r3};
}

// one row's extrinsic constants, and 2 pi f t_c at this lane's bin reduced by the sine
// table's step once for all modes (the FP64 phase adds each mode's reduced part to it)
struct RowC {
  double tc;
  V3 p, q, k;
  double tc_r;  // 2 pi f t_c - tc_k 2 pi / kTrigN, |tc_r| <= pi / kTrigN
  int tc_k;     // low bits of the integer tc_k
  // geocentric orbits: (sqrt3/L) p.U, (sqrt3/L) p.V, the same for q, and k.U, k.V with the
  // detector's constant U, V (row_geo; unused for heliocentric orbits)
  double pU, pV, qU, qV, kU, kV;
};
__device__ inline RowC load_row(const double* __restrict__ pc, const BinPre& b) {
  RowC r;
  r.tc = pc[PC_TC];
  r.p = {pc[PC_P], pc[PC_P + 1], pc[PC_P + 2]};
  r.q = {pc[PC_Q], pc[PC_Q + 1], pc[PC_Q + 2]};
  r.k = {pc[PC_K], pc[PC_K + 1], pc[PC_K + 2]};
  // 2 pi f t_c rounded once, as the reference's own factor
  tab_reduce(2.0 * b.pif * r.tc, r.tc_r, r.tc_k);
  return r;
}
// the row's projections on a geocentric detector's constant U, V, once per row instead of per
// mode and pass
__device__ inline void row_geo(const DetectorDev& det, RowC& r) {
  r.pU = fma(r.p.x, det.sgxs, fma(r.p.y, det.sgys, r.p.z * det.sgzs));
  r.pV = fma(r.p.x, det.sgxc, r.p.y * det.sgyc);
  r.qU = fma(r.q.x, det.sgxs, fma(r.q.y, det.sgys, r.q.z * det.sgzs));
  r.qV = fma(r.q.x, det.sgxc, r.q.y * det.sgyc);
  r.kU = fma(r.k.x, det.gxs, fma(r.k.y, det.gys, r.k.z * det.gzs));
  r.kV = fma(r.k.x, det.gxc, r.k.y * det.gyc);
}

// One mode at one bin: waveform, orbit, and the projections of the three link pairs.
//
// The constellation offsets are xd_i = al_i U + be_i V.  Heliocentric orbits (KIND 0):
// al_i = r' sin kappa_i, be_i = r' cos kappa_i are constants and U, V follow the revolution
// angle; geocentric orbits (KIND 1): al_i, be_i are sin, cos of node i's rotation angle and
// U, V are constants.  The kappa_i are equally spaced by 2 pi / 3, so for pair a with nodes
// (i, j): al_i - al_j = -sqrt3 be_a and be_i - be_j = sqrt3 al_a, and every projection of a
// link vector n_a = (xd_i - xd_j) / L is
//   n_a.v = (sqrt3 / L) (al_a v.V - be_a v.U),      k.xd_a = al_a k.U + be_a k.V:
// six scalar products per mode instead of nine, and two terms per projection.
// KIND: bit 0 the orbit (0 heliocentric, 1 geocentric), bit 1 a duty-cycled detector.
template <int KIND>
constexpr bool kGeo = (KIND & 1) != 0;
template <int KIND>
constexpr bool kDuty = (KIND & 2) != 0;

struct ModeG {
  double A, t;      // amplitude, t_f + t_c
  Cx u, hm, hp, hc; // unit phasor, A u, F+ A u, 2 Fx A u
  Cx z0, z1, z01;   // pair phasors exp(i pi f k.xd_a) of nodes 0, 1, and their product
  double sa, ca;    // revolution angle
  double kx0;
  double np[3], nq[3], kn[3];     // n_a.p, n_a.q, n_a.k
  double al[3], be[3];            // KIND 1: sin, cos of the node rotation angles
  double pU, pV, qU, qV, kU, kV;  // KIND 1: (sqrt3/L) p.U, (sqrt3/L) p.V, same for q; k.U, k.V
};

constexpr double kOmegaYear = 2.0 * PI / YEAR_SI;

// t = t_f + t_c of a mode at this lane's bin, from the mode's node entries n0 (Psi, dPsi/du of
// nodes j, j + 1 at n0[2], n0[3], n0[6], n0[7]) and the Hermite weights of t_f
__device__ inline double mode_time(const double* __restrict__ n0, const BinPre& b, double tc) {
  return fma(b.he1, n0[6] - n0[2], fma(b.he2, n0[3], fma(b.he3, n0[7], tc)));
}

// Bounds of t = t_f + t_c over node interval j of one mode (node entries n: Psi, dPsi/du of
// nodes j, j + 1 at n[2], n[3], n[6], n[7]; frequencies fa < fb).  The fused kernel's t at a bin
// of the interval is N(s) / (2 pi du f), N(s) = 6 s (1 - s) (Psi1 - Psi0) + du [(1 - s)(1 - 3s)
// Psi0' + s (3s - 2) Psi1'], a quadratic in s with N(0) = du Psi0', N(1) = du Psi1': bounded by
// its extremes over [0, 1] times the range of 1 / (2 pi du f).  The bounds are widened by 1 ms
// plus 1e-9 of |t|, which covers the kernel's rounding of t.  NaN entries give NaN bounds.
__device__ inline void interval_t_range(const double* __restrict__ n, double fa, double fb, double du, double tc,
                                        double& tl, double& th) {
  const double dP = n[6] - n[2], c = du * n[3], d1 = du * n[7];
  const double a = 3.0 * (c + d1) - 6.0 * dP, bq = 6.0 * dP - 4.0 * c - 2.0 * d1;
  double lo = fmin(c, d1), hi = fmax(c, d1);
  if (a != 0.0) {
    const double sv = -bq / (2.0 * a);
    if (sv > 0.0 && sv < 1.0) {
      const double v = c - bq * bq / (4.0 * a);
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
  }
  const double ga = 1.0 / (2.0 * PI * du * fa), gb = 1.0 / (2.0 * PI * du * fb);
  tl = tc + fmin(lo * ga, lo * gb);
  th = tc + fmax(hi * ga, hi * gb);
  const double m = 1e-3 + 1e-9 * (fabs(tl) + fabs(th));
  tl -= m;
  th += m;
}

// The interval's t lies inside one off period of the detector's window (w = 0, no bin of it can
// see a nonzero window); false for NaN entries
__device__ inline bool interval_off(const DetectorDev& det, const double* __restrict__ n, double fa, double fb,
                                    double du, double tc) {
  double tl, th;
  interval_t_range(n, fa, fb, du, tc, tl, th);
  const double start = det.duty_t0 + floor((tl - det.duty_t0) * det.duty_inv_period) * det.duty_period;
  return tl - start >= det.duty_on && th - start < det.duty_period;
}

// The interval's t lies on one plateau of the window, clear of its ramps (w = 1 and w' = 0 at
// every bin of it, so the window changes nothing); false for NaN entries
__device__ inline bool interval_on(const DetectorDev& det, const double* __restrict__ n, double fa, double fb,
                                   double du, double tc) {
  double tl, th;
  interval_t_range(n, fa, fb, du, tc, tl, th);
  const double start = det.duty_t0 + floor((tl - det.duty_t0) * det.duty_inv_period) * det.duty_period;
  return tl - start >= det.duty_taper && th - start <= det.duty_on - det.duty_taper;
}

// UPH: 0 forms the unit phasor u; 1 forms it and stores it at uslot (this lane's LDS slot for
// the mode: the gradient kernel's forward pass); 2 takes it from uslot (the adjoint's
// recompute, which then needs neither the phase nor k.x0)
template <int PREC, int KIND, int UPH = 0>
__device__ inline void mode_geom(const DetectorDev& det, const double* __restrict__ mf,
                                 const double* __restrict__ n0, const BinPre& b, const RowC& rc, ModeG& g,
                                 double* __restrict__ uslot = nullptr) {
  const double A0 = n0[0], dA0 = n0[1], P0 = n0[2], dP0 = n0[3];
  const double A1 = n0[4], dA1 = n0[5], P1 = n0[6], dP1 = n0[7];
  g.A = fma(b.hb1, A1 - A0, fma(b.hb2, dA0, fma(b.hb3, dA1, A0)));
  const double dP = P1 - P0;
  const double Psi = fma(b.hb1, dP, fma(b.hb2, dP0, fma(b.hb3, dP1, P0)));
  g.t = mode_time(n0, b, rc.tc);
  double w = 1.0;  // duty window at t (its slope is recomputed by the adjoint)
  if constexpr (kDuty<KIND>) {
    double dw;
    duty_window(det, g.t, w, dw);
  }
  sincos_tab_s(fma(kOmegaYear, g.t, det.rev0), &g.sa, &g.ca);
  const double sa = g.sa, ca = g.ca;
  g.kx0 = AU_SEC * fma(rc.k.x, ca, rc.k.y * sa);
  double kx_0, kx_1;
  if constexpr (!kGeo<KIND>) {
    const double sc = sa * ca;
    const V3 U = {sc, -fma(ca, ca, 1.0), -SQRT3 * sa};
    const V3 V = {-fma(sa, sa, 1.0), sc, -SQRT3 * ca};
    const double pU = dot3(rc.p, U), pV = dot3(rc.p, V);
    const double qU = dot3(rc.q, U), qV = dot3(rc.q, V);
    const double kU = dot3(rc.k, U), kV = dot3(rc.k, V);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g.np[a] = fma(det.hsa[a], pV, -det.hca[a] * pU);
      g.nq[a] = fma(det.hsa[a], qV, -det.hca[a] * qU);
      g.kn[a] = fma(det.hsa[a], kV, -det.hca[a] * kU);
    }
    kx_0 = fma(det.hsk[0], kU, det.hck[0] * kV);
    kx_1 = fma(det.hsk[1], kU, det.hck[1] * kV);
  } else {
    double s1, c1;
    sincos_tab_s(fma(det.omega_rot, g.t, det.rot0), &s1, &c1);
    const double cr = -0.5, sr = 0.5 * SQRT3;  // rotation by 2 pi / 3
    g.al[0] = s1;
    g.be[0] = c1;
    g.al[1] = fma(s1, cr, c1 * sr);
    g.be[1] = fma(c1, cr, -s1 * sr);
    g.al[2] = fma(g.al[1], cr, g.be[1] * sr);
    g.be[2] = fma(g.be[1], cr, -g.al[1] * sr);
    g.pU = rc.pU;
    g.pV = rc.pV;
    g.qU = rc.qU;
    g.qV = rc.qV;
    g.kU = rc.kU;
    g.kV = rc.kV;
    const double kUs = det.c3L * g.kU, kVs = det.c3L * g.kV;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g.np[a] = fma(g.al[a], g.pV, -g.be[a] * g.pU);
      g.nq[a] = fma(g.al[a], g.qV, -g.be[a] * g.qU);
      g.kn[a] = fma(g.al[a], kVs, -g.be[a] * kUs);
    }
    kx_0 = fma(g.al[0], g.kU, g.be[0] * g.kV);
    kx_1 = fma(g.al[1], g.kU, g.be[1] * g.kV);
  }
  {
    double s0, c0, s1, c1;
    if constexpr (Trig<PREC>::kFp32) {
      sincos_f32(b.pif * kx_0, &s0, &c0);
      sincos_f32(b.pif * kx_1, &s1, &c1);
    } else {
      sincos_tab_s(b.pif * kx_0, &s0, &c0);
      sincos_tab_s(b.pif * kx_1, &s1, &c1);
    }
    g.z0 = {c0, s0};
    g.z1 = {c1, s1};
    g.z01 = mul(g.z0, g.z1);
  }
  if constexpr (UPH == 2) {
    const double2 uu = *reinterpret_cast<const double2*>(uslot);
    g.u = {uu.x, uu.y};
  } else {
    double sp, cp;
    if constexpr (Trig<PREC>::kFp32) {
      // the total phase summed and reduced to revolutions in FP64, its sine and cosine in FP32
      sincos_rev_fc(fma(2.0 * b.pif, rc.tc + g.kx0, Psi), &sp, &cp);
    } else {
      // 2 pi f t_c (rounded once, as the reference's own factor, and reduced once per bin and
      // row) plus Psi + 2 pi f k.x0, each reduced on its own and the remainders added: the total
      // is never rounded, so the phase's rounding error does not move with Psi or k.x0, and lnL
      // stays as smooth in the parameters as with three separate factors
      sincos_tab_sum(rc.tc_r, rc.tc_k, fma(2.0 * b.pif, g.kx0, Psi), &sp, &cp);
    }
    g.u = {cp, -sp};
    if constexpr (UPH == 1) *reinterpret_cast<double2*>(uslot) = make_double2(cp, -sp);
  }
  g.hm = scale(g.u, kDuty<KIND> ? g.A * w : g.A);
  g.hp = mul(Cx{mf[MF_FPR], mf[MF_FPI]}, g.hm);
  g.hc = mul(Cx{mf[MF_FCR], mf[MF_FCI]}, g.hm);
}

// sinc(xL +- a) with a = xL k.n and their derivatives (cos a, sin a given)
template <int PREC>
__device__ inline double rcp_prec(double x) {
  if constexpr (Trig<PREC>::kFp32)
    return rcp_f32(x);
  else
    return rcp_d(x);
}
template <int PREC>
__device__ inline void sinc_pair(double xL, double sxL, double cxL, double kn, double ca, double sa, double& sp,
                                 double& sm, double& dsp, double& dsm) {
  const double a = xL * kn;
  const double xp = xL + a, xm = xL - a;
  const double sinp = fma(sxL, ca, cxL * sa), cosp = fma(cxL, ca, -sxL * sa);
  const double sinm = fma(sxL, ca, -cxL * sa), cosm = fma(cxL, ca, sxL * sa);
  const double thr = 1e-2;
  const bool small_p = fabs(xp) < thr, small_m = fabs(xm) < thr;
  if (!small_p && !small_m) {
    const double inv = rcp_prec<PREC>(xp * xm);
    const double ip = xm * inv, im = xp * inv;
    sp = sinp * ip;
    dsp = (cosp - sp) * ip;
    sm = sinm * im;
    dsm = (cosm - sm) * im;
  } else {
    if (small_p) {
      const double x2 = xp * xp;
      sp = 1.0 - x2 * (1.0 / 6.0 - x2 * (1.0 / 120.0 - x2 * (1.0 / 5040.0)));
      dsp = -xp * (1.0 / 3.0 - x2 * (1.0 / 30.0 - x2 * (1.0 / 840.0)));
    } else {
      const double ip = rcp_prec<PREC>(xp);
      sp = sinp * ip;
      dsp = (cosp - sp) * ip;
    }
    if (small_m) {
      const double x2 = xm * xm;
      sm = 1.0 - x2 * (1.0 / 6.0 - x2 * (1.0 / 120.0 - x2 * (1.0 / 5040.0)));
      dsm = -xm * (1.0 / 3.0 - x2 * (1.0 / 30.0 - x2 * (1.0 / 840.0)));
    } else {
      const double im = rcp_prec<PREC>(xm);
      sm = sinm * im;
      dsm = (cosm - sm) * im;
    }
  }
}

// link pair a: link vector n_a = (xd[ia] - xd[ib]) / L with (ia, ib) = (2,1), (0,2), (1,0),
// third node a; links 0:12 1:21 2:23 3:32 4:31 5:13; the (+) link lp = 2, 4, 0 gets
// sinc(xL(1 + k.n_a)), the (-) link lp + 1 gets sinc(xL(1 - k.n_a))

struct PairV {
  double np, nq, kn, Pp, Pc;  // Pp = n.e+.n, Pc = n.ex.n / 2
  Cx z, C;  // pair phasor z_a; link term C = (n.h.n) z_a (without the common (-i xL) exp(-i xL))
  double sp, sm, dsp, dsm;
};

template <int PREC>
__device__ inline void pair_eval(const DetectorDev& det, const BinPre& b, const ModeG& g, int a, PairV& v) {
  v.np = g.np[a];
  v.nq = g.nq[a];
  v.kn = g.kn[a];
  // n.e+.n = (n.p)^2 - (n.q)^2 and n.ex.n / 2 = (n.p)(n.q) (hc carries the factor 2)
  v.Pp = fma(v.np, v.np, -v.nq * v.nq);
  v.Pc = v.np * v.nq;
  const Cx nhn = {fma(v.Pp, g.hp.re, v.Pc * g.hc.re), fma(v.Pp, g.hp.im, v.Pc * g.hc.im)};
  // z_a and exp(i xL k.n_a) = z_ia conj(z_ib), with z_2 = conj(z_0 z_1)
  Cx w;
  if (a == 0) {
    v.z = g.z0;
    w = conjx(mul(g.z01, g.z1));
  } else if (a == 1) {
    v.z = g.z1;
    w = mul(g.z0, g.z01);
  } else {
    v.z = conjx(g.z01);
    w = mul(g.z1, conjx(g.z0));
  }
  v.C = mul(nhn, v.z);
  sinc_pair<PREC>(b.xL, b.sxL, b.cxL, v.kn, w.re, w.im, v.sp, v.sm, v.dsp, v.dsm);
}

struct LinksT {
  Cx l[6];
};

// TDI outputs: A, E from X, Y, Z (tdi.py:1784-1785); T from the differences of the two links
// of each pair: X + Y + Z = pf (1 - D) ((l1 - l0) + (l3 - l2) + (l5 - l4)), so
// T = (X + Y + Z)/sqrt3 = pfT ((l1 - l0) + (l3 - l2) + (l5 - l4)) without the cancellation
// of the three channel sums (the factor pf (1 - D) / sqrt3 is formed from its product form)
__device__ inline void tdi_out(const DetectorDev& det, const TdiC& t, const LinksT& L, Cx out[3]) {
  const Cx* l = L.l;
  const Cx X = mul(t.pf, add(sub(l[5], l[0]), mul(t.D, sub(l[4], l[1]))));
  const Cx Y = mul(t.pf, add(sub(l[1], l[2]), mul(t.D, sub(l[0], l[3]))));
  const Cx Z = mul(t.pf, add(sub(l[3], l[4]), mul(t.D, sub(l[2], l[5]))));
  if (det.orthogonal) {
    const double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637;
    out[0] = scale(sub(Z, X), r2);
    out[1] = scale(add(sub(X, scale(Y, 2.0)), Z), r6);
    out[2] = mul(t.pfT, add(add(sub(l[1], l[0]), sub(l[3], l[2])), sub(l[5], l[4])));
  } else {
    out[0] = X;
    out[1] = Y;
    out[2] = Z;
  }
}

// The six links of one (bin, row) summed over the modes, without the common factor
// (-i xL) exp(-i xL) of every link term.
// nb: node entries (j, j+1) of this bin for mode 0; mode mi's are at nb + mi * mstride
// UPH 1: each mode's unit phasor is stored in u_lds ([mode][lane] double pairs) for the adjoint
template <int PREC, int KIND, int UPH = 0>
__device__ inline void links_row(const DetectorDev& det, const double* __restrict__ pc, const ModeList& modes,
                                 const double* __restrict__ nb, size_t mstride, const BinPre& b, const RowC& rc,
                                 LinksT& L, double* __restrict__ u_lds = nullptr) {
#pragma unroll
  for (int i = 0; i < 6; ++i) L.l[i] = {0.0, 0.0};
  for (int mi = 0; mi < modes.n; ++mi) {
    const double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
    if (b.f > mf[MF_CUT]) continue;
    ModeG g;
    mode_geom<PREC, KIND, UPH>(det, mf, nb + mi * mstride, b, rc, g,
                               UPH ? u_lds + 2 * (mi * TILE + threadIdx.x) : nullptr);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      PairV v;
      pair_eval<PREC>(det, b, g, a, v);
      const int lp = a == 0 ? 2 : (a == 1 ? 4 : 0), lm = lp + 1;
      L.l[lp] = axpy(v.sp, v.C, L.l[lp]);
      L.l[lm] = axpy(v.sm, v.C, L.l[lm]);
    }
  }
}

template <int PREC, int KIND, int UPH = 0>
__device__ inline void response_row(const DetectorDev& det, const double* __restrict__ pc, const ModeList& modes,
                                    const double* __restrict__ nb, size_t mstride, const BinPre& b,
                                    const RowC& rc, const TdiC& tdc, Cx out[3], double* __restrict__ u_lds = nullptr) {
  LinksT L;
  links_row<PREC, KIND, UPH>(det, pc, modes, nb, mstride, b, rc, L, u_lds);
  tdi_out(det, tdc, L, out);
}

// ------------------------------- adjoint ---------------------------------------------------
// Per-(bin, row) adjoint accumulators.  pb, qb, kb: adjoints of the polarization vectors
// and the propagation direction (contracted with their lambda/beta/psi derivatives once per
// row); the rest are parameter adjoints.
struct AdjAcc {
  V3 pb, qb, kb;   // adjoints of p, q, k, summed over the modes
  double gint[4];  // Mc, q, chi1, chi2
  double gh_re;    // sum Re(G h): d lnL / d d_L = -gh_re / d_L
  double mphi;     // sum m phibar: d lnL / d phi_ref = -mphi
  double iota, tc;
};

// wl: adjoints of the six links (the T channel's folded in); mid: mode id
// uslot: this lane's LDS slot of the mode's unit phasor, stored by the forward pass
template <int PREC, int KIND>
__device__ inline void mode_adjoint(const DetectorDev& det, const double* __restrict__ mf, int mid,
                                    const double* __restrict__ n0, const double* __restrict__ g0, const BinPre& b,
                                    const RowC& rc, const Cx wl[6], const double* __restrict__ pc,
                                    AdjAcc& acc, double* __restrict__ uslot) {
  ModeG g;
  mode_geom<PREC, KIND, 2>(det, mf, n0, b, rc, g, uslot);
  // adjoints of n_a.p, n_a.q, n_a.k and of k.xd_a
  double npb[3], nqb[3], knb[3], kxb[3];
  Cx Hp = {0.0, 0.0}, Hc = {0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    PairV v;
    pair_eval<PREC>(det, b, g, a, v);
    const int lp = a == 0 ? 2 : (a == 1 ? 4 : 0), lm = lp + 1;
    const Cx wp = wl[lp], wm = wl[lm];
    const double gp = fma(wp.re, v.C.re, wp.im * v.C.im);  // Re[conj(w) C]
    const double gm = fma(wm.re, v.C.re, wm.im * v.C.im);
    knb[a] = b.xL * (v.dsp * gp - v.dsm * gm);
    const Cx W = {fma(v.sp, wp.re, v.sm * wm.re), fma(v.sp, wp.im, v.sm * wm.im)};
    // through exp(+i pi f k.xd_a): -pi f Im(conj(W) C)
    kxb[a] = -b.pif * (W.re * v.C.im - W.im * v.C.re);
    const Cx V = mul(conjx(W), v.z);  // adjoint of n.h.n
    const double Ppb = V.re * g.hp.re - V.im * g.hp.im;  // Re(V h+)
    const double Pcb = V.re * g.hc.re - V.im * g.hc.im;  // Re(V hc), hc = 2 hx
    const double Ppb2 = 2.0 * Ppb;
    npb[a] = fma(v.np, Ppb2, v.nq * Pcb);
    nqb[a] = fma(v.np, Pcb, -v.nq * Ppb2);
    Hp = axpy(v.Pp, V, Hp);
    Hc = axpy(v.Pc, V, Hc);
  }
  // constellation adjoints: p, q, k directly (accumulated over the modes and contracted with
  // their lambda, beta, psi derivatives once per bin and row), and the time through the orbit
  // angle
  double tgeo;
  double pUb = 0.0, pVb = 0.0, qUb = 0.0, qVb = 0.0, kUb = 0.0, kVb = 0.0;
  if constexpr (!kGeo<KIND>) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pUb = fma(-npb[a], det.hca[a], pUb);
      pVb = fma(npb[a], det.hsa[a], pVb);
      qUb = fma(-nqb[a], det.hca[a], qUb);
      qVb = fma(nqb[a], det.hsa[a], qVb);
      kUb = fma(-knb[a], det.hca[a], fma(kxb[a], det.hsk[a], kUb));
      kVb = fma(knb[a], det.hsa[a], fma(kxb[a], det.hck[a], kVb));
    }
    const double sa = g.sa, ca = g.ca, sc = sa * ca;
    const V3 U = {sc, -fma(ca, ca, 1.0), -SQRT3 * sa};
    const V3 V = {-fma(sa, sa, 1.0), sc, -SQRT3 * ca};
    acc.pb = {fma(pUb, U.x, fma(pVb, V.x, acc.pb.x)), fma(pUb, U.y, fma(pVb, V.y, acc.pb.y)),
              fma(pUb, U.z, fma(pVb, V.z, acc.pb.z))};
    acc.qb = {fma(qUb, U.x, fma(qVb, V.x, acc.qb.x)), fma(qUb, U.y, fma(qVb, V.y, acc.qb.y)),
              fma(qUb, U.z, fma(qVb, V.z, acc.qb.z))};
    acc.kb = {fma(kUb, U.x, fma(kVb, V.x, acc.kb.x)), fma(kUb, U.y, fma(kVb, V.y, acc.kb.y)),
              fma(kUb, U.z, fma(kVb, V.z, acc.kb.z))};
    // dU/dalpha = (c2, s2, -sqrt3 ca), dV/dalpha = (-s2, c2, sqrt3 sa)
    const V3 Ua = {fma(pUb, rc.p.x, fma(qUb, rc.q.x, kUb * rc.k.x)), fma(pUb, rc.p.y, fma(qUb, rc.q.y, kUb * rc.k.y)),
                   fma(pUb, rc.p.z, fma(qUb, rc.q.z, kUb * rc.k.z))};
    const V3 Va = {fma(pVb, rc.p.x, fma(qVb, rc.q.x, kVb * rc.k.x)), fma(pVb, rc.p.y, fma(qVb, rc.q.y, kVb * rc.k.y)),
                   fma(pVb, rc.p.z, fma(qVb, rc.q.z, kVb * rc.k.z))};
    const double c2 = (ca - sa) * (ca + sa), s2 = 2.0 * sc;
    const double ta = fma(c2, Ua.x, fma(s2, Ua.y, fma(-s2, Va.x, c2 * Va.y)));
    tgeo = kOmegaYear * fma(SQRT3, fma(sa, Va.z, -ca * Ua.z), ta);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pUb = fma(-npb[a], g.be[a], pUb);
      pVb = fma(npb[a], g.al[a], pVb);
      qUb = fma(-nqb[a], g.be[a], qUb);
      qVb = fma(nqb[a], g.al[a], qVb);
      const double kns = det.c3L * knb[a];
      kUb = fma(-kns, g.be[a], fma(kxb[a], g.al[a], kUb));
      kVb = fma(kns, g.al[a], fma(kxb[a], g.be[a], kVb));
    }
    acc.pb = {fma(pUb, det.sgxs, fma(pVb, det.sgxc, acc.pb.x)), fma(pUb, det.sgys, fma(pVb, det.sgyc, acc.pb.y)),
              fma(pUb, det.sgzs, acc.pb.z)};
    acc.qb = {fma(qUb, det.sgxs, fma(qVb, det.sgxc, acc.qb.x)), fma(qUb, det.sgys, fma(qVb, det.sgyc, acc.qb.y)),
              fma(qUb, det.sgzs, acc.qb.z)};
    acc.kb = {fma(kUb, det.gxs, fma(kVb, det.gxc, acc.kb.x)), fma(kUb, det.gys, fma(kVb, det.gyc, acc.kb.y)),
              fma(kUb, det.gzs, acc.kb.z)};
    // d al_a / d theta = be_a, d be_a / d theta = -al_a
    const double th = fma(g.pU, pVb, fma(-g.pV, pUb, fma(g.qU, qVb, fma(-g.qV, qUb, fma(g.kU, kVb, -g.kV * kUb)))));
    tgeo = det.omega_rot * th;
  }
  // waveform-side adjoints: Gm = Hp Fp + Hc Fc, contributions Re[Gm dh_m]
  const Cx Fp = {mf[MF_FPR], mf[MF_FPI]}, Fc = {mf[MF_FCR], mf[MF_FCI]};
  const Cx Gm = add(mul(Hp, Fp), mul(Hc, Fc));
  const Cx Gh = mul(Gm, g.hm);
  const double phib = Gh.im;  // adjoint of the total phase Psi + 2 pi f (t_c + k.x0)
  const double kx0b = 2.0 * b.pif * phib;
  acc.kb.x = fma(kx0b * AU_SEC, g.ca, acc.kb.x);
  acc.kb.y = fma(kx0b * AU_SEC, g.sa, acc.kb.y);
  const double kx0_t = AU_SEC * kOmegaYear * fma(-rc.k.x, g.sa, rc.k.y * g.ca);
  const Cx Gu = mul(Gm, g.u);
  double Ab = Gu.re;  // adjoint of the amplitude the response sees (A w with a duty window)
  double tb = fma(kx0b, kx0_t, tgeo);  // adjoint of t = t_f + t_c
  if constexpr (kDuty<KIND>) {
    // h ~ A w(t): the window's slope moves t, A sees w.  The window is recomputed from t here,
    // so that only t and A stay live through the pair loop
    double w, dw;
    duty_window(det, g.t, w, dw);
    tb = fma(Ab * g.A, dw, tb);
    Ab *= w;
  }
  acc.gh_re = fma(Gm.re, g.hm.re, fma(-Gm.im, g.hm.im, acc.gh_re));  // += Re(Gm hm)
  acc.mphi = fma((double)kModeM[mid], phib, acc.mphi);
  {
    const Cx dFp = {mf[MF_DFPR], mf[MF_DFPI]}, dFc = {mf[MF_DFCR], mf[MF_DFCI]};
    const Cx tt = add(mul(Hp, dFp), mul(Hc, dFc));
    acc.iota = fma(tt.re, g.hm.re, fma(-tt.im, g.hm.im, acc.iota));  // += Re(tt hm)
  }
  acc.tc = fma(2.0 * b.pif, phib, acc.tc + tb);
  // intrinsic parameters through the node tables: weights of the eight node entries
  const double cA1 = Ab * b.hb1, cA0 = Ab - cA1, cdA0 = Ab * b.hb2, cdA1 = Ab * b.hb3;
  const double tP = tb * b.he1;
  const double cP1 = fma(phib, b.hb1, tP), cP0 = phib - cP1;
  const double cdP0 = fma(phib, b.hb2, tb * b.he2), cdP1 = fma(phib, b.hb3, tb * b.he3);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double* q0 = g0 + 4 * i;
    const double* q1 = g0 + 16 + 4 * i;
    double s = fma(cA0, q0[0], acc.gint[i]);
    s = fma(cdA0, q0[1], s);
    s = fma(cP0, q0[2], s);
    s = fma(cdP0, q0[3], s);
    s = fma(cA1, q1[0], s);
    s = fma(cdA1, q1[1], s);
    s = fma(cP1, q1[2], s);
    acc.gint[i] = fma(cdP1, q1[3], s);
  }
}

// Cross-lane moves without LDS (gfx950 VALU permutes and DPP).
//   half_swap_sum<32>(a, b): v_permlane32_swap on (a, b) leaves lanes 0-31 with (own a, a of
//     lane + 32) and lanes 32-63 with (b of lane - 32, own b); their sum is a + partner's a in
//     the lower half and b + partner's b in the upper half.
//   half_swap_sum<16>: the same with v_permlane16_swap between rows 2k and 2k + 1 of 16 lanes.
//   dpp_mov<CTRL>: a DPP move of a double (both 32-bit halves) within rows of 16 lanes.
template <int W>
__device__ inline double half_swap_sum(double a, double b) {
  const int alo = __double2loint(a), ahi = __double2hiint(a);
  const int blo = __double2loint(b), bhi = __double2hiint(b);
  if constexpr (W == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
  }
}

template <int CTRL>
__device__ inline double dpp_mov(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum of one value over the wave (every lane gets the total, in a fixed order).
__device__ inline double wave_sum(double v) {
  v = half_swap_sum<32>(v, v);
  v = half_swap_sum<16>(v, v);
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_mov<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  return v;
}

// Sum 12 per-lane values over the wave with 9 half-exchanges and 12 in-row moves instead of
// 12 x 6 exchanges: the first two steps trade halves of the value set (lanes 32-63 keep values
// 6-11, then rows 1 and 3 keep the upper three of their six), the last four reduce the
// remaining three values within each row of 16 lanes (rotations by 8 and 4, then the quad
// permutes lane ^ 2 and lane ^ 1).  Afterwards every lane of row q = lane / 16 holds the totals
// of values 3q .. 3q+2 (in z), summed in a fixed order.
__device__ inline void wave_sum12(const double x[12], double z[3]) {
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) y[i] = half_swap_sum<32>(x[i], x[6 + i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] = half_swap_sum<16>(y[i], y[3 + i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] += dpp_mov<0x128>(z[i]);  // row_ror:8
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] += dpp_mov<0x124>(z[i]);  // row_ror:4
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] += dpp_mov<0x4E>(z[i]);   // quad_perm [2, 3, 0, 1]
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] += dpp_mov<0xB1>(z[i]);   // quad_perm [1, 0, 3, 2]
}

// A row that the fused kernel evaluates on a tile starting at f0: a mode of the row reaches
// f0 and the row's parameters are valid (the two uniform skips of tile_rows).
__device__ inline bool row_live(const double* __restrict__ pconst, int p, double f0) {
  // p is uniform: stated, so that the two loads are scalar (no vector-memory wait per row)
  const double* pc = pconst + (size_t)__builtin_amdgcn_readfirstlane(p) * PC_STRIDE;
  return !(f0 > pc[PC_FCUT]) && pc[PC_VALID] != 0.0;
}

#define PES_LDS(ptr) ((__attribute__((address_space(3))) void*)(ptr))

// Copy node rows r0 .. r0+nj-1 of every mode (value rows of 4 doubles into bn, with GRAD
// gradient rows of 16 into bg, mode after mode) from global memory to LDS with the LDS-DMA
// loads (global_load_lds_dwordx4: each lane moves 16 bytes to the wave-uniform LDS base plus
// 16 x lane, without passing through registers).  Asynchronous: the data are in LDS once
// the issuing wave's vmcnt reaches zero.  Node rows are 32 / 128 bytes and hipMalloc'd, so
// every 16-byte piece is aligned.
template <bool GRAD>
__device__ inline void stage_nodes_dma(const double* __restrict__ nodes, const double* __restrict__ nodes_grad,
                                       size_t r0, int NNS, int nmodes, int nj, double* bn, double* bg, int lane,
                                       int wave) {
  const int cn = nmodes * nj * 2;  // 16-byte pieces of the value rows
  for (int base = wave * 64; base < cn; base += TILE) {
    const int c = base + lane;
    if (c < cn) {
      const int mi = c / (nj * 2), w = c - mi * (nj * 2);
      __builtin_amdgcn_global_load_lds(nodes + (r0 + (size_t)mi * NNS) * 4 + 2 * w, PES_LDS(bn + 2 * base), 16, 0, 0);
    }
  }
  if constexpr (GRAD) {
    const int cg = nmodes * nj * 8;
    for (int base = wave * 64; base < cg; base += TILE) {
      const int c = base + lane;
      if (c < cg) {
        const int mi = c / (nj * 8), w = c - mi * (nj * 8);
        __builtin_amdgcn_global_load_lds(nodes_grad + (r0 + (size_t)mi * NNS) * 16 + 2 * w, PES_LDS(bg + 2 * base),
                                         16, 0, 0);
      }
    }
  }
}

// The rows p0 .. p1-1 of one tile (lane <-> bin) on a detector of orbit kind KIND: lnL and,
// with GRAD, the 11 gradient terms of each row summed over the tile into acc
// [TILE / 64][chunk_rows<GRAD>()][NOUT].  Instantiated per orbit kind so that only that kind's detector
// constants are held in scalar registers.
//
// STAGED: the tile's bins use the nj <= NJ_STAGE consecutive node rows j0 .. j0+nj-1 (bins are
// sorted by frequency, so their node indices are too).  The block copies those node rows of
// every mode (value and, with GRAD, gradient tables) into one of two LDS buffers and the mode
// loops read them there.  The copy for the next evaluated row is issued (LDS-DMA,
// asynchronous) right after the barrier that publishes the current row's copy, so it overlaps
// the current row's arithmetic.  The buffers alternate from evaluated row to evaluated row
// (buf persists across tiles): the barrier of row p also guarantees that every wave has
// finished the previous row, the last reader of the buffer the copy for the next row goes to.
// Without STAGED each lane reads its node rows from global memory (tiles whose node range is
// wider, e.g. the lowest frequencies of short data).
template <int PREC, bool GRAD, int KIND, bool STAGED>
__device__ inline void tile_rows(const DetectorDev& det, int tile, double f0, int p0, int p1,
                                 const double* __restrict__ bin_f, const int* __restrict__ bin_j,
                                 const double* __restrict__ bin_s, const double* __restrict__ bin_rec,
                                 const double* __restrict__ tile_dd, const double* __restrict__ pconst,
                                 const double* __restrict__ nodes, const double* __restrict__ nodes_grad, int NNS,
                                 double du, const ModeList& modes, int lane, int wave, double* acc, int j0,
                                 int nj, double* stage_n, double* stage_g, int& buf,
                                 const unsigned long long* __restrict__ off_bits, int n_tiles, double* lane_lds) {
  constexpr int NOUT = GRAD ? 1 + PES_N_PARAMS : 1;
  const int bi = tile * TILE + threadIdx.x;
  // a duty-cycled detector's tile whose window is zero for every mode of row p throughout
  // (k_duty_off): the response is zero there, as above the row's cuts.  The tile's flag is read
  // only for rows that have such tiles (PC_OFFANY).  p is uniform: scalar loads
  auto tile_off = [&](int p) -> bool {
    if constexpr (kDuty<KIND>) {
      const int pu = __builtin_amdgcn_readfirstlane(p);
      if (pconst[(size_t)pu * PC_STRIDE + PC_OFFANY] == 0.0) return false;
      const size_t w = (size_t)pu * ((n_tiles + 63) >> 6) + (tile >> 6);
      return ((off_bits[w] >> (tile & 63)) & 1ull) != 0;
    } else {
      return false;
    }
  };
  BinPre b;
  bin_prepare(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
  // the bin's weight and data record (w, Re d, Im d per channel).  The forward-only kernel reads
  // it once per tile into the lane's own LDS column and takes it from there after each row's
  // mode loop (1.7% faster than re-reading it through the caches); the gradient kernel, whose
  // LDS holds its row accumulators and the modes' phasors, re-reads it per row from the caches,
  // requested before the mode loops so that their latency is hidden behind them (0.6% faster
  // than an LDS column)
  constexpr bool kRecLds = !GRAD, kRecEarly = GRAD;
  if constexpr (kRecLds) {
    const double2* r2 = reinterpret_cast<const double2*>(bin_rec + (size_t)bi * BIN_REC);
#pragma unroll
    for (int i = 0; i < BIN_REC / 2; ++i) {
      const double2 t = r2[i];
      lane_lds[(2 * i) * TILE + threadIdx.x] = t.x;
      lane_lds[(2 * i + 1) * TILE + threadIdx.x] = t.y;
    }
  }
  constexpr int SN = MAX_MODES * NJ_STAGE * 4, SG = MAX_MODES * NJ_STAGE * 16;  // doubles per buffer
  if constexpr (STAGED) {  // copy for the first evaluated row
    int q = p0;
    while (q < p1 && !(row_live(pconst, q, f0) && !tile_off(q))) ++q;
    if (q < p1)
      stage_nodes_dma<GRAD>(nodes, nodes_grad, (size_t)q * modes.n * NNS + j0, NNS, modes.n, nj, stage_n + buf * SN,
                            stage_g + buf * SG, lane, wave);
  }
  for (int p = p0; p < p1; ++p) {
    const double* pc = pconst + (size_t)p * PC_STRIDE;
    if (f0 > pc[PC_FCUT] || tile_off(p)) {
      if (threadIdx.x == 0) acc[(p - p0) * NOUT] -= tile_dd[tile];
      continue;
    }
    double G[PES_N_PARAMS];
#pragma unroll
    for (int i = 0; i < PES_N_PARAMS; ++i) G[i] = 0.0;
    double v = 0.0;
    if (pc[PC_VALID] != 0.0) {
      const size_t pbase = (size_t)p * modes.n * NNS;
      // this bin's record, not held in registers across the tile's rows: from the LDS column
      // (forward-only) or re-read from the caches (gradient: in the staged path after the barrier
      // and before the next row's copies, so that no wait for those copies comes first)
      double rec[BIN_REC];
      auto load_rec = [&]() {
        if constexpr (kRecLds) {
#pragma unroll
          for (int i = 0; i < BIN_REC; ++i) rec[i] = lane_lds[i * TILE + threadIdx.x];
        } else {
          int off = bi * BIN_REC;
          asm volatile("" : "+v"(off));  // a per-row offset: the loads stay in the row loop
          const double2* r2 = reinterpret_cast<const double2*>(bin_rec + off);
#pragma unroll
          for (int i = 0; i < BIN_REC / 2; ++i) {
            const double2 t = r2[i];
            rec[2 * i] = t.x;
            rec[2 * i + 1] = t.y;
          }
        }
      };
      const double* nb;  // node entries of this lane's bin, mode 0, and the per-mode strides
      const double* gb;
      size_t nstride, gstride;
      if constexpr (STAGED) {
        int qn = p + 1;  // next evaluated row
        while (qn < p1 && !(row_live(pconst, qn, f0) && !tile_off(qn))) ++qn;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies for row p are in LDS
        __syncthreads();                     // ... and every wave's; all waves are past row p - 1
        if constexpr (kRecEarly) load_rec();
        if (qn < p1)
          stage_nodes_dma<GRAD>(nodes, nodes_grad, (size_t)qn * modes.n * NNS + j0, NNS, modes.n, nj,
                                stage_n + (buf ^ 1) * SN, stage_g + (buf ^ 1) * SG, lane, wave);
        nb = stage_n + buf * SN + (b.j - j0) * 4;
        gb = stage_g + buf * SG + (b.j - j0) * 16;
        nstride = (size_t)nj * 4;
        gstride = (size_t)nj * 16;
        buf ^= 1;
      } else {
        if constexpr (kRecEarly) load_rec();
        nb = nodes + (pbase + b.j) * 4;
        gb = nodes_grad + (pbase + b.j) * 16;
        nstride = (size_t)NNS * 4;
        gstride = (size_t)NNS * 16;
      }
      RowC rc = load_row(pc, b);
      if constexpr (kGeo<KIND>) row_geo(det, rc);
      TdiC tdc;
      tdi_consts(det, b, tdc);
      Cx out[3];
      // FP64 in every mode; the gradient kernel keeps each mode's unit phasor for the adjoint
      response_row<0, KIND, GRAD ? 1 : 0>(det, pc, modes, nb, nstride, b, rc, tdc, out, lane_lds);
      // the forward-only kernel (168 registers) takes the record here instead: held across its
      // mode loop, the record's registers were spilled (measured 0.9% slower)
      if constexpr (!kRecEarly) load_rec();
      Cx wch[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        const double wc = rec[3 * ch];
        const double rr = rec[3 * ch + 1] - out[ch].re, ri = rec[3 * ch + 2] - out[ch].im;
        v -= wc * fma(rr, rr, ri * ri);
        wch[ch] = {2.0 * wc * rr, 2.0 * wc * ri};
      }
      if (GRAD) {
        Cx wX, wY, wZ, wT = {0.0, 0.0};
        if (det.orthogonal) {
          const double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637;
          wX = {-r2 * wch[0].re + r6 * wch[1].re, -r2 * wch[0].im + r6 * wch[1].im};
          wY = {-2.0 * r6 * wch[1].re, -2.0 * r6 * wch[1].im};
          wZ = {r2 * wch[0].re + r6 * wch[1].re, r2 * wch[0].im + r6 * wch[1].im};
          wT = mul(conjx(tdc.pfT), wch[2]);
        } else {
          wX = wch[0];
          wY = wch[1];
          wZ = wch[2];
        }
        const Cx cpf = conjx(tdc.pf), cD = conjx(tdc.D);
        const Cx WX = mul(cpf, wX), WY = mul(cpf, wY), WZ = mul(cpf, wZ);
        const Cx DWX = mul(cD, WX), DWY = mul(cD, WY), DWZ = mul(cD, WZ);
        Cx wl[6];
        wl[0] = sub(DWY, WX);
        wl[1] = sub(WY, DWX);
        wl[2] = sub(DWZ, WY);
        wl[3] = sub(WZ, DWY);
        wl[4] = sub(DWX, WZ);
        wl[5] = sub(WX, DWZ);
        // T = pfT sum over pairs of (link lm - link lp): its adjoint moves onto the links
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          wl[2 * a] = sub(wl[2 * a], wT);
          wl[2 * a + 1] = add(wl[2 * a + 1], wT);
        }
        AdjAcc aa;
        aa.pb = aa.qb = aa.kb = V3{0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 4; ++i) aa.gint[i] = 0.0;
        aa.gh_re = aa.mphi = aa.iota = aa.tc = 0.0;
        for (int mi = 0; mi < modes.n; ++mi) {
          const double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
          if (b.f > mf[MF_CUT]) continue;
          mode_adjoint<PREC, KIND>(det, mf, modes.id[mi], nb + mi * nstride, gb + mi * gstride, b, rc, wl, pc,
                                   aa, lane_lds + 2 * (mi * TILE + threadIdx.x));
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) G[i] = aa.gint[i];
        G[4] = -aa.gh_re * pc[PC_INVDL];
        G[5] = -aa.mphi;
        G[6] = aa.iota;
        {  // p, q, k adjoints contracted with their lambda, beta, psi derivatives (row constants)
          const V3 dlp = {pc[PC_DLP], pc[PC_DLP + 1], pc[PC_DLP + 2]};
          const V3 dlq = {pc[PC_DLQ], pc[PC_DLQ + 1], pc[PC_DLQ + 2]};
          const V3 dlk = {pc[PC_DLK], pc[PC_DLK + 1], pc[PC_DLK + 2]};
          G[7] = dot3(aa.pb, dlp) + dot3(aa.qb, dlq) + dot3(aa.kb, dlk);
          const V3 dbp = {pc[PC_DBP], pc[PC_DBP + 1], pc[PC_DBP + 2]};
          const V3 dbq = {pc[PC_DBQ], pc[PC_DBQ + 1], pc[PC_DBQ + 2]};
          const V3 dbk = {pc[PC_DBK], pc[PC_DBK + 1], pc[PC_DBK + 2]};
          G[8] = dot3(aa.pb, dbp) + dot3(aa.qb, dbq) + dot3(aa.kb, dbk);
          G[9] = dot3(aa.pb, rc.q) - dot3(aa.qb, rc.p);
        }
        G[10] = aa.tc;
      }
    }
    if (GRAD) {
      double x[12], z[3];
      x[0] = v;
#pragma unroll
      for (int i = 0; i < PES_N_PARAMS; ++i) x[1 + i] = G[i];
      wave_sum12(x, z);
      if ((lane & 15) == 0) {
        const int q = lane >> 4;
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[(wave * chunk_rows<GRAD>() + p - p0) * NOUT + 3 * q + i] += z[i];
      }
    } else {
      v = wave_sum(v);
      if (lane == 0) acc[(wave * chunk_rows<GRAD>() + p - p0) * NOUT] += v;
    }
  }
}

// ------------------------------- kernels ---------------------------------------------------
// Persistent fused kernel.  Work item = (group of tiles_per_group tiles, chunk of
// chunk_rows<GRAD>() rows);
// lane <-> bin.  A tile whose first frequency lies above every mode's cut of a row has a
// zero response for that row: its lnL contribution is the precomputed -sum w |d|^2 of the
// tile and its gradient contribution is zero, so the row skips it.
// Launch bounds: the gradient kernel runs two waves per SIMD (256 registers of the unified
// file per lane), the forward-only kernel three (168 registers, some spilled; measured 9%
// faster than two for lnL-only batches when three blocks per CU are launched).
// KD: the launch's tiles belong to detectors of one orbit kind (bit 0: heliocentric 0,
// geocentric 1), duty-cycled (bit 1, the duty-window tile variants) or not; each instantiation
// holds that one variant's code (a kernel holding both orbit kinds spilled 106 instead of 59-74
// VGPRs).  items (duty-cycled ranges only; null otherwise): the work items of the range that the
// plain instantiation takes (1) and those the duty one takes (0).
template <int PREC, bool GRAD, int KD>
__global__ __launch_bounds__(TILE, GRAD ? 2 : 3) void k_lnl(const DetectorDev* __restrict__ dets, int n_det, int n_tiles,
                                              const double* __restrict__ bin_f, const int* __restrict__ bin_j,
                                              const double* __restrict__ bin_s, const double* __restrict__ bin_rec,
                                              const double* __restrict__ tile_f0, const double* __restrict__ tile_dd,
                                              int tiles_per_group, const double* __restrict__ pconst,
                                              const double* __restrict__ nodes,
                                              const double* __restrict__ nodes_grad, int NNS, double du,
                                              ModeList modes, int P, int rows, const int* __restrict__ tile_jr,
                                              int tile_begin, int tile_end, int group_offset,
                                              int* __restrict__ work_counter, double* __restrict__ partial,
                                              const unsigned long long* __restrict__ off_bits,
                                              const unsigned char* __restrict__ items) {
  constexpr int NOUT = GRAD ? 1 + PES_N_PARAMS : 1;
  constexpr int CH = chunk_rows<GRAD>();  // 4x the rows without the gradient: 1.3% faster
  __shared__ double acc[TILE / 64][CH][NOUT];
  __shared__ double stage_n[2 * MAX_MODES * NJ_STAGE * 4];
  __shared__ typename std::conditional<GRAD, double[2 * MAX_MODES * NJ_STAGE * 16], double[1]>::type stage_g;
  int buf = 0;
  __shared__ int next_work;
  // each lane's own LDS column: forward-only, the tile's bin records [record entry][lane];
  // gradient, each mode's unit phasor from the forward pass [mode][lane] (pairs), read back by
  // the adjoint instead of its phase and sincos
  __shared__ double lane_lds[GRAD ? 2 * MAX_MODES * TILE : BIN_REC * TILE];
  trig_table_load();
  // the forward-only kernel holds the (uniform) wave index in a scalar register: 1.3% faster
  // there, 0.2% slower with the gradient
  const int lane = threadIdx.x & 63;
  const int wave = GRAD ? (int)(threadIdx.x >> 6) : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // rows (1 <= rows <= CH) per work item, chosen by the host.  This launch covers the tiles
  // [tile_begin, tile_end) in groups of tiles_per_group, whose partial sums go to groups
  // group_offset + 0, 1, ...: one launch (and counter) per tile range of one orbit kind and duty
  // flag, two for a duty-cycled range
  const int n_chunks = (P + rows - 1) / rows;
  const int n_groups_r = (tile_end - tile_begin + tiles_per_group - 1) / tiles_per_group;
  const int n_work = n_chunks * n_groups_r;
  constexpr bool DUTY = kDuty<KD>;
  // the duty pass of a range whose items all run in the plain one: k_duty_items counted none at
  // work_counter[MAX_TILE_RANGES] (every block leaves before its first barrier)
  if (DUTY && items && work_counter[MAX_TILE_RANGES] == 0) return;
  // work items are taken from a counter (zeroed by k_setup earlier on the same stream), so
  // blocks that draw cheap items (tiles above the mode cuts) take more of them; every block
  // leaves once the counter passes n_work
  for (;;) {
    if (threadIdx.x == 0) next_work = atomicAdd(work_counter, 1);
    __syncthreads();
    const int w = next_work;
    if (w >= n_work) break;
    const int gl = w / n_chunks, c = w % n_chunks;
    const int g = group_offset + gl;
    // a duty-cycled range runs twice: items on the window's plateau (items[] = 1, k_duty_items)
    // here in the plain instantiation, the others in the duty one; each writes only its own
    if (items && items[(size_t)__builtin_amdgcn_readfirstlane(g) * n_chunks + __builtin_amdgcn_readfirstlane(c)] !=
                     (DUTY ? 0 : 1)) {
      __syncthreads();  // every wave has read next_work before it is rewritten
      continue;
    }
    const int p0 = c * rows, p1 = min(P, p0 + rows);
    for (int i = threadIdx.x; i < (TILE / 64) * CH * NOUT; i += TILE) (&acc[0][0][0])[i] = 0.0;
    double chunk_cut = 0.0;
    for (int p = p0; p < p1; ++p) chunk_cut = fmax(chunk_cut, pconst[(size_t)p * PC_STRIDE + PC_FCUT]);
    __syncthreads();
    const int t_first = tile_begin + gl * tiles_per_group;
    const int t_end = min(tile_end, t_first + tiles_per_group);
    for (int tile = t_first; tile < t_end; ++tile) {
      const double f0 = tile_f0[tile];
      if (f0 > chunk_cut) {  // no row of the chunk has a mode up here
        if (threadIdx.x == 0) {
          const double dd = tile_dd[tile];
          for (int p = p0; p < p1; ++p) acc[0][p - p0][0] -= dd;
        }
        continue;
      }
      const DetectorDev& det = dets[tile_detector(dets, n_det, tile)];
      const int j0 = tile_jr[2 * tile], nj = tile_jr[2 * tile + 1];
#define PES_TILE_ARGS                                                                                         \
  det, tile, f0, p0, p1, bin_f, bin_j, bin_s, bin_rec, tile_dd, pconst, nodes, nodes_grad, NNS, du, \
      modes, lane, wave, &acc[0][0][0], j0, nj, stage_n, &stage_g[0], buf, off_bits, n_tiles, &lane_lds[0]
      // the host's tile ranges give this launch detectors of one orbit kind (KD bit 0), all
      // duty-cycled or none (bit 1)
      if (nj <= NJ_STAGE)
        tile_rows<PREC, GRAD, KD, true>(PES_TILE_ARGS);
      else
        tile_rows<PREC, GRAD, KD, false>(PES_TILE_ARGS);
#undef PES_TILE_ARGS
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (p1 - p0) * NOUT; i += TILE) {
      const int pl = i / NOUT, o = i % NOUT;
      double s = 0.0;
      for (int k = 0; k < TILE / 64; ++k) s += acc[k][pl][o];
      partial[((size_t)g * NOUT + o) * P + p0 + pl] = s;
    }
    __syncthreads();
  }
}

template <int PREC>
__global__ __launch_bounds__(128) void k_response_t(const DetectorDev* __restrict__ dets, int det_i, const double* __restrict__ bin_f,
                             const int* __restrict__ bin_j, const double* __restrict__ bin_s,
                             const double* __restrict__ pconst, const double* __restrict__ nodes, int NNS, double du,
                             ModeList modes, int P, double* __restrict__ out) {
  trig_table_load();
  const DetectorDev det = dets[det_i];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= det.n_bins || p >= P) return;
  const int bi = det.bin_offset + k;
  BinPre b;
  bin_prepare(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
  const double* pc = pconst + (size_t)p * PC_STRIDE;
  RowC rc = load_row(pc, b);
  row_geo(det, rc);  // used by the geocentric variants only
  TdiC tdc;
  tdi_consts(det, b, tdc);
  Cx o[3];
  const double* nb = nodes + ((size_t)p * modes.n * NNS + b.j) * 4;
  switch (det.kind | (det.duty ? 2 : 0)) {
    case 0: response_row<PREC, 0>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, tdc, o); break;
    case 1: response_row<PREC, 1>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, tdc, o); break;
    case 2: response_row<PREC, 2>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, tdc, o); break;
    default: response_row<PREC, 3>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, tdc, o); break;
  }
  for (int ch = 0; ch < 3; ++ch) {
    const size_t idx = (((size_t)p * 3 + ch) * det.n_bins + k) * 2;
    out[idx] = o[ch].re;
    out[idx + 1] = o[ch].im;
  }
}

#ifndef PES_LNL_FWD_UNIT
// The six single links (12, 21, 23, 32, 31, 13) of P rows on one detector, summed over the
// modes, with the common factor (-i xL) exp(-i xL) applied: out [P][6][n_bins][2]
// (antenna.py:496-543 per link, summed over modes as antenna.py:445-458).
__global__ __launch_bounds__(128) void k_links(const DetectorDev* __restrict__ dets, int det_i,
                                               const double* __restrict__ bin_f, const int* __restrict__ bin_j,
                                               const double* __restrict__ bin_s, const double* __restrict__ pconst,
                                               const double* __restrict__ nodes, int NNS, double du, ModeList modes,
                                               int P, double* __restrict__ out) {
  trig_table_load();
  const DetectorDev det = dets[det_i];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= det.n_bins || p >= P) return;
  const int bi = det.bin_offset + k;
  BinPre b;
  bin_prepare(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
  const double* pc = pconst + (size_t)p * PC_STRIDE;
  RowC rc = load_row(pc, b);
  row_geo(det, rc);  // used by the geocentric variants only
  LinksT L;
  const double* nb = nodes + ((size_t)p * modes.n * NNS + b.j) * 4;
  switch (det.kind | (det.duty ? 2 : 0)) {
    case 0: links_row<0, 0>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, L); break;
    case 1: links_row<0, 1>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, L); break;
    case 2: links_row<0, 2>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, L); break;
    default: links_row<0, 3>(det, pc, modes, nb, (size_t)NNS * 4, b, rc, L); break;
  }
  const Cx cf = {-b.xL * b.sxL, -b.xL * b.cxL};  // (-i xL) exp(-i xL)
  for (int i = 0; i < 6; ++i) {
    const Cx l = mul(cf, L.l[i]);
    const size_t idx = (((size_t)p * 6 + i) * det.n_bins + k) * 2;
    out[idx] = l.re;
    out[idx + 1] = l.im;
  }
}
#endif  // PES_LNL_FWD_UNIT

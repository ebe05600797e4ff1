// Fused per-(bin, row) forward response and adjoint, generic in the bulk real type R.
//
//   R = double : the FP64 path.  The t_c shift is applied as a separate factor
//                exp(-2 pi i f t_c) exactly as the reference does (antenna.py:435-441).
//   R = float  : the mixed path.  FP64 is kept for the phase Psi(f) + 2 pi f t_c,
//                the orbit angle and the barycentric delay 2 pi f k.x0, each reduced
//                mod pi/2 in FP64 before an FP32 sine/cosine; amplitudes, link
//                projections, transfer functions and the TDI combination run in FP32;
//                the residual, lnL and gradient accumulations are FP64.
//
// Algebraic rearrangements relative to antenna.py:496-543 (same values):
//   * sinc(xL(1 +- k.n)): sin(xL +- a) = sin xL cos a +- cos xL sin a, a = xL k.n, with
//     sin xL, cos xL per bin and one sincos per link pair;
//   * exp(-i pi f (L + k.(xi + xj))) = exp(-i pi f L) exp(-2 pi i f k.x0)
//     exp(-i pi f k.xi_det) exp(-i pi f k.xj_det): one large-argument sincos per
//     (bin, mode, row), three small ones, and exp(-i pi f L) per bin;
//   * n.e+.n = (n.p)^2 - (n.q)^2, n.ex.n = 2 (n.p)(n.q) with e+ = pp - qq, ex = pq + qp.
// This header is included inside the anonymous namespace of pespace_amd.hip.

template <class R>
struct Cx {
  R re, im;
};
template <class R>
__device__ inline Cx<R> mul(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R>
__device__ inline Cx<R> add(Cx<R> a, Cx<R> b) {
  return {a.re + b.re, a.im + b.im};
}
template <class R>
__device__ inline Cx<R> sub(Cx<R> a, Cx<R> b) {
  return {a.re - b.re, a.im - b.im};
}
template <class R>
__device__ inline Cx<R> scale(Cx<R> a, R s) {
  return {a.re * s, a.im * s};
}
template <class R>
__device__ inline Cx<R> conjx(Cx<R> a) {
  return {a.re, -a.im};
}
template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
__device__ inline R dot3(const V3<R>& a, const V3<R>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// sin/cos of a large FP64 argument, result in R.  FP32: reduced to revolutions in FP64,
// then the hardware v_sin_f32 / v_cos_f32 (which take revolutions).
__device__ inline void sc_big(double x, double* s, double* c) { sincos_d(x, s, c); }
__device__ inline void sc_big(double x, float* s, float* c) {
  double r = x * 0.15915494309189533577;  // 1/(2 pi)
  r -= rint(r);
  float rf = (float)r;
  *s = __builtin_amdgcn_sinf(rf);
  *c = __builtin_amdgcn_cosf(rf);
}
// sin/cos of a modest argument held in R
__device__ inline void sc_mod(double x, double* s, double* c) { sincos_d(x, s, c); }
__device__ inline void sc_mod(float x, float* s, float* c) {
  float r = x * 0.15915494309189533577f;
  r -= rintf(r);
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}
__device__ inline double rcp(double x) { return 1.0 / x; }
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// per-bin quantities shared by every row and mode of a tile lane
template <class R>
struct BinPre {
  double f, s, du, inv2pidf, pif;
  int j;
  R xL, sxL, cxL;
  Cx<R> eL;        // exp(-i pi f L)
  Cx<R> D, pf;     // TDI delay exp(-2 pi i f L) and generation prefactor
  Cx<R> pfT;       // T-channel factor pf (1 - D) / sqrt(3)
};

// TDI constants are formed in FP64 (1 - D^n cancels at low f) and then rounded to R
template <class R>
__device__ inline void bin_prepare(const DetectorDev& det, double f, double s, int j, double du, BinPre<R>& b) {
  b.f = f;
  b.s = s;
  b.j = j;
  b.du = du;
  b.inv2pidf = 1.0 / (2.0 * PI * du * f);
  b.pif = PI * f;
  const double xL = PI * f * det.L;
  double sx, cx;
  sincos_d(xL, &sx, &cx);
  b.xL = (R)xL;
  b.sxL = (R)sx;
  b.cxL = (R)cx;
  b.eL = {(R)cx, (R)-sx};
  // D = exp(-2 i xL);  1 - D^n = 2 i sin(n xL) exp(-i n xL)
  const double s2 = 2.0 * sx * cx, c2 = cx * cx - sx * sx;
  b.D = {(R)c2, (R)-s2};
  double pfr, pfi;  // prefactor, from its product form
  {
    // 1 - D^2 = 2 i sin(2 xL) exp(-2 i xL) = 2 sin(2xL) (sin(2xL) + i cos(2xL))
    double a_re = 2.0 * s2 * s2, a_im = 2.0 * s2 * c2;
    pfr = a_re;
    pfi = a_im;
    if (det.tdi_gen == 20) {  // (1 - D^2)(1 - D^4), 1 - D^4 = 2 i sin(4 xL) exp(-4 i xL)
      double s8, c8;
      sincos_d(4.0 * xL, &s8, &c8);
      double b_re = 2.0 * s8 * s8, b_im = 2.0 * s8 * c8;
      double r = pfr * b_re - pfi * b_im, i = pfr * b_im + pfi * b_re;
      pfr = r;
      pfi = i;
    }
  }
  b.pf = {(R)pfr, (R)pfi};
  // pf (1 - D) / sqrt(3),  1 - D = 2 i sin(xL) exp(-i xL)
  {
    double o_re = 2.0 * sx * sx, o_im = 2.0 * sx * cx;
    const double r3 = 0.57735026918962576451;
    b.pfT = {(R)((pfr * o_re - pfi * o_im) * r3), (R)((pfr * o_im + pfi * o_re) * r3)};
  }
}

// orbit at time t: link vectors n_a and detector-frame node positions (R); the barycentre
// term k.x0 and its derivatives in FP64; with DT also the time derivatives
struct Bary {
  double kx0, kx0_t, sa, ca;
};
template <class R, bool DT>
__device__ inline void orbit_eval(const DetectorDev& d, double t, const V3<R>& k, V3<R> n[3], V3<R> xd[3],
                                  V3<R> nd[3], V3<R> xdd[3], Bary& bc) {
  const double omega = 2.0 * PI / YEAR_SI;
  double sa, ca;
  sincos_d(omega * t + d.rev0, &sa, &ca);
  bc.sa = sa;
  bc.ca = ca;
  bc.kx0 = AU_SEC * ((double)k.x * ca + (double)k.y * sa);
  bc.kx0_t = AU_SEC * omega * (-(double)k.x * sa + (double)k.y * ca);
  const R saf = (R)sa, caf = (R)ca;
  if (d.kind == 0) {
    const R rp = (R)d.rp;
    R sc = saf * caf, sa2 = (R)1 + saf * saf, ca2 = (R)1 + caf * caf, c2a = caf * caf - saf * saf;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const R ck = (R)d.ck[i], sk = (R)d.sk[i];
      xd[i].x = rp * (sc * sk - sa2 * ck);
      xd[i].y = rp * (sc * ck - ca2 * sk);
      xd[i].z = rp * ((R)(-SQRT3) * (caf * ck + saf * sk));
      if (DT) {
        const R rw = (R)(d.rp * omega);
        xdd[i].x = rw * (c2a * sk - (R)2 * sc * ck);
        xdd[i].y = rw * (c2a * ck + (R)2 * sc * sk);
        xdd[i].z = rw * ((R)(-SQRT3) * (caf * sk - saf * ck));
      }
    }
  } else {
    R s1, c1;
    sc_big(d.omega_rot * t + d.rot0, &s1, &c1);
    const R cr = (R)-0.5, sr = (R)(0.5 * SQRT3);
    R sk[3], ck[3];
    sk[0] = s1;
    ck[0] = c1;
    sk[1] = sk[0] * cr + ck[0] * sr;
    ck[1] = ck[0] * cr - sk[0] * sr;
    sk[2] = sk[1] * cr + ck[1] * sr;
    ck[2] = ck[1] * cr - sk[1] * sr;
    const R rp = (R)d.rp, sb = (R)d.sbet, cb = (R)d.cbet, sl = (R)d.slam, cl = (R)d.clam;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xd[i].x = rp * (sb * cl * sk[i] + sl * ck[i]);
      xd[i].y = rp * (sb * sl * sk[i] - cl * ck[i]);
      xd[i].z = rp * (-cb * sk[i]);
      if (DT) {
        const R rw = (R)(d.rp * d.omega_rot);
        xdd[i].x = rw * (sb * cl * ck[i] - sl * sk[i]);
        xdd[i].y = rw * (sb * sl * ck[i] + cl * sk[i]);
        xdd[i].z = rw * (-cb * ck[i]);
      }
    }
  }
  const R invL = (R)(1.0 / d.L);
  const int ia[3] = {2, 0, 1}, ib[3] = {1, 2, 0};  // n1 = x3 - x2, n2 = x1 - x3, n3 = x2 - x1
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    n[a] = {(xd[ia[a]].x - xd[ib[a]].x) * invL, (xd[ia[a]].y - xd[ib[a]].y) * invL,
            (xd[ia[a]].z - xd[ib[a]].z) * invL};
    if (DT)
      nd[a] = {(xdd[ia[a]].x - xdd[ib[a]].x) * invL, (xdd[ia[a]].y - xdd[ib[a]].y) * invL,
               (xdd[ia[a]].z - xdd[ib[a]].z) * invL};
  }
}

// sinc(x +- a) at x = xL, a = xL k.n, their derivatives, and delta = sinc(x-a) - sinc(x+a).
// delta is the T-channel weight; in FP32 it is taken from
//   delta = 2 x a sum_{m,n>=0} (-1)^{m+n} X^m A^n / ((2m+1)! (2n+1)! (2m+2n+3)),  X = x^2, A = a^2
// (from sinc y = int_0^1 cos(y t) dt, so delta = 2 int_0^1 sin(x t) sin(a t) dt), which has no
// cancellation; used for x < 0.3 where the direct difference loses precision.
template <class R>
__device__ inline void sinc_pair(R xL, R sxL, R cxL, R kn, R& sp, R& sm, R& dsp, R& dsm, R& delta) {
  R a = xL * kn;
  R sa, ca;
  sc_mod(a, &sa, &ca);
  R xp = xL + a, xm = xL - a;
  R sinp = sxL * ca + cxL * sa, cosp = cxL * ca - sxL * sa;
  R sinm = sxL * ca - cxL * sa, cosm = cxL * ca + sxL * sa;
  const R thr = (R)1e-2;
  if (fabs(xp) < thr) {
    R x2 = xp * xp;
    sp = (R)1 - x2 * ((R)(1.0 / 6.0) - x2 * ((R)(1.0 / 120.0) - x2 * (R)(1.0 / 5040.0)));
    dsp = -xp * ((R)(1.0 / 3.0) - x2 * ((R)(1.0 / 30.0) - x2 * (R)(1.0 / 840.0)));
  } else {
    R ip = rcp(xp);
    sp = sinp * ip;
    dsp = (cosp - sp) * ip;
  }
  if (fabs(xm) < thr) {
    R x2 = xm * xm;
    sm = (R)1 - x2 * ((R)(1.0 / 6.0) - x2 * ((R)(1.0 / 120.0) - x2 * (R)(1.0 / 5040.0)));
    dsm = -xm * ((R)(1.0 / 3.0) - x2 * ((R)(1.0 / 30.0) - x2 * (R)(1.0 / 840.0)));
  } else {
    R im = rcp(xm);
    sm = sinm * im;
    dsm = (cosm - sm) * im;
  }
  if (sizeof(R) == sizeof(float) && xL < (R)0.3) {
    const R X = xL * xL, A = a * a;
    // q_m(A) = sum_n c_mn A^n, total degree <= 3
    const R q0 = (R)(1.0 / 3.0) + A * ((R)(-1.0 / 30.0) + A * ((R)(1.0 / 840.0) + A * (R)(-1.0 / 45360.0)));
    const R q1 = (R)(-1.0 / 30.0) + A * ((R)(1.0 / 252.0) + A * (R)(-1.0 / 6480.0));
    const R q2 = (R)(1.0 / 840.0) + A * (R)(-1.0 / 6480.0);
    const R q3 = (R)(-1.0 / 45360.0);
    delta = (R)2 * xL * a * (q0 + X * (q1 + X * (q2 + X * q3)));
  } else {
    delta = sm - sp;
  }
}

template <class R>
struct LinksT {
  Cx<R> l[6];  // 12, 21, 23, 32, 31, 13
  Cx<R> tl;    // sum over pairs of delta_a C_a (T channel, before pf (1-D)/sqrt3)
};

// rows' constants in R
template <class R>
struct RowC {
  double tc;
  V3<R> p, q, k;
};

template <class R>
__device__ inline RowC<R> load_row(const double* __restrict__ pc) {
  RowC<R> r;
  r.tc = pc[PC_TC];
  r.p = {(R)pc[PC_P], (R)pc[PC_P + 1], (R)pc[PC_P + 2]};
  r.q = {(R)pc[PC_Q], (R)pc[PC_Q + 1], (R)pc[PC_Q + 2]};
  r.k = {(R)pc[PC_K], (R)pc[PC_K + 1], (R)pc[PC_K + 2]};
  return r;
}

// One mode at one bin: amplitude A, time t = tf + t_c (FP64) and the phasor
//   u = exp(-i Psi) exp(-2 pi i f t_c) exp(-2 pi i f k.x0(t))
// (FP64: three separate factors as in antenna.py:435, 530-533; FP32: one FP64-reduced phase)
template <class R>
struct ModeState {
  R A;
  double t;
  Cx<R> u;
};

template <class R>
__device__ inline void mode_state(const double* __restrict__ nodes, size_t nbase, const BinPre<R>& b,
                                  const RowC<R>& rc, Cx<R> etc, const Bary* bc_in, ModeState<R>& ms,
                                  double& Psi_out) {
  const double* n0 = nodes + (nbase + b.j) * 4;
  {
    R P0 = (R)n0[0], P1 = (R)n0[4], M0 = (R)(n0[1] * b.du), M1 = (R)(n0[5] * b.du), s = (R)b.s;
    R c2 = (R)3 * (P1 - P0) - (R)2 * M0 - M1;
    R c3 = (R)2 * (P0 - P1) + M0 + M1;
    ms.A = P0 + s * (M0 + s * (c2 + s * c3));
  }
  double P0 = n0[2], P1 = n0[6], M0 = n0[3] * b.du, M1 = n0[7] * b.du, s = b.s;
  double c2 = 3.0 * (P1 - P0) - 2.0 * M0 - M1;
  double c3 = 2.0 * (P0 - P1) + M0 + M1;
  Psi_out = P0 + s * (M0 + s * (c2 + s * c3));
  double dPsids = M0 + s * (2.0 * c2 + s * (3.0 * c3));
  ms.t = dPsids * b.inv2pidf + rc.tc;
  (void)bc_in;
  (void)etc;
}

template <class R>
__device__ inline void mode_phasor(const BinPre<R>& b, const RowC<R>& rc, Cx<R> etc, double Psi, const Bary& bc,
                                   ModeState<R>& ms) {
  if (sizeof(R) == sizeof(double)) {
    R sp, cp, s0, c0;
    sc_big(Psi, &sp, &cp);
    sc_big(-2.0 * b.pif * bc.kx0, &s0, &c0);
    ms.u = mul(mul(Cx<R>{cp, -sp}, etc), Cx<R>{c0, s0});
  } else {
    R sp, cp;
    sc_big(Psi + 2.0 * b.pif * (rc.tc + bc.kx0), &sp, &cp);
    ms.u = {cp, -sp};
  }
}

// forward: add one mode's links
template <class R>
__device__ inline void mode_links(const DetectorDev& det, const double* __restrict__ pc, int mi,
                                  const double* __restrict__ nodes, size_t nbase, const BinPre<R>& b,
                                  const RowC<R>& rc, Cx<R> etc, LinksT<R>& L) {
  const double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
  ModeState<R> ms;
  double Psi;
  mode_state<R>(nodes, nbase, b, rc, etc, nullptr, ms, Psi);
  V3<R> n[3], xd[3], nd_[3], xdd_[3];
  Bary bc;
  orbit_eval<R, false>(det, ms.t, rc.k, n, xd, nd_, xdd_, bc);
  mode_phasor<R>(b, rc, etc, Psi, bc, ms);
  Cx<R> hm = scale(ms.u, ms.A);
  Cx<R> hp = mul(Cx<R>{(R)mf[MF_FPR], (R)mf[MF_FPI]}, hm);
  Cx<R> hc = mul(Cx<R>{(R)mf[MF_FCR], (R)mf[MF_FCI]}, hm);
  Cx<R> ei[3];
  const R mpif = (R)(-b.pif);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R s_, c_;
    sc_mod(mpif * dot3(rc.k, xd[i]), &s_, &c_);
    ei[i] = {c_, s_};
  }
  // pair a: link vector n[a]; (+) and (-) links; nodes (i, j)
  const int lp[3] = {2, 4, 0}, lm[3] = {3, 5, 1};
  const int pi_[3] = {1, 2, 0}, pj_[3] = {2, 0, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R np = dot3(n[a], rc.p), nq = dot3(n[a], rc.q);
    R Pp = np * np - nq * nq, Pc = (R)2 * np * nq;
    Cx<R> nhn = {Pp * hp.re + Pc * hc.re, Pp * hp.im + Pc * hc.im};
    R kn = dot3(rc.k, n[a]);
    R sp, sm, dsp, dsm, dl;
    sinc_pair(b.xL, b.sxL, b.cxL, kn, sp, sm, dsp, dsm, dl);
    Cx<R> E = mul(b.eL, mul(ei[pi_[a]], ei[pj_[a]]));
    Cx<R> C = mul(Cx<R>{b.xL * nhn.im, -b.xL * nhn.re}, E);  // (-i xL) nhn E
    L.l[lp[a]] = add(L.l[lp[a]], scale(C, sp));
    L.l[lm[a]] = add(L.l[lm[a]], scale(C, sm));
    L.tl = add(L.tl, scale(C, dl));
  }
}

// TDI outputs: A, E from X, Y, Z (tdi.py:1784-1785); T from the pair sum, which equals
// (X + Y + Z)/sqrt3 = pf (1 - D) sum_a (sinc_- - sinc_+)_a C_a / sqrt3 without cancellation
template <class R>
__device__ inline void tdi_out(const DetectorDev& det, const BinPre<R>& b, const LinksT<R>& L, Cx<R> out[3]) {
  const Cx<R>* l = L.l;
  Cx<R> X = mul(b.pf, add(sub(l[5], l[0]), mul(b.D, sub(l[4], l[1]))));
  Cx<R> Y = mul(b.pf, add(sub(l[1], l[2]), mul(b.D, sub(l[0], l[3]))));
  Cx<R> Z = mul(b.pf, add(sub(l[3], l[4]), mul(b.D, sub(l[2], l[5]))));
  if (det.orthogonal) {
    const R r2 = (R)0.70710678118654752440, r6 = (R)0.40824829046386301637;
    out[0] = scale(sub(Z, X), r2);
    out[1] = scale(add(sub(X, scale(Y, (R)2)), Z), r6);
    out[2] = mul(b.pfT, L.tl);
  } else {
    out[0] = X;
    out[1] = Y;
    out[2] = Z;
  }
}

template <class R>
__device__ inline Cx<R> row_etc(const BinPre<R>& b, double tc) {
  Cx<R> etc = {(R)1, (R)0};
  if (sizeof(R) == sizeof(double)) {
    R s, c;
    sc_big(-2.0 * b.pif * tc, &s, &c);
    etc = {c, s};
  }
  return etc;
}

template <class R>
__device__ inline void response_row(const DetectorDev& det, const double* __restrict__ pc, const ModeList& modes,
                                    const double* __restrict__ nodes, size_t pbase, int NNS, const BinPre<R>& b,
                                    Cx<R> out[3]) {
  RowC<R> rc = load_row<R>(pc);
  Cx<R> etc = row_etc(b, rc.tc);
  LinksT<R> L;
#pragma unroll
  for (int i = 0; i < 6; ++i) L.l[i] = {(R)0, (R)0};
  L.tl = {(R)0, (R)0};
  for (int mi = 0; mi < modes.n; ++mi) {
    if (b.f > pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]) continue;
    mode_links<R>(det, pc, mi, nodes, pbase + (size_t)mi * NNS, b, rc, etc, L);
  }
  tdi_out(det, b, L, out);
}

// ------------------------------- adjoint ---------------------------------------------------
// wl: link adjoints from the A,E (or X,Y,Z) channels; wT: adjoint of the T pair sum
template <class R>
__device__ inline void mode_adjoint(const DetectorDev& det, const double* __restrict__ pc, int mi, int mid,
                                    const double* __restrict__ nodes, const double* __restrict__ nodes_grad,
                                    size_t nbase, const BinPre<R>& b, const RowC<R>& rc, Cx<R> etc,
                                    const Cx<R> wl[6], Cx<R> wT, double G[PES_N_PARAMS]) {
  const double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
  ModeState<R> ms;
  double Psi;
  mode_state<R>(nodes, nbase, b, rc, etc, nullptr, ms, Psi);
  V3<R> n[3], xd[3], nd[3], xdd[3];
  Bary bc;
  orbit_eval<R, true>(det, ms.t, rc.k, n, xd, nd, xdd, bc);
  mode_phasor<R>(b, rc, etc, Psi, bc, ms);
  Cx<R> hm = scale(ms.u, ms.A);
  Cx<R> Fp = {(R)mf[MF_FPR], (R)mf[MF_FPI]}, Fc = {(R)mf[MF_FCR], (R)mf[MF_FCI]};
  Cx<R> hp = mul(Fp, hm), hc = mul(Fc, hm);
  Cx<R> ei[3];
  const R mpif = (R)(-b.pif);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    R s_, c_;
    sc_mod(mpif * dot3(rc.k, xd[i]), &s_, &c_);
    ei[i] = {c_, s_};
  }
  const V3<R> dlp = {(R)pc[PC_DLP], (R)pc[PC_DLP + 1], (R)pc[PC_DLP + 2]};
  const V3<R> dlq = {(R)pc[PC_DLQ], (R)pc[PC_DLQ + 1], (R)pc[PC_DLQ + 2]};
  const V3<R> dbp = {(R)pc[PC_DBP], (R)pc[PC_DBP + 1], (R)pc[PC_DBP + 2]};
  const V3<R> dbq = {(R)pc[PC_DBQ], (R)pc[PC_DBQ + 1], (R)pc[PC_DBQ + 2]};
  const V3<R> dlk = {(R)pc[PC_DLK], (R)pc[PC_DLK + 1], (R)pc[PC_DLK + 2]};
  const V3<R> dbk = {(R)pc[PC_DBK], (R)pc[PC_DBK + 1], (R)pc[PC_DBK + 2]};
  const int lp[3] = {2, 4, 0}, lm[3] = {3, 5, 1};
  const int pi_[3] = {1, 2, 0}, pj_[3] = {2, 0, 1};
  Cx<R> Hp = {(R)0, (R)0}, Hc = {(R)0, (R)0};
  R tbar = (R)0, lbar = (R)0, bbar = (R)0, psbar = (R)0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    R np = dot3(n[a], rc.p), nq = dot3(n[a], rc.q);
    R Pp = np * np - nq * nq, Pc = (R)2 * np * nq;
    Cx<R> nhn = {Pp * hp.re + Pc * hc.re, Pp * hp.im + Pc * hc.im};
    R kn = dot3(rc.k, n[a]);
    R sp, sm, dsp, dsm, dl;
    sinc_pair(b.xL, b.sxL, b.cxL, kn, sp, sm, dsp, dsm, dl);
    Cx<R> E = mul(b.eL, mul(ei[pi_[a]], ei[pj_[a]]));
    Cx<R> C = mul(Cx<R>{b.xL * nhn.im, -b.xL * nhn.re}, E);
    const Cx<R> wp = wl[lp[a]], wm = wl[lm[a]];
    R gp = wp.re * C.re + wp.im * C.im;  // Re[conj(w) C]
    R gm = wm.re * C.re + wm.im * C.im;
    R gT = wT.re * C.re + wT.im * C.im;
    R Sbar = b.xL * (dsp * gp - dsm * gm - (dsm + dsp) * gT);  // adjoint of k.n_a
    Cx<R> W = {sp * wp.re + sm * wm.re + dl * wT.re, sp * wp.im + sm * wm.im + dl * wT.im};
    Cx<R> cw = conjx(W);
    Cx<R> V = mul(mul(cw, Cx<R>{(R)0, -b.xL}), E);  // conj(W) (-i xL) E
    Cx<R> cwC = mul(cw, C);
    R Tbar = (R)b.pif * cwC.im;  // adjoint of k.(xi_det + xj_det)
    Cx<R> Vhp = mul(V, hp), Vhc = mul(V, hc);
    R npbar = (R)2 * (np * Vhp.re + nq * Vhc.re);
    R nqbar = (R)2 * (np * Vhc.re - nq * Vhp.re);
    Hp.re += V.re * Pp;
    Hp.im += V.im * Pp;
    Hc.re += V.re * Pc;
    Hc.im += V.im * Pc;
    const V3<R>& xi = xd[pi_[a]];
    const V3<R>& xj = xd[pj_[a]];
    V3<R> X = {xi.x + xj.x, xi.y + xj.y, xi.z + xj.z};
    V3<R> Xd = {xdd[pi_[a]].x + xdd[pj_[a]].x, xdd[pi_[a]].y + xdd[pj_[a]].y, xdd[pi_[a]].z + xdd[pj_[a]].z};
    tbar += npbar * dot3(nd[a], rc.p) + nqbar * dot3(nd[a], rc.q) + Sbar * dot3(rc.k, nd[a]) +
            Tbar * dot3(rc.k, Xd);
    lbar += npbar * dot3(n[a], dlp) + nqbar * dot3(n[a], dlq) + Sbar * dot3(dlk, n[a]) + Tbar * dot3(dlk, X);
    bbar += npbar * dot3(n[a], dbp) + nqbar * dot3(n[a], dbq) + Sbar * dot3(dbk, n[a]) + Tbar * dot3(dbk, X);
    psbar += npbar * nq - nqbar * np;
  }
  // waveform-side adjoints: Gm = Hp Fp + Hc Fc, contributions Re[Gm dh_m]
  Cx<R> Gm = add(mul(Hp, Fp), mul(Hc, Fc));
  Cx<R> Gh = mul(Gm, hm);
  const double phibar = (double)Gh.im;  // adjoint of the total phase Psi + 2 pi f (t_c + k.x0)
  const double kx0bar = 2.0 * b.pif * phibar;
  const double tb = (double)tbar + kx0bar * bc.kx0_t;
  G[7] += (double)lbar + kx0bar * AU_SEC * ((double)dlk.x * bc.ca + (double)dlk.y * bc.sa);
  G[8] += (double)bbar + kx0bar * AU_SEC * ((double)dbk.x * bc.ca + (double)dbk.y * bc.sa);
  G[9] += (double)psbar;
  const double m = (double)kModeM[mid];
  G[4] += -(double)Gh.re * pc[PC_INVDL];
  G[5] += -m * phibar;
  {
    Cx<R> dFp = {(R)mf[MF_DFPR], (R)mf[MF_DFPI]}, dFc = {(R)mf[MF_DFCR], (R)mf[MF_DFCI]};
    Cx<R> tt = mul(add(mul(Hp, dFp), mul(Hc, dFc)), hm);
    G[6] += (double)tt.re;
  }
  G[10] += 2.0 * b.pif * phibar + tb;
  Cx<R> Gu = mul(Gm, ms.u);
  const R Abar = Gu.re, Psibar = (R)phibar;
  const double* g0 = nodes_grad + (nbase + b.j) * 16;
  const R s = (R)b.s;
  const R rdu = (R)b.du;
  const R tbi = (R)(tb * b.inv2pidf);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R a0 = (R)g0[4 * i + 0], a1 = (R)g0[16 + 4 * i + 0];
    R am0 = (R)g0[4 * i + 1] * rdu, am1 = (R)g0[16 + 4 * i + 1] * rdu;
    R p0 = (R)g0[4 * i + 2], p1 = (R)g0[16 + 4 * i + 2];
    R pm0 = (R)g0[4 * i + 3] * rdu, pm1 = (R)g0[16 + 4 * i + 3] * rdu;
    R ac2 = (R)3 * (a1 - a0) - (R)2 * am0 - am1, ac3 = (R)2 * (a0 - a1) + am0 + am1;
    R pc2 = (R)3 * (p1 - p0) - (R)2 * pm0 - pm1, pc3 = (R)2 * (p0 - p1) + pm0 + pm1;
    R dA = a0 + s * (am0 + s * (ac2 + s * ac3));
    R dP = p0 + s * (pm0 + s * (pc2 + s * pc3));
    R dPs = pm0 + s * ((R)2 * pc2 + s * ((R)3 * pc3));
    G[i] += (double)(Abar * dA + Psibar * dP + tbi * dPs);
  }
}

// ------------------------------- kernels ---------------------------------------------------
template <class R, bool GRAD>
__global__ __launch_bounds__(TILE) void k_lnl(const DetectorDev* __restrict__ dets, int n_det, int n_tiles,
                                              const double* __restrict__ bin_f, const int* __restrict__ bin_j,
                                              const double* __restrict__ bin_s, const double* __restrict__ bin_w,
                                              const double* __restrict__ dre, const double* __restrict__ dim, int NB,
                                              const double* __restrict__ pconst, const double* __restrict__ nodes,
                                              const double* __restrict__ nodes_grad, int NNS, double du,
                                              ModeList modes, int P, double* __restrict__ partial) {
  constexpr int NOUT = GRAD ? 1 + PES_N_PARAMS : 1;
  __shared__ double acc[TILE / 64][CHUNK][NOUT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n_chunks = (P + CHUNK - 1) / CHUNK;
  const int n_groups = (n_tiles + TILES_PER_GROUP - 1) / TILES_PER_GROUP;
  const int n_work = n_chunks * n_groups;
  for (int w = blockIdx.x; w < n_work; w += gridDim.x) {
    const int g = w / n_chunks, c = w % n_chunks;
    const int p0 = c * CHUNK, p1 = min(P, p0 + CHUNK);
    for (int i = threadIdx.x; i < (TILE / 64) * CHUNK * NOUT; i += TILE) (&acc[0][0][0])[i] = 0.0;
    __syncthreads();
    const int t_end = min(n_tiles, (g + 1) * TILES_PER_GROUP);
    for (int tile = g * TILES_PER_GROUP; tile < t_end; ++tile) {
      const DetectorDev det = dets[tile_detector(dets, n_det, tile)];
      const int bi = tile * TILE + threadIdx.x;
      BinPre<R> b;
      bin_prepare<R>(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
      double wc[3];
      Cplx dc[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        wc[ch] = bin_w[ch * NB + bi];
        dc[ch] = {dre[ch * NB + bi], dim[ch * NB + bi]};
      }
      for (int p = p0; p < p1; ++p) {
        const double* pc = pconst + (size_t)p * PC_STRIDE;
        double G[PES_N_PARAMS];
#pragma unroll
        for (int i = 0; i < PES_N_PARAMS; ++i) G[i] = 0.0;
        double v = 0.0;
        if (pc[PC_VALID] != 0.0) {
          const size_t pbase = (size_t)p * modes.n * NNS;
          Cx<R> out[3];
          response_row<R>(det, pc, modes, nodes, pbase, NNS, b, out);
          Cx<R> wch[3];
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            double rr = dc[ch].re - (double)out[ch].re, ri = dc[ch].im - (double)out[ch].im;
            v -= wc[ch] * (rr * rr + ri * ri);
            wch[ch] = {(R)(2.0 * wc[ch] * rr), (R)(2.0 * wc[ch] * ri)};
          }
          if (GRAD) {
            Cx<R> wX, wY, wZ, wT = {(R)0, (R)0};
            if (det.orthogonal) {
              const R r2 = (R)0.70710678118654752440, r6 = (R)0.40824829046386301637;
              wX = {-r2 * wch[0].re + r6 * wch[1].re, -r2 * wch[0].im + r6 * wch[1].im};
              wY = {(R)-2 * r6 * wch[1].re, (R)-2 * r6 * wch[1].im};
              wZ = {r2 * wch[0].re + r6 * wch[1].re, r2 * wch[0].im + r6 * wch[1].im};
              wT = mul(conjx(b.pfT), wch[2]);
            } else {
              wX = wch[0];
              wY = wch[1];
              wZ = wch[2];
            }
            Cx<R> cpf = conjx(b.pf), cD = conjx(b.D);
            Cx<R> WX = mul(cpf, wX), WY = mul(cpf, wY), WZ = mul(cpf, wZ);
            Cx<R> DWX = mul(cD, WX), DWY = mul(cD, WY), DWZ = mul(cD, WZ);
            Cx<R> wl[6];
            wl[0] = sub(DWY, WX);
            wl[1] = sub(WY, DWX);
            wl[2] = sub(DWZ, WY);
            wl[3] = sub(WZ, DWY);
            wl[4] = sub(DWX, WZ);
            wl[5] = sub(WX, DWZ);
            RowC<R> rc = load_row<R>(pc);
            Cx<R> etc = row_etc(b, rc.tc);
            for (int mi = 0; mi < modes.n; ++mi) {
              if (b.f > pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]) continue;
              mode_adjoint<R>(det, pc, mi, modes.id[mi], nodes, nodes_grad, pbase + (size_t)mi * NNS, b, rc, etc,
                              wl, wT, G);
            }
          }
        }
        v = wave_sum(v);
        if (lane == 0) acc[wave][p - p0][0] += v;
        if (GRAD) {
#pragma unroll
          for (int i = 0; i < PES_N_PARAMS; ++i) {
            double s = wave_sum(G[i]);
            if (lane == 0) acc[wave][p - p0][1 + i] += s;
          }
        }
      }
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

template <class R>
__global__ void k_response_t(const DetectorDev* __restrict__ dets, int det_i, const double* __restrict__ bin_f,
                             const int* __restrict__ bin_j, const double* __restrict__ bin_s,
                             const double* __restrict__ pconst, const double* __restrict__ nodes, int NNS, double du,
                             ModeList modes, int P, double* __restrict__ out) {
  const DetectorDev det = dets[det_i];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int p = blockIdx.y;
  if (k >= det.n_bins || p >= P) return;
  int bi = det.bin_offset + k;
  BinPre<R> b;
  bin_prepare<R>(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
  const double* pc = pconst + (size_t)p * PC_STRIDE;
  Cx<R> o[3];
  response_row<R>(det, pc, modes, nodes, (size_t)p * modes.n * NNS, NNS, b, o);
  for (int ch = 0; ch < 3; ++ch) {
    size_t idx = (((size_t)p * 3 + ch) * det.n_bins + k) * 2;
    out[idx] = (double)o[ch].re;
    out[idx + 1] = (double)o[ch].im;
  }
}

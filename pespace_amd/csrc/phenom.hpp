// IMRPhenomD (2,2) and PhenomHM-style higher modes on the device.
//
// The reference imports its waveforms from packages absent here (tiwave,
// ripplegw; pespace/waveform/interface_ripple.py:16-24) and fixes only the
// container contract: per mode {plus, cross, tf} with t_c = 0, h = A exp(-i Psi),
// tf = (1/2pi) dPsi/df (interface_ripple.py:42-51, antenna.py:435-449).
//
// Model: Khan et al. 2016 (arXiv:1508.07253) / Husa et al. 2016 (arXiv:1508.07250)
// for the 22 mode, with the TaylorF2 phasing coefficients of the reference's
// get_PN_d_phase (pespace/waveform/interface_lalsim.py:353-433) through 3.5PN,
// and London et al. 2018 (arXiv:1708.00404)'s frequency-map construction for
// (2,1),(3,3),(3,2),(4,4),(4,3).  Ringdown frequencies come from a Kerr QNM table
// computed with Leaver's continued fraction (qnm_table.inc).
//
// Everything is templated on the scalar type so that one code path yields plain
// values (double), first derivatives (Dual<double,N>) and the mixed second
// derivatives the interpolation nodes need (Dual<Dual<double,1>,4>).
#pragma once
#include "dual.hpp"
#include "qnm_table.inc"

namespace pes {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double MRSUN_SI = 1.476625038050124729627979840144936351e3;
constexpr double MTSUN_SI = 4.925490947641266978197229498498379006e-6;
constexpr double PC_SI = 3.085677581491367e16;
constexpr double C_SI = 299792458.0;
constexpr double YEAR_SI = 31557600.0;
constexpr double DAY_SI = 86400.0;
constexpr double AU_SI = 149597870700.0;
constexpr double EULER_GAMMA = 0.5772156649015329;
constexpr double AMP_F_INS = 0.014;
constexpr double PHI_F_INS = 0.018;
constexpr double MF_CUT = 0.2;
constexpr double HM_MAP_JOIN = 0.014;
constexpr double TAPER = 0.1;

// ---- calibrated coefficients (Khan et al. 2016 table V) --------------------------------
// order: l00, l10, (xi) l01 l11 l21, (xi^2) l02 l12 l22, (xi^3) l03 l13 l23
enum FitId {
  F_RHO1, F_RHO2, F_RHO3, F_V2, F_G1, F_G2, F_G3, F_S1, F_S2, F_S3, F_S4, F_B1, F_B2, F_B3,
  F_A1, F_A2, F_A3, F_A4, F_A5, N_FIT
};
__constant__ static const double kFits[N_FIT][11] = {
    {3931.8979897196696, -17395.758706812805, 3132.375545898835, 343965.86092361377, -1.2162565819981997e6,
     -70698.00600428853, 1.383907177859705e6, -3.9662761890979446e6, -60017.52423652596, 803515.1181825735,
     -2.091710365941658e6},
    {-40105.47653771657, 112253.0169706701, 23561.696065836168, -3.476180699403351e6, 1.137593670849482e7,
     754313.1127166454, -1.308476044625268e7, 3.6444584853928134e7, 596226.612472288, -7.4277901143564405e6,
     1.8928977514040343e7},
    {83208.35471266537, -191237.7264145924, -210916.2454782992, 8.71797508352568e6, -2.6914942420669552e7,
     -1.9889806527362722e6, 3.0888029960154563e7, -8.390870279256162e7, -1.4535031953446497e6,
     1.7063528990822166e7, -4.2748659731120914e7},
    {0.8149838730507785, 2.5747553517454658, 1.1610198035496786, -2.3627771785551537, 6.771038707057573,
     0.7570782938606834, -2.7256896890432474, 7.1140380397149965, 0.1766934149293479, -0.7978690983168183,
     2.1162391502005153},
    {0.006927402739328343, 0.03020474290328911, 0.006308024337706171, -0.12074130661131138, 0.26271598905781324,
     0.0034151773647198794, -0.10779338611188374, 0.27098966966891747, 0.0007374185938559283,
     -0.02749621038376281, 0.0733150789135702},
    {1.010344404799477, 0.0008993122007234548, 0.283949116804459, -4.049752962958005, 13.207828172665366,
     0.10396278486805426, -7.025059158961947, 24.784892370130475, 0.03093202475605892, -2.6924023896851663,
     9.609374464684983},
    {1.3081615607036106, -0.005537729694807678, -0.06782917938621007, -0.6689834970767117, 3.403147966134083,
     -0.05296577374411866, -0.9923793203111362, 4.820681208409587, -0.006134139870393713, -0.38429253308696365,
     1.7561754421985984},
    {2096.551999295543, 1463.7493168261553, 1312.5493286098522, 18307.330017082117, -43534.1440746107,
     -833.2889543511114, 32047.31997183187, -108609.45037520859, 452.25136398112204, 8353.439546391714,
     -44531.3250037322},
    {-10114.056472621156, -44631.01109458185, -6541.308761668722, -266959.23419307504, 686328.3229317984,
     3405.6372187679685, -437507.7208209015, 1.6318171307344697e6, -7462.648563007646, -114585.25177153319,
     674402.4689098676},
    {22933.658273436497, 230960.00814979506, 14961.083974183695, 1.1940181342318142e6, -3.1042239693052764e6,
     -3038.166617199259, 1.8720322849093592e6, -7.309145012085539e6, 42738.22871475411, 467502.018616601,
     -3.064853498512499e6},
    {-14621.71522218357, -377812.8579387104, -9608.682631509726, -1.7108925257214056e6, 4.332924601416521e6,
     -22366.683262266528, -2.5019716386377467e6, 1.0274495902259542e7, -85360.30079034246, -570025.3441737515,
     4.396844346849777e6},
    {97.89747327985583, -42.659730877489224, 153.48421037904913, -1417.0620760768954, 2752.8614143665027,
     138.7406469558649, -1433.6585075135881, 2857.7418952430758, 41.025109467376126, -423.680737974639,
     850.3594335657173},
    {-3.282701958759534, -9.051384468245866, -12.415449742258042, 55.4716447709787, -106.05109938966335,
     -11.953044553690658, 76.80704618365418, -155.33172948098394, -3.4129261592393263, 25.572377569952536,
     -54.408036707740465},
    {-0.000025156429818799565, 0.000019750256942201327, -0.000018370671469295915, 0.000021886317041311973,
     0.00008250240316860033, 7.157371250566708e-6, -0.000055780000112270685, 0.00019142082884072178,
     5.447166261464217e-6, -0.00003220610095021982, 0.00007974016714984341},
    {43.31514709695348, 638.6332679188081, -32.85768747216059, 2415.8938269370315, -5766.875169379177,
     -61.85459307173841, 2953.967762459948, -8986.29057591497, -21.571435779762044, 981.2158224673428,
     -3239.5664895930286},
    {-0.07020209449091723, -0.16269798450687084, -0.1872514685185499, 1.138313650449945, -2.8334196304430046,
     -0.17137955686840617, 1.7197549338119527, -4.539717148261272, -0.049983437357548705, 0.6062072055948309,
     -1.682769616644546},
    {9.5988072383479, -397.05438595557433, 16.202126189517813, -1574.8286986717037, 3600.3410843831093,
     27.092429659075467, -1786.482357315139, 5152.919378666511, 11.175710130033895, -577.7999423177481,
     1808.730762932043},
    {-0.02989487384493607, 1.4022106448583738, -0.07356049468633846, 0.8337006542278661, 0.2240008282397391,
     -0.055202870001177226, 0.5667186343606578, 0.7186931973380503, -0.015507437354325743, 0.15750322779277187,
     0.21076815715176228},
    {0.9974408278363099, -0.007884449714907203, -0.059046901195591035, 1.3958712396764088, -4.516631601676276,
     -0.05585343136869692, 1.7516580039343603, -5.990208965347804, -0.017945336522161195, 0.5965097794825992,
     -2.0608879367971804},
};

template <class S>
PES_HD S fit_eval(int id, const S& eta, const S& chi) {
  const double* c = kFits[id];
  S xi = chi - 1.0;
  S e2 = eta * eta;
  S t1 = c[2] + c[3] * eta + c[4] * e2;
  S t2 = c[5] + c[6] * eta + c[7] * e2;
  S t3 = c[8] + c[9] * eta + c[10] * e2;
  return c[0] + c[1] * eta + xi * (t1 + xi * (t2 + xi * t3));
}

// Catmull-Rom interpolation of the Leaver QNM table (same scheme as the oracle).
template <class S>
PES_HD S qnm_interp(const double* vals, const S& a) {
  S x = (a - kQnmAmin) * (1.0 / kQnmStep);
  int i = (int)floor(value_of(x));
  i = i < 1 ? 1 : (i > kQnmN - 3 ? kQnmN - 3 : i);
  S t = x - (double)i;
  double p0 = vals[i], p1 = vals[i + 1];
  double m0 = 0.5 * (vals[i + 1] - vals[i - 1]);
  double m1 = 0.5 * (vals[i + 2] - vals[i]);
  double c2 = 3.0 * (p1 - p0) - 2.0 * m0 - m1;
  double c3 = 2.0 * (p0 - p1) + m0 + m1;
  return p0 + t * (m0 + t * (c2 + t * c3));
}

// mode ids shared by host and device: 0:22 1:21 2:33 3:32 4:44 5:43
constexpr int kModeL[6] = {2, 2, 3, 3, 4, 4};
constexpr int kModeM[6] = {2, 1, 3, 2, 4, 3};

// ---- per-binary f-independent setup (one instance per particle) -------------------------
enum SetupId {
  SU_MSEC, SU_AMPSCALE, SU_ETA, SU_DELTA, SU_CHIS, SU_CHIA, SU_FRD, SU_FDM, SU_FMAX,
  SU_A23, SU_A1, SU_A43, SU_A53, SU_A2, SU_RHO1, SU_RHO2, SU_RHO3, SU_G1, SU_G2, SU_G3,
  SU_D0, SU_D1, SU_D2, SU_D3, SU_D4,
  SU_P0, SU_P2, SU_P3, SU_P4, SU_P5L, SU_P6, SU_P6L, SU_P7,
  SU_S1, SU_S2, SU_S3, SU_S4, SU_B1, SU_B2, SU_B3, SU_AL1, SU_AL2, SU_AL3, SU_AL4, SU_AL5,
  SU_C1INT, SU_C2INT, SU_C1MRD, SU_C2MRD, SU_T0, SU_PHIREF0, SU_MFREF,
  SU_FRD21, SU_FRD33, SU_FRD32, SU_FRD44, SU_FRD43,
  N_SETUP
};

// Coefficient store of Phenom: an array (written by phenom_setup) or, in the node kernels,
// SetupView: a read-only view of one row of the setup table [N_SETUP][5] (value and d/dtheta
// for Mc, q, chi1, chi2), building each coefficient as the requested dual on access, so the
// kernels hold no per-lane copy of the ~60 coefficients.
template <class S>
struct SetupView;
template <>
struct SetupView<Dual<double, 1>> {
  const double* su;
  PES_HD Dual<double, 1> operator[](int i) const {
    Dual<double, 1> r;
    r.v = su[i * 5];
    r.d[0] = 0.0;
    return r;
  }
};
template <>
struct SetupView<Dual<Dual<double, 1>, 4>> {
  const double* su;
  PES_HD Dual<Dual<double, 1>, 4> operator[](int i) const {
    Dual<Dual<double, 1>, 4> r;
    r.v.v = su[i * 5];
    r.v.d[0] = 0.0;
    for (int k = 0; k < 4; ++k) {
      r.d[k].v = su[i * 5 + 1 + k];
      r.d[k].d[0] = 0.0;
    }
    return r;
  }
};

template <class S, class Store = S[N_SETUP]>
struct Phenom {
  Store c;

  // -- amplitude (reduced ansatz) --
  PES_HD S amp_ins(const S& f) const {
    S f13 = dcbrt(f);
    S f23 = f13 * f13;
    S f43 = f * f13;
    S f53 = f * f23;
    S f2 = f * f;
    S f73 = f2 * f13;
    S f83 = f2 * f23;
    return 1.0 + f23 * c[SU_A23] + f * c[SU_A1] + f43 * c[SU_A43] + f53 * c[SU_A53] + f2 * c[SU_A2] +
           f73 * c[SU_RHO1] + f83 * c[SU_RHO2] + f2 * f * c[SU_RHO3];
  }
  PES_HD S amp_mrd(const S& f) const {
    S fdg = c[SU_FDM] * c[SU_G3];
    S dfr = f - c[SU_FRD];
    return dexp(-(dfr * c[SU_G2]) / fdg) * (fdg * c[SU_G1]) / (dfr * dfr + fdg * fdg);
  }
  PES_HD S amp_int(const S& f) const {
    return c[SU_D0] + f * (c[SU_D1] + f * (c[SU_D2] + f * (c[SU_D3] + f * c[SU_D4])));
  }
  PES_HD S amplitude(const S& f) const {
    double fv = value_of(f);
    S red = fv <= AMP_F_INS ? amp_ins(f) : (fv >= value_of(c[SU_FMAX]) ? amp_mrd(f) : amp_int(f));
    S pre = dsqrt(c[SU_ETA] * (2.0 / 3.0)) * pow(PI, -1.0 / 6.0);
    return red * pre * dpow(f, -7.0 / 6.0);
  }
  // -- phase --
  PES_HD S phi_ins(const S& f) const {
    S f13 = dcbrt(f);
    S f23 = f13 * f13;
    S lf = dlog(f);
    S tf2 = c[SU_P0] / (f23 * f13 * f23) + c[SU_P2] / f + c[SU_P3] / f23 + c[SU_P4] / f13 + c[SU_P5L] * lf +
            c[SU_P6] * f13 + c[SU_P6L] * f13 * (lf + log(PI)) + c[SU_P7] * f23;
    S ps = c[SU_S1] * f + 0.75 * c[SU_S2] * (f * f13) + 0.6 * c[SU_S3] * (f * f23) + 0.5 * c[SU_S4] * (f * f);
    return ps / c[SU_ETA] - tf2;
  }
  PES_HD S phi_int_raw(const S& f) const {
    return (c[SU_B1] * f - c[SU_B3] / (3.0 * f * f * f) + c[SU_B2] * dlog(f)) / c[SU_ETA];
  }
  PES_HD S phi_mrd_raw(const S& f) const {
    return (c[SU_AL1] * f - c[SU_AL2] / f + (4.0 / 3.0) * c[SU_AL3] * dpow(f, 0.75) +
            c[SU_AL4] * datan((f - c[SU_AL5] * c[SU_FRD]) / c[SU_FDM])) /
           c[SU_ETA];
  }
  PES_HD S phase(const S& f) const {
    double fv = value_of(f);
    if (fv < PHI_F_INS) return phi_ins(f);
    if (fv >= 0.5 * value_of(c[SU_FRD])) return phi_mrd_raw(f) + c[SU_C1MRD] + c[SU_C2MRD] * f;
    return phi_int_raw(f) + c[SU_C1INT] + c[SU_C2INT] * f;
  }
  // aligned 22 phase: Psi(f_ref) = 0, dPsi/dMf(f_peak) = 0
  PES_HD S psi22(const S& f) const { return phase(f) - c[SU_PHIREF0] - c[SU_T0] * (f - c[SU_MFREF]); }

  PES_HD S frd_mode(int mode) const {
    switch (mode) {
      case 1: return c[SU_FRD21];
      case 2: return c[SU_FRD33];
      case 3: return c[SU_FRD32];
      case 4: return c[SU_FRD44];
      case 5: return c[SU_FRD43];
      default: return c[SU_FRD];
    }
  }
  // f22(f) frequency map (mode != 0)
  PES_HD S fmap(int mode, const S& f) const {
    if (mode == 0) return f;
    const double m = (double)kModeM[mode];
    S fr = frd_mode(mode);
    S rho = c[SU_FRD] / fr;
    const double fi = 0.5 * m * HM_MAP_JOIN;
    double fv = value_of(f);
    if (fv < fi) return (2.0 / m) * f;
    if (fv >= value_of(fr)) return rho * f;
    S w = fr - fi;
    S t = (f - fi) / w;
    S M0 = (2.0 / m) * w;
    S M1 = rho * w;
    S c2 = 3.0 * (c[SU_FRD] - HM_MAP_JOIN) - 2.0 * M0 - M1;
    S c3 = 2.0 * (HM_MAP_JOIN - c[SU_FRD]) + M0 + M1;
    return HM_MAP_JOIN + t * (M0 + t * (c2 + t * c3));
  }
  // signed leading-order PN amplitude ratio
  PES_HD S beta(int mode, const S& f) const {
    if (mode == 0) return lift<S>(1.0);
    const double m = (double)kModeM[mode];
    const double fi = 0.5 * m * HM_MAP_JOIN;
    S vraw = dcbrt(f * (2.0 * PI / m));
    const double vi = cbrt(fi * (2.0 * PI / m));
    // C1 saturation above f_i: v = vi (2 - exp(1 - vraw/vi))
    S v = value_of(f) < fi ? vraw : vi * (2.0 - dexp(1.0 - vraw * (1.0 / vi)));
    const S& d = c[SU_DELTA];
    const S& eta = c[SU_ETA];
    switch (mode) {
      case 1: return (sqrt(2.0) / 3.0) * (d * v - 1.5 * (c[SU_CHIA] + d * c[SU_CHIS]) * v * v);
      case 2: return (sqrt(2.0 / 3.0) * 0.75 * sqrt(15.0 / 14.0)) * d * v;
      case 3: return ((1.0 / 3.0) * sqrt(5.0 / 7.0)) * (1.0 - 3.0 * eta) * v * v;
      case 4: return (sqrt(0.5) * (8.0 / 9.0) * sqrt(5.0 / 7.0)) * (1.0 - 3.0 * eta) * v * v;
      default: return (sqrt(2.0 / 3.0) * (9.0 / (4.0 * sqrt(70.0)))) * d * (1.0 - 2.0 * eta) * v * v * v;
    }
  }
  // amplitude (m s, no 1/dL) and phase of a mode at Mf
  // amplitude with a C1 raised-cosine taper over the last TAPER of the mode's range
  PES_HD void amp_phase(int mode, const S& Mf, S& A, S& P) const {
    S f22 = fmap(mode, Mf);
    A = beta(mode, Mf) * amplitude(f22) * c[SU_AMPSCALE];
    S x = Mf / mode_cut_s(mode);
    const double x0 = 1.0 - TAPER;
    double xv = value_of(x);
    if (xv >= 1.0) {
      A = A * 0.0;
    } else if (xv > x0) {
      S w = 0.5 * (1.0 + dcos((x - x0) * (PI / TAPER)));
      A = A * w;
    }
    P = (0.5 * (double)kModeM[mode]) * psi22(f22);
  }
  // mode cut (Mf) as a differentiable quantity
  PES_HD S mode_cut_s(int mode) const {
    if (mode == 0) return lift<S>(MF_CUT);
    return MF_CUT * frd_mode(mode) / c[SU_FRD];
  }
  // Mf beyond which the mode is zero (f22 = 0.2)
  PES_HD double mode_cut(int mode) const {
    if (mode == 0) return MF_CUT;
    return MF_CUT * value_of(frd_mode(mode)) / value_of(c[SU_FRD]);
  }
};

// ---- setup from physical parameters -------------------------------------------------
template <class S>
PES_HD S final_spin(const S& eta, const S& d, const S& chi1, const S& chi2) {
  S m1 = 0.5 * (1.0 + d), m2 = 0.5 * (1.0 - d);
  S s = m1 * m1 * chi1 + m2 * m2 * chi2;
  S e2 = eta * eta, e3 = e2 * eta;
  S s2 = s * s, s3 = s2 * s;
  return eta * (3.4641016151377544 - 4.399247300629289 * eta + 9.397292189321194 * e2 -
                13.180949901606242 * e3 +
                s * ((1.0 / eta - 0.0850917821418767 - 5.837029316602263 * eta) +
                     (0.1014665242971878 - 2.0967746996832157 * eta) * s +
                     (-1.3546806617824356 + 4.108962025369336 * eta) * s2 +
                     (-0.8676969352555539 + 2.064046835273906 * eta) * s3));
}

template <class S>
PES_HD S radiated_energy(const S& eta, const S& d, const S& chi1, const S& chi2) {
  S m1 = 0.5 * (1.0 + d), m2 = 0.5 * (1.0 - d);
  S s = (m1 * m1 * chi1 + m2 * m2 * chi2) / (m1 * m1 + m2 * m2);
  S e2 = eta * eta, e3 = e2 * eta, e4 = e3 * eta;
  return ((0.055974469826360077 * eta + 0.5809510763115132 * e2 - 0.9606726679372312 * e3 +
           3.352411249771192 * e4) *
          (1.0 + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * e2) * s)) /
         (1.0 + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * e2) * s);
}

// Gaussian elimination with partial pivoting (by value) on a 5x5 system.
template <class S>
PES_HD void solve5(S a[5][5], S b[5], S x[5]) {
  for (int k = 0; k < 5; ++k) {
    int piv = k;
    double best = fabs(value_of(a[k][k]));
    for (int r = k + 1; r < 5; ++r) {
      double v = fabs(value_of(a[r][k]));
      if (v > best) { best = v; piv = r; }
    }
    if (piv != k) {
      for (int cc = 0; cc < 5; ++cc) { S t = a[k][cc]; a[k][cc] = a[piv][cc]; a[piv][cc] = t; }
      S t = b[k]; b[k] = b[piv]; b[piv] = t;
    }
    for (int r = k + 1; r < 5; ++r) {
      S fct = a[r][k] / a[k][k];
      for (int cc = k; cc < 5; ++cc) a[r][cc] = a[r][cc] - fct * a[k][cc];
      b[r] = b[r] - fct * b[k];
    }
  }
  for (int k = 4; k >= 0; --k) {
    S acc = b[k];
    for (int cc = k + 1; cc < 5; ++cc) acc = acc - a[k][cc] * x[cc];
    x[k] = acc / a[k][k];
  }
}

// Build all setup coefficients from (Mc [Msun], q = m2/m1 <= 1, chi1, chi2) and f_ref [Hz].
// S is double or Dual<double,4> (derivatives w.r.t. the four intrinsic parameters).
template <class S>
PES_HD void phenom_setup(Phenom<S>& P, const S& Mc, const S& q, const S& chi1_in, const S& chi2_in, double f_ref) {
  // bilby convention: m1 >= m2, chi1 on m1
  S m1 = Mc * dpow(1.0 + q, 0.2) / dpow(q, 0.6);
  S m2 = m1 * q;
  S chi1 = chi1_in, chi2 = chi2_in;
  if (value_of(q) > 1.0) {  // keep m1 >= m2
    S t = m1; m1 = m2; m2 = t;
    S u = chi1; chi1 = chi2; chi2 = u;
  }
  S M = m1 + m2;
  S eta = m1 * m2 / (M * M);
  if (value_of(eta) > 0.25) eta = lift<S>(0.25);
  // delta = (m1 - m2) / M = sqrt(1 - 4 eta), formed from the masses: the same value (without the
  // cancellation of 1 - 4 eta near equal masses) and a finite derivative at q = 1
  S delta = (m1 - m2) / M;
  S chis = 0.5 * (chi1 + chi2), chia = 0.5 * (chi1 - chi2);
  S chipn = chis * (1.0 - 76.0 * eta / 113.0) + delta * chia;
  S* c = P.c;
  for (int i = 0; i < N_SETUP; ++i) c[i] = lift<S>(0.0);
  c[SU_MSEC] = M * MTSUN_SI;
  c[SU_AMPSCALE] = (M * MRSUN_SI) * (M * MTSUN_SI);
  c[SU_ETA] = eta;
  c[SU_DELTA] = delta;
  c[SU_CHIS] = chis;
  c[SU_CHIA] = chia;
  S af = final_spin(eta, delta, chi1, chi2);
  S erad = radiated_energy(eta, delta, chi1, chi2);
  S inv1me = 1.0 / (1.0 - erad);
  c[SU_FRD] = qnm_interp(kQnmFring[0], af) * inv1me;
  c[SU_FDM] = qnm_interp(kQnmFdamp[0], af) * inv1me;
  c[SU_FRD21] = qnm_interp(kQnmFring[1], af) * inv1me;
  c[SU_FRD33] = qnm_interp(kQnmFring[2], af) * inv1me;
  c[SU_FRD32] = qnm_interp(kQnmFring[3], af) * inv1me;
  c[SU_FRD44] = qnm_interp(kQnmFring[4], af) * inv1me;
  c[SU_FRD43] = qnm_interp(kQnmFring[5], af) * inv1me;
  // fits
  c[SU_RHO1] = fit_eval(F_RHO1, eta, chipn);
  c[SU_RHO2] = fit_eval(F_RHO2, eta, chipn);
  c[SU_RHO3] = fit_eval(F_RHO3, eta, chipn);
  c[SU_G1] = fit_eval(F_G1, eta, chipn);
  c[SU_G2] = fit_eval(F_G2, eta, chipn);
  c[SU_G3] = fit_eval(F_G3, eta, chipn);
  c[SU_S1] = fit_eval(F_S1, eta, chipn);
  c[SU_S2] = fit_eval(F_S2, eta, chipn);
  c[SU_S3] = fit_eval(F_S3, eta, chipn);
  c[SU_S4] = fit_eval(F_S4, eta, chipn);
  c[SU_B1] = fit_eval(F_B1, eta, chipn);
  c[SU_B2] = fit_eval(F_B2, eta, chipn);
  c[SU_B3] = fit_eval(F_B3, eta, chipn);
  c[SU_AL1] = fit_eval(F_A1, eta, chipn);
  c[SU_AL2] = fit_eval(F_A2, eta, chipn);
  c[SU_AL3] = fit_eval(F_A3, eta, chipn);
  c[SU_AL4] = fit_eval(F_A4, eta, chipn);
  c[SU_AL5] = fit_eval(F_A5, eta, chipn);
  S v2 = fit_eval(F_V2, eta, chipn);
  // PN amplitude coefficients (Khan et al. 2016 eq. 29)
  {
    const double pi = PI;
    S dp1 = delta + 1.0;
    S c12 = chi1 * chi1, c22 = chi2 * chi2;
    S e2 = eta * eta, e3 = e2 * eta;
    c[SU_A23] = ((-969.0 + 1804.0 * eta) * pow(pi, 2.0 / 3.0)) / 672.0;
    c[SU_A1] = ((chi1 * (81.0 * dp1 - 44.0 * eta) + chi2 * (81.0 - 81.0 * delta - 44.0 * eta)) * pi) / 48.0;
    c[SU_A43] = ((-27312085.0 - 10287648.0 * c22 - 10287648.0 * c12 * dp1 + 10287648.0 * c22 * delta +
                  24.0 * (-1975055.0 + 857304.0 * c12 - 994896.0 * chi1 * chi2 + 857304.0 * c22) * eta +
                  35371056.0 * e2) *
                 pow(pi, 4.0 / 3.0)) /
                8.128512e6;
    c[SU_A53] = (pow(pi, 5.0 / 3.0) *
                 (chi2 * (-285197.0 * (delta - 1.0) + 4.0 * (-91902.0 + 1579.0 * delta) * eta - 35632.0 * e2) +
                  chi1 * (285197.0 * dp1 - 4.0 * (91902.0 + 1579.0 * delta) * eta - 35632.0 * e2) +
                  42840.0 * (-1.0 + 4.0 * eta) * pi)) /
                32256.0;
    c[SU_A2] =
        -(pi * pi *
          (-336.0 * (-3248849057.0 + 2943675504.0 * c12 - 3339284256.0 * chi1 * chi2 + 2943675504.0 * c22) * e2 -
           324322727232.0 * e3 -
           7.0 * (-177520268561.0 + 107414046432.0 * c22 + 107414046432.0 * c12 * dp1 -
                  107414046432.0 * c22 * delta + 11087290368.0 * (chi1 + chi2 + chi1 * delta - chi2 * delta) * pi) +
           12.0 * eta *
               (-545384828789.0 - 176491177632.0 * chi1 * chi2 + 202603761360.0 * c22 +
                77616.0 * c12 * (2610335.0 + 995766.0 * delta) - 77287373856.0 * c22 * delta +
                5841690624.0 * (chi1 + chi2) * pi + 21384760320.0 * pi * pi))) /
        6.0085960704e10;
  }
  // TaylorF2 coefficients (interface_lalsim.py:353-433), with their -3/(128 eta) factor
  {
    const double pi = PI;
    const double p13 = cbrt(pi), p23 = p13 * p13, p43 = pi * p13, p53 = p43 * p13;
    S e2 = eta * eta, e3 = e2 * eta;
    S ca = chia, cs = chis;
    S ca2 = ca * ca, cs2 = cs * cs;
    S cf = -3.0 / 128.0 / eta;
    c[SU_P0] = cf * (1.0 / p53);
    c[SU_P2] = (37.15 / 7.56 + 55.0 / 9.0 * eta) * cf / pi;
    c[SU_P3] = (-16.0 * pi + 113.0 / 3.0 * delta * ca + (113.0 / 3.0 - 76.0 / 3.0 * eta) * cs) * cf / p23;
    c[SU_P4] = (152.93365 / 5.08032 + 271.45 / 5.04 * eta + 308.5 / 7.2 * e2 + (-405.0 / 8.0 + 200.0 * eta) * ca2 -
                405.0 / 4.0 * delta * ca * cs + (-405.0 / 8.0 + 5.0 / 2.0 * eta) * cs2) *
               cf / p13;
    c[SU_P5L] = ((386.45 / 7.56 - 65.0 / 9.0 * eta) * pi + (-732.985 / 2.268 - 140.0 / 9.0 * eta) * delta * ca +
                 (-732.985 / 2.268 + 2426.0 / 8.1 * eta + 340.0 / 9.0 * e2) * cs) *
                cf;
    c[SU_P6] = ((11583.231236531 / 4.694215680 - 640.0 / 3.0 * pi * pi - 684.8 / 2.1 * EULER_GAMMA +
                 (-15737.765635 / 3.048192 + 225.5 / 1.2 * pi * pi) * eta + 76.055 / 1.728 * e2 -
                 127.825 / 1.296 * e3 - log(2.0) * 1369.6 / 2.1) +
                2270.0 / 3.0 * pi * delta * ca + (2270.0 / 3.0 - 520.0 * eta) * pi * cs +
                (755.15 / 1.44 - 822.5 / 1.8 * eta) * delta * cs * ca +
                (755.15 / 2.88 - 2632.45 / 2.52 * eta - 480.0 * e2) * ca2 +
                (755.15 / 2.88 - 2324.15 / 5.04 * eta + 1255.0 / 9.0 * e2) * cs2) *
               p13 * cf;
    c[SU_P6L] = (-684.8 / 6.3 * p13) * cf;
    c[SU_P7] = ((770.96675 / 2.54016 + 378.515 / 1.512 * eta - 740.45 / 7.56 * e2) * pi +
                (-25150.083775 / 3.048192 + 26804.935 / 6.048 * eta - 198.5 / 4.8 * e2) * delta * ca +
                (-25150.083775 / 3.048192 + 105666.55595 / 7.62048 * eta - 1042.165 / 3.024 * e2 +
                 534.5 / 3.6 * e3) *
                    cs -
                1140.0 * pi * delta * ca * cs + (-570.0 + 2240.0 * eta) * pi * ca2 +
                (-570.0 + 40.0 * eta) * pi * cs2 + (14585.0 / 8.0 - 215.0 / 2.0 * eta) * delta * ca * cs2 +
                (14585.0 / 8.0 - 7270.0 * eta + 80.0 * e2) * ca2 * cs +
                (1458.5 / 2.4 - 2380.0 * eta) * delta * ca2 * ca +
                (1458.5 / 2.4 - 475.0 / 6.0 * eta + 100.0 / 3.0 * e2) * cs2 * cs) *
               p23 * cf;
  }
  // amplitude peak and intermediate collocation
  {
    S g2 = c[SU_G2], g3 = c[SU_G3];
    if (value_of(g2) <= 1.0)
      c[SU_FMAX] = c[SU_FRD] + c[SU_FDM] * (dsqrt(1.0 - g2 * g2) - 1.0) * g3 / g2;
    else
      c[SU_FMAX] = c[SU_FRD] - c[SU_FDM] * g3 / g2;
    if (value_of(c[SU_FMAX]) < 0.0) c[SU_FMAX] = -c[SU_FMAX];
    // the delta coefficients are set below; evaluate boundary values with f-duals
    typedef Dual<S, 1> SF;
    const double f1 = AMP_F_INS;
    S f3 = c[SU_FMAX];
    S f2 = 0.5 * (f1 + f3);
    // promote coefficients to SF is avoided: evaluate ansatz and derivative by hand
    Phenom<SF> Q;
    for (int i = 0; i < N_SETUP; ++i) {
      Q.c[i].v = c[i];
      Q.c[i].d[0] = lift<S>(0.0);
    }
    SF x1;
    x1.v = lift<S>(f1);
    x1.d[0] = lift<S>(1.0);
    SF ins = Q.amp_ins(x1);
    SF x3;
    x3.v = f3;
    x3.d[0] = lift<S>(1.0);
    SF mrd = Q.amp_mrd(x3);
    S A[5][5], b[5], x[5];
    S f1s = lift<S>(f1);
    S pw1[5], pw2[5], pw3[5];
    pw1[0] = lift<S>(1.0); pw2[0] = lift<S>(1.0); pw3[0] = lift<S>(1.0);
    for (int k = 1; k < 5; ++k) { pw1[k] = pw1[k - 1] * f1s; pw2[k] = pw2[k - 1] * f2; pw3[k] = pw3[k - 1] * f3; }
    for (int k = 0; k < 5; ++k) {
      A[0][k] = pw1[k];
      A[1][k] = k == 0 ? lift<S>(0.0) : (double)k * pw1[k - 1];
      A[2][k] = pw2[k];
      A[3][k] = pw3[k];
      A[4][k] = k == 0 ? lift<S>(0.0) : (double)k * pw3[k - 1];
    }
    b[0] = ins.v; b[1] = ins.d[0]; b[2] = v2; b[3] = mrd.v; b[4] = mrd.d[0];
    solve5(A, b, x);
    c[SU_D0] = x[0]; c[SU_D1] = x[1]; c[SU_D2] = x[2]; c[SU_D3] = x[3]; c[SU_D4] = x[4];
  }
  // phase C1 connection
  {
    typedef Dual<S, 1> SF;
    Phenom<SF> Q;
    c[SU_C1INT] = c[SU_C2INT] = c[SU_C1MRD] = c[SU_C2MRD] = lift<S>(0.0);
    c[SU_T0] = c[SU_PHIREF0] = c[SU_MFREF] = lift<S>(0.0);
    for (int i = 0; i < N_SETUP; ++i) {
      Q.c[i].v = c[i];
      Q.c[i].d[0] = lift<S>(0.0);
    }
    SF fa;
    fa.v = lift<S>(PHI_F_INS);
    fa.d[0] = lift<S>(1.0);
    SF pi_ = Q.phi_ins(fa);
    SF pn_ = Q.phi_int_raw(fa);
    S c2int = pi_.d[0] - pn_.d[0];
    S c1int = pi_.v - pn_.v - c2int * PHI_F_INS;
    SF fb;
    fb.v = 0.5 * c[SU_FRD];
    fb.d[0] = lift<S>(1.0);
    SF pn2 = Q.phi_int_raw(fb);
    SF pm2 = Q.phi_mrd_raw(fb);
    S pn2v = pn2.v + c1int + c2int * fb.v;
    S pn2d = pn2.d[0] + c2int;
    S c2mrd = pn2d - pm2.d[0];
    S c1mrd = pn2v - pm2.v - c2mrd * fb.v;
    c[SU_C1INT] = c1int; c[SU_C2INT] = c2int; c[SU_C1MRD] = c1mrd; c[SU_C2MRD] = c2mrd;
    Q.c[SU_C1INT].v = c1int; Q.c[SU_C2INT].v = c2int; Q.c[SU_C1MRD].v = c1mrd; Q.c[SU_C2MRD].v = c2mrd;
    SF fm;
    fm.v = c[SU_FMAX];
    fm.d[0] = lift<S>(1.0);
    c[SU_T0] = Q.phase(fm).d[0];
    S mfref = f_ref > 0.0 ? f_ref * c[SU_MSEC] : c[SU_FMAX];
    c[SU_MFREF] = mfref;
    SF fr;
    fr.v = mfref;
    fr.d[0] = lift<S>(0.0);
    c[SU_PHIREF0] = Q.phase(fr).v;
  }
}

// ---- spin-weighted harmonics (s = -2), theta part: -2Y_lm = y_lm(theta) e^{i m phi} ----
// Goldberg et al. 1967 general formula, written with cos/sin(theta/2) powers.
template <class S>
PES_HD S sw_y(int l, int m, const S& theta) {
  const int s = -2;
  auto fact = [](int n) {
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
  };
  auto comb = [&](int n, int k) { return (k < 0 || k > n) ? 0.0 : fact(n) / (fact(k) * fact(n - k)); };
  double pre = ((m % 2) ? -1.0 : 1.0) *
               sqrt(fact(l + m) * fact(l - m) * (2 * l + 1) / (4.0 * PI * fact(l + s) * fact(l - s)));
  S h = 0.5 * theta;
  S c = dcos(h);
  S sn = dsin(h);
  S tot = lift<S>(0.0);
  for (int r = 0; r <= l - s; ++r) {
    int k = r + s - m;
    if (k < 0 || k > l + s) continue;
    int p = 2 * r + s - m;
    S term = lift<S>(comb(l - s, r) * comb(l + s, k) * (((l - r - s) % 2) ? -1.0 : 1.0));
    for (int i = 0; i < p; ++i) term = term * c;
    for (int i = 0; i < 2 * l - p; ++i) term = term * sn;
    tot = tot + term;
  }
  return pre * tot;
}

}  // namespace pes

// MI355X (gfx950) kernels and C-ABI for the FD lnL(+grad) hot path.
//
// Pipeline per batch of P parameter rows (one launch each, all on one stream):
//   k_setup     one thread per row: PhenomD/HM coefficients (+ d/dtheta_int) and
//               the extrinsic constants (polarization basis, harmonic factors)
//   k_nodes     one thread per (row, mode, node): A, dA/du, Psi, dPsi/du on the
//               log-spaced interpolation nodes (u = ln f)
//   k_lnl_*     persistent fused kernel: lane <-> frequency bin, loop over the rows
//               of a chunk; cubic-Hermite node -> bin interpolation, t_c shift,
//               orbit, six single links, TDI, residual, FP64 reduction
//   k_reduce    fixed-order sum of the per-(bin group, row) partials
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pespace_amd.h"
#include "response.hpp"
#include "mathfn.hpp"

using namespace pes;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(PES_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#include "kernel_common.hpp"

__device__ inline bool finite_d(double x) { return isfinite(x); }

// unit complex factor of a mode: conj(phase of leading PN H_lm) * exp(-i (m-2) pi/8)
__device__ inline Cplx mode_factor(int id) {
  const double m = (double)kModeM[id];
  double ph = -(m - 2.0) * PI / 8.0;
  Cplx u;
  switch (id) {
    case 1: u = {0.0, -1.0}; break;   // (2,1): -i
    case 2: u = {0.0, 1.0}; break;    // (3,3): +i
    case 4: u = {-1.0, 0.0}; break;   // (4,4): -1
    case 5: u = {0.0, -1.0}; break;   // (4,3): -i
    default: u = {1.0, 0.0}; break;   // (2,2), (3,2)
  }
  double s, c;
  sincos(ph, &s, &c);
  return cmul(u, {c, s});
}

#ifndef PES_LNL_FWD_UNIT
// ---- k_setup ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_setup(const double* __restrict__ params, int P, ModeList modes, double f_ref,
                        double* __restrict__ setup, double* __restrict__ pconst, int* __restrict__ work_counter) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0)  // work queues of the fused kernel launches that follow on this stream
    for (int i = 0; i < N_WORK_COUNTERS; ++i) work_counter[i] = 0;
  if (p >= P) return;
  const double* th = params + (size_t)p * PES_N_PARAMS;
  double Mc = th[0], q = th[1], c1 = th[2], c2 = th[3], dL = th[4], phi = th[5], iota = th[6];
  double lam = th[7], bet = th[8], psi = th[9], tc = th[10];
  bool valid = true;
  for (int i = 0; i < PES_N_PARAMS; ++i) valid = valid && finite_d(th[i]);
  valid = valid && Mc > 0.0 && q > 0.0 && q <= 1.0 && fabs(c1) < 1.0 && fabs(c2) < 1.0 && dL > 0.0;
  double* pc = pconst + (size_t)p * PC_STRIDE;
  if (!valid) {  // harmless stand-in values; the row's lnL is set to -inf by k_reduce
    Mc = 1.0e6; q = 1.0; c1 = 0.0; c2 = 0.0; dL = 1.0e4; phi = 0.0; iota = 0.0; lam = 0.0; bet = 0.0; psi = 0.0;
    tc = 0.0;
  }
  typedef Dual<double, 4> D4;
  D4 dMc = lift<D4>(Mc), dq = lift<D4>(q), dc1 = lift<D4>(c1), dc2 = lift<D4>(c2);
  dMc.d[0] = 1.0; dq.d[1] = 1.0; dc1.d[2] = 1.0; dc2.d[3] = 1.0;
  Phenom<D4> ph;
  phenom_setup(ph, dMc, dq, dc1, dc2, f_ref);
  double* su = setup + (size_t)p * N_SETUP * 5;
  for (int i = 0; i < N_SETUP; ++i) {
    su[i * 5 + 0] = ph.c[i].v;
    for (int k = 0; k < 4; ++k) su[i * 5 + 1 + k] = ph.c[i].d[k];
  }
  pc[PC_TC] = tc;
  pc[PC_DL] = dL;
  pc[PC_INVDL] = 1.0 / dL;
  pc[PC_VALID] = valid ? 1.0 : 0.0;
  pc[PC_OFFANY] = 0.0;  // set by k_duty_off
  double sl, cl, sb, cb, sp, cp;
  sincos(lam, &sl, &cl);
  sincos(bet, &sb, &cb);
  sincos(psi, &sp, &cp);
  // antenna.py:70-83, 112-114 and their lambda / beta derivatives
  double pv[3] = {sl * cp - cl * sb * sp, -cl * cp - sl * sb * sp, cb * sp};
  double qv[3] = {-sl * sp - cl * sb * cp, cl * sp - sl * sb * cp, cb * cp};
  double kv[3] = {-cb * cl, -cb * sl, -sb};
  double dlp[3] = {cl * cp + sl * sb * sp, sl * cp - cl * sb * sp, 0.0};
  double dbp[3] = {-cl * cb * sp, -sl * cb * sp, -sb * sp};
  double dlq[3] = {-cl * sp + sl * sb * cp, -sl * sp - cl * sb * cp, 0.0};
  double dbq[3] = {-cl * cb * cp, -sl * cb * cp, -sb * cp};
  double dlk[3] = {cb * sl, -cb * cl, 0.0};
  double dbk[3] = {sb * cl, sb * sl, -cb};
  for (int i = 0; i < 3; ++i) {
    pc[PC_P + i] = pv[i]; pc[PC_Q + i] = qv[i]; pc[PC_K + i] = kv[i];
    pc[PC_DLP + i] = dlp[i]; pc[PC_DLQ + i] = dlq[i]; pc[PC_DBP + i] = dbp[i];
    pc[PC_DBQ + i] = dbq[i]; pc[PC_DLK + i] = dlk[i]; pc[PC_DBK + i] = dbk[i];
  }
  double dist = dL * 1.0e6 * PC_SI;
  double msec = ph.c[SU_MSEC].v;
  double az = 0.5 * PI - phi;  // lalsimulation azimuth convention
  typedef Dual<double, 1> D1;
  D1 di = lift<D1>(iota);
  di.d[0] = 1.0;
  double fcut = 0.0;
  for (int mi = 0; mi < modes.n; ++mi) {
    int id = modes.id[mi];
    int l = kModeL[id], m = kModeM[id];
    D1 y = sw_y(l, m, di);
    D1 ym = sw_y(l, -m, di);
    double sgn = (l % 2) ? -1.0 : 1.0;
    double se, ce;
    sincos(-(double)m * az, &se, &ce);
    Cplx e = {ce, se};
    Cplx u = mode_factor(id);
    Cplx eu = cmul(e, u);
    double kp = 0.5 * (y.v + sgn * ym.v), dkp = 0.5 * (y.d[0] + sgn * ym.d[0]);
    double kc = 0.5 * (sgn * ym.v - y.v), dkc = 0.5 * (sgn * ym.d[0] - y.d[0]);  // times i
    double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
    Cplx Fp = cscale(eu, kp / dist);
    // the cross factor is stored doubled: the link term's n.e_x.n = 2 (n.p)(n.q) keeps its 2 there
    Cplx Fc = cmul({0.0, 2.0 * kc / dist}, eu);
    Cplx dFp = cscale(eu, dkp / dist);
    Cplx dFc = cmul({0.0, 2.0 * dkc / dist}, eu);
    mf[MF_FPR] = Fp.re; mf[MF_FPI] = Fp.im; mf[MF_FCR] = Fc.re; mf[MF_FCI] = Fc.im;
    mf[MF_CUT] = ph.mode_cut(id) / msec;
    fcut = fmax(fcut, mf[MF_CUT]);
    mf[MF_DFPR] = dFp.re; mf[MF_DFPI] = dFp.im; mf[MF_DFCR] = dFc.re; mf[MF_DFCI] = dFc.im;
  }
  pc[PC_FCUT] = fcut;
}

// ---- k_nodes ---------------------------------------------------------------------------
// node table: [P][n_modes][NNS][4] = (A, dA/du, Psi, dPsi/du); NNS = n_nodes + 1 (padding)
// grad table: [P][n_modes][NNS][16] = d/dtheta_i of the same four, i = Mc, q, chi1, chi2
template <bool GRAD>
__global__ __launch_bounds__(128) void k_nodes(const double* __restrict__ setup, int P, ModeList modes, int n_nodes, double u0, double du,
                        const double* __restrict__ node_f, double* __restrict__ nodes,
                        double* __restrict__ nodes_grad) {
  int jn = blockIdx.x * blockDim.x + threadIdx.x;
  int pm = blockIdx.y;
  int p = pm / modes.n, mi = pm % modes.n;
  const int NNS = n_nodes + 1;
  if (jn >= NNS || p >= P) return;
  size_t row = ((size_t)p * modes.n + mi) * NNS + jn;
  if (jn == n_nodes) {  // padding node (read as the right end of the last direct-mode bin)
    for (int k = 0; k < 4; ++k) nodes[row * 4 + k] = 0.0;
    if (GRAD)
      for (int k = 0; k < 16; ++k) nodes_grad[row * 16 + k] = 0.0;
    return;
  }
  double f = node_f ? node_f[jn] : exp(u0 + du * (double)jn);
  const double* su = setup + (size_t)p * N_SETUP * 5;
  int id = modes.id[mi];
  if (!GRAD) {
    typedef Dual<double, 1> S;
    Phenom<S, SetupView<S>> ph{SetupView<S>{su}};
    S Mf;
    Mf.v = f * su[SU_MSEC * 5];
    Mf.d[0] = Mf.v;  // dMf/du
    S A, Ps;
    ph.amp_phase(id, Mf, A, Ps);
    nodes[row * 4 + 0] = A.v;
    nodes[row * 4 + 1] = A.d[0];
    nodes[row * 4 + 2] = Ps.v;
    nodes[row * 4 + 3] = Ps.d[0];
  } else {
    typedef Dual<double, 1> D1;
    typedef Dual<D1, 4> S;
    Phenom<S, SetupView<S>> ph{SetupView<S>{su}};
    S Mf;
    Mf.v.v = f * su[SU_MSEC * 5];
    Mf.v.d[0] = Mf.v.v;
    for (int k = 0; k < 4; ++k) {
      Mf.d[k].v = f * su[SU_MSEC * 5 + 1 + k];
      Mf.d[k].d[0] = Mf.d[k].v;
    }
    S A, Ps;
    ph.amp_phase(id, Mf, A, Ps);
    nodes[row * 4 + 0] = A.v.v;
    nodes[row * 4 + 1] = A.v.d[0];
    nodes[row * 4 + 2] = Ps.v.v;
    nodes[row * 4 + 3] = Ps.v.d[0];
    for (int k = 0; k < 4; ++k) {
      nodes_grad[row * 16 + k * 4 + 0] = A.d[k].v;
      nodes_grad[row * 16 + k * 4 + 1] = A.d[k].d[0];
      nodes_grad[row * 16 + k * 4 + 2] = Ps.d[k].v;
      nodes_grad[row * 16 + k * 4 + 3] = Ps.d[k].d[0];
    }
  }
}
#endif  // PES_LNL_FWD_UNIT

// ---- Hermite interpolation between node j and j+1 ---------------------------------------
struct Herm {
  double v, ds;  // value and d/ds
};
__device__ inline Herm hermite(double P0, double D0, double P1, double D1, double du, double s) {
  double M0 = D0 * du, M1 = D1 * du;
  double c2 = 3.0 * (P1 - P0) - 2.0 * M0 - M1;
  double c3 = 2.0 * (P0 - P1) + M0 + M1;
  Herm h;
  h.v = P0 + s * (M0 + s * (c2 + s * c3));
  h.ds = M0 + s * (2.0 * c2 + s * (3.0 * c3));
  return h;
}

struct BinCtx {
  double f, s, du;
  int j;
};

// Waveform of one mode at one bin: amplitude, phase and tf (t_c = 0).
__device__ inline void mode_at_bin(const double* __restrict__ nodes, size_t base, const BinCtx& b, double& A,
                                   double& Psi, double& tf) {
  const double* n0 = nodes + (base + b.j) * 4;
  Herm ha = hermite(n0[0], n0[1], n0[4], n0[5], b.du, b.s);
  Herm hp = hermite(n0[2], n0[3], n0[6], n0[7], b.du, b.s);
  A = ha.v;
  Psi = hp.v;
  tf = hp.ds / (2.0 * PI * b.du * b.f);
}


struct BinData {
  double f, s, w[3];
  Cplx d[3];
  int j;
};

#include "hotpath.hpp"

#ifndef PES_LNL_FWD_UNIT
// One block per (64 rows, output): each of the 4 waves sums every 4th bin group in order, then
// the four wave sums are added in a fixed order, so results do not depend on scheduling.
constexpr int REDUCE_WAVES = 4;  // the final sum below adds exactly four wave sums
__global__ __launch_bounds__(64 * REDUCE_WAVES) void k_reduce(const double* __restrict__ partial, int n_groups, int P,
                                                           int nout, const double* __restrict__ pconst,
                                                           double* __restrict__ lnl, double* __restrict__ grad) {
  __shared__ double part[REDUCE_WAVES][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  const int o = blockIdx.y;
  double s = 0.0;
  if (p < P) {
#pragma unroll 8
    for (int g = w; g < n_groups; g += REDUCE_WAVES) s += partial[((size_t)g * nout + o) * P + p];
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && p < P) {
    const double t = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    const bool valid = pconst[(size_t)p * PC_STRIDE + PC_VALID] != 0.0;
    if (o == 0)
      lnl[p] = valid ? t : -INFINITY;
    else if (grad)
      grad[(size_t)p * PES_N_PARAMS + (o - 1)] = valid ? t : 0.0;
  }
}

// ---- work accounting for the profiler: per row, the number of (bin, mode) pairs below the
// mode cuts and of bins below the largest cut (the bins on which the fused kernel has work)
__device__ inline int count_le(const double* __restrict__ f, int n, double x) {
  int lo = 0, hi = n;  // first index with f > x (f ascending)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (f[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
__global__ __launch_bounds__(64) void k_live(const DetectorDev* __restrict__ dets, int n_det,
                                             const double* __restrict__ bin_f, const double* __restrict__ pconst,
                                             ModeList modes, int P, double* __restrict__ live) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double* pc = pconst + (size_t)p * PC_STRIDE;
  double bm = 0.0, br = 0.0;
  if (pc[PC_VALID] != 0.0) {
    for (int d = 0; d < n_det; ++d) {
      const double* f = bin_f + dets[d].bin_offset;
      const int n = dets[d].n_bins;
      for (int mi = 0; mi < modes.n; ++mi) bm += (double)count_le(f, n, pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]);
      br += (double)count_le(f, n, pc[PC_FCUT]);
    }
  }
  live[2 * (size_t)p] = bm;
  live[2 * (size_t)p + 1] = br;
}

// profiler only: per row, the (bin, mode) of one duty-cycled detector, below the mode's cut,
// whose observation window is zero at t = t_f + t_c (they need no work; the fused kernel skips
// the tiles where that holds for every mode, k_duty_off), subtracted from k_live's count.  One lane per bin, rows over the
// grid's y dimension; the counts are whole numbers, so the atomic sums are exact.
__global__ __launch_bounds__(256) void k_live_off(const DetectorDev* __restrict__ dets, int det_i,
                                                  const double* __restrict__ bin_f, const int* __restrict__ bin_j,
                                                  const double* __restrict__ bin_s, const double* __restrict__ pconst,
                                                  const double* __restrict__ nodes, int NNS, double du, ModeList modes,
                                                  int P, double* __restrict__ live) {
  const DetectorDev& det = dets[det_i];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    const double* pc = pconst + (size_t)p * PC_STRIDE;
    double off = 0.0;
    if (k < det.n_bins && pc[PC_VALID] != 0.0) {
      const int bi = det.bin_offset + k;
      BinPre b;
      bin_prepare(det, bin_f[bi], bin_s[bi], bin_j[bi], du, b);
      for (int mi = 0; mi < modes.n; ++mi) {
        if (b.f > pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]) continue;
        const double* n0 = nodes + (((size_t)p * modes.n + mi) * NNS + b.j) * 4;
        double w, dw;
        duty_window(det, mode_time(n0, b, pc[PC_TC]), w, dw);
        if (w == 0.0) off += 1.0;
      }
    }
    off = wave_sum(off);  // every lane of the block reaches this
    if ((threadIdx.x & 63) == 0 && off != 0.0) atomicAdd(&live[2 * (size_t)p], -off);
  }
}

// Per (tile, row): bit 1 when the tile belongs to a duty-cycled detector and every mode's t over
// the tile's node intervals lies inside an off period of its window (modes whose cut lies below
// the tile's first frequency aside), so that the response is zero over the whole tile: the fused
// kernel then treats it as it treats a tile above the row's cuts (lnL adds -sum w |d|^2 of the
// tile, the gradient nothing).  One wave per 64 tiles writes the row's word with a vote (no
// clearing needed) and marks a row with any such tile (PC_OFFANY, cleared by k_setup), so that
// the kernel reads the flags of those rows only; rows over the grid's y dimension.  enable = 0
// (direct evaluation at every bin, no node intervals) writes zeros.
__global__ __launch_bounds__(64) void k_duty_off(const DetectorDev* __restrict__ dets, int n_det, int n_tiles,
                                                 const double* __restrict__ tile_f0, const int* __restrict__ tile_jr,
                                                 double* __restrict__ pconst, const double* __restrict__ nodes,
                                                 int NNS, double u0, double du, ModeList modes, int P, int words,
                                                 int enable, unsigned long long* __restrict__ off_bits) {
  const int tile = blockIdx.x * 64 + threadIdx.x;
  for (int p = blockIdx.y; p < P; p += gridDim.y) {
    const double* pc = pconst + (size_t)p * PC_STRIDE;
    bool off = false;
    if (enable && tile < n_tiles && pc[PC_VALID] != 0.0) {
      const DetectorDev& det = dets[tile_detector(dets, n_det, tile)];
      if (det.duty) {
        off = true;
        const double f0 = tile_f0[tile], tc = pc[PC_TC];
        const int j0 = tile_jr[2 * tile], j1 = j0 + tile_jr[2 * tile + 1];
        for (int mi = 0; mi < modes.n && off; ++mi) {
          if (f0 > pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]) continue;
          const double* nm = nodes + ((size_t)p * modes.n + mi) * NNS * 4;
          for (int j = j0; j < j1 && off; ++j)
            off = interval_off(det, nm + (size_t)j * 4, exp(u0 + du * j), exp(u0 + du * (j + 1)), du, tc);
        }
      }
    }
    const unsigned long long m = __ballot(off);
    if (threadIdx.x == 0) {
      off_bits[(size_t)p * words + blockIdx.x] = m;
      if (m != 0ull) pconst[(size_t)p * PC_STRIDE + PC_OFFANY] = 1.0;  // every writer writes 1
    }
  }
}

// Per work item (tile group x row chunk) of a duty-cycled detectors' tile range: 1 when no
// (tile, row) of it needs the window: over every mode's node intervals of the tile the window is
// flat 1 (interval_on; modes whose cut lies below the tile's first frequency aside), or the row
// evaluates nothing there (the tile above its cuts, an invalid row).  Such items run in the plain
// instantiation of the fused kernel, whose response and gradient equal the duty one's there; the
// others in the duty instantiation.  The item index is that of the fused kernel (global group
// g = group_offset + gl, chunk c: g n_chunks + c).  One block of 64 per item; enable = 0 (no node
// intervals: direct evaluation at every bin) sends every item to the duty instantiation.
__global__ __launch_bounds__(64) void k_duty_items(const DetectorDev* __restrict__ dets, int n_det,
                                                   const double* __restrict__ tile_f0, const int* __restrict__ tile_jr,
                                                   const double* __restrict__ pconst, const double* __restrict__ nodes,
                                                   int NNS, double u0, double du, ModeList modes, int P, int rows,
                                                   int n_chunks, int tile_begin, int tile_end, int tiles_per_group,
                                                   int group_offset, int enable, unsigned char* __restrict__ items,
                                                   int* __restrict__ n_duty_items) {
  const int gl = blockIdx.x / n_chunks, c = blockIdx.x % n_chunks;
  const int t_first = tile_begin + gl * tiles_per_group, t_end = min(tile_end, t_first + tiles_per_group);
  const int p0 = c * rows, p1 = min(P, p0 + rows);
  const int nt = t_end - t_first, np = p1 - p0;
  int ok = enable;
  for (int k = threadIdx.x; ok && k < nt * np; k += blockDim.x) {
    const int tile = t_first + k % nt, p = p0 + k / nt;
    const double* pc = pconst + (size_t)p * PC_STRIDE;
    const double f0 = tile_f0[tile];
    if (pc[PC_VALID] == 0.0 || f0 > pc[PC_FCUT]) continue;
    const DetectorDev& det = dets[tile_detector(dets, n_det, tile)];
    const int j0 = tile_jr[2 * tile], j1 = j0 + tile_jr[2 * tile + 1];
    const double tc = pc[PC_TC];
    for (int mi = 0; mi < modes.n && ok; ++mi) {
      if (f0 > pc[PC_MODE + mi * PC_MODE_STRIDE + MF_CUT]) continue;
      const double* nm = nodes + ((size_t)p * modes.n + mi) * NNS * 4;
      for (int j = j0; j < j1 && ok; ++j)
        ok = det.duty && interval_on(det, nm + (size_t)j * 4, exp(u0 + du * j), exp(u0 + du * (j + 1)), du, tc);
    }
  }
  ok = __syncthreads_and(ok);
  if (threadIdx.x == 0) {
    items[(size_t)(group_offset + gl) * n_chunks + c] = ok ? 1 : 0;
    if (!ok) atomicAdd(n_duty_items, 1);  // the range's duty pass ends at once when none need it
  }
}

// ---- response / waveform output kernels (parity checks) -------------------------------------
__global__ __launch_bounds__(128) void k_waveform(const DetectorDev* __restrict__ dets, int det_i, const double* __restrict__ bin_f,
                           const int* __restrict__ bin_j, const double* __restrict__ bin_s,
                           const double* __restrict__ pconst, const double* __restrict__ nodes, int NNS, double du,
                           ModeList modes, int P, double* __restrict__ out) {
  const DetectorDev det = dets[det_i];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int p = blockIdx.y;
  if (k >= det.n_bins || p >= P) return;
  int bi = det.bin_offset + k;
  BinCtx b;
  b.f = bin_f[bi];
  b.s = bin_s[bi];
  b.j = bin_j[bi];
  b.du = du;
  const double* pc = pconst + (size_t)p * PC_STRIDE;
  for (int mi = 0; mi < modes.n; ++mi) {
    const double* mf = pc + PC_MODE + mi * PC_MODE_STRIDE;
    double A, Psi, tf;
    mode_at_bin(nodes, (size_t)p * modes.n * NNS + (size_t)mi * NNS, b, A, Psi, tf);
    if (b.f > mf[MF_CUT]) A = 0.0;
    double sp, cp;
    sincos(Psi, &sp, &cp);
    Cplx h = {A * cp, -A * sp};
    Cplx hp = cmul({mf[MF_FPR], mf[MF_FPI]}, h);
    Cplx hc = cmul({0.5 * mf[MF_FCR], 0.5 * mf[MF_FCI]}, h);  // stored doubled
    size_t idx = (((size_t)p * modes.n + mi) * det.n_bins + k) * 5;
    out[idx] = hp.re;
    out[idx + 1] = hp.im;
    out[idx + 2] = hc.re;
    out[idx + 3] = hc.im;
    out[idx + 4] = tf;
  }
}


// ---- response from a caller-supplied waveform container ------------------------------------
// wf: [n_modes][n_bins][5] = plus (re,im), cross (re,im), tf with t_c = 0 (the reference's
// container contract, interface_ripple.py:42-51); antenna.py:429-460 + tdi.py:1764-1793.
struct Extrinsic {
  Vec3 p, q, k;
  double tc;
};
__global__ __launch_bounds__(128) void k_response_container(const DetectorDev* __restrict__ dets, int det_i, const double* __restrict__ bin_f,
                                     int n_modes, const double* __restrict__ wf, Extrinsic ex,
                                     double* __restrict__ out) {
  const DetectorDev det = dets[det_i];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= det.n_bins) return;
  double f = bin_f[det.bin_offset + k];
  Cplx etc;
  sincos(-2.0 * PI * f * ex.tc, &etc.im, &etc.re);
  Links L = {};
  for (int m = 0; m < n_modes; ++m) {
    const double* w = wf + ((size_t)m * det.n_bins + k) * 5;
    Cplx hp = cmul({w[0], w[1]}, etc);
    Cplx hc = cmul({w[2], w[3]}, etc);
    add_mode_links(det, f, w[4] + ex.tc, hp, hc, ex.p, ex.q, ex.k, L);
  }
  Cplx o[3];
  tdi_combine(det, f, L, o);
  for (int ch = 0; ch < 3; ++ch) {
    out[((size_t)ch * det.n_bins + k) * 2] = o[ch].re;
    out[((size_t)ch * det.n_bins + k) * 2 + 1] = o[ch].im;
  }
}

// ---- the six single links of a caller-supplied container (antenna.py:403-460), their TDI
// combination (tdi.py:1705-1793, the JAX twin _tdi_jax.py:70-102), and the constellation
// vectors (orbit.py:260-322, 363-419) ----------------------------------------------------------
// links: [6][n_bins][2] in the order 12, 21, 23, 32, 31, 13 (antenna.py:24-31)
__global__ __launch_bounds__(128) void k_links_container(const DetectorDev* __restrict__ dets, int det_i,
                                                         const double* __restrict__ bin_f, int n_modes,
                                                         const double* __restrict__ wf, Extrinsic ex,
                                                         double* __restrict__ out) {
  const DetectorDev det = dets[det_i];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= det.n_bins) return;
  double f = bin_f[det.bin_offset + k];
  Cplx etc;
  sincos(-2.0 * PI * f * ex.tc, &etc.im, &etc.re);
  Links L = {};
  for (int m = 0; m < n_modes; ++m) {
    const double* w = wf + ((size_t)m * det.n_bins + k) * 5;
    Cplx hp = cmul({w[0], w[1]}, etc);
    Cplx hc = cmul({w[2], w[3]}, etc);
    add_mode_links(det, f, w[4] + ex.tc, hp, hc, ex.p, ex.q, ex.k, L);
  }
  const Cplx l[6] = {L.l12, L.l21, L.l23, L.l32, L.l31, L.l13};
  for (int i = 0; i < 6; ++i) {
    out[((size_t)i * det.n_bins + k) * 2] = l[i].re;
    out[((size_t)i * det.n_bins + k) * 2 + 1] = l[i].im;
  }
}

// links [P][6][n_bins][2] -> channels [P][3][n_bins][2]
__global__ __launch_bounds__(128) void k_tdi_from_links(const DetectorDev* __restrict__ dets, int det_i,
                                                        const double* __restrict__ bin_f, int P,
                                                        const double* __restrict__ links, double* __restrict__ out) {
  const DetectorDev det = dets[det_i];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= det.n_bins || p >= P) return;
  const double f = bin_f[det.bin_offset + k];
  const double* lp = links + (size_t)p * 6 * det.n_bins * 2;
  auto get = [&](int i) { return Cplx{lp[((size_t)i * det.n_bins + k) * 2], lp[((size_t)i * det.n_bins + k) * 2 + 1]}; };
  Links L = {get(0), get(1), get(2), get(3), get(4), get(5)};
  Cplx o[3];
  tdi_combine(det, f, L, o);
  for (int ch = 0; ch < 3; ++ch) {
    const size_t idx = (((size_t)p * 3 + ch) * det.n_bins + k) * 2;
    out[idx] = o[ch].re;
    out[idx + 1] = o[ch].im;
  }
}

// constellation vectors at n times: out [n][6][3] = n1, n2, n3 (unit link vectors) and x1, x2, x3
// (node positions in the SSB frame, seconds)
__global__ __launch_bounds__(128) void k_orbit_vectors(DetectorDev det, const double* __restrict__ t, int n,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Vec3 nv[3], xv[3];
  constellation(det, t[i], nv, xv);
  double* o = out + (size_t)i * 18;
  for (int a = 0; a < 3; ++a) {
    o[3 * a] = nv[a].x;
    o[3 * a + 1] = nv[a].y;
    o[3 * a + 2] = nv[a].z;
    o[9 + 3 * a] = xv[a].x;
    o[9 + 3 * a + 1] = xv[a].y;
    o[9 + 3 * a + 2] = xv[a].z;
  }
}

// ---- time domain: single links of a sampled waveform and the Michelson TDI of sampled links ----
// (the reference's TDResponseModelCornish2003, antenna.py:566-800, and TDMichelsonConstantEqualArm,
// tdi.py:1388-1600, with its linear interpolation kernel, utils.py:83-100)
struct TdWave {  // h+, hx sampled at t0 + m dt, m = 0 .. n - 1; zero outside
  const double* hp;
  const double* hc;
  double t0, inv_dt;
  int n;
};
// h+, hx at the fractional sample x = (t - t0) / dt: linear between samples, zero outside (and
// for NaN).  In two steps, so that a thread forms all its positions before it loads, and its
// twelve loads are all issued before the first is waited for: td_pos clamps the position into
// the samples (the loads need no branch) and notes whether it was inside; td_lerp interpolates
// the loaded pair, or gives zero outside.  x = n - 1 exactly reads the last pair at fraction 1.
struct TdPos {
  int j;
  double fr;
  bool inside;
};
__device__ inline TdPos td_pos(const TdWave& w, double x) {
  const double xc = fmin(fmax(x, 0.0), (double)(w.n - 1));  // NaN -> 0
  const int j = min((int)xc, w.n - 2);                      // xc >= 0: truncation is floor
  return {j, xc - (double)j, x >= 0.0 && x <= (double)(w.n - 1)};
}
__device__ inline double td_lerp(const TdPos& q, double v0, double v1) {
  return q.inside ? v0 + (v1 - v0) * q.fr : 0.0;
}
// A position x in a block's LDS window of the waveform that starts at sample xlo (k_links_td,
// STAGED): the window index of the sample below x and the fraction (both exact: x - xlo is); a
// position outside the waveform's samples [0, xmax] is given the index zero_slot of two zero
// entries past the window, so that its interpolation reads zero without a select per value
// (its fraction is finite: zero times it is zero)
__device__ inline TdPos td_win_pos(double x, double xlo, double xmax, int wlen, int zero_slot) {
  const double xw = x - xlo;
  const int j = min(max((int)xw, 0), wlen - 2);
  const bool inside = x >= 0.0 && x <= xmax;
  return {inside ? j : zero_slot, xw - (double)j, true};
}

// y_rs(t) = n.[H(t - L - k.x_s) - H(t - k.x_r)].n / (2 (1 - k.n_sr)) at t_i = t0 + i dt, with the
// constellation at t_i; out [6][n] in the order 12, 21, 23, 32, 31, 13 (link_rs: sent from s).
// One thread per sample.  The geometry takes the closed forms of the fused kernel's mode_geom
// (node offsets xd_i = al_i U + be_i V; for link vector n_a = (xd_a+2 - xd_a+1) / L every
// projection is n_a.v = (sqrt3 / L)(al_a v.V - be_a v.U)): one sincos of the revolution angle
// (and, KIND 1, of node 1's rotation angle; the table sincos, its entry read from constant
// memory), six scalar products, no library sincos and no division (the three denominators by
// rcp_d).  The waveform positions are formed in sample units
// (x = (t - t_w0) / dt - (L + k.x_a) / dt).  Neighbouring threads read neighbouring samples.
//
// STAGED: the block first copies the stretch of h+, hx its samples can reach into LDS (all loads
// issued before the first wait), and the twelve interpolation reads per sample come from LDS
// (the gathers through L1, 192 B per sample, cost more than the kernel's stores).  A node lies
// within |x_a| <= AU + 2 L of the barycentre, so every position of a block whose samples have
// base positions in [b_lo, b_hi] lies in [b_lo - R - L/dt_w, b_hi + R], R = (AU + 2 L) / dt_w
// (td_links_window); the host stages windows of at most 256 KW samples (KW = 2 or 6: 5 s
// sampling needs ~470, 1 s ~1300), wider ones read global memory (KW = 0).
constexpr int TD_LINKS_WMAX = 1536;
__host__ __device__ inline void td_links_window(const DetectorDev& det, const TdWave& w, double t0, double dt,
                                                int i0, int& lo, int& len) {
  const double b_lo = (t0 + dt * (double)i0 - w.t0) * w.inv_dt;
  const double b_hi = (t0 + dt * (double)(i0 + 255) - w.t0) * w.inv_dt;
  const double R = (AU_SEC + 2.0 * det.L) * w.inv_dt, Ls = det.L * w.inv_dt;
  lo = (int)floor(b_lo - R - Ls) - 2;
  len = (int)floor(b_hi + R) + 4 - lo;
}

// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b mod 8), each with its own L2:
// xcd_block gives XCD x a contiguous run of the logical blocks instead (a bijection of
// [0, nb)), so that neighbouring blocks, whose waveform windows overlap, share one L2.
__device__ inline int xcd_block(int b, int nb) {
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <int KIND, int KW>
__global__ __launch_bounds__(256) void k_links_td(DetectorDev det, TdWave w, Extrinsic ex, double t0, double dt, int n,
                                                  double* __restrict__ out) {
  const int lb = KW > 0 ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int i = lb * blockDim.x + threadIdx.x;
  constexpr bool STAGED = KW > 0;
  constexpr int WMAX = 256 * KW;
  // the window, then two zero entries that positions outside the waveform's samples read
  __shared__ double win_p[STAGED ? WMAX + 2 : 1], win_c[STAGED ? WMAX + 2 : 1];
  // STAGED: the window loads are issued first and land while the geometry is computed
  double vp[STAGED ? KW : 1], vc[STAGED ? KW : 1];
  int wlo = 0, wlen = 0;
  if constexpr (STAGED) {
    td_links_window(det, w, t0, dt, lb * blockDim.x, wlo, wlen);
    wlen = min(max(wlen, 2), WMAX);  // the host checked the length; LDS reads stay inside
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int q = threadIdx.x + 256 * k, s = wlo + q;
      const bool ok = q < wlen && s >= 0 && s < w.n;
      vp[k] = ok ? w.hp[s] : 0.0;
      vc[k] = ok ? w.hc[s] : 0.0;
    }
  } else if (i >= n) {
    return;
  }
  // (a thread past the last sample computes a geometry it never stores: it takes part in the
  // block's barrier)
  const double t = t0 + dt * (double)i;
  double sa, ca;
  sincos_tab_gmem(fma(kOmegaYear, t, det.rev0), &sa, &ca);
  const double kx0 = AU_SEC * fma(ex.k.x, ca, ex.k.y * sa);  // k.x0, x0 = AU (cos, sin, 0)
  double np[3], nq[3], kn[3], kx[3];                         // n_a.p, n_a.q, n_a.k; k.x_i
  if constexpr (KIND == 0) {
    const double sc = sa * ca;
    const V3 U = {sc, -fma(ca, ca, 1.0), -SQRT3 * sa};
    const V3 V = {-fma(sa, sa, 1.0), sc, -SQRT3 * ca};
    const V3 p = {ex.p.x, ex.p.y, ex.p.z}, q = {ex.q.x, ex.q.y, ex.q.z}, k = {ex.k.x, ex.k.y, ex.k.z};
    const double pU = dot3(p, U), pV = dot3(p, V), qU = dot3(q, U), qV = dot3(q, V);
    const double kU = dot3(k, U), kV = dot3(k, V);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      np[a] = fma(det.hsa[a], pV, -det.hca[a] * pU);
      nq[a] = fma(det.hsa[a], qV, -det.hca[a] * qU);
      kn[a] = fma(det.hsa[a], kV, -det.hca[a] * kU);
      kx[a] = kx0 + fma(det.hsk[a], kU, det.hck[a] * kV);
    }
  } else {
    double s1, c1;
    sincos_tab_gmem(fma(det.omega_rot, t, det.rot0), &s1, &c1);
    const double cr = -0.5, sr = 0.5 * SQRT3;  // rotation by 2 pi / 3
    double al[3], be[3];
    al[0] = s1;
    be[0] = c1;
    al[1] = fma(s1, cr, c1 * sr);
    be[1] = fma(c1, cr, -s1 * sr);
    al[2] = fma(al[1], cr, be[1] * sr);
    be[2] = fma(be[1], cr, -al[1] * sr);
    const double pU = fma(ex.p.x, det.sgxs, fma(ex.p.y, det.sgys, ex.p.z * det.sgzs));
    const double pV = fma(ex.p.x, det.sgxc, ex.p.y * det.sgyc);
    const double qU = fma(ex.q.x, det.sgxs, fma(ex.q.y, det.sgys, ex.q.z * det.sgzs));
    const double qV = fma(ex.q.x, det.sgxc, ex.q.y * det.sgyc);
    const double kU = fma(ex.k.x, det.gxs, fma(ex.k.y, det.gys, ex.k.z * det.gzs));
    const double kV = fma(ex.k.x, det.gxc, ex.k.y * det.gyc);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      np[a] = fma(al[a], pV, -be[a] * pU);
      nq[a] = fma(al[a], qV, -be[a] * qU);
      kn[a] = det.c3L * fma(al[a], kV, -be[a] * kU);
      kx[a] = kx0 + fma(al[a], kU, be[a] * kV);
    }
  }
  const double base = (t - w.t0) * w.inv_dt, Ls = det.L * w.inv_dt;
  TdPos qs[3], qr[3];  // waveform positions sent (t - L - k.x_a) and received (t - k.x_a)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double xr = fma(-kx[a], w.inv_dt, base);
    if constexpr (STAGED) {
      // every position of the block lies in its window (the host checked the window's length)
      const double xlo = (double)wlo, xmax = (double)(w.n - 1);
      qs[a] = td_win_pos(xr - Ls, xlo, xmax, wlen, WMAX);
      qr[a] = td_win_pos(xr, xlo, xmax, wlen, WMAX);
    } else {
      qs[a] = td_pos(w, xr - Ls);
      qr[a] = td_pos(w, xr);
    }
  }
  if constexpr (STAGED) {
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int q = threadIdx.x + 256 * k;
      if (q < wlen) {
        win_p[q] = vp[k];
        win_c[q] = vc[k];
      }
    }
    if (threadIdx.x < 2) {
      win_p[WMAX + threadIdx.x] = 0.0;
      win_c[WMAX + threadIdx.x] = 0.0;
    }
    __syncthreads();
    if (i >= n) return;
  }
  double hs[3][4], hr[3][4];  // h+ at j, j + 1 and hx at j, j + 1
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if constexpr (STAGED) {
      const int js = qs[a].j, jr = qr[a].j;
      hs[a][0] = win_p[js];
      hs[a][1] = win_p[js + 1];
      hs[a][2] = win_c[js];
      hs[a][3] = win_c[js + 1];
      hr[a][0] = win_p[jr];
      hr[a][1] = win_p[jr + 1];
      hr[a][2] = win_c[jr];
      hr[a][3] = win_c[jr + 1];
    } else {
      hs[a][0] = w.hp[qs[a].j];
      hs[a][1] = w.hp[qs[a].j + 1];
      hs[a][2] = w.hc[qs[a].j];
      hs[a][3] = w.hc[qs[a].j + 1];
      hr[a][0] = w.hp[qr[a].j];
      hr[a][1] = w.hp[qr[a].j + 1];
      hr[a][2] = w.hc[qr[a].j];
      hr[a][3] = w.hc[qr[a].j + 1];
    }
  }
  double Pp[3], Pc[3], sp[3], sc[3], rp[3], rc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Pp[a] = fma(np[a], np[a], -nq[a] * nq[a]);  // n.e+.n
    Pc[a] = 2.0 * np[a] * nq[a];                // n.ex.n
    if constexpr (STAGED) {  // outside positions read the zero entries
      sp[a] = fma(hs[a][1] - hs[a][0], qs[a].fr, hs[a][0]);
      sc[a] = fma(hs[a][3] - hs[a][2], qs[a].fr, hs[a][2]);
      rp[a] = fma(hr[a][1] - hr[a][0], qr[a].fr, hr[a][0]);
      rc[a] = fma(hr[a][3] - hr[a][2], qr[a].fr, hr[a][2]);
    } else {
      sp[a] = td_lerp(qs[a], hs[a][0], hs[a][1]);
      sc[a] = td_lerp(qs[a], hs[a][2], hs[a][3]);
      rp[a] = td_lerp(qr[a], hr[a][0], hr[a][1]);
      rc[a] = td_lerp(qr[a], hr[a][2], hr[a][3]);
    }
  }
  // link (r, s) lies along n_c with c the third node; 1 - k.n_sr = 1 + k.n_c when s -> r runs
  // against n_c (links 12, 23, 31) and 1 - k.n_c along it (21, 32, 13)
  double up[3], dn[3];  // 0.5 / (1 + k.n_c), 0.5 / (1 - k.n_c)
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double h = 0.5 * rcp_d(fma(-kn[c], kn[c], 1.0));
    up[c] = (1.0 - kn[c]) * h;
    dn[c] = (1.0 + kn[c]) * h;
  }
  auto link = [&](int r, int s, int c, double f) {
    return f * fma(Pp[c], sp[s] - rp[r], Pc[c] * (sc[s] - rc[r]));
  };
  const size_t N = (size_t)n;
  __builtin_nontemporal_store(link(0, 1, 2, up[2]), out + i);
  __builtin_nontemporal_store(link(1, 0, 2, dn[2]), out + N + i);
  __builtin_nontemporal_store(link(1, 2, 0, up[0]), out + 2 * N + i);
  __builtin_nontemporal_store(link(2, 1, 0, dn[0]), out + 3 * N + i);
  __builtin_nontemporal_store(link(2, 0, 1, up[1]), out + 4 * N + i);
  __builtin_nontemporal_store(link(0, 2, 1, dn[1]), out + 5 * N + i);
}

// Integer and fractional parts of the delays m L / dt on the extended link grid: D^m y at output
// sample i is y[i + jb[m]] + (y[i + jb[m] + 1] - y[i + jb[m]]) fr[m], with jb[m] + fr[m] =
// added - m L / dt (jb[0] = added, fr[0] = 0)
struct TdDelays {
  int jb[8];
  double fr[8];
};

// TDI of links sampled on the extended grid (n_ext = n_out + added samples): output sample i at
// link sample i + added; X = P(D) [y13 - y12 + D (y31 - y21)], Y, Z cyclic, P = 1 - D^2 (1.5,
// MAXD = 3) or (1 - D^2)(1 - D^4) (2.0, MAXD = 7); out [3][n_out] (A, E, T or X, Y, Z).  Every
// link is needed at every delay 0..MAXD (as y_a / y_b of one channel at even m, as y_c / y_e of
// another at odd m); each delayed value is added into the three channel sums as it is read.
constexpr int TD_CHAN[3][4] = {{5, 0, 4, 1}, {1, 2, 0, 3}, {3, 4, 2, 5}};  // (a, b, c, e) per X, Y, Z
__host__ __device__ constexpr double td_coef(int c, int l, int m) {
  // coefficient of D^m y_l in channel c: p_k (y_a - y_b) at m = k, p_k (y_c - y_e) at m = k + 1,
  // p = 1, -1, -1, 1 at k = 0, 2, 4, 6
  const int k = m & ~1;
  const double p = (k == 0 || k == 6) ? 1.0 : -1.0;
  const int* ch = TD_CHAN[c];
  return (m % 2 == 0) ? p * ((ch[0] == l) - (ch[1] == l)) : p * ((ch[2] == l) - (ch[3] == l));
}

template <bool ORTH>
__device__ inline void td_store(double* __restrict__ out, size_t N, int i, double X, double Y, double Z) {
  if (ORTH) {  // products with 1/sqrt2, 1/sqrt6, 1/sqrt3 rounded (no IEEE division)
    constexpr double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637, r3 = 0.57735026918962576451;
    __builtin_nontemporal_store((Z - X) * r2, out + i);
    __builtin_nontemporal_store((X - 2.0 * Y + Z) * r6, out + N + i);
    __builtin_nontemporal_store((X + Y + Z) * r3, out + 2 * N + i);
  } else {
    __builtin_nontemporal_store(X, out + i);
    __builtin_nontemporal_store(Y, out + N + i);
    __builtin_nontemporal_store(Z, out + 2 * N + i);
  }
}

// The same with the links staged in LDS: a block of TD_BLOCK x TD_OPT outputs reads the window
// of link samples [i0 + jb_min, i0 + TD_BLOCK TD_OPT - 1 + added] of each link once, coalesced,
// and takes every delayed value from LDS (the global path reads each sample 2 MAXD + 1 times
// through L1); each thread computes TD_OPT outputs TD_BLOCK apart, so that more window loads are
// in flight per thread before the barrier.  Used when the window fits (added - jb_min <=
// TD_WIN_EXTRA: delays up to ~256 samples).
#ifndef TD_OPT
#define TD_OPT 2
#endif
constexpr int TD_BLOCK = 256, TD_WIN_EXTRA = 256;
template <int MAXD, bool ORTH>
__global__ __launch_bounds__(TD_BLOCK) void k_tdi_td_lds(const double* __restrict__ links, int n_ext, int n_out,
                                                         TdDelays dl, double* __restrict__ out) {
  extern __shared__ double win[];  // [6][W]
  const int jb_min = dl.jb[MAXD];
  const int W = TD_BLOCK * TD_OPT + dl.jb[0] - jb_min + 1;
  const int i0 = blockIdx.x * (TD_BLOCK * TD_OPT), tid = threadIdx.x;
  const int g0 = i0 + jb_min;  // first link sample of the window
  // this thread's window samples (at most KW per link): every load is issued before the first is
  // waited for (a load-then-store loop waits for each load in turn)
  constexpr int KW = (TD_BLOCK * TD_OPT + TD_WIN_EXTRA + 1 + TD_BLOCK - 1) / TD_BLOCK;
  double v[6][KW];
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    const double* y = links + (size_t)l * n_ext;
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int w = tid + k * TD_BLOCK;
      v[l][k] = (w < W && g0 + w < n_ext) ? y[g0 + w] : 0.0;
    }
  }
#pragma unroll
  for (int l = 0; l < 6; ++l) {
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      const int w = tid + k * TD_BLOCK;
      if (w < W) win[l * W + w] = v[l][k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int o = 0; o < TD_OPT; ++o) {
    const int t = o * TD_BLOCK + tid;  // this output's place in the block
    const int i = i0 + t;
    if (i >= n_out) return;
    double X = 0.0, Y = 0.0, Z = 0.0;
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const double* y = win + l * W + t - jb_min;  // y[j] = link sample i + j
#pragma unroll
      for (int m = 0; m <= MAXD; ++m) {
        double v;
        if (m == 0) {
          v = y[dl.jb[0]];
        } else {
          const double a = y[dl.jb[m]], b = y[dl.jb[m] + 1];
          v = a + (b - a) * dl.fr[m];
        }
        if (td_coef(0, l, m) != 0.0) X += td_coef(0, l, m) * v;
        if (td_coef(1, l, m) != 0.0) Y += td_coef(1, l, m) * v;
        if (td_coef(2, l, m) != 0.0) Z += td_coef(2, l, m) * v;
      }
    }
    td_store<ORTH>(out, (size_t)n_out, i, X, Y, Z);
  }
}

template <int MAXD, bool ORTH>
__global__ __launch_bounds__(256) void k_tdi_td(const double* __restrict__ links, int n_ext, int n_out, TdDelays dl,
                                                double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  double X = 0.0, Y = 0.0, Z = 0.0;  // links 0:12 1:21 2:23 3:32 4:31 5:13
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    const double* y = links + (size_t)l * n_ext + i;
#pragma unroll
    for (int m = 0; m <= MAXD; ++m) {
      double v;
      if (m == 0) {
        v = y[dl.jb[0]];
      } else {
        const double a = y[dl.jb[m]], b = y[dl.jb[m] + 1];
        v = a + (b - a) * dl.fr[m];
      }
      if (td_coef(0, l, m) != 0.0) X += td_coef(0, l, m) * v;
      if (td_coef(1, l, m) != 0.0) Y += td_coef(1, l, m) * v;
      if (td_coef(2, l, m) != 0.0) Z += td_coef(2, l, m) * v;
    }
  }
  td_store<ORTH>(out, (size_t)n_out, i, X, Y, Z);
}

// ---- standalone Whittle lnL: -2 df sum_c sum_k |d - h|^2 / S (likelihood_ti.py:58-67) --------
constexpr int WHITTLE_BLOCKS = 1024;
__global__ __launch_bounds__(TILE) void k_whittle_partial(const double* __restrict__ data, const double* __restrict__ resp,
                                  const double* __restrict__ psd, int n_chan, int n, double* __restrict__ work) {
  __shared__ double red[TILE / 64];
  double v = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    for (int c = 0; c < n_chan; ++c) {
      size_t i = (size_t)c * n + k;
      double rr = data[2 * i] - resp[2 * i], ri = data[2 * i + 1] - resp[2 * i + 1];
      v += (rr * rr + ri * ri) / psd[i];
    }
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < TILE / 64; ++w) s += red[w];
    work[blockIdx.x] = s;
  }
}
__global__ void k_whittle_final(const double* __restrict__ work, int nb, double df, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += work[b];
    out[0] = -2.0 * df * s;
  }
}

#endif  // PES_LNL_FWD_UNIT

#define PES_LNL_PARAMS                                                                                       \
  const DetectorDev *dets, int n_det, int n_tiles, const double *bin_f, const int *bin_j, const double *bin_s, \
      const double *bin_rec, const double *tile_f0,                                                       \
      const double *tile_dd, int tiles_per_group, const double *pconst, const double *nodes,                 \
      const double *nodes_grad, int NNS, double du, ModeList modes, int P, int rows, const int *tile_jr,     \
      int tile_begin, int tile_end, int group_offset, int *work_counter, double *partial,                    \
      const unsigned long long *off_bits, const unsigned char *items
#define PES_LNL_NAMES                                                                                     \
  dets, n_det, n_tiles, bin_f, bin_j, bin_s, bin_rec, tile_f0, tile_dd, tiles_per_group, pconst, \
      nodes, nodes_grad, NNS, du, modes, P, rows, tile_jr, tile_begin, tile_end, group_offset, work_counter, partial, \
      off_bits, items

}  // namespace

// The forward-only instantiations of the fused kernel live in their own compilation unit
// (lnl_fwd.hip, this file compiled with PES_LNL_FWD_UNIT defined), so that they can be
// scheduled with their own compiler options (three waves per SIMD against two for the
// gradient kernel).  This launcher is that unit's only export.
namespace pes {
hipError_t launch_lnl_fwd(int precision, int kd, int grid, hipStream_t st, PES_LNL_PARAMS);
}

#ifdef PES_LNL_FWD_UNIT
hipError_t pes::launch_lnl_fwd(int precision, int kd, int grid, hipStream_t st, PES_LNL_PARAMS) {
  // lnL alone is the FP64 forward in every precision mode (the modes differ in the adjoint only)
  (void)precision;
  switch (kd) {
    case 0: hipLaunchKernelGGL((k_lnl<0, false, 0>), dim3(grid), dim3(TILE), 0, st, PES_LNL_NAMES); break;
    case 1: hipLaunchKernelGGL((k_lnl<0, false, 1>), dim3(grid), dim3(TILE), 0, st, PES_LNL_NAMES); break;
    case 2: hipLaunchKernelGGL((k_lnl<0, false, 2>), dim3(grid), dim3(TILE), 0, st, PES_LNL_NAMES); break;
    default: hipLaunchKernelGGL((k_lnl<0, false, 3>), dim3(grid), dim3(TILE), 0, st, PES_LNL_NAMES); break;
  }
  return hipGetLastError();
}
#else
// ================================= host side ================================================
struct pes_context {
  int device = 0;
  ModeList modes{};
  double f_ref = 0.0;
  bool direct = false;
  int n_nodes = 0;
  double f_node_min = 0.0, f_node_max = 0.0, u0 = 0.0, du = 1.0;
  bool finalized = false;
  bool any_duty = false;  // a detector has a duty cycle (selects the duty-window kernel variants)
  struct HostDet {
    DetectorDev d;
    std::vector<double> f, w, dre, dim;  // w, dre, dim: [3][n]
  };
  std::vector<HostDet> hdets;
  // device constant data
  DetectorDev* d_dets = nullptr;
  int n_tiles = 0, NB = 0;
  double *d_bin_f = nullptr, *d_bin_s = nullptr;
  double* d_bin_rec = nullptr;  // per bin: (w, Re d, Im d) of the three channels and one pad, 80 B
  int* d_bin_j = nullptr;
  double *d_tile_f0 = nullptr, *d_tile_dd = nullptr;  // first frequency and sum w |d|^2 per tile
  int* d_tile_jr = nullptr;  // per tile: first node index and number of node rows its bins use
  double* d_node_f = nullptr;  // direct mode only
  // scratch, grown on demand; while pinned (a captured HIP graph holds its pointers) it is
  // never reallocated and a call that would need more is rejected
  int cap_P = 0;
  bool cap_grad = false;
  bool pinned = false;
  size_t cap_partial = 0;  // doubles
  // schedule of the fused kernel (pes_context_set_schedule)
  int tiles_per_group = TILES_PER_GROUP;
  int blocks_per_cu = BLOCKS_PER_CU;
  int rows_override = 0;  // parameter rows per work item (0: chosen per batch)
  double *d_setup = nullptr, *d_pconst = nullptr, *d_nodes = nullptr, *d_nodes_grad = nullptr,
         *d_partial = nullptr;
  int* d_work_counter = nullptr;
  unsigned long long* d_off_bits = nullptr;  // [cap_P][off_words] duty off-tile flags (duty contexts)
  int n_cu = 256;
  // optional timing of the fused kernel (HIP events on the launch stream)
  bool profiling = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  double* d_live = nullptr;  // [cap_live][2] per-row work counts, filled by pes_profile_work
  int cap_live = 0;
  // the context scratch is refilled by every batch set up through it (scratch_seq counts them);
  // a profiled pes_lnl_batch records its row count and that batch's number, so that
  // pes_profile_work can count its work from the scratch while it still holds the batch
  unsigned long long scratch_seq = 0, live_seq = 0;
  int live_P = 0;
};

extern "C" const char* pes_last_error(void) { return g_err.c_str(); }

extern "C" int pes_context_create(pes_context** out, int device, int n_modes, const int* mode_ids, double f_ref,
                                  int n_nodes, double f_node_min, double f_node_max) {
  if (!out) return fail(PES_ERR_INVALID, "out is NULL");
  if (n_modes < 1 || n_modes > 6) return fail(PES_ERR_INVALID, "n_modes must be in [1, 6]");
  if (n_nodes == 1 || n_nodes < 0) return fail(PES_ERR_INVALID, "n_nodes must be 0 (direct) or >= 2");
  if (n_nodes >= 2 && !(f_node_min > 0.0 && f_node_max > f_node_min))
    return fail(PES_ERR_INVALID, "node range must satisfy 0 < f_node_min < f_node_max");
  pes_context* c = new pes_context();
  c->device = device;
  c->modes.n = n_modes;
  for (int i = 0; i < n_modes; ++i) {
    if (mode_ids[i] < 0 || mode_ids[i] > 5) {
      delete c;
      return fail(PES_ERR_INVALID, "mode id out of range [0, 5]");
    }
    for (int k = 0; k < i; ++k)
      if (mode_ids[k] == mode_ids[i]) {
        delete c;
        return fail(PES_ERR_INVALID, "duplicate mode id");
      }
    c->modes.id[i] = mode_ids[i];
  }
  c->f_ref = f_ref;
  c->direct = (n_nodes == 0);
  c->n_nodes = n_nodes;
  c->f_node_min = f_node_min;
  c->f_node_max = f_node_max;
  if (!c->direct) {
    c->u0 = std::log(f_node_min);
    c->du = (std::log(f_node_max) - c->u0) / (double)(n_nodes - 1);
  }
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete c;
    return fail(PES_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
  *out = c;
  return PES_OK;
}

// Orbit constants of one detector (heliocentric: orbit.py:189-228; geocentric: orbit.py:112-139
// with the rotation rate and reference source given).
static int fill_orbit(DetectorDev& d, int orbit_kind, double arm_length_m, double rotation_initial,
                      double revolution_initial, double omega_rotation, double lambda_ref, double beta_ref) {
  if (orbit_kind != 0 && orbit_kind != 1) return fail(PES_ERR_INVALID, "orbit_kind must be 0 or 1");
  if (!(arm_length_m > 0.0) || !std::isfinite(arm_length_m)) return fail(PES_ERR_INVALID, "arm length must be positive");
  if (!std::isfinite(rotation_initial) || !std::isfinite(revolution_initial))
    return fail(PES_ERR_INVALID, "initial orbit phases must be finite");
  if (orbit_kind == 1 && !(std::isfinite(omega_rotation) && std::isfinite(lambda_ref) && std::isfinite(beta_ref)))
    return fail(PES_ERR_INVALID, "geocentric rotation rate and reference source must be finite");
  orbit_constants(d, orbit_kind, arm_length_m, rotation_initial, revolution_initial, omega_rotation, lambda_ref,
                  beta_ref);
  return PES_OK;
}

extern "C" int pes_context_add_detector(pes_context* ctx, int orbit_kind, double arm_length_m,
                                        double rotation_initial, double revolution_initial, double omega_rotation,
                                        double lambda_ref, double beta_ref, int tdi_generation, int orthogonal,
                                        int n_bins, const double* f, const double* weight, double df,
                                        const double* data_re, const double* data_im, const double* psd,
                                        const int* channel_mask) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (ctx->finalized) return fail(PES_ERR_STATE, "context already finalized");
  if (tdi_generation != 15 && tdi_generation != 20) return fail(PES_ERR_INVALID, "tdi_generation must be 15 or 20");
  if (n_bins < 1 || !f || !data_re || !data_im || !psd || !channel_mask)
    return fail(PES_ERR_INVALID, "detector arrays must be non-NULL and n_bins >= 1");
  pes_context::HostDet h;
  DetectorDev& d = h.d;
  std::memset(&d, 0, sizeof(d));
  const int orc = fill_orbit(d, orbit_kind, arm_length_m, rotation_initial, revolution_initial, omega_rotation,
                             lambda_ref, beta_ref);
  if (orc != PES_OK) return orc;
  d.tdi_gen = tdi_generation;
  d.orthogonal = orthogonal ? 1 : 0;
  d.n_bins = n_bins;
  h.f.assign(f, f + n_bins);
  h.w.assign(3 * (size_t)n_bins, 0.0);
  h.dre.assign(data_re, data_re + 3 * (size_t)n_bins);
  h.dim.assign(data_im, data_im + 3 * (size_t)n_bins);
  for (int k = 0; k < n_bins; ++k) {
    if (!(f[k] > 0.0) || !std::isfinite(f[k])) return fail(PES_ERR_INVALID, "frequencies must be positive");
    if (k > 0 && !(f[k] > f[k - 1])) return fail(PES_ERR_INVALID, "frequencies must be strictly increasing");
    double dfk = weight ? weight[k] : df;
    for (int ch = 0; ch < 3; ++ch) {
      if (!channel_mask[ch]) continue;
      double s = psd[(size_t)ch * n_bins + k];
      if (!(s > 0.0)) return fail(PES_ERR_INVALID, "PSD must be positive on used channels");
      h.w[(size_t)ch * n_bins + k] = 2.0 * dfk / s;  // likelihood_ti.py:63-67: -2 df |d-h|^2 / S
    }
  }
  ctx->hdets.push_back(std::move(h));
  return PES_OK;
}

extern "C" int pes_context_set_duty_cycle(pes_context* ctx, int det_index, double period, double on_duration,
                                          double start_time, double taper) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (ctx->finalized) return fail(PES_ERR_STATE, "context already finalized");
  if (det_index < 0 || det_index >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "no such detector");
  DetectorDev& d = ctx->hdets[det_index].d;
  if (period == 0.0) {  // always observing
    d.duty = 0;
    return PES_OK;
  }
  if (!(period > 0.0) || !(on_duration > 0.0) || !(on_duration <= period) || !std::isfinite(start_time) ||
      !(taper >= 0.0) || !(2.0 * taper <= on_duration) || !std::isfinite(period))
    return fail(PES_ERR_INVALID,
                "duty cycle: need 0 < on_duration <= period, a finite start time and 0 <= taper <= on_duration / 2");
  d.duty = 1;
  d.duty_period = period;
  d.duty_on = on_duration;
  d.duty_t0 = start_time;
  d.duty_taper = taper;
  d.duty_inv_period = 1.0 / period;
  d.duty_inv_taper = taper > 0.0 ? 1.0 / taper : 0.0;
  return PES_OK;
}

// The fused kernel's launches: maximal runs of consecutive detectors with the same duty flag,
// each split into groups of tiles_per_group tiles (a group never straddles two runs), with the
// partial sums of run r's groups at group_offset(r) + 0, 1, ...
struct TileRange {
  int tile_begin, tile_end, duty, kind, group_offset;
};
// runs of consecutive detectors with the same orbit kind and duty flag: each is launched in the
// fused kernel's instantiation for that kind (only its code, so that its register allocation is
// not that of the other kind's code as well)
static int tile_ranges(const pes_context* ctx, TileRange* out) {
  int n = 0, groups = 0;
  for (const auto& h : ctx->hdets) {
    const int t0 = h.d.bin_offset / TILE, t1 = t0 + h.d.n_tiles, duty = h.d.duty ? 1 : 0, kind = h.d.kind;
    if (n > 0 && out[n - 1].duty == duty && out[n - 1].kind == kind && out[n - 1].tile_end == t0) {
      out[n - 1].tile_end = t1;
    } else {
      out[n++] = {t0, t1, duty, kind, 0};
    }
  }
  for (int r = 0; r < n; ++r) {
    out[r].group_offset = groups;
    groups += (out[r].tile_end - out[r].tile_begin + ctx->tiles_per_group - 1) / ctx->tiles_per_group;
  }
  return n;
}

// 64-bit words per row of the duty off-tile flags (one bit per tile of the context; the fused
// kernel forms the same count from n_tiles)
static int off_words_of(const pes_context* ctx) { return (ctx->n_tiles + 63) / 64; }

static int n_groups_of(const pes_context* ctx) {
  TileRange r[MAX_TILE_RANGES];
  const int n = tile_ranges(ctx, r);
  return r[n - 1].group_offset +
         (r[n - 1].tile_end - r[n - 1].tile_begin + ctx->tiles_per_group - 1) / ctx->tiles_per_group;
}

// ---- the entry point that allocates nothing ---------------------------------------------------
// Its scratch (work counter, per-row setup and constants, node tables, per-tile-group partials)
// lives in this file-scope device array, of which every GPU holds its own copy until the process
// ends.  Host code never uses the array's name as an address: its device address is taken once
// per device with hipGetSymbolAddress (at pes_context_finalize, with the device current) and
// kept in g_noalloc_base.
static __device__ __align__(256) unsigned char scratch[PES_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static char* g_noalloc_base[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on `device` (which must be current)
static int noalloc_base(int device, char** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(PES_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_base[device]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_base[device] = static_cast<char*>(p);
  }
  *out = g_noalloc_base[device];
  return PES_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets of the scratch buffers of a batch of P rows inside the array, and their total
static size_t noalloc_layout(const pes_context* ctx, int P, bool grad, size_t off[8]) {
  const size_t NNS = (size_t)ctx->n_nodes + 1, rows = (size_t)P, nm = (size_t)ctx->modes.n;
  size_t o = 0;
  off[0] = o; o += align256(N_WORK_COUNTERS * sizeof(int));                // work counters
  off[1] = o; o += align256(rows * N_SETUP * 5 * sizeof(double));          // setup
  off[2] = o; o += align256(rows * PC_STRIDE * sizeof(double));            // per-row constants
  off[3] = o; o += align256(rows * nm * NNS * 4 * sizeof(double));         // node tables
  off[4] = o; if (grad) o += align256(rows * nm * NNS * 16 * sizeof(double));  // their gradients
  off[5] = o; o += align256((size_t)n_groups_of(ctx) * (grad ? 1 + PES_N_PARAMS : 1) * rows * sizeof(double));
  off[6] = o; if (ctx->any_duty) o += align256(rows * off_words_of(ctx) * sizeof(unsigned long long));  // off tiles
  off[7] = o; if (ctx->any_duty) o += align256((size_t)n_groups_of(ctx) * rows);  // work items: plain or duty
  return o;
}

static void free_dev(void* p) {
  if (p) (void)hipFree(p);
}

extern "C" int pes_context_finalize(pes_context* ctx) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (ctx->finalized) return fail(PES_ERR_STATE, "context already finalized");
  if (ctx->hdets.empty()) return fail(PES_ERR_STATE, "no detector added");
  {  // runs of consecutive detectors with the same orbit kind and duty flag: a fused-kernel launch each
    int runs = 0;
    for (size_t i = 0; i < ctx->hdets.size(); ++i) {
      const auto& a = ctx->hdets[i].d;
      runs += (i == 0 || (a.duty != 0) != (ctx->hdets[i - 1].d.duty != 0) || a.kind != ctx->hdets[i - 1].d.kind)
                  ? 1
                  : 0;
    }
    if (runs > MAX_TILE_RANGES)
      return fail(PES_ERR_INVALID, "too many alternations of orbit kinds or of duty-cycled and continuous "
                                   "detectors (order the detectors so that at most 16 runs remain)");
  }
  HIP_TRY(hipSetDevice(ctx->device));
  long long total = 0;
  for (auto& h : ctx->hdets) total += (long long)(h.d.n_bins + TILE - 1) / TILE * TILE;
  // bin indices times the record length (BIN_REC doubles) are formed in 32-bit integers
  if (total > (long long)(INT32_MAX / BIN_REC))
    return fail(PES_ERR_INVALID, "too many bins in the network (at most 214,748,364 after padding to tiles)");
  int off = 0;
  for (auto& h : ctx->hdets) {
    h.d.bin_offset = off;
    h.d.n_tiles = (h.d.n_bins + TILE - 1) / TILE;
    off += h.d.n_tiles * TILE;
  }
  ctx->NB = off;
  ctx->n_tiles = off / TILE;
  const int NB = ctx->NB;
  std::vector<double> bf(NB), bs(NB, 0.0), bw(3 * (size_t)NB, 0.0), dre(3 * (size_t)NB, 0.0), dim(3 * (size_t)NB, 0.0);
  std::vector<int> bj(NB, 0);
  for (auto& h : ctx->hdets) {
    const int n = h.d.n_bins, o = h.d.bin_offset;
    for (int k = 0; k < h.d.n_tiles * TILE; ++k) {
      int src = k < n ? k : n - 1;  // padding lanes repeat the last bin with zero weight
      bf[o + k] = h.f[src];
      for (int ch = 0; ch < 3; ++ch) {
        if (k < n) {
          bw[(size_t)ch * NB + o + k] = h.w[(size_t)ch * n + k];
          dre[(size_t)ch * NB + o + k] = h.dre[(size_t)ch * n + k];
          dim[(size_t)ch * NB + o + k] = h.dim[(size_t)ch * n + k];
        }
      }
    }
  }
  if (ctx->direct) {
    ctx->n_nodes = NB;
    ctx->du = 1.0;
    for (int k = 0; k < NB; ++k) {
      bj[k] = k;
      bs[k] = 0.0;
    }
  } else {
    // every bin must lie inside the node range: the cubic Hermite would extrapolate outside it
    const double tol = 1e-6;  // in node intervals (a relative frequency tolerance of ~1e-6 du)
    for (int k = 0; k < NB; ++k) {
      const double x = (std::log(bf[k]) - ctx->u0) / ctx->du;
      if (!(x >= -tol && x <= (double)(ctx->n_nodes - 1) + tol)) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "bin frequency %.17g Hz lies outside the interpolation node range [%.17g, %.17g] Hz: "
                      "widen f_node_min / f_node_max to cover every detector grid",
                      bf[k], ctx->f_node_min, ctx->f_node_max);
        return fail(PES_ERR_INVALID, msg);
      }
    }
    for (int k = 0; k < NB; ++k) {
      double x = (std::log(bf[k]) - ctx->u0) / ctx->du;
      double jf = std::floor(x);
      if (jf < 0.0) jf = 0.0;
      if (jf > (double)(ctx->n_nodes - 2)) jf = (double)(ctx->n_nodes - 2);
      bj[k] = (int)jf;
      bs[k] = x - jf;
    }
  }
  size_t nd = ctx->hdets.size();
  std::vector<DetectorDev> dd(nd);
  ctx->any_duty = false;
  for (size_t i = 0; i < nd; ++i) {
    dd[i] = ctx->hdets[i].d;
    ctx->any_duty = ctx->any_duty || dd[i].duty != 0;
  }
  HIP_TRY(hipMalloc(&ctx->d_dets, nd * sizeof(DetectorDev)));
  HIP_TRY(hipMemcpy(ctx->d_dets, dd.data(), nd * sizeof(DetectorDev), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&ctx->d_bin_f, NB * sizeof(double)));
  HIP_TRY(hipMalloc(&ctx->d_bin_s, NB * sizeof(double)));
  HIP_TRY(hipMalloc(&ctx->d_bin_j, NB * sizeof(int)));
  HIP_TRY(hipMalloc(&ctx->d_bin_rec, BIN_REC * (size_t)NB * sizeof(double)));
  HIP_TRY(hipMemcpy(ctx->d_bin_f, bf.data(), NB * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->d_bin_s, bs.data(), NB * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->d_bin_j, bj.data(), NB * sizeof(int), hipMemcpyHostToDevice));
  {
    std::vector<double> rec(BIN_REC * (size_t)NB, 0.0);
    for (size_t k = 0; k < (size_t)NB; ++k)
      for (int ch = 0; ch < 3; ++ch) {
        rec[k * BIN_REC + 3 * ch] = bw[(size_t)ch * NB + k];
        rec[k * BIN_REC + 3 * ch + 1] = dre[(size_t)ch * NB + k];
        rec[k * BIN_REC + 3 * ch + 2] = dim[(size_t)ch * NB + k];
      }
    HIP_TRY(hipMemcpy(ctx->d_bin_rec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  {
    const int nt = ctx->n_tiles;
    std::vector<double> tf0(nt), tdd(nt, 0.0);
    for (int t = 0; t < nt; ++t) {
      tf0[t] = bf[(size_t)t * TILE];
      double s = 0.0;
      for (int ch = 0; ch < 3; ++ch)
        for (int k = 0; k < TILE; ++k) {
          const size_t i = (size_t)ch * NB + (size_t)t * TILE + k;
          s += bw[i] * (dre[i] * dre[i] + dim[i] * dim[i]);
        }
      tdd[t] = s;
    }
    HIP_TRY(hipMalloc(&ctx->d_tile_f0, nt * sizeof(double)));
    HIP_TRY(hipMalloc(&ctx->d_tile_dd, nt * sizeof(double)));
    HIP_TRY(hipMemcpy(ctx->d_tile_f0, tf0.data(), nt * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->d_tile_dd, tdd.data(), nt * sizeof(double), hipMemcpyHostToDevice));
    // node rows j .. j+1 of every bin of a tile (node indices are non-decreasing with f)
    std::vector<int> tjr(2 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
      const int jf = bj[(size_t)t * TILE], jl = bj[(size_t)t * TILE + TILE - 1];
      tjr[2 * (size_t)t] = jf;
      tjr[2 * (size_t)t + 1] = jl - jf + 2;
    }
    HIP_TRY(hipMalloc(&ctx->d_tile_jr, 2 * (size_t)nt * sizeof(int)));
    HIP_TRY(hipMemcpy(ctx->d_tile_jr, tjr.data(), 2 * (size_t)nt * sizeof(int), hipMemcpyHostToDevice));
  }
  if (ctx->direct) ctx->d_node_f = ctx->d_bin_f;
  HIP_TRY(hipMalloc(&ctx->d_work_counter, N_WORK_COUNTERS * sizeof(int)));  // one per fused-kernel launch
  {  // the device address of the no-allocation entry point's scratch, once per device
    char* base = nullptr;
    const int rc = noalloc_base(ctx->device, &base);
    if (rc != PES_OK) return rc;
  }
  ctx->finalized = true;
  return PES_OK;
}

extern "C" void pes_context_destroy(pes_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  free_dev(ctx->d_dets);
  free_dev(ctx->d_bin_f);
  free_dev(ctx->d_bin_s);
  free_dev(ctx->d_bin_j);
  free_dev(ctx->d_bin_rec);
  free_dev(ctx->d_tile_f0);
  free_dev(ctx->d_work_counter);
  free_dev(ctx->d_tile_dd);
  free_dev(ctx->d_tile_jr);
  free_dev(ctx->d_setup);
  free_dev(ctx->d_pconst);
  free_dev(ctx->d_nodes);
  free_dev(ctx->d_nodes_grad);
  free_dev(ctx->d_off_bits);
  free_dev(ctx->d_partial);
  free_dev(ctx->d_live);
  for (auto e : ctx->ev_start) (void)hipEventDestroy(e);
  for (auto e : ctx->ev_stop) (void)hipEventDestroy(e);
  delete ctx;
}

extern "C" int pes_detector_bins(const pes_context* ctx, int det) {
  if (!ctx || det < 0 || det >= (int)ctx->hdets.size()) return -1;
  return ctx->hdets[det].d.n_bins;
}


// (re)allocate scratch for P rows; allocation happens only when P, the gradient flag or the
// number of tile groups grows
// The partial-sum buffer ([groups][12][cap] doubles) carries, for duty contexts, one byte per
// work item (tile group x row chunk) after the sums: whether the item runs in the plain or the
// duty instantiation (k_duty_items)
static size_t items_offset_doubles(const pes_context* ctx, size_t cap) {
  return (size_t)n_groups_of(ctx) * (1 + PES_N_PARAMS) * cap;
}
static size_t partial_doubles(const pes_context* ctx, size_t cap) {
  size_t n = items_offset_doubles(ctx, cap);
  if (ctx->any_duty) n += ((size_t)n_groups_of(ctx) * cap + 7) / 8;
  return n;
}

static int ensure_scratch(pes_context* ctx, int P, bool grad) {
  {
    const size_t need = partial_doubles(ctx, (size_t)(P > ctx->cap_P ? P : ctx->cap_P));
    const bool grow = need > ctx->cap_partial || P > ctx->cap_P || (grad && !ctx->cap_grad);
    if (grow && ctx->pinned)
      return fail(PES_ERR_STATE,
                  "scratch is pinned (a captured HIP graph holds its pointers) and this call needs more: "
                  "reserve the largest batch, the gradient and the schedule before pinning");
    if (need > ctx->cap_partial) {
      free_dev(ctx->d_partial);
      ctx->d_partial = nullptr;
      HIP_TRY(hipMalloc(&ctx->d_partial, need * sizeof(double)));
      ctx->cap_partial = need;
    }
  }
  if (P <= ctx->cap_P && (!grad || ctx->cap_grad)) return PES_OK;
  int cap = P > ctx->cap_P ? P : ctx->cap_P;
  bool g = grad || ctx->cap_grad;
  ++ctx->scratch_seq;  // the rows and node tables of earlier batches are gone
  free_dev(ctx->d_setup);
  free_dev(ctx->d_pconst);
  free_dev(ctx->d_nodes);
  free_dev(ctx->d_nodes_grad);
  free_dev(ctx->d_off_bits);
  ctx->d_setup = ctx->d_pconst = ctx->d_nodes = ctx->d_nodes_grad = nullptr;
  ctx->d_off_bits = nullptr;
  size_t NNS = (size_t)ctx->n_nodes + 1;
  HIP_TRY(hipMalloc(&ctx->d_setup, (size_t)cap * N_SETUP * 5 * sizeof(double)));
  HIP_TRY(hipMalloc(&ctx->d_pconst, (size_t)cap * PC_STRIDE * sizeof(double)));
  HIP_TRY(hipMalloc(&ctx->d_nodes, (size_t)cap * ctx->modes.n * NNS * 4 * sizeof(double)));
  if (g) HIP_TRY(hipMalloc(&ctx->d_nodes_grad, (size_t)cap * ctx->modes.n * NNS * 16 * sizeof(double)));
  if (ctx->any_duty)
    HIP_TRY(hipMalloc(&ctx->d_off_bits, (size_t)cap * off_words_of(ctx) * sizeof(unsigned long long)));
  ctx->cap_P = cap;
  ctx->cap_grad = g;
  return PES_OK;
}

// Device buffers a batch of P rows works in besides its inputs and outputs.
struct BatchScratch {
  double *setup, *pconst, *nodes, *nodes_grad, *partial;
  int* work_counter;
  unsigned long long* off_bits;  // duty contexts only
  unsigned char* items;          // duty contexts only: [groups][chunks] plain (1) or duty (0)
};

static BatchScratch context_scratch(const pes_context* ctx) {
  unsigned char* items =
      ctx->any_duty && ctx->d_partial
          ? reinterpret_cast<unsigned char*>(ctx->d_partial + items_offset_doubles(ctx, (size_t)ctx->cap_P))
          : nullptr;
  return {ctx->d_setup,   ctx->d_pconst,       ctx->d_nodes,    ctx->d_nodes_grad,
          ctx->d_partial, ctx->d_work_counter, ctx->d_off_bits, items};
}

static int check_batch_args(const pes_context* ctx, const double* d_params, int P) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (!ctx->finalized) return fail(PES_ERR_STATE, "context not finalized");
  if (P < 0) return fail(PES_ERR_INVALID, "P must be >= 0");
  if (P > 0 && !d_params) return fail(PES_ERR_INVALID, "params is NULL");
  return PES_OK;
}

// k_setup (which also zeroes the work counters) and k_nodes for P >= 1 rows
static int launch_setup_nodes(const pes_context* ctx, const BatchScratch& sc, const double* d_params, int P, bool grad,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_setup, dim3((P + 63) / 64), dim3(64), 0, st, d_params, P, ctx->modes, ctx->f_ref, sc.setup,
                     sc.pconst, sc.work_counter);
  HIP_TRY(hipGetLastError());
  const int NNS = ctx->n_nodes + 1;
  dim3 g((NNS + 127) / 128, P * ctx->modes.n);
  if (grad)
    hipLaunchKernelGGL(k_nodes<true>, g, dim3(128), 0, st, sc.setup, P, ctx->modes, ctx->n_nodes, ctx->u0, ctx->du,
                       ctx->d_node_f, sc.nodes, sc.nodes_grad);
  else
    hipLaunchKernelGGL(k_nodes<false>, g, dim3(128), 0, st, sc.setup, P, ctx->modes, ctx->n_nodes, ctx->u0, ctx->du,
                       ctx->d_node_f, sc.nodes, nullptr);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

static int prepare(pes_context* ctx, const double* d_params, int P, bool grad, hipStream_t st) {
  int rc = check_batch_args(ctx, d_params, P);
  if (rc != PES_OK) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  rc = ensure_scratch(ctx, P > 0 ? P : 1, grad);
  if (rc != PES_OK) return rc;
  if (P == 0) return PES_OK;
  ++ctx->scratch_seq;
  return launch_setup_nodes(ctx, context_scratch(ctx), d_params, P, grad, st);
}

// rows per work item of the fused kernel: the kernel's chunk, fewer when the grid has too few
// tile groups to give every persistent block about four work items (small grids, e.g. 16k bins)
static int rows_per_item(const pes_context* ctx, int P, bool grad, int grid) {
  const int max_rows = grad ? chunk_rows<true>() : chunk_rows<false>();
  if (ctx->rows_override > 0) return std::min(ctx->rows_override, max_rows);
  const long long fit = (long long)P * n_groups_of(ctx) / (4LL * grid);
  return (int)std::max(8LL, std::min((long long)max_rows, fit));
}

// The fused kernel and the reduction, on scratch already set up by launch_setup_nodes.
static int launch_lnl_reduce(pes_context* ctx, const BatchScratch& sc, int P, double* d_lnl, double* d_grad,
                             int precision, hipStream_t st, bool profile) {
  const bool grad = d_grad != nullptr;
  const int NNS = ctx->n_nodes + 1;
  const int n_groups = n_groups_of(ctx);
  int grid = ctx->n_cu * ctx->blocks_per_cu;
  const int rows = rows_per_item(ctx, P, grad, grid);
  const int n_chunks = (P + rows - 1) / rows;
  const int nout = grad ? 1 + PES_N_PARAMS : 1;
  // one launch per run of detectors with the same duty flag (two for a duty-cycled run, below):
  // the duty-cycled detectors' tiles run the DUTY instantiation, the others the plain one (whose
  // register allocation is that of a kernel without the window code)
  TileRange ranges[MAX_TILE_RANGES];
  const int n_ranges = tile_ranges(ctx, ranges);
  const int off_words = off_words_of(ctx);
  if (ctx->any_duty) {  // tiles that a duty cycle's off periods cover for every mode of a row
    hipLaunchKernelGGL(k_duty_off, dim3(off_words, std::min(P, 65535)), dim3(64), 0, st, ctx->d_dets,
                       (int)ctx->hdets.size(), ctx->n_tiles, ctx->d_tile_f0, ctx->d_tile_jr, sc.pconst, sc.nodes, NNS,
                       ctx->u0, ctx->du, ctx->modes, P, off_words, ctx->direct ? 0 : 1, sc.off_bits);
    HIP_TRY(hipGetLastError());
  }
  const int n_det = (int)ctx->hdets.size();
  // a duty-cycled range's work items on the window's plateau (for every row and mode of the item)
  // run in the plain instantiation, the others in the duty one: two launches of the range
  for (int r = 0; r < n_ranges; ++r) {
    if (!ranges[r].duty) continue;
    const int n_work =
        (ranges[r].tile_end - ranges[r].tile_begin + ctx->tiles_per_group - 1) / ctx->tiles_per_group * n_chunks;
    hipLaunchKernelGGL(k_duty_items, dim3(n_work), dim3(64), 0, st, ctx->d_dets, n_det, ctx->d_tile_f0,
                       ctx->d_tile_jr, sc.pconst, sc.nodes, NNS, ctx->u0, ctx->du, ctx->modes, P, rows, n_chunks,
                       ranges[r].tile_begin, ranges[r].tile_end, ctx->tiles_per_group, ranges[r].group_offset,
                       ctx->direct ? 0 : 1, sc.items, sc.work_counter + 2 * MAX_TILE_RANGES + r);
    HIP_TRY(hipGetLastError());
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (profile) {
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
  }
  for (int r = 0; r < n_ranges; ++r) {
    const int tile_begin = ranges[r].tile_begin, tile_end = ranges[r].tile_end;
    const int group_offset = ranges[r].group_offset;
    const int n_work = (tile_end - tile_begin + ctx->tiles_per_group - 1) / ctx->tiles_per_group * n_chunks;
    const int grid_r = std::min(grid, n_work);
    for (int pass = 0; pass < (ranges[r].duty ? 2 : 1); ++pass) {
      // pass 0: the plain instantiation (all items of a plain range; the plateau items of a
      // duty-cycled one), pass 1: the duty instantiation (the other items); both for the range's
      // orbit kind only
      const int kd = ranges[r].kind | (pass == 1 ? 2 : 0);
      const unsigned char* items = ranges[r].duty ? sc.items : nullptr;
      int* counter = sc.work_counter + (pass == 0 ? r : MAX_TILE_RANGES + r);
#define PES_LNL_ARGS                                                                                         \
  ctx->d_dets, n_det, ctx->n_tiles, ctx->d_bin_f, ctx->d_bin_j, ctx->d_bin_s, ctx->d_bin_rec,                 \
      ctx->d_tile_f0, ctx->d_tile_dd, ctx->tiles_per_group, sc.pconst, sc.nodes, sc.nodes_grad, NNS, ctx->du, \
      ctx->modes, P, rows, ctx->d_tile_jr, tile_begin, tile_end, group_offset, counter, sc.partial,       \
      sc.off_bits, items
#define PES_LNL_GRAD(PR, KD) hipLaunchKernelGGL((k_lnl<PR, true, KD>), dim3(grid_r), dim3(TILE), 0, st, PES_LNL_ARGS)
      if (!grad) {
        HIP_TRY(launch_lnl_fwd(precision, kd, grid_r, st, PES_LNL_ARGS));
      } else if (precision == 0) {
        switch (kd) {
          case 0: PES_LNL_GRAD(0, 0); break;
          case 1: PES_LNL_GRAD(0, 1); break;
          case 2: PES_LNL_GRAD(0, 2); break;
          default: PES_LNL_GRAD(0, 3); break;
        }
      } else {
        switch (kd) {
          case 0: PES_LNL_GRAD(1, 0); break;
          case 1: PES_LNL_GRAD(1, 1); break;
          case 2: PES_LNL_GRAD(1, 2); break;
          default: PES_LNL_GRAD(1, 3); break;
        }
      }
#undef PES_LNL_GRAD
#undef PES_LNL_ARGS
      HIP_TRY(hipGetLastError());
    }
  }
  if (profile) {
    HIP_TRY(hipEventRecord(e1, st));
    ctx->ev_start.push_back(e0);
    ctx->ev_stop.push_back(e1);
    ctx->live_P = P;  // the work is counted by pes_profile_work, outside the caller's timed steps
    ctx->live_seq = ctx->scratch_seq;
  }
  hipLaunchKernelGGL(k_reduce, dim3((P + 63) / 64, nout), dim3(64 * REDUCE_WAVES), 0, st, sc.partial, n_groups, P,
                     nout, sc.pconst, d_lnl, d_grad);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" size_t pes_lnl_batch_noalloc_bytes(const pes_context* ctx, int P, int grad) {
  if (!ctx || !ctx->finalized || P < 1) return 0;
  size_t off[8];
  return noalloc_layout(ctx, P, grad != 0, off);
}

extern "C" int pes_lnl_batch_noalloc(pes_context* ctx, const double* d_params, int P, double* d_lnl, double* d_grad,
                                     int precision, void* stream) {
  if (precision != PES_PREC_FP64 && precision != PES_PREC_MIXED)
    return fail(PES_ERR_INVALID, "precision must be PES_PREC_FP64 or PES_PREC_MIXED");
  if (P > 0 && !d_lnl) return fail(PES_ERR_INVALID, "lnl output is NULL");
  int rc = check_batch_args(ctx, d_params, P);
  if (rc != PES_OK || P == 0) return rc;
  const bool grad = d_grad != nullptr;
  size_t off[8];
  const size_t need = noalloc_layout(ctx, P, grad, off);
  if (need > (size_t)PES_NOALLOC_SCRATCH_BYTES) {
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "pes_lnl_batch_noalloc: %d rows%s need %zu bytes of scratch, more than the fixed %zu "
                  "(split the batch, or use pes_lnl_batch)",
                  P, grad ? " with the gradient" : "", need, (size_t)PES_NOALLOC_SCRATCH_BYTES);
    return fail(PES_ERR_INVALID, msg);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  char* base = nullptr;
  rc = noalloc_base(ctx->device, &base);
  if (rc != PES_OK) return rc;
  BatchScratch sc;
  sc.work_counter = reinterpret_cast<int*>(base + off[0]);
  sc.setup = reinterpret_cast<double*>(base + off[1]);
  sc.pconst = reinterpret_cast<double*>(base + off[2]);
  sc.nodes = reinterpret_cast<double*>(base + off[3]);
  sc.nodes_grad = grad ? reinterpret_cast<double*>(base + off[4]) : nullptr;
  sc.partial = reinterpret_cast<double*>(base + off[5]);
  sc.off_bits = ctx->any_duty ? reinterpret_cast<unsigned long long*>(base + off[6]) : nullptr;
  sc.items = ctx->any_duty ? reinterpret_cast<unsigned char*>(base + off[7]) : nullptr;
  hipStream_t st = (hipStream_t)stream;
  rc = launch_setup_nodes(ctx, sc, d_params, P, grad, st);
  if (rc != PES_OK) return rc;
  return launch_lnl_reduce(ctx, sc, P, d_lnl, d_grad, precision, st, false);
}

extern "C" int pes_lnl_batch(pes_context* ctx, const double* d_params, int P, double* d_lnl, double* d_grad,
                             int precision, void* stream) {
  if (precision != PES_PREC_FP64 && precision != PES_PREC_MIXED)
    return fail(PES_ERR_INVALID, "precision must be PES_PREC_FP64 or PES_PREC_MIXED");
  if (P > 0 && !d_lnl) return fail(PES_ERR_INVALID, "lnl output is NULL");
  hipStream_t st = (hipStream_t)stream;
  bool grad = d_grad != nullptr;
  int rc = prepare(ctx, d_params, P, grad, st);
  if (rc != PES_OK || P == 0) return rc;
  return launch_lnl_reduce(ctx, context_scratch(ctx), P, d_lnl, d_grad, precision, st, ctx->profiling);
}

extern "C" int pes_response(pes_context* ctx, int det, const double* d_params, int P, double* d_out, int precision,
                            void* stream) {
  if (precision != PES_PREC_FP64 && precision != PES_PREC_MIXED)
    return fail(PES_ERR_INVALID, "precision must be PES_PREC_FP64 or PES_PREC_MIXED");
  if (!ctx || det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (P > 0 && !d_out) return fail(PES_ERR_INVALID, "out is NULL");
  hipStream_t st = (hipStream_t)stream;
  int rc = prepare(ctx, d_params, P, false, st);
  if (rc != PES_OK || P == 0) return rc;
  int n = ctx->hdets[det].d.n_bins;
#define PES_RESP_ARGS                                                                                          \
  dim3((n + 127) / 128, P), dim3(128), 0, st, ctx->d_dets, det, ctx->d_bin_f, ctx->d_bin_j, ctx->d_bin_s,     \
      ctx->d_pconst, ctx->d_nodes, ctx->n_nodes + 1, ctx->du, ctx->modes, P, d_out
  // the response is the FP64 forward in every precision mode
  hipLaunchKernelGGL(k_response_t<0>, PES_RESP_ARGS);
#undef PES_RESP_ARGS
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_waveform(pes_context* ctx, int det, const double* d_params, int P, double* d_out, void* stream) {
  if (!ctx || det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (P > 0 && !d_out) return fail(PES_ERR_INVALID, "out is NULL");
  hipStream_t st = (hipStream_t)stream;
  int rc = prepare(ctx, d_params, P, false, st);
  if (rc != PES_OK || P == 0) return rc;
  int n = ctx->hdets[det].d.n_bins;
  hipLaunchKernelGGL(k_waveform, dim3((n + 127) / 128, P), dim3(128), 0, st, ctx->d_dets, det, ctx->d_bin_f,
                     ctx->d_bin_j, ctx->d_bin_s, ctx->d_pconst, ctx->d_nodes, ctx->n_nodes + 1, ctx->du, ctx->modes,
                     P, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

static Extrinsic extrinsic_of(double lam, double beta, double psi, double tc);

extern "C" int pes_response_from_waveform(pes_context* ctx, int det, int n_modes, const double* d_wf, double lam,
                                          double beta, double psi, double tc, double* d_out, void* stream) {
  if (!ctx || !ctx->finalized) return fail(PES_ERR_STATE, "context missing or not finalized");
  if (det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (n_modes < 1 || !d_wf || !d_out) return fail(PES_ERR_INVALID, "waveform/out must be non-NULL, n_modes >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  const Extrinsic ex = extrinsic_of(lam, beta, psi, tc);
  int n = ctx->hdets[det].d.n_bins;
  hipLaunchKernelGGL(k_response_container, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, ctx->d_dets,
                     det, ctx->d_bin_f, n_modes, d_wf, ex, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_links(pes_context* ctx, int det, const double* d_params, int P, double* d_out, void* stream) {
  if (!ctx || det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (P > 0 && !d_out) return fail(PES_ERR_INVALID, "out is NULL");
  hipStream_t st = (hipStream_t)stream;
  int rc = prepare(ctx, d_params, P, false, st);
  if (rc != PES_OK || P == 0) return rc;
  const int n = ctx->hdets[det].d.n_bins;
  hipLaunchKernelGGL(k_links, dim3((n + 127) / 128, P), dim3(128), 0, st, ctx->d_dets, det, ctx->d_bin_f,
                     ctx->d_bin_j, ctx->d_bin_s, ctx->d_pconst, ctx->d_nodes, ctx->n_nodes + 1, ctx->du, ctx->modes,
                     P, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

static Extrinsic extrinsic_of(double lam, double beta, double psi, double tc) {
  Extrinsic ex;
  double sl = std::sin(lam), cl = std::cos(lam), sb = std::sin(beta), cb = std::cos(beta);
  double sp = std::sin(psi), cp = std::cos(psi);
  ex.p = {sl * cp - cl * sb * sp, -cl * cp - sl * sb * sp, cb * sp};   // antenna.py:70-76
  ex.q = {-sl * sp - cl * sb * cp, cl * sp - sl * sb * cp, cb * cp};   // antenna.py:77-83
  ex.k = {-cb * cl, -cb * sl, -sb};                                    // antenna.py:112-114
  ex.tc = tc;
  return ex;
}

extern "C" int pes_links_from_waveform(pes_context* ctx, int det, int n_modes, const double* d_wf, double lam,
                                       double beta, double psi, double tc, double* d_out, void* stream) {
  if (!ctx || !ctx->finalized) return fail(PES_ERR_STATE, "context missing or not finalized");
  if (det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (n_modes < 1 || !d_wf || !d_out) return fail(PES_ERR_INVALID, "waveform/out must be non-NULL, n_modes >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  const int n = ctx->hdets[det].d.n_bins;
  hipLaunchKernelGGL(k_links_container, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, ctx->d_dets, det,
                     ctx->d_bin_f, n_modes, d_wf, extrinsic_of(lam, beta, psi, tc), d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_tdi_from_links(pes_context* ctx, int det, int P, const double* d_links, double* d_out,
                                  void* stream) {
  if (!ctx || !ctx->finalized) return fail(PES_ERR_STATE, "context missing or not finalized");
  if (det < 0 || det >= (int)ctx->hdets.size()) return fail(PES_ERR_INVALID, "bad detector index");
  if (P < 0) return fail(PES_ERR_INVALID, "P must be >= 0");
  if (P > 0 && (!d_links || !d_out)) return fail(PES_ERR_INVALID, "links/out must be non-NULL");
  if (P == 0) return PES_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  const int n = ctx->hdets[det].d.n_bins;
  hipLaunchKernelGGL(k_tdi_from_links, dim3((n + 127) / 128, P), dim3(128), 0, (hipStream_t)stream, ctx->d_dets,
                     det, ctx->d_bin_f, P, d_links, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_orbit_vectors(int orbit_kind, double arm_length_m, double rotation_initial,
                                 double revolution_initial, double omega_rotation, double lambda_ref, double beta_ref,
                                 const double* d_t, int n, double* d_out, void* stream) {
  if (n < 0) return fail(PES_ERR_INVALID, "n must be >= 0");
  if (n > 0 && (!d_t || !d_out)) return fail(PES_ERR_INVALID, "times/out must be non-NULL");
  DetectorDev d;
  std::memset(&d, 0, sizeof(d));
  const int rc = fill_orbit(d, orbit_kind, arm_length_m, rotation_initial, revolution_initial, omega_rotation,
                            lambda_ref, beta_ref);
  if (rc != PES_OK || n == 0) return rc;
  hipLaunchKernelGGL(k_orbit_vectors, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, d, d_t, n, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_td_links(int orbit_kind, double arm_length_m, double rotation_initial, double revolution_initial,
                            double omega_rotation, double lambda_ref, double beta_ref, const double* d_hp,
                            const double* d_hc, double wave_t0, double wave_dt, int wave_n, double lam, double beta,
                            double psi, double t0, double dt, int n, double* d_out, void* stream) {
  if (n < 0 || wave_n < 2) return fail(PES_ERR_INVALID, "td links: n >= 0 and at least two waveform samples");
  if (!(wave_dt > 0.0) || !(dt > 0.0)) return fail(PES_ERR_INVALID, "td links: sample spacings must be positive");
  if (n > 0 && (!d_hp || !d_hc || !d_out)) return fail(PES_ERR_INVALID, "td links: arrays must be non-NULL");
  DetectorDev d;
  std::memset(&d, 0, sizeof(d));
  const int rc = fill_orbit(d, orbit_kind, arm_length_m, rotation_initial, revolution_initial, omega_rotation,
                            lambda_ref, beta_ref);
  if (rc != PES_OK || n == 0) return rc;
  const TdWave w{d_hp, d_hc, wave_t0, 1.0 / wave_dt, wave_n};
  const Extrinsic ex = extrinsic_of(lam, beta, psi, 0.0);
  // every block's waveform window has length floor(b_hi + R) - floor(b_lo - R - L / dt_w) + 6
  // (td_links_window), at most 255 dt / dt_w + 2 R + L / dt_w + 8: stage it in LDS when that fits
  const double span = 255.0 * dt / wave_dt + 2.0 * (AU_SEC + 3.0 * d.L) / wave_dt + 8.0;  // >= every block's
  const int kw = span <= 512.0 ? 2 : (span <= (double)TD_LINKS_WMAX ? 6 : 0);
  const dim3 grid((n + 255) / 256), block(256);
  hipStream_t st = (hipStream_t)stream;
#define PES_TDL(K, W) hipLaunchKernelGGL((k_links_td<K, W>), grid, block, 0, st, d, w, ex, t0, dt, n, d_out)
  if (d.kind == 0) {
    if (kw == 2) PES_TDL(0, 2); else if (kw == 6) PES_TDL(0, 6); else PES_TDL(0, 0);
  } else {
    if (kw == 2) PES_TDL(1, 2); else if (kw == 6) PES_TDL(1, 6); else PES_TDL(1, 0);
  }
#undef PES_TDL
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_td_tdi_added_samples(double arm_length_m, double dt, int tdi_generation) {
  if (!(arm_length_m > 0.0) || !(dt > 0.0) || (tdi_generation != 15 && tdi_generation != 20)) return -1;
  const int max_delay = tdi_generation == 20 ? 7 : 3;
  return (int)std::floor(max_delay * (arm_length_m / C_SI) / dt) + 1;
}

extern "C" int pes_td_tdi(double arm_length_m, double dt, int tdi_generation, int orthogonal, const double* d_links,
                          int n_ext, double* d_out, void* stream) {
  const int added = pes_td_tdi_added_samples(arm_length_m, dt, tdi_generation);
  if (added < 0) return fail(PES_ERR_INVALID, "td tdi: arm length and dt must be positive, generation 15 or 20");
  const int n_out = n_ext - added;
  if (n_out < 0) return fail(PES_ERR_INVALID, "td tdi: fewer link samples than the delays need");
  if (n_out > 0 && (!d_links || !d_out)) return fail(PES_ERR_INVALID, "td tdi: arrays must be non-NULL");
  if (n_out == 0) return PES_OK;
  // delays m L / dt split once: the index stays inside [0, n_ext - 1] for every output sample
  // (added - m L / dt > 0 up to rounding, clamped; the pair read at m >= 1 ends at i + added)
  const int maxd = tdi_generation == 20 ? 7 : 3;
  const double r = (arm_length_m / C_SI) / dt;
  TdDelays dl;
  std::memset(&dl, 0, sizeof(dl));
  dl.jb[0] = added;
  for (int m = 1; m <= maxd; ++m) {
    const double c = (double)added - m * r;
    const int jb = std::min(std::max((int)std::floor(c), 0), added - 1);
    dl.jb[m] = jb;
    dl.fr[m] = c - (double)jb;
  }
  hipStream_t st = (hipStream_t)stream;
  const int span = added - dl.jb[maxd];  // window beyond the block: jb[0] - jb_min
  if (span <= TD_WIN_EXTRA) {
    // LDS window of each link: TD_BLOCK TD_OPT + span + 1 samples (<= 36.9 KB for the six links
    // at TD_OPT = 2); every read index t - jb_min + jb[m] (+ 1), t < TD_BLOCK TD_OPT, lies in
    // [0, TD_BLOCK TD_OPT - 1 + span]
    const dim3 grid((n_out + TD_BLOCK * TD_OPT - 1) / (TD_BLOCK * TD_OPT)), block(TD_BLOCK);
    const size_t lds = 6 * (size_t)(TD_BLOCK * TD_OPT + span + 1) * sizeof(double);
    if (maxd == 3 && orthogonal) hipLaunchKernelGGL((k_tdi_td_lds<3, true>), grid, block, lds, st, d_links, n_ext, n_out, dl, d_out);
    else if (maxd == 3) hipLaunchKernelGGL((k_tdi_td_lds<3, false>), grid, block, lds, st, d_links, n_ext, n_out, dl, d_out);
    else if (orthogonal) hipLaunchKernelGGL((k_tdi_td_lds<7, true>), grid, block, lds, st, d_links, n_ext, n_out, dl, d_out);
    else hipLaunchKernelGGL((k_tdi_td_lds<7, false>), grid, block, lds, st, d_links, n_ext, n_out, dl, d_out);
  } else {
    const dim3 grid((n_out + 255) / 256), block(256);
    if (maxd == 3 && orthogonal) hipLaunchKernelGGL((k_tdi_td<3, true>), grid, block, 0, st, d_links, n_ext, n_out, dl, d_out);
    else if (maxd == 3) hipLaunchKernelGGL((k_tdi_td<3, false>), grid, block, 0, st, d_links, n_ext, n_out, dl, d_out);
    else if (orthogonal) hipLaunchKernelGGL((k_tdi_td<7, true>), grid, block, 0, st, d_links, n_ext, n_out, dl, d_out);
    else hipLaunchKernelGGL((k_tdi_td<7, false>), grid, block, 0, st, d_links, n_ext, n_out, dl, d_out);
  }
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_context_reserve(pes_context* ctx, int P, int grad) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (!ctx->finalized) return fail(PES_ERR_STATE, "context not finalized");
  if (P < 1) return fail(PES_ERR_INVALID, "P must be >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  return ensure_scratch(ctx, P, grad != 0);
}

extern "C" int pes_context_pin_scratch(pes_context* ctx, int pinned) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  ctx->pinned = pinned != 0;
  return PES_OK;
}

extern "C" int pes_whittle_lnl(const double* d_data, const double* d_resp, const double* d_psd, int n_chan, int n,
                               double df, double* d_work, double* d_out, void* stream) {
  if (n_chan < 1 || n < 1 || !d_data || !d_resp || !d_psd || !d_work || !d_out)
    return fail(PES_ERR_INVALID, "whittle: arrays must be non-NULL, n_chan >= 1, n >= 1");
  int nb = (n + TILE - 1) / TILE;
  if (nb > WHITTLE_BLOCKS) nb = WHITTLE_BLOCKS;
  hipLaunchKernelGGL(k_whittle_partial, dim3(nb), dim3(TILE), 0, (hipStream_t)stream, d_data, d_resp, d_psd, n_chan,
                     n, d_work);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_whittle_final, dim3(1), dim3(64), 0, (hipStream_t)stream, d_work, nb, df, d_out);
  HIP_TRY(hipGetLastError());
  return PES_OK;
}

extern "C" int pes_context_set_schedule(pes_context* ctx, int tiles_per_group, int blocks_per_cu) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (tiles_per_group < 1 || tiles_per_group > 4096 || blocks_per_cu < 1 || blocks_per_cu > 64)
    return fail(PES_ERR_INVALID, "schedule: tiles_per_group in [1, 4096], blocks_per_cu in [1, 64]");
  ctx->tiles_per_group = tiles_per_group;
  ctx->blocks_per_cu = blocks_per_cu;
  return PES_OK;
}

extern "C" int pes_context_set_rows_per_item(pes_context* ctx, int rows) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  if (rows < 0 || rows > chunk_rows<false>())
    return fail(PES_ERR_INVALID, "rows per work item: 0 (automatic) or 1 .. 160");
  ctx->rows_override = rows;
  return PES_OK;
}

extern "C" int pes_context_get_rows_per_item(const pes_context* ctx, int* rows) {
  if (!ctx || !rows) return fail(PES_ERR_INVALID, "NULL argument");
  *rows = ctx->rows_override;
  return PES_OK;
}

extern "C" int pes_context_get_schedule(const pes_context* ctx, int* tiles_per_group, int* blocks_per_cu) {
  if (!ctx || !tiles_per_group || !blocks_per_cu) return fail(PES_ERR_INVALID, "NULL argument");
  *tiles_per_group = ctx->tiles_per_group;
  *blocks_per_cu = ctx->blocks_per_cu;
  return PES_OK;
}

extern "C" int pes_context_set_profiling(pes_context* ctx, int enable) {
  if (!ctx) return fail(PES_ERR_INVALID, "ctx is NULL");
  ctx->profiling = enable != 0;
  return PES_OK;
}

extern "C" int pes_profile_read(pes_context* ctx, int* n_launches, double* total_ms) {
  if (!ctx || !n_launches || !total_ms) return fail(PES_ERR_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(ctx->device));
  double tot = 0.0;
  for (size_t i = 0; i < ctx->ev_start.size(); ++i) {
    HIP_TRY(hipEventSynchronize(ctx->ev_stop[i]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_start[i], ctx->ev_stop[i]));
    tot += ms;
    (void)hipEventDestroy(ctx->ev_start[i]);
    (void)hipEventDestroy(ctx->ev_stop[i]);
  }
  *n_launches = (int)ctx->ev_start.size();
  *total_ms = tot;
  ctx->ev_start.clear();
  ctx->ev_stop.clear();
  return PES_OK;
}

extern "C" int pes_profile_work(pes_context* ctx, double* bin_mode_rows, double* bin_rows) {
  if (!ctx || !bin_mode_rows || !bin_rows) return fail(PES_ERR_INVALID, "NULL argument");
  *bin_mode_rows = 0.0;
  *bin_rows = 0.0;
  const int P = ctx->live_P;
  if (P == 0) return PES_OK;
  if (ctx->live_seq != ctx->scratch_seq)
    return fail(PES_ERR_STATE,
                "pes_profile_work: a later evaluation has reused the scratch of the profiled batch (count its "
                "work before the next batch or output call on the context)");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  if (P > ctx->cap_live) {
    free_dev(ctx->d_live);
    ctx->d_live = nullptr;
    ctx->cap_live = 0;
    HIP_TRY(hipMalloc(&ctx->d_live, (size_t)P * 2 * sizeof(double)));
    ctx->cap_live = P;
  }
  const BatchScratch sc = context_scratch(ctx);
  const int n_det = (int)ctx->hdets.size();
  hipLaunchKernelGGL(k_live, dim3((P + 63) / 64), dim3(64), 0, nullptr, ctx->d_dets, n_det, ctx->d_bin_f, sc.pconst,
                     ctx->modes, P, ctx->d_live);
  HIP_TRY(hipGetLastError());
  for (int d = 0; d < n_det; ++d) {
    if (!ctx->hdets[d].d.duty) continue;
    const dim3 g_off((ctx->hdets[d].d.n_bins + 255) / 256, std::min(P, 65535));
    hipLaunchKernelGGL(k_live_off, g_off, dim3(256), 0, nullptr, ctx->d_dets, d, ctx->d_bin_f, ctx->d_bin_j,
                       ctx->d_bin_s, sc.pconst, sc.nodes, ctx->n_nodes + 1, ctx->du, ctx->modes, P, ctx->d_live);
    HIP_TRY(hipGetLastError());
  }
  std::vector<double> h((size_t)P * 2);
  HIP_TRY(hipMemcpy(h.data(), ctx->d_live, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int p = 0; p < P; ++p) {
    *bin_mode_rows += h[2 * (size_t)p];
    *bin_rows += h[2 * (size_t)p + 1];
  }
  return PES_OK;
}
#endif  // PES_LNL_FWD_UNIT

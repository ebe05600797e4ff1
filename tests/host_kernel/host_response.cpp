// Host build of the fused kernel's per-(bin, row) device functions (bin_prepare, load_row,
// tdi_consts, response_row, mode_adjoint) for the sanitizer check of tests/test_host_kernel.py.
// Reads one detector, the rows' constant blocks and node tables from a binary file, writes the
// response A/E/T (or X/Y/Z) of every row and bin, and exercises the adjoint of every mode with
// unit link adjoints (its values are not compared; ASan / UBSan watch every access and operation).
#include <hip/hip_runtime.h>  // the host stand-in (tests/host_kernel/stub)

#include <cstdio>
#include <vector>

#include "../../include/pespace_amd.h"
#include "response.hpp"
#include "mathfn.hpp"

using namespace pes;

namespace {
#include "kernel_common.hpp"
#include "hotpath.hpp"

template <class T>
bool rd(std::FILE* fp, T* p, size_t n) { return std::fread(p, sizeof(T), n, fp) == n; }

int run(int argc, char** argv) {
  if (argc != 3) return 2;
  std::FILE* fp = std::fopen(argv[1], "rb");
  if (!fp) return 3;
  int hdr[13];  // kind, tdi_gen, orthogonal, n_bins, n_modes, NNS, P, then the 6 mode ids
  double geo[7];  // arm [m], rot0, rev0, omega_rot, lambda_ref, beta_ref, du
  if (!rd(fp, hdr, 13) || !rd(fp, geo, 7)) return 4;
  const int kind = hdr[0], n = hdr[3], nm = hdr[4], NNS = hdr[5], P = hdr[6];
  ModeList modes{};
  modes.n = nm;
  for (int i = 0; i < nm; ++i) modes.id[i] = hdr[7 + i];
  DetectorDev det;
  std::memset(&det, 0, sizeof det);
  orbit_constants(det, kind, geo[0], geo[1], geo[2], geo[3], geo[4], geo[5]);
  det.tdi_gen = hdr[1];
  det.orthogonal = hdr[2];
  det.n_bins = n;
  const double du = geo[6];
  std::vector<double> f(n), s(n), pc((size_t)P * PC_STRIDE), nodes((size_t)P * nm * NNS * 4);
  std::vector<int> j(n);
  if (!rd(fp, f.data(), n) || !rd(fp, s.data(), n) || !rd(fp, j.data(), n) || !rd(fp, pc.data(), pc.size()) ||
      !rd(fp, nodes.data(), nodes.size()))
    return 5;
  std::fclose(fp);
  blockDim.x = 1;
  threadIdx.x = 0;
  trig_table_load();
  std::vector<double> out((size_t)P * 3 * n * 2);
  std::vector<double> u_col((size_t)2 * MAX_MODES * TILE);  // the gradient kernel's per-lane phasor column
  double adj_check = 0.0;
  for (int p = 0; p < P; ++p) {
    const double* pcp = pc.data() + (size_t)p * PC_STRIDE;
    for (int k = 0; k < n; ++k) {
      BinPre b;
      bin_prepare(det, f[k], s[k], j[k], du, b);
      RowC rc = load_row(pcp, b);
      row_geo(det, rc);
      TdiC tdc;
      tdi_consts(det, b, tdc);
      Cx o[3];
      const double* nb = nodes.data() + ((size_t)p * nm * NNS + b.j) * 4;
      // the forward pass stores each mode's unit phasor in the lane's column, as in the gradient
      // kernel, and the adjoint below reads it back
      if (kind == 0)
        response_row<0, 0, 1>(det, pcp, modes, nb, (size_t)NNS * 4, b, rc, tdc, o, u_col.data());
      else
        response_row<0, 1, 1>(det, pcp, modes, nb, (size_t)NNS * 4, b, rc, tdc, o, u_col.data());
      for (int c = 0; c < 3; ++c) {
        out[(((size_t)p * 3 + c) * n + k) * 2] = o[c].re;
        out[(((size_t)p * 3 + c) * n + k) * 2 + 1] = o[c].im;
      }
      // the adjoint of every mode, both precision modes, with unit link adjoints and zero node gradients
      Cx wl[6];
      for (int i = 0; i < 6; ++i) wl[i] = {1.0, 0.5};
      const double g0[32] = {};
      for (int mi = 0; mi < nm; ++mi) {
        const double* mf = pcp + PC_MODE + mi * PC_MODE_STRIDE;
        if (b.f > mf[MF_CUT]) continue;
        AdjAcc a0, a1;
        std::memset(&a0, 0, sizeof a0);
        std::memset(&a1, 0, sizeof a1);
        double* us = u_col.data() + 2 * ((size_t)mi * TILE + threadIdx.x);
        if (kind == 0) {
          mode_adjoint<0, 0>(det, mf, modes.id[mi], nb + (size_t)mi * NNS * 4, g0, b, rc, wl, pcp, a0, us);
          mode_adjoint<1, 0>(det, mf, modes.id[mi], nb + (size_t)mi * NNS * 4, g0, b, rc, wl, pcp, a1, us);
        } else {
          mode_adjoint<0, 1>(det, mf, modes.id[mi], nb + (size_t)mi * NNS * 4, g0, b, rc, wl, pcp, a0, us);
          mode_adjoint<1, 1>(det, mf, modes.id[mi], nb + (size_t)mi * NNS * 4, g0, b, rc, wl, pcp, a1, us);
        }
        adj_check += a0.tc + a1.tc + a0.iota + a0.mphi;
      }
    }
  }
  std::FILE* fo = std::fopen(argv[2], "wb");
  if (!fo) return 6;
  std::fwrite(out.data(), sizeof(double), out.size(), fo);
  std::fclose(fo);
  std::printf("adjoint checksum %.17g\n", adj_check);
  return std::isfinite(adj_check) ? 0 : 7;
}
}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }

// Host build of interval_off and interval_on (the bounds by which the duty cycle's off periods
// and plateaus are found per tile: k_duty_off, k_duty_items) against brute force, for
// tests/test_host_kernel.py.
//
// For random node intervals (Psi, dPsi/du at both ends, chosen so that t_f + t_c starts and ends
// near the window's edges and ramp ends, with and without an interior extremum) and several
// windows, every interval that interval_off (interval_on) marks is sampled at 4001 points s in
// [0, 1] with the fused kernel's own t (bin_prepare's Hermite weights at f = f_j e^{s du},
// mode_time) and the kernel's window (duty_window): each sample must see w = 0 (w = 1 and
// w' = 0).  The same is done for weakened bounds (the endpoints only, no margin), which must be
// caught on some intervals: the sampling would see a bound that misses the interior extremum.
// Prints "marked <n> marked_on <n> weak_misses <n> weak_on_misses <n>" and exits 1 on a
// violation of a real bound.
#include <hip/hip_runtime.h>  // the host stand-in (tests/host_kernel/stub)

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../include/pespace_amd.h"
#include "response.hpp"
#include "mathfn.hpp"

using namespace pes;

namespace {
#include "kernel_common.hpp"
#include "hotpath.hpp"

// the weakened bound: t at the two ends only, no margin
bool interval_off_weak(const DetectorDev& det, const double* n, double fa, double fb, double du, double tc) {
  const double ta = tc + n[3] / (2.0 * PI * fa), tb = tc + n[7] / (2.0 * PI * fb);
  const double tl = std::fmin(ta, tb), th = std::fmax(ta, tb);
  const double start = det.duty_t0 + std::floor((tl - det.duty_t0) * det.duty_inv_period) * det.duty_period;
  return tl - start >= det.duty_on && th - start < det.duty_period;
}

bool interval_on_weak(const DetectorDev& det, const double* n, double fa, double fb, double du, double tc) {
  const double ta = tc + n[3] / (2.0 * PI * fa), tb = tc + n[7] / (2.0 * PI * fb);
  const double tl = std::fmin(ta, tb), th = std::fmax(ta, tb);
  const double start = det.duty_t0 + std::floor((tl - det.duty_t0) * det.duty_inv_period) * det.duty_period;
  return tl - start >= det.duty_taper && th - start <= det.duty_on - det.duty_taper;
}

// some sample of the interval sees a nonzero window (want_off) or a window other than a flat 1
bool any_bad(const DetectorDev& det, const double* n, double fa, double du, double tc, int j, bool want_off) {
  for (int k = 0; k <= 4000; ++k) {
    const double s = k / 4000.0;
    BinPre b;
    bin_prepare(det, fa * std::exp(s * du), s, j, du, b);
    double w, dw;
    duty_window(det, mode_time(n, b, tc), w, dw);
    if (want_off ? w != 0.0 : (w != 1.0 || dw != 0.0)) return true;
  }
  return false;
}
}  // namespace

int main() {
  DetectorDev det;
  std::memset(&det, 0, sizeof det);
  orbit_constants(det, 1, 1.7e8, 0.0, 0.0, 2.0 * PI / (3.65 * 86400.0), 2.1, -0.47);
  blockDim.x = 1;
  threadIdx.x = 0;
  trig_table_load();
  const double windows[3][4] = {{15778800.0, 7889400.0, 3944700.0, 86400.0},  // 3 months on / off, 1-day ramps
                                {1.0e8, 5.0e7, 523000.0, 1500.0},
                                {86400.0, 43200.0, 1000.0, 0.0}};               // no ramps
  std::mt19937_64 rng(20261203);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const double du = std::log(600.0) / 1024.0;
  long marked = 0, marked_on = 0, weak_misses = 0, weak_on_misses = 0;
  for (const auto& win : windows) {
    det.duty = 1;
    det.duty_period = win[0];
    det.duty_on = win[1];
    det.duty_t0 = win[2];
    det.duty_taper = win[3];
    det.duty_inv_period = 1.0 / win[0];
    det.duty_inv_taper = win[3] > 0.0 ? 1.0 / win[3] : 0.0;
    for (int it = 0; it < 20000; ++it) {
      const int j = (int)(U(rng) * 1023.0);
      const double fa = 1e-4 * std::exp(du * j), fb = fa * std::exp(du);
      const double tc = 3.0e7 * U(rng);
      // t - t_c at the two ends near an edge of the window or an end of a ramp
      const double edges[4] = {win[1], win[0], win[3], win[1] - win[3]};
      const double edge = edges[(int)(U(rng) * 4.0) & 3] + win[2] - tc;
      const double span = win[3] + 50.0;
      const double ta = edge + span * (2.0 * U(rng) - 1.0), tb = edge + span * (2.0 * U(rng) - 1.0);
      double n[8] = {0.0};
      n[3] = 2.0 * PI * fa * ta;  // dPsi/du = 2 pi f (t - t_c)
      n[7] = 2.0 * PI * fb * tb;
      n[2] = 1.0e5 * U(rng);
      // Psi1 - Psi0 near the trapezoid value, perturbed so that t may have an interior extremum
      n[6] = n[2] + 0.5 * du * (n[3] + n[7]) + du * 2.0 * PI * fa * span * (U(rng) - 0.5);
      if (interval_off(det, n, fa, fb, du, tc)) {
        ++marked;
        if (any_bad(det, n, fa, du, tc, j, true)) {
          std::printf("VIOLATION off window %g j %d tc %.17g ta %.17g tb %.17g\n", win[0], j, tc, ta, tb);
          return 1;
        }
      }
      if (interval_on(det, n, fa, fb, du, tc)) {
        ++marked_on;
        if (any_bad(det, n, fa, du, tc, j, false)) {
          std::printf("VIOLATION on window %g j %d tc %.17g ta %.17g tb %.17g\n", win[0], j, tc, ta, tb);
          return 1;
        }
      }
      if (interval_off_weak(det, n, fa, fb, du, tc) && any_bad(det, n, fa, du, tc, j, true)) ++weak_misses;
      if (interval_on_weak(det, n, fa, fb, du, tc) && any_bad(det, n, fa, du, tc, j, false)) ++weak_on_misses;
    }
  }
  std::printf("marked %ld marked_on %ld weak_misses %ld weak_on_misses %ld\n", marked, marked_on, weak_misses,
              weak_on_misses);
  return 0;
}

// Definitions shared by the kernels of pespace_amd.hip and hotpath.hpp: tile and chunk sizes,
// the per-row constant block, and the tile -> detector map.  Included inside the anonymous
// namespace of pespace_amd.hip.
#pragma once

constexpr int TILE = 256;      // bins per tile = threads per block
// parameter rows per work item of the gradient kernel (48: the most that leaves two blocks per
// CU the LDS for the staged node rows, the accumulators and the forward pass's unit phasors)
// and of the lnL-only one (160: 1.1% faster than 256)
template <bool GRAD>
constexpr int chunk_rows() { return GRAD ? 48 : 160; }
// default schedule: tiles per work item (8: the fastest in every timed schedule table of the
// one-year network, 4-7% ahead of 32 on 512-row samples) ...
constexpr int TILES_PER_GROUP = 8;
constexpr int BLOCKS_PER_CU = 4;     // ... and persistent blocks launched per CU
constexpr int MAX_MODES = 6;         // modes per context (ModeList)
// node rows per tile that the fused kernel stages in LDS (tiles spanning more read them from
// global memory): a 256-bin tile of a one-year grid spans at most 1.3e-3 Hz / f + 1 of the 1024
// log-spaced node intervals over 1e-4 - 0.06 Hz, under 16 above 1e-4 Hz
constexpr int NJ_STAGE = 16;
constexpr int BIN_REC = 10;          // doubles per bin of the fused kernel's weight and data record
constexpr int MAX_TILE_RANGES = 16;  // runs of detectors by duty flag per context
// fused-kernel launches per batch: one per range, and a second for a duty-cycled range (its work
// items on the window's plateau run in the plain instantiation), each with its own work counter;
// a third set counts a duty-cycled range's items that need the window (k_duty_items)
constexpr int N_WORK_COUNTERS = 3 * MAX_TILE_RANGES;

// ---- per-row constant block ------------------------------------------------------------
enum PConst {
  PC_TC = 0, PC_INVDL, PC_VALID, PC_DL,
  PC_P = 4, PC_Q = 7, PC_K = 10,
  PC_DLP = 13, PC_DLQ = 16, PC_DBP = 19, PC_DBQ = 22, PC_DLK = 25, PC_DBK = 28,
  PC_MODE = 32,  // per mode: Fp re,im, 2 Fc re,im, cut_hz, dFp/diota re,im, 2 dFc/diota re,im
  PC_MODE_STRIDE = 9,
  PC_FCUT = 32 + 6 * 9,  // largest mode cut [Hz]: the response is zero above it
  PC_OFFANY,             // 1: a tile of a duty-cycled detector lies in an off period for this row
  PC_STRIDE = 32 + 6 * 9 + 2
};
enum { MF_FPR = 0, MF_FPI, MF_FCR, MF_FCI, MF_CUT, MF_DFPR, MF_DFPI, MF_DFCR, MF_DFCI };

// detector owning tile `tile` (detectors are concatenated in tile order)
__device__ inline int tile_detector(const DetectorDev* dets, int n_det, int tile) {
  int acc = 0;
  for (int d = 0; d < n_det; ++d) {
    acc += dets[d].n_tiles;
    if (tile < acc) return d;
  }
  return n_det - 1;
}

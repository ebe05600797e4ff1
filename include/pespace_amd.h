/*
 * pespace_amd C-ABI: MI355X (gfx950) frequency-domain lnL(+grad) hot path.
 *
 * The reference (nnrui/pespace) is pure Python with taichi/JAX backends and has
 * no native FFI; each entry point below replaces one stage of its FD path and
 * is bound from Python with ctypes (pespace_amd/_lib.py).  All device buffers
 * passed to the evaluation entry points are caller-owned; the context owns the
 * constant detector data it uploads and its scratch buffers.  No entry point
 * synchronises the host inside an evaluation, so evaluations may be captured in
 * a HIP graph once the context's scratch has been sized by a first call.
 *
 * Errors: every function returns a status (PES_OK = 0); pes_last_error()
 * returns the message of the last failure on the calling thread.
 */
#ifndef PESPACE_AMD_H
#define PESPACE_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PES_OK 0
#define PES_ERR_INVALID 1
#define PES_ERR_HIP 2
#define PES_ERR_STATE 3

/* parameter row layout of every batch (reference likelihood namedtuple order,
 * pespace/inference/likelihood_jax.py:119-127 with waveform.required_params first) */
#define PES_N_PARAMS 11
#define PES_P_CHIRP_MASS 0          /* solar masses                  */
#define PES_P_MASS_RATIO 1          /* m2/m1 in (0, 1]               */
#define PES_P_CHI_1 2
#define PES_P_CHI_2 3
#define PES_P_LUMINOSITY_DISTANCE 4 /* Mpc                           */
#define PES_P_REFERENCE_PHASE 5
#define PES_P_INCLINATION 6
#define PES_P_ECLIPTIC_LONGITUDE 7
#define PES_P_ECLIPTIC_LATITUDE 8
#define PES_P_POLARIZATION 9
#define PES_P_COALESCENCE_TIME 10   /* seconds                       */

/* precision modes of pes_lnl_batch / pes_response.  lnL and the response are the FP64 forward in
 * both modes (identical results).  PES_PREC_MIXED changes the gradient only: the adjoint's
 * recompute of each mode takes the pair phasors and the total-phase phasor (phase reduced in
 * FP64) from the FP32 hardware sine/cosine and the sinc denominators from the FP32 reciprocal,
 * so the gradient carries FP32-level relative errors (~1e-7).  Any other value is rejected
 * with PES_ERR_INVALID. */
#define PES_PREC_FP64 0
#define PES_PREC_MIXED 1

typedef struct pes_context pes_context;

const char* pes_last_error(void);

/* Waveform model + interpolation grid.
 * Replaces the waveform object of interface_ripple.py:27-90 / tiwave
 * (modes: ids 0:(2,2) 1:(2,1) 2:(3,3) 3:(3,2) 4:(4,4) 5:(4,3)).
 * n_nodes > 1: amplitude and phase are evaluated on n_nodes log-spaced nodes in
 *              [f_node_min, f_node_max] and cubic-Hermite interpolated to the bins;
 * n_nodes == 0: direct evaluation at every bin (the reference's behaviour,
 *              interface_ripple.py:149-151).
 * f_ref: reference frequency [Hz]; <= 0 means the amplitude peak. */
int pes_context_create(pes_context** out, int device, int n_modes, const int* mode_ids, double f_ref,
                       int n_nodes, double f_node_min, double f_node_max);

/* One detector: orbit, TDI combination, grid, data and PSD (host arrays, copied).
 * Replaces InterferometerAntenna + TDIChannelData setup (antenna.py:128-202,
 * tdi.py:585-649, 771-837) and the static data of likelihood_jax.py:132-205.
 * orbit_kind: 0 heliocentric (KaplerianHeliocentric, orbit.py:155-241), 1 geocentric
 *   (KeplerianGeocentric, orbit.py:67-152).  arm_length_m, rotation_initial,
 *   revolution_initial: both kinds; omega_rotation [rad/s], lambda_ref, beta_ref [rad]: the
 *   geocentric rotation rate and reference source (orbit.py:77-85), ignored for kind 0.
 * tdi_generation: 15 or 20 (tdi.py:1614-1628); orthogonal: 1 A,E,T / 0 X,Y,Z.
 * f[n_bins]: bin frequencies, strictly increasing, any range (the transfer functions take
 * every pi f L).  weight[n_bins]: per-bin frequency weight df_k (may be NULL: then the scalar
 * df is used for every bin, tdi.py:130).
 * data_re/data_im/psd: [3][n_bins] channel-major; channel_mask[3]: which of the
 * three outputs enter lnL (likelihood_ti.py:61, interface_bilby.py:70-75). */
int pes_context_add_detector(pes_context* ctx, int orbit_kind, double arm_length_m, double rotation_initial,
                             double revolution_initial, double omega_rotation, double lambda_ref, double beta_ref,
                             int tdi_generation, int orthogonal, int n_bins, const double* f, const double* weight,
                             double df, const double* data_re, const double* data_im, const double* psd,
                             const int* channel_mask);

/* Observation duty cycle of detector det_index (in order of addition; before finalize): the
 * detector observes for on_duration seconds of every period, the windows starting at
 * start_time + n period, with C1 smoothstep ramps of taper seconds inside each window (taper 0:
 * a hard window; period 0: always observing, the default).  Each mode's response at a bin is
 * weighted by the window at its time t = t_f + t_c (the stationary-phase time of that
 * frequency), in the likelihood, its gradient and every response output.  The reference models
 * no duty cycle; this serves the benchmark network's TianQin (3 months on, 3 months off). */
int pes_context_set_duty_cycle(pes_context* ctx, int det_index, double period, double on_duration, double start_time,
                               double taper);

/* Upload detector data; must be called once after the last add_detector.  With
 * interpolation nodes (n_nodes > 1) every bin of every detector must lie inside
 * [f_node_min, f_node_max] (no extrapolation): otherwise PES_ERR_INVALID. */
int pes_context_finalize(pes_context* ctx);

void pes_context_destroy(pes_context* ctx);

/* Batched lnL (and gradient) for P parameter rows.  A context serves one stream at a time:
 * its scratch and work queue are shared by the calls made on it.
 * Replaces compute_whittle_likelihood_ti (likelihood_ti.py:17-69) summed over the
 * detectors (interface_bilby.py:97-116), vmapped over particles
 * (sampler/base.py:27-35).  d_params: device [P][11]; d_lnl: device [P];
 * d_grad: device [P][11] or NULL (no gradient).  Invalid rows (q outside (0,1],
 * |chi| >= 1, dL <= 0, non-finite) give lnL = -inf and a zero gradient.
 * stream: hipStream_t (NULL = default stream). */
int pes_lnl_batch(pes_context* ctx, const double* d_params, int P, double* d_lnl, double* d_grad, int precision,
                  void* stream);

/* pes_lnl_batch without any device allocation.  No counterpart in the reference: for callers
 * that cannot allocate (code under HIP-graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as pes_lnl_batch, but the per-batch scratch
 * (work counter, per-row setup, node tables and their gradients, per-tile-group partials) lives
 * in a fixed array of PES_NOALLOC_SCRATCH_BYTES that every GPU holds once for the whole process;
 * pes_lnl_batch_noalloc_bytes(ctx, P, grad) gives what a batch needs (0 for a bad argument) and
 * a batch that needs more is rejected with PES_ERR_INVALID before anything is launched.  Every
 * call on a GPU shares that GPU's array and leaves its contents behind (each call rewrites all it
 * reads): calls that overlap on one GPU race, so serialise them (one stream, or events).  The
 * context's profiling (pes_context_set_profiling) does not apply to this entry point.
 * Example: six modes, 1024 nodes, a 1-yr LISA grid and the gradient take ~1.2 MB per row. */
#define PES_NOALLOC_SCRATCH_BYTES (64u << 20)
int pes_lnl_batch_noalloc(pes_context* ctx, const double* d_params, int P, double* d_lnl, double* d_grad,
                          int precision, void* stream);
size_t pes_lnl_batch_noalloc_bytes(const pes_context* ctx, int P, int grad);

/* Detector response h_c(f) for P rows on detector det: device out [P][3][n_bins][2]
 * (re, im), computed by the same device code as pes_lnl_batch at the given
 * precision.  Replaces update_detector_response + tdi_response_numpy
 * (antenna.py:204-229, 305-313; tdi.py:1764-1793). */
int pes_response(pes_context* ctx, int det, const double* d_params, int P, double* d_out, int precision,
                 void* stream);

/* Waveform container on detector det's bins: device out [P][n_modes][n_bins][5] =
 * plus (re, im), cross (re, im), tf, with t_c = 0 (interface_ripple.py:42-51). */
int pes_waveform(pes_context* ctx, int det, const double* d_params, int P, double* d_out, void* stream);

/* Response of detector det to a caller-supplied waveform container (one parameter
 * set): device d_wf [n_modes][n_bins][5] = plus (re, im), cross (re, im), tf with
 * t_c = 0; device d_out [3][n_bins][2].  The drop-in for
 * _FDResponseModelMarset2018Taichi.update_single_link_response followed by
 * update_tdi_response (antenna.py:403-460, tdi.py:1764-1793). */
int pes_response_from_waveform(pes_context* ctx, int det, int n_modes, const double* d_wf, double lam, double beta,
                               double psi, double tc, double* d_out, void* stream);

/* The six single links of P rows on detector det, summed over the modes: device out
 * [P][6][n_bins][2] (re, im) in the order link12, link21, link23, link32, link31, link13,
 * from the same device code as pes_lnl_batch (FP64).  Replaces the single-link stage
 * _FDResponseModelMarset2018Taichi.update_single_link_response with a hip waveform
 * (antenna.py:403-460, 496-543; the links field of antenna.py:399-401). */
int pes_links(pes_context* ctx, int det, const double* d_params, int P, double* d_out, void* stream);

/* The six single links of detector det for a caller-supplied waveform container (one
 * parameter set): d_wf as for pes_response_from_waveform, device out [6][n_bins][2].
 * Replaces update_single_link_response (antenna.py:403-460) and the JAX twin's
 * get_response_jit (_antenna_jax.py:82-195). */
int pes_links_from_waveform(pes_context* ctx, int det, int n_modes, const double* d_wf, double lam, double beta,
                            double psi, double tc, double* d_out, void* stream);

/* TDI combination of detector det from single links: device links [P][6][n_bins][2] ->
 * device out [P][3][n_bins][2] (A, E, T = (X+Y+Z)/sqrt3 or X, Y, Z, the reference's own
 * formulas).  Replaces update_tdi_response (tdi.py:1705-1793) and the JAX twin's
 * get_combination_jit (_tdi_jax.py:70-102). */
int pes_tdi_from_links(pes_context* ctx, int det, int P, const double* d_links, double* d_out, void* stream);

/* Constellation vectors at n times (device d_t[n]): device out [n][6][3] = n1, n2, n3 (unit
 * link vectors) and x1, x2, x3 (node positions, SSB frame, seconds) of an orbit given by the
 * same parameters as pes_context_add_detector.  Replaces get_constellation_vectors
 * (orbit.py:260-322 geocentric, 363-419 heliocentric; _orbit_jax.py). */
int pes_orbit_vectors(int orbit_kind, double arm_length_m, double rotation_initial, double revolution_initial,
                      double omega_rotation, double lambda_ref, double beta_ref, const double* d_t, int n,
                      double* d_out, void* stream);

/* Time domain.  pes_td_links: the six single links of a detector (orbit as in
 * pes_context_add_detector) at the n times t0 + i dt for a waveform h+, hx sampled at
 * wave_t0 + m wave_dt (device d_hp, d_hc [wave_n], linearly interpolated, zero outside), as
 * y_rs(t) = n.[H(t - L - k.x_s) - H(t - k.x_r)].n / (2 (1 - k.n_sr)) with the constellation at t:
 * device out [6][n] (link12, link21, link23, link32, link31, link13).  Replaces the reference's
 * time-domain response model (TDResponseModelCornish2003, antenna.py:566-800, commented out there).
 * pes_td_tdi: Michelson TDI 1.5 / 2.0 of links sampled at spacing dt (device links [6][n_ext]):
 * output sample i is the combination at link index i + added, added =
 * pes_td_tdi_added_samples(...) = floor(max_delay L / dt) + 1 (max_delay 3 or 7), each delay
 * D^m y = y(t - m L) linearly interpolated; device out [3][n_ext - added] (A, E, T or X, Y, Z).
 * Replaces TDMichelsonConstantEqualArm.update_tdi_response (tdi.py:1388-1600). */
int pes_td_links(int orbit_kind, double arm_length_m, double rotation_initial, double revolution_initial,
                 double omega_rotation, double lambda_ref, double beta_ref, const double* d_hp, const double* d_hc,
                 double wave_t0, double wave_dt, int wave_n, double lam, double beta, double psi, double t0, double dt,
                 int n, double* d_out, void* stream);
int pes_td_tdi_added_samples(double arm_length_m, double dt, int tdi_generation);
int pes_td_tdi(double arm_length_m, double dt, int tdi_generation, int orthogonal, const double* d_links, int n_ext,
               double* d_out, void* stream);

/* Scratch for HIP-graph capture.  pes_context_reserve sizes the context's scratch for batches
 * of up to P rows (with the gradient tables if grad != 0) under the current schedule.  While
 * pes_context_pin_scratch(ctx, 1) holds, the scratch is never reallocated: a call that would
 * need more (a larger batch, a first gradient call, a schedule with more tile groups) fails
 * with PES_ERR_STATE instead of freeing buffers that a captured graph still uses; unpin with
 * pes_context_pin_scratch(ctx, 0).  No counterpart in the reference (it has no graphs). */
int pes_context_reserve(pes_context* ctx, int P, int grad);
int pes_context_pin_scratch(pes_context* ctx, int pinned);

/* Standalone Whittle lnL = -2 df sum_c sum_k |d - h|^2 / S on device arrays
 * (d_data, d_resp: [n_chan][n][2]; d_psd: [n_chan][n]); the result is written to
 * device d_out[0].  d_work: caller-owned device workspace of at least
 * PES_WHITTLE_WORK doubles.  Replaces compute_whittle_likelihood_ti
 * (likelihood_ti.py:17-69) and noise_weighted_inner_product (likelihood_jax.py:21-42),
 * with a fixed-order (deterministic) reduction instead of atomic_add. */
#define PES_WHITTLE_WORK 1024
int pes_whittle_lnl(const double* d_data, const double* d_resp, const double* d_psd, int n_chan, int n, double df,
                    double* d_work, double* d_out, void* stream);

/* Schedule of the fused kernel: work items of tiles_per_group 256-bin tiles x 48 rows (160 for
 * lnL without gradient), taken from a queue by blocks_per_cu x (number of CUs) persistent
 * blocks of 256 threads.  Defaults
 * 8 and 4; pespace_amd.core.LikelihoodCore.autotune picks them by timing.  Changing the
 * schedule, a larger batch or a first gradient call may reallocate scratch at the next call:
 * reserve and pin the scratch (below) before capturing a HIP graph.
 * No counterpart in the reference. */
int pes_context_set_schedule(pes_context* ctx, int tiles_per_group, int blocks_per_cu);
int pes_context_get_schedule(const pes_context* ctx, int* tiles_per_group, int* blocks_per_cu);

/* Parameter rows per work item of the fused kernel (the rows a block evaluates over one tile
 * group): 0 (the default) picks them per batch (the kernel's chunk, 48 with the gradient and 160
 * without, fewer when the grid has too few tile groups); 1..160 fixes them (capped at the
 * kernel's chunk).  A tunable of the project's autotuner, beside the tile-group size and the
 * blocks per CU; no counterpart in the reference. */
int pes_context_set_rows_per_item(pes_context* ctx, int rows);
int pes_context_get_rows_per_item(const pes_context* ctx, int* rows);

/* Timing of the fused lnL kernel: when enabled, pes_lnl_batch records a HIP event
 * pair around the fused kernel on its stream.  pes_profile_read waits for the
 * recorded pairs, returns their count and summed duration (ms) and clears them.
 * No counterpart in the reference (it has no per-stage timing). */
int pes_context_set_profiling(pes_context* ctx, int enable);
int pes_profile_read(pes_context* ctx, int* n_launches, double* total_ms);

/* Work of the last profiled pes_lnl_batch: the number of (bin, mode, row) triples with the bin
 * at or below the mode's cut and, for a duty-cycled detector, a nonzero observation window at
 * the mode's t_f + t_c (where the fused kernel needs to evaluate a mode; it skips the tiles of
 * 256 bins where the window is zero for every mode), and of (bin, row) pairs at or below the row's
 * largest cut.  Counted when called, from the context's scratch, which must still hold that
 * batch: PES_ERR_STATE once another batch or output call has set up rows through the context
 * (a HIP graph replay also rewrites the scratch without being seen here).  Synchronises the
 * device.  No counterpart in the reference. */
int pes_profile_work(pes_context* ctx, double* bin_mode_rows, double* bin_rows);

/* Number of bins of detector det (after finalize). */
int pes_detector_bins(const pes_context* ctx, int det);

#ifdef __cplusplus
}
#endif
#endif

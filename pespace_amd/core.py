"""Thin host object over one C-ABI context: detectors in, batched lnL(+grad) out.

Device memory and streams come from PyTorch (ROCm); the compute is the HIP
library.  Parameter rows follow ``PARAM_NAMES``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import PespaceError, c_int_p, check, dptr, lib

PARAM_NAMES = ("chirp_mass", "mass_ratio", "chi_1", "chi_2", "luminosity_distance", "reference_phase",
               "inclination", "ecliptic_longitude", "ecliptic_latitude", "polarization", "coalescence_time")
N_PARAMS = len(PARAM_NAMES)
MODE_IDS = {"22": 0, "21": 1, "33": 2, "32": 3, "44": 4, "43": 5}
PRECISION = {"fp64": 0, "mixed": 1}
# geocentric orbit defaults of the reference (TianQin: orbit.py:82-84)
GEO_OMEGA_ROTATION = 2 * np.pi / (3.65 * 86400.0)
GEO_LAMBDA_REF = 120.5 / 180 * np.pi
GEO_BETA_REF = -4.7 / 180 * np.pi


def _mode_key(m):
    if isinstance(m, str):
        return m
    return f"{m[0]}{m[1]}"


class LikelihoodCore:
    """One waveform model + interpolation grid + detectors, on one GPU."""

    def __init__(self, modes=("22",), reference_frequency=0.0, n_nodes=1024, f_node_min=None, f_node_max=None,
                 device=0):
        if not torch.cuda.is_available():
            raise PespaceError("pespace_amd's hip backend needs a GPU (torch.cuda.is_available() is False)")
        self.modes = tuple(_mode_key(m) for m in modes)
        for m in self.modes:
            if m not in MODE_IDS:
                raise ValueError(f"unsupported mode {m}; available {tuple(MODE_IDS)}")
        self.reference_frequency = float(reference_frequency)
        self.n_nodes = int(n_nodes)
        self.f_node_min = f_node_min
        self.f_node_max = f_node_max
        self.device = int(device)
        self._dets = []
        self._ctx = None
        self._n_bins = []

    # -- detectors ---------------------------------------------------------------------
    def add_detector(self, orbit_kind, arm_length, rotation_initial, revolution_initial, tdi_generation, orthogonal,
                     frequencies, data, psd, channel_mask, df=None, weight=None, omega_rotation=GEO_OMEGA_ROTATION,
                     lambda_ref=GEO_LAMBDA_REF, beta_ref=GEO_BETA_REF, duty_cycle=None):
        """data: complex (3, N); psd: (3, N) (unused channels may hold ones).  omega_rotation,
        lambda_ref, beta_ref: rotation rate and reference source of a geocentric orbit
        (orbit.py:77-85; TianQin's by default), unused for heliocentric ones.  duty_cycle:
        None (always observing) or (period, on_duration, start_time, taper) in seconds: each
        mode's response is weighted by the observation window at its time t_f + t_c (no
        counterpart in the reference)."""
        if self._ctx is not None:
            raise PespaceError("detectors must be added before the first evaluation")
        if duty_cycle is not None:
            duty_cycle = tuple(float(x) for x in duty_cycle)
            if len(duty_cycle) != 4:
                raise ValueError("duty_cycle is (period, on_duration, start_time, taper)")
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        data = np.asarray(data, dtype=np.complex128)
        psd = np.ascontiguousarray(psd, dtype=np.float64)
        if data.shape != (3, f.size) or psd.shape != (3, f.size):
            raise ValueError("data and psd must have shape (3, n_bins)")
        self._dets.append(dict(kind=int(orbit_kind), arm=float(arm_length), rot0=float(rotation_initial),
                               rev0=float(revolution_initial), omega=float(omega_rotation), lam=float(lambda_ref),
                               bet=float(beta_ref), gen=int(tdi_generation), orth=int(bool(orthogonal)),
                               f=f, dre=np.ascontiguousarray(data.real), dim=np.ascontiguousarray(data.imag),
                               psd=psd, mask=np.asarray(channel_mask, dtype=np.int32),
                               df=0.0 if df is None else float(df),
                               weight=None if weight is None else np.ascontiguousarray(weight, dtype=np.float64),
                               duty=duty_cycle))
        return len(self._dets) - 1

    def _build(self):
        if self._ctx is not None:
            return
        L = lib()
        if not self._dets:
            raise PespaceError("no detector added")
        fmin = min(float(d["f"].min()) for d in self._dets)
        fmax = max(float(d["f"].max()) for d in self._dets)
        f_lo = self.f_node_min if self.f_node_min is not None else fmin
        f_hi = self.f_node_max if self.f_node_max is not None else fmax
        ids = (ctypes.c_int * len(self.modes))(*[MODE_IDS[m] for m in self.modes])
        h = ctypes.c_void_p()
        check(L.pes_context_create(ctypes.byref(h), self.device, len(self.modes), ids, self.reference_frequency,
                                   self.n_nodes, f_lo, f_hi))
        try:
            for i, d in enumerate(self._dets):
                mask = np.ascontiguousarray(d["mask"], dtype=np.int32)
                w = d["weight"]
                check(L.pes_context_add_detector(
                    h, d["kind"], d["arm"], d["rot0"], d["rev0"], d["omega"], d["lam"], d["bet"], d["gen"],
                    d["orth"], d["f"].size, dptr(d["f"]),
                    dptr(w) if w is not None else None, d["df"], dptr(d["dre"]), dptr(d["dim"]), dptr(d["psd"]),
                    mask.ctypes.data_as(c_int_p)))
                if d["duty"] is not None:
                    check(L.pes_context_set_duty_cycle(h, i, *d["duty"]))
            check(L.pes_context_finalize(h))
        except Exception:
            L.pes_context_destroy(h)
            raise
        self._ctx = h
        self._n_bins = [L.pes_detector_bins(h, i) for i in range(len(self._dets))]

    def close(self):
        if self._ctx is not None:
            lib().pes_context_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- evaluation ----------------------------------------------------------------------
    def _params(self, params):
        p = torch.as_tensor(params, dtype=torch.float64)
        if p.dim() == 1:
            p = p.unsqueeze(0)
        if p.dim() != 2 or p.shape[1] != N_PARAMS:
            raise ValueError(f"params must have shape (P, {N_PARAMS})")
        return p.to(device=f"cuda:{self.device}").contiguous()

    @staticmethod
    def _check_out(t, name, shape, dev):
        """A caller-supplied output buffer must be a contiguous float64 tensor of exactly ``shape``
        on the evaluation device: the kernels write through its raw pointer."""
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if t.dtype != torch.float64:
            raise ValueError(f"{name} must be float64, got {t.dtype}")
        if t.device != dev:
            raise ValueError(f"{name} must be on {dev}, got {t.device}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def lnl(self, params, grad=False, precision="fp64", out=None, grad_out=None):
        if precision not in PRECISION:
            raise ValueError(f"precision must be one of {tuple(PRECISION)}, got {precision!r}")
        self._build()
        p = self._params(params)
        P = p.shape[0]
        dev = p.device
        if out is not None:
            self._check_out(out, "out", (P,), dev)
        lnl = out if out is not None else torch.empty(P, dtype=torch.float64, device=dev)
        g = None
        if grad:
            if grad_out is not None:
                self._check_out(grad_out, "grad_out", (P, N_PARAMS), dev)
            g = grad_out if grad_out is not None else torch.empty(P, N_PARAMS, dtype=torch.float64, device=dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(lib().pes_lnl_batch(self._ctx, p.data_ptr(), P, lnl.data_ptr(), g.data_ptr() if g is not None else None,
                                  PRECISION[precision], stream))
        return (lnl, g) if grad else lnl

    # -- scratch for HIP-graph capture ---------------------------------------------------------
    def reserve(self, P, grad=True):
        """Size the context's scratch for batches of up to ``P`` rows (with the gradient tables if
        ``grad``) under the current schedule."""
        self._build()
        check(lib().pes_context_reserve(self._ctx, int(P), int(bool(grad))))

    def pin_scratch(self, pinned=True):
        """While pinned the scratch is never reallocated: a call that would need more (a larger
        batch, the first gradient call, a schedule with more tile groups) raises instead of
        freeing buffers that a captured HIP graph still uses."""
        self._build()
        check(lib().pes_context_pin_scratch(self._ctx, int(bool(pinned))))

    # -- schedule of the fused kernel ------------------------------------------------------
    def set_schedule(self, tiles_per_group, blocks_per_cu, rows_per_item=None):
        """Work items of ``tiles_per_group`` 256-bin tiles x ``rows_per_item`` parameter rows
        (0: chosen per batch, 48 with the gradient and 160 without unless the grid is small;
        None: unchanged), ``blocks_per_cu`` persistent blocks per CU."""
        self._build()
        check(lib().pes_context_set_schedule(self._ctx, int(tiles_per_group), int(blocks_per_cu)))
        if rows_per_item is not None:
            check(lib().pes_context_set_rows_per_item(self._ctx, int(rows_per_item)))

    def schedule(self):
        """(tiles_per_group, blocks_per_cu)."""
        self._build()
        t, b = ctypes.c_int(), ctypes.c_int()
        check(lib().pes_context_get_schedule(self._ctx, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def rows_per_item(self):
        """The fixed rows per work item, or 0 when they are chosen per batch."""
        self._build()
        r = ctypes.c_int()
        check(lib().pes_context_get_rows_per_item(self._ctx, ctypes.byref(r)))
        return r.value

    def autotune(self, params, grad=True, precision="fp64", candidates=None, reps=2, warmup=1):
        """Time the fused kernel on ``params`` for each candidate (tiles_per_group, blocks_per_cu)
        or (tiles_per_group, blocks_per_cu, rows_per_item) with HIP events, keep the fastest,
        and return it with every candidate's mean time (ms).

        The default candidates cover work items of 8-64 tiles and 2-8 persistent blocks per CU
        (the gradient kernel holds two blocks of four waves per CU, the lnL-only kernel three;
        blocks beyond those wait and find the work queue drained), with the rows per item chosen
        per batch.
        """
        self._build()
        p = self._params(params)
        P = p.shape[0]
        cands = candidates or [(t, b) for t in (8, 16, 32, 64) for b in (2, 3, 4, 8)]
        lnl = torch.empty(P, dtype=torch.float64, device=p.device)
        g = torch.empty(P, N_PARAMS, dtype=torch.float64, device=p.device) if grad else None
        times = {}
        for cand in cands:
            t, b = cand[0], cand[1]
            self.set_schedule(t, b, cand[2] if len(cand) > 2 else 0)
            for _ in range(warmup):
                self.lnl(p, grad=grad, precision=precision, out=lnl, grad_out=g)
            self.set_profiling(True)
            for _ in range(reps):
                self.lnl(p, grad=grad, precision=precision, out=lnl, grad_out=g)
            n, ms = self.read_profile()
            self.set_profiling(False)
            times[tuple(cand)] = ms / max(n, 1)
        best = min(times, key=times.get)
        self.set_schedule(best[0], best[1], best[2] if len(best) > 2 else 0)
        return best, times

    def set_profiling(self, enable=True):
        self._build()
        check(lib().pes_context_set_profiling(self._ctx, int(bool(enable))))

    def read_profile(self):
        """(launches, total_ms) of the fused kernel since the last read."""
        n = ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().pes_profile_read(self._ctx, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def profile_work(self):
        """(bin_mode_rows, bin_rows) of the last profiled evaluation: the triples on which a mode
        needs evaluating (bin at or below the mode's cut and, for a duty-cycled detector, inside
        an observation window) and the (bin, row) pairs at or below the row's largest mode cut.
        Counted from the context's scratch: call it before the next evaluation on this core."""
        bm = ctypes.c_double()
        br = ctypes.c_double()
        check(lib().pes_profile_work(self._ctx, ctypes.byref(bm), ctypes.byref(br)))
        return bm.value, br.value

    @property
    def n_bins(self):
        self._build()
        return list(self._n_bins)

    def response(self, det, params, precision="fp64"):
        """Complex (P, 3, N) TDI response of detector det."""
        self._build()
        p = self._params(params)
        n = self._n_bins[det]
        out = torch.empty(p.shape[0], 3, n, 2, dtype=torch.float64, device=p.device)
        check(lib().pes_response(self._ctx, det, p.data_ptr(), p.shape[0], out.data_ptr(), PRECISION[precision],
                                 torch.cuda.current_stream(p.device).cuda_stream))
        return torch.view_as_complex(out)

    def links(self, det, params):
        """Complex (P, 6, N) single links (12, 21, 23, 32, 31, 13) of detector det, summed over the
        modes (antenna.py:496-543), from the same device code as the fused kernel."""
        self._build()
        p = self._params(params)
        n = self._n_bins[det]
        out = torch.empty(p.shape[0], 6, n, 2, dtype=torch.float64, device=p.device)
        check(lib().pes_links(self._ctx, det, p.data_ptr(), p.shape[0], out.data_ptr(),
                              torch.cuda.current_stream(p.device).cuda_stream))
        return torch.view_as_complex(out)

    def tdi_from_links(self, det, links):
        """TDI channels (P, 3, N) (A, E, T or X, Y, Z) of detector det from single links (P, 6, N)
        or (6, N) (tdi.py:1705-1793)."""
        self._build()
        n = self._n_bins[det]
        dev = torch.device(f"cuda:{self.device}")
        lk = torch.as_tensor(links, device=dev).to(torch.complex128)
        squeeze = lk.dim() == 2
        if squeeze:
            lk = lk[None]
        if lk.dim() != 3 or lk.shape[1] != 6 or lk.shape[2] != n:
            raise ValueError(f"links must have shape (P, 6, {n}) or (6, {n})")
        lr = torch.view_as_real(lk.contiguous()).contiguous()
        out = torch.empty(lk.shape[0], 3, n, 2, dtype=torch.float64, device=dev)
        check(lib().pes_tdi_from_links(self._ctx, det, lk.shape[0], lr.data_ptr(), out.data_ptr(),
                                       torch.cuda.current_stream(dev).cuda_stream))
        out = torch.view_as_complex(out)
        return out[0] if squeeze else out

    def _container(self, det, waveform):
        n = self._n_bins[det]
        modes = [waveform] if {"plus", "cross", "tf"} <= set(waveform.keys()) else list(waveform.values())
        dev = torch.device(f"cuda:{self.device}")
        wf = torch.empty(len(modes), n, 5, dtype=torch.float64, device=dev)
        for i, m in enumerate(modes):
            hp = torch.as_tensor(m["plus"], device=dev)
            hc = torch.as_tensor(m["cross"], device=dev)
            tf = torch.as_tensor(m["tf"], device=dev, dtype=torch.float64)
            if hp.shape[-1] != n or hc.shape[-1] != n or tf.shape[-1] != n:
                raise ValueError(f"waveform arrays must have length {n}")
            hp = hp.to(torch.complex128)
            hc = hc.to(torch.complex128)
            wf[i, :, 0] = hp.real
            wf[i, :, 1] = hp.imag
            wf[i, :, 2] = hc.real
            wf[i, :, 3] = hc.imag
            wf[i, :, 4] = tf
        return wf, len(modes), dev

    def links_from_container(self, det, waveform, lam, beta, psi, tc):
        """Complex (6, N) single links of detector det for a reference-style container."""
        self._build()
        wf, n_modes, dev = self._container(det, waveform)
        out = torch.empty(6, self._n_bins[det], 2, dtype=torch.float64, device=dev)
        check(lib().pes_links_from_waveform(self._ctx, det, n_modes, wf.data_ptr(), float(lam), float(beta),
                                            float(psi), float(tc), out.data_ptr(),
                                            torch.cuda.current_stream(dev).cuda_stream))
        return torch.view_as_complex(out)

    def waveform(self, det, params):
        """(P, n_modes, N, 5): plus re/im, cross re/im, tf on detector det's bins."""
        self._build()
        p = self._params(params)
        n = self._n_bins[det]
        out = torch.empty(p.shape[0], len(self.modes), n, 5, dtype=torch.float64, device=p.device)
        check(lib().pes_waveform(self._ctx, det, p.data_ptr(), p.shape[0], out.data_ptr(),
                                 torch.cuda.current_stream(p.device).cuda_stream))
        return out

    def response_from_container(self, det, waveform, lam, beta, psi, tc):
        """Complex (3, N) response of detector det to a reference-style container.

        waveform: {"plus","cross","tf"} or a mode-keyed dict of those (numpy or torch).
        """
        self._build()
        wf, n_modes, dev = self._container(det, waveform)
        n = self._n_bins[det]
        out = torch.empty(3, n, 2, dtype=torch.float64, device=dev)
        check(lib().pes_response_from_waveform(self._ctx, det, n_modes, wf.data_ptr(), float(lam), float(beta),
                                               float(psi), float(tc), out.data_ptr(),
                                               torch.cuda.current_stream(dev).cuda_stream))
        return torch.view_as_complex(out)


def orbit_vectors(orbit_kind, arm_length, rotation_initial, revolution_initial, time,
                  omega_rotation=GEO_OMEGA_ROTATION, lambda_ref=GEO_LAMBDA_REF, beta_ref=GEO_BETA_REF, device=0):
    """Constellation vectors at the given times on the GPU: a (..., 6, 3) float64 tensor holding
    n1, n2, n3 (unit link vectors) and x1, x2, x3 (node positions in the SSB frame, seconds)
    (orbit.py:260-322, 363-419)."""
    if not torch.cuda.is_available():
        raise PespaceError("pespace_amd's hip backend needs a GPU (torch.cuda.is_available() is False)")
    dev = torch.device(f"cuda:{int(device)}")
    t = torch.as_tensor(time, dtype=torch.float64, device=dev)
    shape = tuple(t.shape)
    t = t.reshape(-1).contiguous()
    out = torch.empty(t.numel(), 6, 3, dtype=torch.float64, device=dev)
    check(lib().pes_orbit_vectors(int(orbit_kind), float(arm_length), float(rotation_initial),
                                  float(revolution_initial), float(omega_rotation), float(lambda_ref),
                                  float(beta_ref), t.data_ptr(), t.numel(), out.data_ptr(),
                                  torch.cuda.current_stream(dev).cuda_stream))
    return out.reshape(shape + (6, 3))


def _device(device):
    if not torch.cuda.is_available():
        raise PespaceError("pespace_amd's hip backend needs a GPU (torch.cuda.is_available() is False)")
    return torch.device(f"cuda:{int(device)}") if not isinstance(device, torch.device) else device


def td_links(orbit_kind, arm_length, rotation_initial, revolution_initial, hp, hc, wave_t0, wave_dt, lam, beta, psi,
             t0, dt, n, omega_rotation=GEO_OMEGA_ROTATION, lambda_ref=GEO_LAMBDA_REF, beta_ref=GEO_BETA_REF,
             out=None, device=0):
    """Time-domain single links on the GPU: a (6, n) float64 tensor (link12, link21, link23,
    link32, link31, link13) at the times t0 + i dt, for h+, hx sampled at wave_t0 + m wave_dt
    (linearly interpolated, zero outside the samples), with the constellation at each sample
    time (the reference's commented TDResponseModelCornish2003, antenna.py:681-800)."""
    dev = hp.device if isinstance(hp, torch.Tensor) and hp.is_cuda else _device(device)
    hp = torch.as_tensor(hp, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    hc = torch.as_tensor(hc, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    if hp.numel() != hc.numel():
        raise ValueError(f"h+ and hx differ in length ({hp.numel()} vs {hc.numel()})")
    n = int(n)
    if out is None:
        out = torch.empty(6, n, dtype=torch.float64, device=dev)
    else:
        LikelihoodCore._check_out(out, "out", (6, n), dev)
    check(lib().pes_td_links(int(orbit_kind), float(arm_length), float(rotation_initial), float(revolution_initial),
                             float(omega_rotation), float(lambda_ref), float(beta_ref), hp.data_ptr(), hc.data_ptr(),
                             float(wave_t0), float(wave_dt), hp.numel(), float(lam), float(beta), float(psi),
                             float(t0), float(dt), n, out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream))
    return out


def td_added_samples(arm_length, dt, tdi_generation):
    """Link samples the time-domain TDI needs before its first output: int(max_delay L // dt + 1),
    max_delay 3 (TDI 1.5) or 7 (TDI 2.0) (tdi.py:1445-1455)."""
    gen = {"1.5": 15, "2.0": 20, 15: 15, 20: 20}.get(tdi_generation)
    if gen is None:
        raise ValueError(f"Unsupported generation {tdi_generation}.")
    added = lib().pes_td_tdi_added_samples(float(arm_length), float(dt), gen)
    if added < 0:
        raise ValueError("arm length and dt must be positive")
    return added


def td_tdi(arm_length, dt, tdi_generation, orthogonal, links, out=None):
    """Time-domain Michelson TDI on the GPU: (3, n_ext - added) float64 (A, E, T or X, Y, Z) from
    (6, n_ext) links sampled at spacing dt, output sample i at link sample i + added
    (TDMichelsonConstantEqualArm.update_tdi_response, tdi.py:1457-1600)."""
    if not (isinstance(links, torch.Tensor) and links.is_cuda):
        links = torch.as_tensor(links, dtype=torch.float64, device=_device(0))
    links = links.to(torch.float64).contiguous()
    if links.dim() != 2 or links.shape[0] != 6:
        raise ValueError(f"links must be (6, n_ext), got {tuple(links.shape)}")
    added = td_added_samples(arm_length, dt, tdi_generation)
    n_out = links.shape[1] - added
    if n_out < 0:
        raise ValueError(f"{links.shape[1]} link samples, fewer than the {added} the delays need")
    gen = 15 if str(tdi_generation) in ("1.5", "15") else 20
    if out is None:
        out = torch.empty(3, n_out, dtype=torch.float64, device=links.device)
    else:
        LikelihoodCore._check_out(out, "out", (3, n_out), links.device)
    check(lib().pes_td_tdi(float(arm_length), float(dt), gen, int(bool(orthogonal)), links.data_ptr(),
                           links.shape[1], out.data_ptr(), torch.cuda.current_stream(links.device).cuda_stream))
    return out


def whittle_lnl(data, response, psd, df):
    """-2 df sum |d - h|^2 / S over channels and bins, on the GPU (deterministic reduction).

    data, response: complex (n_chan, N) tensors; psd: real (n_chan, N).
    """
    dev = data.device
    d = torch.view_as_real(data.to(torch.complex128).contiguous()).contiguous()
    r = torch.view_as_real(response.to(device=dev, dtype=torch.complex128).contiguous()).contiguous()
    s = psd.to(device=dev, dtype=torch.float64).contiguous()
    n_chan, n = s.shape
    work = torch.empty(1024, dtype=torch.float64, device=dev)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    check(lib().pes_whittle_lnl(d.data_ptr(), r.data_ptr(), s.data_ptr(), n_chan, n, float(df), work.data_ptr(),
                                out.data_ptr(), torch.cuda.current_stream(dev).cuda_stream))
    return out

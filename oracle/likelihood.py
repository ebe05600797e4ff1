"""Oracle: end-to-end lnL, A/E/T response and finite-difference gradient.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Mirrors the reference's orchestration: the waveform is computed once per
parameter set and shared by every detector (pespace/inference/likelihood_jax.py:231-258,
interface_bilby.py:97-114); each detector adds its Whittle lnL over the
requested channels.
"""

from __future__ import annotations

import numpy as np

from . import detector as det
from . import phenom as ph

# parameter order of a batch row: the reference's likelihood namedtuple order,
# waveform.required_params + (lambda, beta, psi, t_c)  (likelihood_jax.py:119-127)
PARAM_NAMES = ("chirp_mass", "mass_ratio", "chi_1", "chi_2", "luminosity_distance", "reference_phase",
               "inclination", "ecliptic_longitude", "ecliptic_latitude", "polarization", "coalescence_time")


def as_dict(row):
    return {k: float(v) for k, v in zip(PARAM_NAMES, row)}


class Detector:
    """Plain container for one detector's grid, data and PSD (oracle side)."""

    def __init__(self, f, data, psd, df, orbit, tdi_generation="1.5", channels=("A", "E", "T"),
                 orthogonal=True, duty=None):
        self.f = np.asarray(f, dtype=np.float64)
        self.data = data
        self.psd = psd
        self.df = df
        self.orbit = orbit  # (kind, arm_length[m], rot0, rev0)
        self.tdi_generation = tdi_generation
        self.channels = tuple(channels)
        self.orthogonal = orthogonal
        self.duty = duty  # None or (period, on_duration, start_time, taper)


def waveform(params, f, modes, f_ref, interp=None):
    if interp is None:
        wf = ph.waveform_direct(params, f, modes, f_ref)
        if len(modes) == 1:
            return {"22": wf} if modes[0] == (2, 2) else {f"{modes[0][0]}{modes[0][1]}": wf}
        return wf
    fmin_n, fmax_n, n_nodes = interp
    return ph.waveform_interpolated(params, f, modes, f_ref, fmin_n, fmax_n, n_nodes)


def response(params, d, modes=((2, 2),), f_ref=0.0, interp=None, wf=None):
    if wf is None:
        wf = waveform(params, d.f, modes, f_ref, interp)
    links = det.single_links(wf, params["ecliptic_longitude"], params["ecliptic_latitude"],
                             params["polarization"], params["coalescence_time"], d.f, d.orbit,
                             getattr(d, "duty", None))
    return det.tdi_combination(links, d.f, d.orbit[1], d.tdi_generation, d.orthogonal)


def lnl(params, detectors, modes=((2, 2),), f_ref=0.0, interp=None):
    """Sum over detectors of the Whittle lnL (likelihood_ti.py:58-67 per detector)."""
    tot = 0.0
    for d in detectors:
        if interp is None:
            wf = None
        else:
            wf = waveform(params, d.f, modes, f_ref, interp)
        resp = response(params, d, modes, f_ref, interp, wf)
        tot += det.whittle_lnl(d.channels, d.data, resp, d.psd, d.df)
    return tot


def lnl_batch(rows, detectors, modes=((2, 2),), f_ref=0.0, interp=None):
    return np.array([lnl(as_dict(r), detectors, modes, f_ref, interp) for r in np.atleast_2d(rows)])


def fd_gradient(row, detectors, modes=((2, 2),), f_ref=0.0, interp=None, rel_step=1e-6, abs_floor=1e-6):
    """Central finite differences, step rel_step*|theta| (abs_floor for theta ~ 0)."""
    row = np.asarray(row, dtype=np.float64)
    g = np.zeros_like(row)
    for i in range(row.size):
        h = rel_step * abs(row[i]) if row[i] != 0.0 else abs_floor
        rp, rm = row.copy(), row.copy()
        rp[i] += h
        rm[i] -= h
        g[i] = (lnl(as_dict(rp), detectors, modes, f_ref, interp)
                - lnl(as_dict(rm), detectors, modes, f_ref, interp)) / (2 * h)
    return g

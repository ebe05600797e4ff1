"""Time-domain single links and Michelson TDI (the oracle's restatement).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): tests/ use it as the checker of the HIP
time-domain path; the product never does.

The reference's time-domain path is marked incomplete: the Cornish response model is commented
out (pespace/detector/antenna.py:566-800) and TDMichelsonConstantEqualArm
(pespace/detector/tdi.py:1388-1600) has no caller.  What they define, restated here:

* links (antenna.py:681-800): at each sample time t, with the constellation at t,
  y_rs(t) = n.[H(t - L - k.x_s) - H(t - k.x_r)].n / (2 (1 - k.n_sr)), link_rs sent from s and
  received at r, H = h+ e+ + hx ex; the waveform is sampled at t_w0 + m dt and linearly
  interpolated (utils.py:83-100, antenna.py:617-628), zero outside its samples (the model's
  _ensure_waveform_length pads it, antenna.py:630-648);
* TDI (tdi.py:1457-1600): X = P(D) [y13 - y12 + D (y31 - y21)], Y, Z cyclic, P = 1 - D^2 (1.5)
  or 1 - D^2 - D^4 + D^6 (2.0), D^m y(t) = y(t - m L) by linear interpolation between samples;
  the links are sampled on an extended grid that starts ``added`` samples before the output,
  added = int(max_delay L // dt + 1) with max_delay 3 (1.5) or 7 (2.0) (tdi.py:1445-1455).
  Output sample i reads D^m y at the fractional link index i + added - m L / dt, whose integer
  part and fraction are split once per delay (i is an integer).  (The
  reference's loop, tdi.py:1482-1497, takes the base index i + floor((max_delay - m) L / dt)
  with fraction 1 - frac(m L / dt); that equals i + added - m L / dt except where
  frac(m L / dt) > frac(max_delay L / dt), where it reads one sample early.  The restatement
  follows the delay's definition.)
"""

from __future__ import annotations

import numpy as np

from .detector import C_SI, GEO_BETA_REF, GEO_LAMBDA_REF, GEO_OMEGA_ROT, constellation, polarization_basis

LINK_NAMES = ("link12", "link21", "link23", "link32", "link31", "link13")
# link -> (receiver, sender), 0-based, and the arm vector n_c (c = the third node) it runs along
LINK_RS = ((0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2))
MAX_DELAY = {"1.5": 3, "2.0": 7}
P_COEF = {"1.5": (1.0, 0.0, -1.0), "2.0": (1.0, 0.0, -1.0, 0.0, -1.0, 0.0, 1.0)}


def added_samples(arm_length, dt, generation="1.5"):
    """tdi.py:1445-1450."""
    return int(np.floor(MAX_DELAY[generation] * (arm_length / C_SI) / dt)) + 1


def _interp(y, x):
    """Linear interpolation of samples y at fractional indices x; 0 outside [0, len - 1]."""
    ok = (x >= 0) & (x <= y.size - 1)
    jj = np.clip(np.floor(np.where(ok, x, 0.0)), 0, y.size - 2).astype(np.int64)
    fr = x - jj
    return np.where(ok, y[jj] + (y[jj + 1] - y[jj]) * fr, 0.0)


def links(orbit, hp, hc, wave_t0, wave_dt, lam, beta, psi, t0, dt, n):
    """(6, n) links at t0 + i dt for h+, hx sampled at wave_t0 + m wave_dt; orbit as in
    detector.single_links: (kind, arm[m], rot0, rev0[, omega_rot, lambda_ref, beta_ref])."""
    kind, arm, rot0, rev0 = orbit[:4]
    geo = tuple(orbit[4:]) or (GEO_OMEGA_ROT, GEO_LAMBDA_REF, GEO_BETA_REF)
    L = arm / C_SI
    t = t0 + dt * np.arange(n, dtype=np.float64)
    n1, n2, n3, x1, x2, x3 = constellation(kind, arm, rot0, rev0, t, *geo)
    nv, xv = (n1, n2, n3), (x1, x2, x3)
    p, q, k = polarization_basis(lam, beta, psi)
    hp, hc = np.asarray(hp, np.float64), np.asarray(hc, np.float64)
    kx = [k @ x for x in xv]
    send = [(_interp(hp, (t - L - kx[a] - wave_t0) / wave_dt), _interp(hc, (t - L - kx[a] - wave_t0) / wave_dt))
            for a in range(3)]
    rece = [(_interp(hp, (t - kx[a] - wave_t0) / wave_dt), _interp(hc, (t - kx[a] - wave_t0) / wave_dt))
            for a in range(3)]
    out = np.zeros((6, n))
    for li, (r, s) in enumerate(LINK_RS):
        c = 3 - r - s
        na = nv[c]
        n_sr = (xv[r] - xv[s]) / L  # = +- n_c
        npp, nqq = p @ na, q @ na
        Pp, Pc = npp * npp - nqq * nqq, 2.0 * npp * nqq
        out[li] = 0.5 * (Pp * (send[s][0] - rece[r][0]) + Pc * (send[s][1] - rece[r][1])) / (1.0 - k @ n_sr)
    return out


def tdi(link_samples, arm_length, dt, generation="1.5", orthogonal=True):
    """(3, n_ext - added) channels from (6, n_ext) links (order LINK_NAMES)."""
    y = dict(zip(LINK_NAMES, np.asarray(link_samples, np.float64)))
    n_ext = link_samples.shape[1]
    added = added_samples(arm_length, dt, generation)
    n_out = n_ext - added
    r = (arm_length / C_SI) / dt
    i = np.arange(n_out)

    def D(name, m):
        """y at the fractional index i + added - m r, split once into an integer base and a
        fraction (i is an integer); the base lies in [0, added - 1] for m >= 1 (added - m r > 0
        up to rounding), so every read is inside the samples."""
        if m == 0:
            return y[name][i + added]
        c = added - m * r
        jb = min(max(int(np.floor(c)), 0), added - 1)
        fr = c - jb
        a, b = y[name][i + jb], y[name][i + jb + 1]
        return a + (b - a) * fr

    def chan(a, b, c, e):
        v = np.zeros(n_out)
        for kk, pk in enumerate(P_COEF[generation]):
            if pk:
                v += pk * (D(a, kk) - D(b, kk) + D(c, kk + 1) - D(e, kk + 1))
        return v

    X = chan("link13", "link12", "link31", "link21")
    Y = chan("link21", "link23", "link12", "link32")
    Z = chan("link32", "link31", "link23", "link13")
    if orthogonal:
        return np.stack([(Z - X) / np.sqrt(2.0), (X - 2.0 * Y + Z) / np.sqrt(6.0), (X + Y + Z) / np.sqrt(3.0)])
    return np.stack([X, Y, Z])

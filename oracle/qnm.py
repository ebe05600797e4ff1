"""Kerr quasi-normal-mode frequencies by Leaver's continued-fraction method.

TEST INFRASTRUCTURE (see ``oracle/__init__.py``), plus the generator of the
committed QNM table ``pespace_amd/data/qnm_table.json``.

The PhenomD / PhenomHM ringdowns need the fundamental (n = 0) spin-weight -2
QNM frequency of each (l, m) mode as a function of the remnant spin.  The
reference holds no such data (its waveforms come from the absent tiwave /
ripplegw packages), so the table is computed here from the Teukolsky equation
with Leaver's method (E. W. Leaver, Proc. R. Soc. Lond. A 402, 285 (1985),
eqs. 18-21 for the angular and 25-26 for the radial three-term recurrences),
in Leaver's units 2M = 1, time dependence exp(-i w t).

Pins: the a = 0 values reproduce the published Schwarzschild frequencies
M w_l0 = 0.373672-0.088962i (l=2), 0.599443-0.092703i (l=3),
0.809178-0.094164i (l=4); the a > 0 values agree with the Berti-Cardoso-Will
fits (Phys. Rev. D 73, 064030, table VIII) within their stated few-percent
accuracy; and m -> -m with a -> -a leaves w invariant (tests/test_qnm.py).
"""

from __future__ import annotations

import cmath
import json
import math

import numpy as np

S = -2  # spin weight of the radiative Weyl scalar


def _angular_cf(A, c, l, m, nmax=80):
    """Residual of the angular three-term recurrence (Leaver eqs. 18-21)."""
    k1 = 0.5 * abs(m - S)
    k2 = 0.5 * abs(m + S)

    def alpha(n):
        return -2.0 * (n + 1) * (n + 2 * k1 + 1)

    def beta(n):
        return (n * (n - 1) + 2 * n * (k1 + k2 + 1 - 2 * c)
                - (2 * c * (2 * k1 + S + 1) - (k1 + k2) * (k1 + k2 + 1))
                - (c * c + S * (S + 1) + A))

    def gamma(n):
        return 2 * c * (n + k1 + k2 + S)

    # Leaver's inverted fraction at index N = l - max(|m|, |s|) (his eq. 14): the
    # eigenvalue that continues to l(l+1) - s(s+1) at c = 0 is the root of
    # beta_N - [alpha_{N-1} gamma_N / (beta_{N-1} - ...)] - [alpha_N gamma_{N+1} / (beta_{N+1} - ...)]
    N = l - max(abs(m), abs(S))
    up = 0.0 + 0.0j
    for n in range(nmax, N, -1):
        up = alpha(n - 1) * gamma(n) / (beta(n) - up)
    down = 0.0 + 0.0j
    for n in range(0, N):
        down = alpha(n) * gamma(n + 1) / (beta(n) - down)
    return beta(N) - down - up


def _radial_cf(w, A, a, m, nmax=600):
    """Residual of the radial three-term recurrence (Leaver eqs. 25-26), 2M = 1."""
    b = math.sqrt(1.0 - 4.0 * a * a)
    sig = (w / 2 - a * m)
    c0 = 1 - S - 1j * w - (2j / b) * sig
    c1 = -4 + 2j * w * (2 + b) + (4j / b) * sig
    c2 = S + 3 - 3j * w - (2j / b) * sig
    c3 = (w * w * (4 + 2 * b - a * a) - 2 * a * m * w - S - 1 + (2 + b) * 1j * w - A
          + ((4 * w + 2j) / b) * sig)
    c4 = S + 1 - 2 * w * w - (2 * S + 3) * 1j * w - ((4 * w + 2j) / b) * sig

    def alpha(n):
        return n * n + (c0 + 1) * n + c0

    def beta(n):
        return -2 * n * n + (c1 + 2) * n + c3

    def gamma(n):
        return n * n + (c2 - 3) * n + c4 - c2 + 2

    # Nollert-style tail estimate: remainder R_N ~ -1 + sqrt(-2iw/N) ... use a
    # plain deep truncation, which converges for the fundamental mode.
    tail = 0.0 + 0.0j
    for n in range(nmax, 0, -1):
        tail = alpha(n - 1) * gamma(n) / (beta(n) - tail)
    return (beta(0) - tail) / beta(0)


def _solve_A(c, l, m, A0):
    """Newton-secant for the angular separation constant at c = a w."""
    x0, x1 = A0, A0 * (1 + 1e-6) + 1e-6
    f0, f1 = _angular_cf(x0, c, l, m), _angular_cf(x1, c, l, m)
    for _ in range(100):
        if f1 == f0:
            break
        x2 = x1 - f1 * (x1 - x0) / (f1 - f0)
        x0, f0 = x1, f1
        x1, f1 = x2, _angular_cf(x2, c, l, m)
        if abs(x1 - x0) < 1e-14 * max(1.0, abs(x1)):
            break
    return x1


def _solve_mode(a, l, m, w0, A0):
    """Solve the coupled angular + radial problem for w (Leaver units)."""
    state = {"A": A0}

    def F(w):
        state["A"] = _solve_A(a * w, l, m, state["A"])
        return _radial_cf(w, state["A"], a, m)

    x0, x1 = w0, w0 * (1 + 1e-5)
    f0, f1 = F(x0), F(x1)
    for _ in range(200):
        if f1 == f0:
            break
        x2 = x1 - f1 * (x1 - x0) / (f1 - f0)
        x0, f0 = x1, f1
        x1, f1 = x2, F(x2)
        if abs(x1 - x0) < 1e-13:
            break
    F(x1)
    return x1, state["A"]


SCHWARZSCHILD = {2: 0.373672 - 0.088962j, 3: 0.599443 - 0.092703j, 4: 0.809178 - 0.094164j}


def qnm_branch(l, m, spins):
    """Return M*w (complex, Re>0, Im<0) for each dimensionless spin in ``spins``.

    ``spins`` must be sorted by |a| outward from 0 on one side (all >= 0 or all <= 0);
    the branch is followed by continuation from the Schwarzschild value.
    """
    w = 2 * SCHWARZSCHILD[l]
    A = l * (l + 1) - S * (S + 1)
    w, A = _solve_mode(0.0, l, m, w, A)
    out = []
    prev = [(0.0, w)]
    for chi in spins:
        a = 0.5 * chi
        if len(prev) >= 2:  # linear extrapolation in a for the initial guess
            (a1, w1), (a2, w2) = prev[-2], prev[-1]
            guess = w2 + (w2 - w1) * (a - a2) / (a2 - a1)
        else:
            guess = prev[-1][1]
        w, A = _solve_mode(a, l, m, guess, A)
        prev.append((a, w))
        out.append(w / 2.0)
    return out


MODES = ((2, 2), (2, 1), (3, 3), (3, 2), (4, 4), (4, 3))
TABLE_AMIN = -0.995
TABLE_STEP = 0.005
TABLE_N = 399  # a_i = TABLE_AMIN + i * TABLE_STEP, i = 0..398 -> [-0.995, 0.995]


def build_table(path=None):
    """Compute Re/Im(M w)/(2 pi) for the six modes on the uniform spin grid."""
    grid = TABLE_AMIN + TABLE_STEP * np.arange(TABLE_N)
    pos = [float(a) for a in grid if a > 1e-12]
    neg = [float(a) for a in grid[::-1] if a < -1e-12]
    table = {}
    for (l, m) in MODES:
        wpos = qnm_branch(l, m, pos)
        wneg = qnm_branch(l, m, neg)
        w0 = _solve_mode(0.0, l, m, 2 * SCHWARZSCHILD[l], l * (l + 1) - 2)[0] / 2
        vals = {}
        for a, w in zip(pos, wpos):
            vals[round(a, 6)] = w
        for a, w in zip(neg, wneg):
            vals[round(a, 6)] = w
        vals[0.0] = w0
        ws = [vals[round(float(a), 6) if abs(a) > 1e-12 else 0.0] for a in grid]
        table[f"{l}{m}"] = {
            "fring": [float(w.real / (2 * math.pi)) for w in ws],
            "fdamp": [float(-w.imag / (2 * math.pi)) for w in ws],
        }
    doc = {
        "description": "Kerr n=0 spin-weight -2 QNM frequencies from Leaver's continued "
                       "fraction: fring = Re(M_f w)/2pi, fdamp = -Im(M_f w)/2pi, in units of "
                       "the remnant mass, on the uniform remnant-spin grid a_i = amin + i*step.",
        "amin": TABLE_AMIN, "step": TABLE_STEP, "n": TABLE_N,
        "modes": table,
    }
    if path is not None:
        with open(path, "w") as fh:
            json.dump(doc, fh, indent=0)
    return doc


def interp_table(vals, a):
    """Catmull-Rom (C1 cubic Hermite, centred-difference slopes) on the uniform grid.

    Identical scheme to the device code: cell i = floor((a-amin)/step) clamped to
    [1, n-3], slopes m_i = (v[i+1]-v[i-1])/2 in grid units.
    """
    vals = np.asarray(vals, dtype=np.float64)
    x = (np.asarray(a, dtype=np.float64) - TABLE_AMIN) / TABLE_STEP
    i = np.clip(np.floor(x).astype(np.int64), 1, TABLE_N - 3)
    t = x - i
    p0, p1 = vals[i], vals[i + 1]
    m0 = 0.5 * (vals[i + 1] - vals[i - 1])
    m1 = 0.5 * (vals[i + 2] - vals[i])
    c2 = 3.0 * (p1 - p0) - 2.0 * m0 - m1
    c3 = 2.0 * (p0 - p1) + m0 + m1
    return p0 + t * (m0 + t * (c2 + t * c3))


if __name__ == "__main__":  # regenerate the committed table
    import os
    import sys
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(
        os.path.dirname(__file__), "..", "pespace_amd", "data", "qnm_table.json")
    build_table(out)
    print("wrote", out)

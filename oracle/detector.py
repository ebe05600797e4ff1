"""Oracle: detector side of the FD hot path (grid, PSD, noise, orbit, response,
TDI, Whittle lnL), restated in numpy FP64.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function cites the reference lines it restates.  Array layout is the
reference's: one complex array per channel / link over the frequency bins.
"""

from __future__ import annotations

import numpy as np

# --- constants: pespace/utils/constants.py:29-57 ---------------------------
C_SI = 299792458.0
MRSUN_SI = 1.476625038050124729627979840144936351e3
MTSUN_SI = 4.925490947641266978197229498498379006e-6
MSUN_SI = 1.988409870698050731911960804878414216e30
YEAR_SI = 31557600.0
DAY_SI = 86400.0
AU_SI = 149597870700.0
PC_SI = 3.085677581491367e16
PI = 3.141592653589793238462643383279502884
EULER_GAMMA = 0.5772156649015329
AU_SEC = AU_SI / C_SI  # pespace/detector/orbit.py:12
SQRT3 = np.sqrt(3.0)  # orbit.py:13


# --- grid: pespace/detector/tdi.py:129-159 (DataInformation.__post_init__) --
def data_grid(duration, delta_time, minimum_frequency, maximum_frequency):
    """Return (frequencies, mask, n_t, df) exactly as DataInformation does.

    tdi.py:130  delta_frequency = 1/duration
    tdi.py:131  time_series_length = int(duration // delta_time + 1)
    tdi.py:136  full grid = rfftfreq(time_series_length, delta_time)
    tdi.py:141  mask = (f >= fmin) * (f <= fmax)
    """
    df = 1 / duration
    n_t = int(duration // delta_time + 1)
    f_full = np.fft.rfftfreq(n_t, delta_time)
    mask = (f_full >= minimum_frequency) * (f_full <= maximum_frequency)
    return f_full[mask], mask, n_t, df, f_full


# --- PSD: pespace/detector/noise.py:71-207, 231-265, 268-284 ---------------
PSD_MODELS = {
    # name: (kind, OMS level, acc level, arm length [s])   noise.py:268-284
    "LISA_SciRDv1": ("analytic", (15.0e-12) ** 2, (3.0e-15) ** 2, 2.5e9 / C_SI),
    "Taiji_TDC": ("analytic", (8.0e-12) ** 2, (3.0e-15) ** 2, 3.0e9 / C_SI),
    "Tianqin_GWSpace": ("tianqin", (1.0e-12) ** 2, (1.0e-15) ** 2, 1.7e8 / C_SI),
}


def _s_oms(kind, P_o, f):
    if kind == "analytic":  # noise.py:85-89
        return P_o * (1.0 + (2.0e-3 / f) ** 4) * (2.0 * PI * f / C_SI) ** 2
    return P_o * (2.0 * PI * f / C_SI) ** 2  # noise.py:245


def _s_acc(kind, P_a, f):
    if kind == "analytic":  # noise.py:105-110
        return P_a * (1.0 + (0.4e-3 / f) ** 2) * (1.0 + (f / 8e-3) ** 4) / (2 * PI * f * C_SI) ** 2
    return P_a * (1.0 + 0.1e-3 / f) / (2 * PI * f * C_SI) ** 2  # noise.py:261-265


def psd(model, f, channels, scaling=False, tdi_generation="1.5"):
    """noise.py:112-207: dict channel -> S_c(f)."""
    kind, P_o, P_a, L = PSD_MODELS[model]
    S_oms = _s_oms(kind, P_o, f)
    S_acc = _s_acc(kind, P_a, f)
    if scaling:  # noise.py:154-157
        fac = (PC_SI / (MRSUN_SI * MTSUN_SI)) ** 2
        S_oms = S_oms * fac
        S_acc = S_acc * fac
    if tdi_generation == "1.5":
        pre = 1.0
    elif tdi_generation == "2.0":
        pre = 4.0 * np.sin(4 * PI * f * L) ** 2  # noise.py:162
    else:
        raise ValueError(tdi_generation)
    out = {}
    for c in channels:
        if c in ("X", "Y", "Z"):  # noise.py:168-176
            s = 16 * np.sin(2 * PI * f * L) ** 2 * (S_oms + (3 + np.cos(4 * PI * f * L)) * S_acc)
        elif c in ("A", "E"):  # noise.py:177-190
            s = 8 * np.sin(2 * PI * f * L) ** 2 * (
                (2 + np.cos(2 * PI * f * L)) * S_oms
                + (6 + 4 * np.cos(2 * PI * f * L) + 2 * np.cos(4 * PI * f * L)) * S_acc
            )
        elif c == "T":  # noise.py:191-200
            s = (
                32 * np.sin(2 * PI * f * L) ** 2 * np.sin(PI * f * L) ** 2
                * (S_oms + 4 * np.sin(PI * f * L) ** 2 * S_acc)
            )
        else:
            raise ValueError(c)
        out[c] = s * pre
    return out


# --- noise: pespace/detector/tdi.py:896-916 ---------------------------------
def noise_realization(seed, channels, n_t, mask, psd_dict, df):
    """Exact numpy RNG call sequence of get_fd_noise_realization."""
    rng = np.random.default_rng(seed=seed)
    var = 0.5 / df ** 0.5
    n_full = int(mask.shape[0])
    out = {}
    for c in channels:
        re, im = rng.normal(0, var, (2, n_full))
        re[0] = 0.0
        im[0] = 0.0
        if np.mod(n_t, 2) == 0:
            im[-1] = 0.0
            re[-1] = 0.0
        out[c] = (re + 1j * im)[mask] * np.asarray(psd_dict[c]) ** 0.5
    return out


# --- orbits: pespace/detector/orbit.py ------------------------------------
ORBITS = {
    # orbit.py:442-446
    "LISA_analytic": ("helio", 2.5e9, 0.0, -PI / 9),
    "Taiji_analytic": ("helio", 3.0e9, 0.0, PI / 9),
    "Tianqin_analytic": ("geo", 1.7e8, 0.0, 0.0),
}
GEO_OMEGA_ROT = 2 * PI / (3.65 * DAY_SI)  # orbit.py:82
GEO_LAMBDA_REF = 120.5 / 180 * PI  # orbit.py:83
GEO_BETA_REF = -4.7 / 180 * PI  # orbit.py:84


def constellation(kind, arm_length, rot0, rev0, t, omega_rotation=GEO_OMEGA_ROT, lambda_ref=GEO_LAMBDA_REF,
                  beta_ref=GEO_BETA_REF):
    """Return (n1, n2, n3, x1, x2, x3), each (3, ...) arrays, at times t.

    helio: orbit.py:189-228 (params), 381-419 (vectors)
    geo:   orbit.py:128-139 (params), 278-322 (vectors); omega_rotation, lambda_ref and
           beta_ref are the constructor's arguments (orbit.py:77-85, TianQin defaults)
    """
    t = np.asarray(t, dtype=np.float64)
    L = arm_length / C_SI
    if kind == "helio":
        r_prime = L / (2 * np.sqrt(3))
        omega = 2 * PI / YEAR_SI
        kap = [rot0, rot0 + 2 * PI / 3, rot0 + 2 * PI / 3 + 2 * PI / 3]
        alpha = omega * t + rev0
        ca, sa = np.cos(alpha), np.sin(alpha)
        xd = []
        for k in kap:
            ck, sk = np.cos(k), np.sin(k)
            xd.append(r_prime * np.array([
                sa * ca * sk - (1 + sa * sa) * ck,
                sa * ca * ck - (1 + ca * ca) * sk,
                -SQRT3 * (ca * ck + sa * sk),
            ]))
        x0 = np.array([ca, sa, np.zeros_like(ca)]) * AU_SEC
    elif kind == "geo":
        r = L / np.sqrt(3)
        sl, cl = np.sin(lambda_ref), np.cos(lambda_ref)
        sb, cb = np.sin(beta_ref), np.cos(beta_ref)
        alpha = (2 * PI / YEAR_SI) * t + rev0
        k1 = omega_rotation * t + rot0
        kap = [k1, k1 + 2 * PI / 3, k1 + 2 * PI / 3 + 2 * PI / 3]
        xd = []
        for k in kap:
            ck, sk = np.cos(k), np.sin(k)
            xd.append(r * np.array([sb * cl * sk + sl * ck, sb * sl * sk - cl * ck, -cb * sk]))
        x0 = np.array([np.cos(alpha), np.sin(alpha), np.zeros_like(alpha)]) * AU_SEC
    else:
        raise ValueError(kind)
    n1 = (xd[2] - xd[1]) / L
    n2 = (xd[0] - xd[2]) / L
    n3 = (xd[1] - xd[0]) / L
    return n1, n2, n3, xd[0] + x0, xd[1] + x0, xd[2] + x0


# --- polarization: pespace/detector/antenna.py:42-114 ----------------------
def polarization_basis(lam, beta, psi):
    """p, q of antenna.py:70-83 and k of antenna.py:112-114."""
    sl, cl, sb, cb, sp, cp = np.sin(lam), np.cos(lam), np.sin(beta), np.cos(beta), np.sin(psi), np.cos(psi)
    p = np.array([sl * cp - cl * sb * sp, -cl * cp - sl * sb * sp, cb * sp])
    q = np.array([-sl * sp - cl * sb * cp, cl * sp - sl * sb * cp, cb * cp])
    k = np.array([-cb * cl, -cb * sl, -sb])
    return p, q, k


def polarization_tensors(lam, beta, psi):
    """e+ = pp - qq, ex = pq + qp (antenna.py:84-87)."""
    p, q, _ = polarization_basis(lam, beta, psi)
    return np.outer(p, p) - np.outer(q, q), np.outer(p, q) + np.outer(q, p)


def _sinc(x):
    """pespace/utils/utils.py:60-79: sin(x)/x, 1 at x == 0."""
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    nz = x != 0.0
    out[nz] = np.sin(x[nz]) / x[nz]
    return out


def duty_window(t, duty):
    """Observation window of a duty-cycled detector (not in the reference; the benchmark
    network's TianQin observes 3 months on, 3 months off).  duty = (period, on_duration,
    start_time, taper): s = (t - start_time) mod period; inside [0, on_duration) the window
    rises over the first taper seconds and falls over the last taper seconds by the C1
    smoothstep S(x) = x^2 (3 - 2x) of the distance into the window over taper; 0 outside."""
    period, on, t0, taper = (float(x) for x in duty)
    t = np.asarray(t, dtype=np.float64)
    u = (t - t0) / period
    s = (u - np.floor(u)) * period
    x = np.minimum(s, on - s) / taper if taper > 0 else np.ones_like(s)
    x = np.clip(x, 0.0, 1.0)
    return np.where(s < on, x * x * (3.0 - 2.0 * x), 0.0)


def single_links(waveform, lam, beta, psi, tc, f, orbit, duty=None):
    """Marsat-Baker FD single-link response, summed over modes.

    waveform: {"plus","cross","tf"} or a mode-keyed dict of those (antenna.py:438-458)
    orbit:    (kind, arm_length[m], rot0, rev0) or, geocentric, (kind, arm_length, rot0, rev0,
              omega_rotation, lambda_ref, beta_ref)
    duty:     None, or a duty cycle (duty_window): each mode weighted by the window at tf + tc
    Returns dict link -> complex[N].  antenna.py:429-460 (driver), 496-543 (links).
    """
    kind, arm, rot0, rev0 = orbit[:4]
    extra = tuple(orbit[4:])
    eplus, ecross = polarization_tensors(lam, beta, psi)
    _, _, k = polarization_basis(lam, beta, psi)
    f = np.asarray(f, dtype=np.float64)
    cexp_tshift = np.exp(1j * (-2.0 * PI * f * tc))  # antenna.py:435
    modes = [waveform] if set(waveform.keys()) == {"plus", "cross", "tf"} else list(waveform.values())
    names = ("link12", "link21", "link23", "link32", "link31", "link13")
    links = {n: np.zeros(f.shape, dtype=np.complex128) for n in names}
    L = arm / C_SI
    for wf in modes:
        hp = wf["plus"] * cexp_tshift  # antenna.py:447-449
        hc = wf["cross"] * cexp_tshift
        tf = wf["tf"] + tc
        if duty is not None:
            w = duty_window(tf, duty)
            hp, hc = hp * w, hc * w
        n1, n2, n3, x1, x2, x3 = constellation(kind, arm, rot0, rev0, tf, *extra)  # :496

        def nhn(n):  # antenna.py:499-512
            return (np.einsum("in,ij,jn->n", n, eplus, n) * hp
                    + np.einsum("in,ij,jn->n", n, ecross, n) * hc)

        n1hn1, n2hn2, n3hn3 = nhn(n1), nhn(n2), nhn(n3)
        k_n1, k_n2, k_n3 = k @ n1, k @ n2, k @ n3  # :514-516
        k_x1_x2 = k @ (x1 + x2)  # :518-520
        k_x2_x3 = k @ (x2 + x3)
        k_x3_x1 = k @ (x3 + x1)
        pi_f_L = PI * f * L  # :522
        sinc32 = _sinc(pi_f_L * (1.0 - k_n1))
        sinc23 = _sinc(pi_f_L * (1.0 + k_n1))
        sinc13 = _sinc(pi_f_L * (1.0 - k_n2))
        sinc31 = _sinc(pi_f_L * (1.0 + k_n2))
        sinc21 = _sinc(pi_f_L * (1.0 - k_n3))
        sinc12 = _sinc(pi_f_L * (1.0 + k_n3))
        common = -PI * f * 1j  # :530
        exp12 = np.exp(common * (L + k_x1_x2))
        exp23 = np.exp(common * (L + k_x2_x3))
        exp31 = np.exp(common * (L + k_x3_x1))
        pre = -pi_f_L * 1j  # :535
        links["link12"] += sinc12 * (pre * n3hn3) * exp12  # :536-541
        links["link21"] += sinc21 * (pre * n3hn3) * exp12
        links["link23"] += sinc23 * (pre * n1hn1) * exp23
        links["link32"] += sinc32 * (pre * n1hn1) * exp23
        links["link31"] += sinc31 * (pre * n2hn2) * exp31
        links["link13"] += sinc13 * (pre * n2hn2) * exp31
    return links


def tdi_combination(links, f, arm_length, generation="1.5", orthogonal=True):
    """FD Michelson TDI: tdi.py:1683-1703 (cache), 1725-1762 (X,Y,Z), 1767-1793 (A,E,T)."""
    L = arm_length / C_SI
    D = np.exp(-2.0 * PI * np.asarray(f) * L * 1j)  # tdi.py:1684-1690
    if generation == "1.5":
        pf = 1.0 - D ** 2  # tdi.py:1694
    elif generation == "2.0":
        pf = 1.0 - D ** 2 - D ** 4 + D ** 6  # tdi.py:1696-1701
    else:
        raise ValueError(generation)
    l = links
    X = pf * (l["link13"] - l["link12"] + D * (l["link31"] - l["link21"]))
    Y = pf * (l["link21"] - l["link23"] + D * (l["link12"] - l["link32"]))
    Z = pf * (l["link32"] - l["link31"] + D * (l["link23"] - l["link13"]))
    if orthogonal:
        return {"A": (Z - X) / np.sqrt(2.0), "E": (X - 2.0 * Y + Z) / np.sqrt(6.0),
                "T": (X + Y + Z) / np.sqrt(3.0)}
    return {"X": X, "Y": Y, "Z": Z}


def whittle_lnl(channels, data, response, psd_dict, df):
    """pespace/inference/likelihood_ti.py:58-67: -2 df sum |d-h|^2/S.

    ``df`` may be a scalar (uniform grid) or a per-bin weight array (log grid).
    """
    tot = np.zeros_like(np.asarray(psd_dict[channels[0]], dtype=np.float64))
    for c in channels:
        r = np.asarray(data[c]) - np.asarray(response[c])
        tot = tot + (r.real * r.real + r.imag * r.imag) / np.asarray(psd_dict[c])
    return float(np.sum(-2.0 * np.asarray(df) * tot))

"""Pins of the oracle's detector side (grid, PSD, noise, orbit, response, TDI, lnL).

Each test checks a property the mathematics or the reference's own notebooks fix,
chosen so a dropped term, a wrong sign/index or a swapped link fails.
"""

import numpy as np
import pytest

from oracle import detector as det
from oracle import likelihood as olk

YEAR = det.YEAR_SI


def test_grid_sizes_match_reference_configs():
    # examples/_checks/likelihood_jax.ipynb cells 1-2: T = 655360 s, dt = 5 s, [1e-4, 0.1] -> 65471 bins
    f, mask, n_t, df, _ = det.data_grid(655360.0, 5.0, 1e-4, 0.1)
    assert f.size == 65471 and df == 1 / 655360.0
    # 1-yr, dt = 10 s, [1e-4, 0.05]: n_t = int(T // dt + 1) = 3155761 -> 1574725 bins
    f, mask, n_t, df, full = det.data_grid(YEAR, 10.0, 1e-4, 0.05)
    assert n_t == 3155761 and f.size == 1574725
    # grid spacing 1/(n_t dt) differs from df = 1/T (tdi.py:129-145)
    assert abs((full[1] - full[0]) * YEAR - n_t * 10.0 / (n_t * 10.0) * YEAR / (n_t * 10.0)) < 1.0
    assert full[1] - full[0] == 1.0 / (n_t * 10.0)


@pytest.mark.parametrize("model", list(det.PSD_MODELS))
@pytest.mark.parametrize("gen", ["1.5", "2.0"])
def test_psd_symmetries(model, gen):
    f = np.geomspace(1e-5, 0.5, 4001)
    p = det.psd(model, f, ("X", "Y", "Z", "A", "E", "T"), False, gen)
    # psd.ipynb cell 5: X = Y = Z and A = E exactly
    assert np.array_equal(p["X"], p["Y"]) and np.array_equal(p["X"], p["Z"])
    assert np.array_equal(p["A"], p["E"])
    # A/E/T is an orthonormal rotation of X/Y/Z for a noise covariance with equal autos S and
    # equal cross-spectra C: S_A = S_E = S - C, S_T = S + 2C  ->  S_A + S_E + S_T = 3 S
    np.testing.assert_allclose(p["A"] + p["E"] + p["T"], 3 * p["X"], rtol=1e-9)


def test_psd_generation_factor_and_scaling():
    f = np.geomspace(1e-4, 0.1, 1001)
    L = 2.5e9 / det.C_SI
    a = det.psd("LISA_SciRDv1", f, ("A", "T"), False, "1.5")
    b = det.psd("LISA_SciRDv1", f, ("A", "T"), False, "2.0")
    np.testing.assert_allclose(b["A"], a["A"] * 4 * np.sin(4 * np.pi * f * L) ** 2, rtol=1e-14)
    c = det.psd("LISA_SciRDv1", f, ("A",), True, "1.5")
    np.testing.assert_allclose(c["A"], a["A"] * (det.PC_SI / (det.MRSUN_SI * det.MTSUN_SI)) ** 2, rtol=1e-14)


def test_noise_statistics_and_edges():
    f, mask, n_t, df, full = det.data_grid(2 ** 16 * 5.0, 5.0, 0.0, 0.1)  # keep DC and Nyquist
    p = {c: np.ones_like(f) for c in ("A", "E")}
    n = det.noise_realization(3, ("A", "E"), n_t, mask, p, df)
    assert n["A"][0] == 0 and n["E"][0] == 0  # DC zeroed (tdi.py:906-907)
    assert n_t % 2 == 1  # odd length: the last rfft bin is not a Nyquist bin and is kept
    # E|n|^2 = 2 * (0.5/sqrt(df))^2 = 0.5/df per bin
    v = np.mean(np.abs(n["A"][1:-1]) ** 2) * df
    assert abs(v - 0.5) < 0.02
    again = det.noise_realization(3, ("A", "E"), n_t, mask, p, df)
    assert np.array_equal(again["E"], n["E"])


@pytest.mark.parametrize("name", list(det.ORBITS))
def test_orbit_geometry(name):
    kind, arm, rot0, rev0 = det.ORBITS[name]
    t = np.linspace(0, YEAR, 97)
    n1, n2, n3, x1, x2, x3 = det.constellation(kind, arm, rot0, rev0, t)
    L = arm / det.C_SI
    for n in (n1, n2, n3):
        np.testing.assert_allclose(np.linalg.norm(n, axis=0), 1.0, atol=5e-3 if kind == "helio" else 1e-12)
    # link vectors are the edges of the triangle: n1 = (x3-x2)/L ... (orbit.py:413-415)
    np.testing.assert_allclose(n1, (x3 - x2) / L, atol=1e-9)
    np.testing.assert_allclose(n1 + n2 + n3, 0.0, atol=1e-14)
    # centroid = AU (cos a, sin a, 0), a = Omega t + rev0 (orbit_model.ipynb cell 8)
    c = (x1 + x2 + x3) / 3
    a = 2 * np.pi / YEAR * t + rev0
    np.testing.assert_allclose(c[0], det.AU_SEC * np.cos(a), atol=1e-9)
    np.testing.assert_allclose(c[1], det.AU_SEC * np.sin(a), atol=1e-9)
    # plane normal (orbit_model.ipynb cells 9, 21)
    nrm = np.cross(n3.T, -n2.T)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    if kind == "helio":
        ref = np.stack([np.cos(a + np.pi) * np.cos(np.pi / 6), np.sin(a + np.pi) * np.cos(np.pi / 6),
                        np.full_like(a, np.sin(np.pi / 6))], axis=1)
    else:
        lr, br = det.GEO_LAMBDA_REF, det.GEO_BETA_REF
        ref = np.tile([np.cos(br) * np.cos(lr), np.cos(br) * np.sin(lr), np.sin(br)], (t.size, 1))
    cosang = np.abs(np.sum(nrm * ref, axis=1))
    np.testing.assert_allclose(cosang, 1.0, atol=1e-12)


def test_polarization_tensors():
    rng = np.random.default_rng(1)
    for _ in range(20):
        lam, beta, psi = rng.uniform(0, 2 * np.pi), rng.uniform(-1.5, 1.5), rng.uniform(0, np.pi)
        ep, ec = det.polarization_tensors(lam, beta, psi)
        _, _, k = det.polarization_basis(lam, beta, psi)
        for e in (ep, ec):
            assert abs(np.trace(e)) < 1e-14
            np.testing.assert_allclose(e, e.T, atol=1e-15)
            np.testing.assert_allclose(e @ k, 0.0, atol=1e-14)
        assert abs(np.sum(ep * ep) - 2) < 1e-13 and abs(np.sum(ec * ec) - 2) < 1e-13
        assert abs(np.sum(ep * ec)) < 1e-13
        ep2, ec2 = det.polarization_tensors(lam, beta, psi + np.pi / 2)
        np.testing.assert_allclose(ep2, -ep, atol=1e-14)
        np.testing.assert_allclose(ec2, -ec, atol=1e-14)
        # k = -(source direction) (antenna.py:112-114)
        np.testing.assert_allclose(k, -np.array([np.cos(beta) * np.cos(lam), np.cos(beta) * np.sin(lam),
                                                 np.sin(beta)]), atol=1e-15)


def _synthetic(f, amp=1.0, tf=1.0):
    # conventions_difference.ipynb cell 6: amp = 1, phase = 0, tf = const
    return {"plus": amp * np.ones_like(f) + 0j, "cross": 0.3j * amp * np.ones_like(f), "tf": tf * np.ones_like(f)}


def test_links_linear_and_paired():
    f = np.geomspace(1e-4, 0.1, 501)
    orb = det.ORBITS["LISA_analytic"]
    w1, w2 = _synthetic(f), _synthetic(f, amp=2.5)
    w2["cross"] = 0.0 * w2["cross"] - 0.7
    a = det.single_links(w1, 1.0, 0.3, 0.2, 1e5, f, orb)
    b = det.single_links(w2, 1.0, 0.3, 0.2, 1e5, f, orb)
    s = det.single_links({"plus": w1["plus"] + w2["plus"], "cross": w1["cross"] + w2["cross"], "tf": w1["tf"]},
                         1.0, 0.3, 0.2, 1e5, f, orb)
    for k in a:
        np.testing.assert_allclose(s[k], a[k] + b[k], rtol=1e-12, atol=1e-30)
    # link12 / link21 = sinc(pi f L (1 + k.n3)) / sinc(pi f L (1 - k.n3)) (shared n3 h n3 and delay)
    _, _, k = det.polarization_basis(1.0, 0.3, 0.2)
    n1, n2, n3, *_ = det.constellation(orb[0], orb[1], orb[2], orb[3], np.full_like(f, 1.0 + 1e5))
    xL = np.pi * f * orb[1] / det.C_SI
    kn3 = k @ n3
    np.testing.assert_allclose(a["link12"] / a["link21"], np.sinc(xL * (1 + kn3) / np.pi) / np.sinc(xL * (1 - kn3) / np.pi),
                               rtol=1e-10)
    # multi-mode containers sum (antenna.py:446-458)
    m = det.single_links({"22": w1, "33": w2}, 1.0, 0.3, 0.2, 1e5, f, orb)
    for kk in a:
        np.testing.assert_allclose(m[kk], a[kk] + b[kk], rtol=1e-12, atol=1e-30)


def test_tdi_identities():
    f = np.geomspace(1e-4, 0.1, 301)
    orb = det.ORBITS["Taiji_analytic"]
    links = det.single_links(_synthetic(f), 0.4, -0.7, 1.1, 3e6, f, orb)
    x15 = det.tdi_combination(links, f, orb[1], "1.5", False)
    x20 = det.tdi_combination(links, f, orb[1], "2.0", False)
    L = orb[1] / det.C_SI
    D = np.exp(-2j * np.pi * f * L)
    # TDI 2.0 = (1 - D^4) x TDI 1.5, i.e. prefactor (1 - D^2)(1 - D^4) (tdi.py:1696-1701)
    for c in "XYZ":
        np.testing.assert_allclose(x20[c], (1 - D ** 4) * x15[c], rtol=1e-10, atol=1e-300)
    aet = det.tdi_combination(links, f, orb[1], "1.5", True)
    # A, E, T is an orthonormal map of X, Y, Z
    lhs = sum(np.abs(aet[c]) ** 2 for c in "AET")
    rhs = sum(np.abs(x15[c]) ** 2 for c in "XYZ")
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10)
    np.testing.assert_allclose(aet["A"], (x15["Z"] - x15["X"]) / np.sqrt(2), rtol=1e-15)


def test_whittle_lnl_closed_forms():
    rng = np.random.default_rng(4)
    n = 1000
    chans = ("A", "E")
    d = {c: rng.normal(size=n) + 1j * rng.normal(size=n) for c in chans}
    h = {c: rng.normal(size=n) + 1j * rng.normal(size=n) for c in chans}
    s = {c: rng.uniform(1, 2, n) for c in chans}
    df = 0.25
    assert det.whittle_lnl(chans, d, d, s, df) == 0.0
    zero = {c: np.zeros(n, complex) for c in chans}
    np.testing.assert_allclose(det.whittle_lnl(chans, d, zero, s, df),
                               -2 * df * sum(np.sum(np.abs(d[c]) ** 2 / s[c]) for c in chans), rtol=1e-13)
    # dd + hh + dh decomposition (likelihood.ipynb cell 6)
    dd = det.whittle_lnl(chans, d, zero, s, df)
    hh = det.whittle_lnl(chans, zero, h, s, df)
    dh = 4 * df * sum(np.sum((np.conj(d[c]) * h[c]).real / s[c]) for c in chans)
    np.testing.assert_allclose(det.whittle_lnl(chans, d, h, s, df), dd + hh + dh, rtol=1e-12)


def test_uniform_weights_reduce_to_reference_form():
    rng = np.random.default_rng(2)
    n = 500
    d = {"A": rng.normal(size=n) + 1j * rng.normal(size=n)}
    h = {"A": rng.normal(size=n) + 1j * rng.normal(size=n)}
    s = {"A": rng.uniform(1, 2, n)}
    a = det.whittle_lnl(("A",), d, h, s, 0.125)
    b = det.whittle_lnl(("A",), d, h, s, np.full(n, 0.125))
    np.testing.assert_allclose(a, b, rtol=1e-14)


def test_lnl_of_noiseless_injection_is_zero():
    f, mask, n_t, df, _ = det.data_grid(2 ** 14 * 5.0, 5.0, 1e-4, 0.05)
    psd = det.psd("LISA_SciRDv1", f, ("A", "E", "T"), False, "1.5")
    orb = det.ORBITS["LISA_analytic"]
    p = dict(chirp_mass=1e6, mass_ratio=0.5, chi_1=0.2, chi_2=-0.1, luminosity_distance=3e4, reference_phase=0.3,
             inclination=1.0, ecliptic_longitude=2.0, ecliptic_latitude=0.4, polarization=1.2,
             coalescence_time=4e4)
    d0 = olk.Detector(f, None, psd, df, orb)
    resp = olk.response(p, d0, ((2, 2),), 0.0)
    d = olk.Detector(f, resp, psd, df, orb)
    assert olk.lnl(p, [d]) == 0.0


def test_duty_window_closed_forms():
    """The observation window of a duty-cycled detector: plateau 1, off 0, smoothstep ramps
    S(x) = x^2 (3 - 2x) (S(1/4) = 5/32, S(1/2) = 1/2), periodic, C1 at the ramp ends."""
    period, on, t0, taper = 100.0, 60.0, 7.0, 10.0
    duty = (period, on, t0, taper)
    for k in (-3, 0, 5):
        base = t0 + k * period
        np.testing.assert_allclose(det.duty_window(base + np.array([10.0, 30.0, 50.0]), duty), 1.0)
        np.testing.assert_array_equal(det.duty_window(base + np.array([60.5, 75.0, 99.9]), duty), 0.0)
        np.testing.assert_allclose(det.duty_window(base + np.array([2.5, 5.0, 57.5, 55.0]), duty),
                                   [5 / 32, 0.5, 5 / 32, 0.5], rtol=1e-12)
    # slopes: zero at both ends of each ramp (C1), 1.5 / taper at mid-ramp
    h = 1e-5
    for tt, slope in ((t0 + 1e-3, 0.0), (t0 + 10.0, 0.0), (t0 + 5.0, 1.5 / taper), (t0 + 55.0, -1.5 / taper),
                      (t0 + 60.0 - 1e-3, 0.0)):
        d = (det.duty_window(tt + h, duty) - det.duty_window(tt - h, duty)) / (2 * h)
        assert abs(d - slope) < 1e-3 / taper, (tt, d, slope)
    # no ramps: a hard window
    np.testing.assert_array_equal(det.duty_window(np.array([t0, t0 + 59.9, t0 + 60.0]), (period, on, t0, 0.0)),
                                  [1.0, 1.0, 0.0])


def test_duty_cycled_links_are_the_windowed_links():
    """With a duty cycle each mode's links carry the window at its own t_f + t_c; a window that
    is 1 everywhere (on = period, no ramps) changes nothing."""
    rng = np.random.default_rng(2)
    f = np.linspace(1e-3, 1e-2, 50)
    wf = {m: {"plus": rng.standard_normal(50) + 1j * rng.standard_normal(50),
              "cross": rng.standard_normal(50) + 1j * rng.standard_normal(50),
              "tf": np.linspace(-3e5, 1e3, 50) * (i + 1)} for i, m in enumerate(("22", "33"))}
    orb = det.ORBITS["LISA_analytic"]
    args = (0.3, 0.2, 0.1, 1e5, f, orb)
    plain = det.single_links(wf, *args)
    np.testing.assert_array_equal(det.single_links(wf, *args, (1e6, 1e6, 0.0, 0.0))["link12"], plain["link12"])
    duty = (2e5, 1.2e5, 3e4, 2e4)
    got = det.single_links(wf, *args, duty)
    ref = {n: 0.0 for n in plain}
    for m, w in wf.items():
        scaled = dict(w)
        win = det.duty_window(w["tf"] + 1e5, duty)
        scaled["plus"], scaled["cross"] = w["plus"] * win, w["cross"] * win
        one = det.single_links({m: scaled}, *args)
        for n in ref:
            ref[n] = ref[n] + one[n]
    for n in plain:
        np.testing.assert_allclose(got[n], ref[n], rtol=1e-14, atol=0)

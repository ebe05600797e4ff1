"""Pins of the oracle's time-domain links and TDI (oracle/td.py).

* links of a monochromatic plane wave, sampled finely: y_rs(t) = Re[l_rs(f0) exp(2 pi i f0 t)]
  with l_rs the FD single link of the frozen constellation at t (itself pinned against the
  two-pulse formula), to within the bound of linear interpolation, dt^2 (2 pi f0)^2 / 8 per
  pulse; a one-sample offset, a swapped sender/receiver or a wrong sign fails by orders of
  magnitude more;
* links of a waveform linear in time: interpolation is exact and the two pulses give the closed
  form y_rs = -(L / 2) n.(dH/dt).n for every link (the 1 / (1 - k.n_sr) factor cancels
  exactly, so a wrong denominator or sign fails);
* zero signal where both pulses fall outside the waveform's samples;
* TDI of random link samples against the delay operator applied by numpy.interp on the time
  axis (an independent evaluation of y(t - m L)), with L / dt fractional and integral;
* TDI of tone links against the FD combination (D = exp(-2 pi i f L)) within the interpolation
  bound, for both generations and both channel sets.
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import td as otd

C_SI = 299792458.0
ORBITS = [odet.ORBITS["LISA_analytic"], odet.ORBITS["Tianqin_analytic"],
          ("geo", 1.7e8, 0.4, 1.1, 2 * np.pi / (5 * 86400.0), 0.7, 0.3)]
SKIES = [(1.375, -1.2108, 2.659), (4.0, 1.1, 1.9)]


def _tone_wave(f0, tc, hp_hat, hc_hat, t_lo, t_hi, wdt):
    tw = np.arange(t_lo, t_hi, wdt)
    hp = np.real(hp_hat * np.exp(2j * np.pi * f0 * (tw - tc)))
    hc = np.real(hc_hat * np.exp(2j * np.pi * f0 * (tw - tc)))
    return tw[0], hp, hc


@pytest.mark.parametrize("orbit", range(len(ORBITS)))
@pytest.mark.parametrize("sky", SKIES)
def test_td_links_of_a_tone_are_the_fd_links(orbit, sky):
    orb = ORBITS[orbit]
    lam, beta, psi = sky
    f0, tc, wdt = 5e-3, 1234.5, 0.02
    hp_hat, hc_hat = 0.8 - 0.3j, 0.2 + 1.1j
    t0, dt, n = 2.0e6, 7.3, 300
    wt0, hp, hc = _tone_wave(f0, tc, hp_hat, hc_hat, t0 - 700.0, t0 + dt * n + 700.0, wdt)
    y = otd.links(orb, hp, hc, wt0, wdt, lam, beta, psi, t0, dt, n)
    t = t0 + dt * np.arange(n)
    wf = {"plus": np.full(n, hp_hat), "cross": np.full(n, hc_hat), "tf": t - tc}
    fd = odet.single_links(wf, lam, beta, psi, tc, np.full(n, f0), orb)
    # interpolation bound per link and sample: 0.5 (|n.e+.n| + |n.ex.n|) 2 eps_h / |1 - k.n_sr|,
    # eps_h = dt^2 (2 pi f0)^2 |h| / 8 for each interpolated pulse
    eps_h = wdt ** 2 * (2 * np.pi * f0) ** 2 / 8 * max(abs(hp_hat), abs(hc_hat))
    p, q, k = odet.polarization_basis(lam, beta, psi)
    nv = odet.constellation(orb[0], orb[1], orb[2], orb[3], t, *orb[4:])[:3]
    scale = max(np.max(np.abs(fd[name])) for name in otd.LINK_NAMES)
    worst = 0.0
    for li, (name, (r, s)) in enumerate(zip(otd.LINK_NAMES, otd.LINK_RS)):
        na = nv[3 - r - s]
        P_sum = np.abs((p @ na) ** 2 - (q @ na) ** 2) + np.abs(2 * (p @ na) * (q @ na))
        bound = P_sum * eps_h / np.abs(1 - np.abs(k @ na))
        ref = np.real(fd[name] * np.exp(2j * np.pi * f0 * t))
        err = np.abs(y[li] - ref)
        assert np.all(err <= bound + 1e-12 * scale), (name, np.max(err - bound))
        worst = max(worst, np.max(bound))
    # the bound is well below a one-sample offset's error, ~ 2 pi f0 wdt |y|
    assert worst < 0.1 * 2 * np.pi * f0 * wdt * scale, (worst, scale)


def test_a_one_sample_offset_fails_the_tone_pin():
    orb = ORBITS[0]
    lam, beta, psi = SKIES[0]
    f0, tc, wdt = 5e-3, 1234.5, 0.05
    t0, dt, n = 2.0e6, 7.3, 50
    wt0, hp, hc = _tone_wave(f0, tc, 1.0, 0.5j, t0 - 700.0, t0 + dt * n + 700.0, wdt)
    good = otd.links(orb, hp, hc, wt0, wdt, lam, beta, psi, t0, dt, n)
    shifted = otd.links(orb, hp, hc, wt0 + wdt, wdt, lam, beta, psi, t0, dt, n)
    eps_h = wdt ** 2 * (2 * np.pi * f0) ** 2 / 8
    assert np.max(np.abs(good - shifted)) > 100 * eps_h


@pytest.mark.parametrize("orbit", range(len(ORBITS)))
@pytest.mark.parametrize("sky", SKIES)
def test_td_links_of_a_linear_waveform_closed_form(orbit, sky):
    orb = ORBITS[orbit]
    lam, beta, psi = sky
    L = orb[1] / C_SI
    bp, bc = 3.0e-3, -1.7e-3  # dh+/dt, dhx/dt
    wdt = 2.0
    tw = np.arange(-2000.0, 2.0e5, wdt)
    hp, hc = 0.7 + bp * tw, -0.2 + bc * tw
    t0, dt, n = 1000.0, 97.0, 1500
    y = otd.links(orb, hp, hc, tw[0], wdt, lam, beta, psi, t0, dt, n)
    t = t0 + dt * np.arange(n)
    nv = odet.constellation(orb[0], orb[1], orb[2], orb[3], t, *orb[4:])[:3]
    p, q, _ = odet.polarization_basis(lam, beta, psi)
    for li, (r, s) in enumerate(otd.LINK_RS):
        na = nv[3 - r - s]
        Pp = (p @ na) ** 2 - (q @ na) ** 2
        Pc = 2 * (p @ na) * (q @ na)
        ref = -0.5 * L * (Pp * bp + Pc * bc)
        # rounding: the pulses are differenced at |h| ~ 600 x the result
        np.testing.assert_allclose(y[li], ref, rtol=0, atol=1e-10 * np.max(np.abs(hp)) * L)


def test_td_links_vanish_outside_the_waveform():
    orb = ORBITS[0]
    wdt = 1.0
    hp = np.ones(100)
    hc = np.ones(100)
    y = otd.links(orb, hp, hc, 0.0, wdt, 0.3, 0.2, 0.1, 1.0e5, 10.0, 20)
    assert np.all(y == 0.0)


def _delay_by_interp(y, t_grid, t):
    return np.interp(t, t_grid, y, left=np.nan, right=np.nan)


@pytest.mark.parametrize("gen", ["1.5", "2.0"])
@pytest.mark.parametrize("orthogonal", [True, False])
@pytest.mark.parametrize("arm", [2.5e9, 8 * C_SI * 2.0])  # L / dt fractional; L / dt = 4 exactly
def test_td_tdi_of_random_samples_is_the_delay_operator(gen, orthogonal, arm):
    dt = 2.0
    L = arm / C_SI
    added = otd.added_samples(arm, dt, gen)
    assert added == int(otd.MAX_DELAY[gen] * L // dt + 1)
    rng = np.random.default_rng(11)
    n_ext = 400
    links = rng.standard_normal((6, n_ext))
    out = otd.tdi(links, arm, dt, gen, orthogonal)
    assert out.shape == (3, n_ext - added)
    tg = 100.0 + dt * np.arange(n_ext)
    t = tg[added:]
    y = dict(zip(otd.LINK_NAMES, links))
    P = [1.0, 0.0, -1.0] if gen == "1.5" else [1.0, 0.0, -1.0, 0.0, -1.0, 0.0, 1.0]

    def D(name, m):
        return _delay_by_interp(y[name], tg, t - m * L)

    chans = []
    for a, b, c, e in (("link13", "link12", "link31", "link21"), ("link21", "link23", "link12", "link32"),
                       ("link32", "link31", "link23", "link13")):
        chans.append(sum(P[m] * (D(a, m) - D(b, m) + D(c, m + 1) - D(e, m + 1)) for m in range(len(P))))
    X, Y, Z = chans
    ref = (np.stack([(Z - X) / np.sqrt(2), (X - 2 * Y + Z) / np.sqrt(6), (X + Y + Z) / np.sqrt(3)])
           if orthogonal else np.stack([X, Y, Z]))
    assert np.all(np.isfinite(ref))  # every delayed read lies inside the link samples
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("gen", ["1.5", "2.0"])
@pytest.mark.parametrize("orthogonal", [True, False])
def test_td_tdi_of_tone_links_is_the_fd_combination(gen, orthogonal):
    arm, dt, f = 2.5e9, 0.01, 2e-3
    rng = np.random.default_rng(5)
    yh = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    n_ext = 8000
    tg = 3.0e5 + dt * np.arange(n_ext)
    links = np.real(yh[:, None] * np.exp(2j * np.pi * f * tg))
    out = otd.tdi(links, arm, dt, gen, orthogonal)
    added = otd.added_samples(arm, dt, gen)
    fd = odet.tdi_combination({nm: np.array([yh[i]]) for i, nm in enumerate(otd.LINK_NAMES)}, np.array([f]),
                              arm, gen, orthogonal)
    t = tg[added:]
    n_terms = 4 * (2 if gen == "1.5" else 4)
    bound = n_terms * dt ** 2 * (2 * np.pi * f) ** 2 / 8 * np.max(np.abs(yh)) * 2
    for ci, c in enumerate(fd):
        ref = np.real(fd[c][0] * np.exp(2j * np.pi * f * t))
        assert np.max(np.abs(out[ci] - ref)) <= bound, c
        # sharp: a one-sample offset moves the channel by ~ 2 pi f dt |c| (T, which cancels to
        # O(f L)^2, is the smallest)
        assert bound < 0.2 * 2 * np.pi * f * dt * np.abs(fd[c][0]), c


def test_td_tdi_rejects_too_few_samples():
    out = otd.tdi(np.zeros((6, 10)), 2.5e9, 5.0, "1.5")
    assert out.shape == (3, 4)  # added = floor(3 L / 5) + 1 = 6

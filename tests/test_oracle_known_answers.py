"""Known-answer pins of the oracle's detector response and of its TaylorF2 / PN terms.

These compare the oracle with independent forms of the same physics, chosen so that a flipped
sign, a swapped link or a wrong coefficient fails:

* single links: the time-domain two-pulse response of a frozen constellation to a
  monochromatic plane wave (the definition the Marsat-Baker FD links are derived from) against
  the oracle's FD links at that frequency, for every link, both orbit kinds and several sky
  positions.  Pins the delay-phase sign exp(-i pi f (L + k.(x_i + x_j))), the t_c-shift sign
  exp(-2 pi i f t_c), the sinc argument of each link (1 +- k.n), the -i pi f L prefactor and the
  link labels (link_rs: sent from s, received at r; the reference's commented time-domain model,
  antenna.py:566-800, uses the same labels);
* TDI: the Michelson X, Y, Z as delay-operator combinations of time-domain links
  (X = (1 - D^2) [y13 + D y31 - y12 - D y21], D y(t) = y(t - L)) against the oracle's FD TDI;
* TaylorF2 spin terms: the 1.5PN spin-orbit coefficient against Poisson & Will's
  beta = (1/12) sum_i chi_i (113 (m_i/M)^2 + 75 eta), the 2PN spin-spin coefficient against
  Arun et al. (2009)'s sigma for aligned black-hole spins (kappa = 1), both written in the
  individual spins, and the test-mass limit (eta -> 0) of the spin terms through 3.5PN against
  t(f) derived from the exact circular-equatorial Kerr energy and the black-hole-perturbation
  flux (Tagoshi, Shibata, Tanaka & Sasaki 1996);
* PN amplitude: the non-spinning PhenomD inspiral amplitude coefficients against the
  stationary-phase amplitude |H_22(x)| sqrt(R(x)) re-expanded, with H_22 from Blanchet et al.
  (2008) and R from the PN energy and flux (Blanchet, Living Reviews 2014).
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import phenom as ph

PI = np.pi
C_SI = 299792458.0
EG = 0.5772156649015329


# --------------------------------------------------------------------------------------------
# single links and TDI against the time-domain two-pulse formula
# --------------------------------------------------------------------------------------------
# link name -> (receiver, sender), 1-based (antenna.py:24-31; the time-domain model's labels)
LINKS = {"link12": (1, 2), "link21": (2, 1), "link23": (2, 3), "link32": (3, 2), "link31": (3, 1), "link13": (1, 3)}


def _td_link(Hhat, f0, tc, xr, xs, L, k):
    """Complex amplitude at frequency f0 of the time-domain one-way link response
    y(t) = n.[H(t - L - k.x_s) - H(t - k.x_r)].n / (2 (1 - k.n)),  n = (x_r - x_s)/L,
    to the plane wave H_ij(t, x) = Hhat_ij exp(2 pi i f0 (t - k.x - t_c)) (frozen geometry).
    The two pulses nearly cancel at low f (their phases reach ~1e4 rad, their difference
    ~pi f L), so the difference is formed in 40-digit arithmetic from the float64 inputs."""
    import mpmath as mp
    mp.mp.dps = 40
    n = (xr - xs) / L
    nHn = complex(n @ Hhat @ n)
    kxs = mp.fsum(mp.mpf(float(a)) * mp.mpf(float(b)) for a, b in zip(k, xs))
    kxr = mp.fsum(mp.mpf(float(a)) * mp.mpf(float(b)) for a, b in zip(k, xr))
    kn = mp.fsum(mp.mpf(float(a)) * mp.mpf(float(b)) for a, b in zip(k, n))
    w = -2 * mp.pi * mp.mpf(float(f0)) * mp.mpc(0, 1)
    send = mp.exp(w * (mp.mpf(float(L)) + kxs + mp.mpf(float(tc))))
    rece = mp.exp(w * (kxr + mp.mpf(float(tc))))
    return nHn * complex((send - rece) / (2 * (1 - kn)))


@pytest.mark.parametrize("orbit", ["LISA_analytic", "Tianqin_analytic"])
@pytest.mark.parametrize("sky", [(1.375, -1.2108, 2.659), (0.3, 0.5, 0.1), (4.0, 1.1, 1.9)])
def test_single_links_match_time_domain_two_pulse_response(orbit, sky):
    lam, beta, psi = sky
    kind, arm, rot0, rev0 = odet.ORBITS[orbit]
    L = arm / C_SI
    rng = np.random.default_rng(3)
    f = np.array([2e-4, 3e-3, 1.7e-2, 4.1e-2])  # pi f L from 5e-4 to ~1 (LISA)
    hp = rng.standard_normal(f.size) + 1j * rng.standard_normal(f.size)
    hc = rng.standard_normal(f.size) + 1j * rng.standard_normal(f.size)
    # the constellation is frozen at t0 = tf + tc; t_c is kept short so that the oracle's own
    # float64 factor exp(-2 pi i f t_c) stays exact to ~1e-13 (at t_c ~ 3e6 s and f ~ 0.04 Hz
    # its phase alone, ~1e6 rad, rounds to ~1e-10 rad)
    tc, t0 = 2345.6, 1.9e7  # 2 pi f t_c is no multiple of pi at the test frequencies
    wf = {"plus": hp, "cross": hc, "tf": np.full(f.size, t0 - tc)}
    fd = odet.single_links(wf, lam, beta, psi, tc, f, odet.ORBITS[orbit])
    eplus, ecross = odet.polarization_tensors(lam, beta, psi)
    _, _, k = odet.polarization_basis(lam, beta, psi)
    n1, n2, n3, x1, x2, x3 = odet.constellation(kind, arm, rot0, rev0, t0)
    x = {1: x1, 2: x2, 3: x3}
    for i in range(f.size):
        Hhat = hp[i] * eplus + hc[i] * ecross
        td = {name: _td_link(Hhat, f[i], tc, x[r], x[s], L, k) for name, (r, s) in LINKS.items()}
        scale = max(abs(v) for v in td.values())  # a link whose n.H.n is small keeps absolute errors
        for name, v in td.items():
            assert abs(fd[name][i] - v) <= 1e-12 * scale, (name, f[i], fd[name][i], v)


@pytest.mark.parametrize("gen", ["1.5", "2.0"])
def test_tdi_is_the_delay_operator_combination_of_links(gen):
    """X(t) = P(D) [y13(t) + y31(t - L) - y12(t) - y21(t - L)] with P = 1 - D^2 (1.5) or
    (1 - D^2)(1 - D^4) (2.0), Y, Z cyclic; on a tone y(t - nL) = y_hat exp(-2 pi i f n L)."""
    kind, arm, rot0, rev0 = odet.ORBITS["LISA_analytic"]
    L = arm / C_SI
    rng = np.random.default_rng(4)
    f = np.array([1e-3, 1e-2, 3e-2])
    links = {n: rng.standard_normal(f.size) + 1j * rng.standard_normal(f.size) for n in LINKS}
    out = odet.tdi_combination(links, f, arm, gen, orthogonal=False)

    def delay(y, n):
        return y * np.exp(-2j * PI * f * n * L)

    def P(y):
        y1 = y - delay(y, 2)
        return y1 if gen == "1.5" else y1 - delay(y1, 4)

    l = links
    X = P(l["link13"] + delay(l["link31"], 1) - l["link12"] - delay(l["link21"], 1))
    Y = P(l["link21"] + delay(l["link12"], 1) - l["link23"] - delay(l["link32"], 1))
    Z = P(l["link32"] + delay(l["link23"], 1) - l["link31"] - delay(l["link13"], 1))
    for c, ref in (("X", X), ("Y", Y), ("Z", Z)):
        np.testing.assert_allclose(out[c], ref, rtol=1e-13, atol=0)


def test_a_flipped_delay_sign_fails_the_two_pulse_pin():
    """The pin is sharp: the same test with exp(+i pi f (...)) (a plausible slip) is far off."""
    lam, beta, psi = 0.3, 0.5, 0.1
    kind, arm, rot0, rev0 = odet.ORBITS["LISA_analytic"]
    L = arm / C_SI
    f0, tc, t0 = 1.7e-2, 2345.6, 1.9e7
    wf = {"plus": np.array([1.0 + 0.5j]), "cross": np.array([-0.3 + 1.0j]), "tf": np.array([t0 - tc])}
    fd = odet.single_links(wf, lam, beta, psi, tc, np.array([f0]), odet.ORBITS["LISA_analytic"])["link12"][0]
    eplus, ecross = odet.polarization_tensors(lam, beta, psi)
    _, _, k = odet.polarization_basis(lam, beta, psi)
    _, _, _, x1, x2, _ = odet.constellation(kind, arm, rot0, rev0, t0)
    Hhat = wf["plus"][0] * eplus + wf["cross"][0] * ecross
    td = _td_link(Hhat, f0, tc, x1, x2, L, k)
    flipped = _td_link(Hhat, -f0, -tc, x1, x2, L, k)  # conjugate propagation phases
    assert abs(fd - td) < 1e-12 * abs(td)
    assert abs(fd - flipped) > 0.1 * abs(td)


# --------------------------------------------------------------------------------------------
# truncated power series (float coefficients), for the PN derivations below
# --------------------------------------------------------------------------------------------
N_SER = 10


def smul(a, b):
    return np.convolve(a, b)[:N_SER]


def spow(a, alpha):
    """(a0 (1 + u))^alpha with a0 = a[0] = 1, by the binomial series in u."""
    assert a[0] == 1.0
    u = np.array(a, dtype=float)
    u[0] = 0.0
    out = np.zeros(N_SER)
    out[0] = 1.0
    term = np.zeros(N_SER)
    term[0] = 1.0
    coef = 1.0
    for j in range(1, N_SER):
        coef *= (alpha - j + 1) / j
        term = smul(term, u)
        out += coef * term
    return out


def sinv(a):
    return spow(a, -1.0)


def ser(*pairs):
    out = np.zeros(N_SER)
    for p, c in pairs:
        out[p] += c
    return out


def kerr_R(a):
    """R(x) = -eps'(x) / (x F_hat(x)) for a test mass on a circular equatorial Kerr orbit, as a
    series in x = (M Omega)^(1/3) (spin-dependent part exact through x^7; the spin-independent
    log of the x^6 flux term is left out: only spin parts are compared).

    eps: E/mu = (1 - 2 v^2 + a v^3) / sqrt(1 - 3 v^2 + 2 a v^3) - 1 with v = sqrt(M/r) and
    M Omega = v^3 / (1 + a v^3), i.e. v = x (1 - a x^3)^(-1/3).
    F_hat: Tagoshi, Shibata, Tanaka & Sasaki (1996) eq. (1), in v = (M Omega)^(1/3) (= x)."""
    v = smul(ser((1, 1.0)), spow(ser((0, 1.0), (3, -a)), -1.0 / 3.0))  # v(x)
    v2 = smul(v, v)
    v3 = smul(v2, v)
    num = ser((0, 1.0)) - 2.0 * v2 + a * v3
    den = spow(ser((0, 1.0)) - 3.0 * v2 + 2.0 * a * v3, -0.5)
    eps = smul(num, den) - ser((0, 1.0))
    deps = np.array([k * eps[k] for k in range(1, N_SER)] + [0.0])  # d eps / dx
    mdeps_over_x = -np.append(deps[1:], 0.0)  # -eps'(x) / x
    F = ser((0, 1.0), (2, -1247.0 / 336.0), (3, 4 * PI - 11.0 / 4.0 * a), (4, -44711.0 / 9072.0 + 33.0 / 16.0 * a * a),
            (5, -8191.0 * PI / 672.0 - 59.0 / 16.0 * a), (6, -65.0 * PI / 6.0 * a + 611.0 / 504.0 * a * a),
            (7, -16285.0 * PI / 504.0 + 162035.0 / 3888.0 * a + 65.0 * PI / 8.0 * a * a - 71.0 / 24.0 * a ** 3))
    return smul(mdeps_over_x, sinv(F))


def _tau_from_R(R):
    """t(x) = -(5 M / (256 eta)) x^-8 sum_k tau_k x^k from dt/dx = (5 M / (32 eta)) x^-9 R(x)."""
    return {k: 8.0 * R[k] / (8.0 - k) for k in range(0, 8)}


def _tau_from_psi(psi):
    """The same bracket from the TaylorF2 phase Psi = (3 / (128 eta x^5)) sum_k psi_k x^k
    (psi_5 multiplying 1 + 3 ln x): dPsi/dMf = 2 pi t(f)."""
    tau = {k: psi[k] * (5.0 - k) / 5.0 for k in (0, 1, 2, 3, 4, 6, 7)}
    tau[5] = -0.6 * psi[5]
    return tau


def _psi_of(eta, chi1, chi2, delta):
    p = ph.tf2_coeffs(eta, chi1, chi2, delta)
    cf = -3.0 / 128.0 / eta
    psi = {k: p[k] * PI ** ((5.0 - k) / 3.0) / cf for k in (0, 1, 2, 3, 4, 6, 7)}
    psi[5] = p["5l"] / cf  # the coefficient of the 1 + 3 ln x factor (its f-dependence: ln f)
    return psi


def _m(q):
    m1 = 1.0 / (1.0 + q)
    m2 = q * m1
    eta = m1 * m2
    return m1, m2, eta, m1 - m2


@pytest.mark.parametrize("q,c1,c2", [(0.6, 0.75, 0.62), (0.15, -0.8, 0.3), (0.93, 0.2, -0.95), (1.0, 0.5, 0.5)])
def test_taylorf2_spin_orbit_15pn_is_poisson_will_beta(q, c1, c2):
    m1, m2, eta, d = _m(q)
    psi = _psi_of(eta, c1, c2, d)
    beta = (c1 * (113.0 * m1 * m1 + 75.0 * eta) + c2 * (113.0 * m2 * m2 + 75.0 * eta)) / 12.0
    assert abs(psi[3] - (-16.0 * PI + 4.0 * beta)) < 1e-12 * (16 * PI)


@pytest.mark.parametrize("q,c1,c2", [(0.6, 0.75, 0.62), (0.15, -0.8, 0.3), (0.93, 0.2, -0.95), (1.0, 0.5, 0.5)])
def test_taylorf2_spin_spin_2pn_is_arun_sigma(q, c1, c2):
    """psi_4 = 15293365/508032 + 27145 eta/504 + 3085 eta^2/72 - 10 sigma with, for aligned
    black-hole spins, sigma = (eta/48)(721 - 247) chi1 chi2 + sum_i (m_i/M)^2 (1/96)
    ((720 kappa - 1) - (240 kappa - 7)) chi_i^2, kappa = 1 (Arun, Buonanno, Faye & Ochsner 2009)."""
    m1, m2, eta, d = _m(q)
    psi = _psi_of(eta, c1, c2, d)
    sigma = eta / 48.0 * (721.0 - 247.0) * c1 * c2 + (m1 * m1 * c1 * c1 + m2 * m2 * c2 * c2) * (719.0 - 233.0) / 96.0
    book = 15293365.0 / 508032.0 + 27145.0 * eta / 504.0 + 3085.0 * eta ** 2 / 72.0 - 10.0 * sigma
    assert abs(psi[4] - book) < 1e-12 * abs(book)


@pytest.mark.parametrize("a", [-0.9, -0.3, 0.4, 0.95])
def test_taylorf2_spin_terms_test_mass_kerr_limit(a):
    """eta -> 0 with the heavy body's spin a: the spin-dependent parts of the oracle's t(f)
    coefficients against those derived from the Kerr energy and flux: in full through 3PN
    (spin-orbit at 1.5, 2.5 and 3PN, spin-squared at 2 and 3PN), and the part odd in a at 3.5PN
    (spin-orbit and spin-cubed).  The 3.5PN spin-squared term (the -570 pi (chi_a^2 + chi_s^2)
    - 1140 pi delta chi_a chi_s terms of interface_lalsim.py:353-433, which the oracle restates)
    does not reach the test-mass value: it equals the value with no q^2 term in the v^7 flux,
    where Tagoshi et al. give 65 pi q^2 / 8; that term is left unpinned."""
    eta = 1e-12
    d = np.sqrt(1.0 - 4.0 * eta)

    def spin_parts(sign):
        tc = _tau_from_psi(_psi_of(eta, sign * a, 0.0, d))
        tc0 = _tau_from_psi(_psi_of(eta, 0.0, 0.0, d))
        tk = _tau_from_R(kerr_R(sign * a))
        tk0 = _tau_from_R(kerr_R(0.0))
        return {k: tc[k] - tc0[k] for k in tc}, {k: tk[k] - tk0[k] for k in tk}

    code, kerr = spin_parts(1.0)
    code_m, kerr_m = spin_parts(-1.0)
    for k in (3, 4, 5, 6):
        assert abs(code[k] - kerr[k]) < 1e-6 * max(1.0, abs(kerr[k])), (k, code[k], kerr[k])
    odd_code, odd_kerr = 0.5 * (code[7] - code_m[7]), 0.5 * (kerr[7] - kerr_m[7])
    assert abs(odd_code - odd_kerr) < 1e-6 * max(1.0, abs(odd_kerr)), (odd_code, odd_kerr)
    # the one term left unpinned, measured: the even (spin-squared) 3.5PN parts differ by
    # 8 * (65 pi / 8) a^2 in tau_7 (= 8 R_7)
    even_code, even_kerr = 0.5 * (code[7] + code_m[7]), 0.5 * (kerr[7] + kerr_m[7])
    assert abs((even_code - even_kerr) - 65.0 * PI * a * a) < 1e-6 * 65.0 * PI


def test_taylorf2_nonspinning_test_mass_limit_from_kerr_series():
    """The same derivation at a = 0 reproduces the non-spinning 1PN, 1.5PN and 2.5PN t(f)
    coefficients of the oracle at eta -> 0 (a check of the derivation itself)."""
    eta = 1e-12
    tau_code = _tau_from_psi(_psi_of(eta, 0.0, 0.0, np.sqrt(1.0 - 4.0 * eta)))
    tau_kerr = _tau_from_R(kerr_R(0.0))
    for k in (0, 2, 3, 4, 5):
        assert abs(tau_code[k] - tau_kerr[k]) < 1e-6 * max(1.0, abs(tau_kerr[k])), (k, tau_code[k], tau_kerr[k])


def _pn_R(eta):
    """R(x) = -E'(x) / (x F_hat(x)) from the non-spinning PN energy and flux through 2.5PN
    (Blanchet, Living Rev. Relativ. 17, 2 (2014), eqs. 233 and 314), E = -(mu x / 2)(1 + ...)."""
    e = ser((0, 1.0), (1, -0.75 - eta / 12.0), (2, -27.0 / 8.0 + 19.0 * eta / 8.0 - eta * eta / 24.0))
    # E(x) ~ -x/2 e(x) in x = v^2; as a series in v: E(v) = -(v^2/2)(1 + e1 v^2 + e2 v^4)
    Ev = np.zeros(N_SER)
    Ev[2], Ev[4], Ev[6] = -0.5, -0.5 * e[1], -0.5 * e[2]
    dE = np.array([k * Ev[k] for k in range(1, N_SER)] + [0.0])
    mdE_over_v = -np.append(dE[1:], 0.0)
    F = ser((0, 1.0), (2, -1247.0 / 336.0 - 35.0 * eta / 12.0), (3, 4.0 * PI),
            (4, -44711.0 / 9072.0 + 9271.0 * eta / 504.0 + 65.0 * eta * eta / 18.0),
            (5, (-8191.0 / 672.0 - 583.0 * eta / 24.0) * PI))
    return smul(mdE_over_v, sinv(F))


@pytest.mark.parametrize("q", [1.0, 0.6, 0.2])
def test_pn_amplitude_nonspinning_through_25pn(q):
    """|h~22(f)| / (leading f^-7/6 term) = |H_22(v)| sqrt(R(v)) re-expanded: H_22 real part
    (Blanchet, Faye, Iyer & Sinha 2008) 1 + v^2 (-107/42 + 55 eta/42) + 2 pi v^3
    + v^4 (-2173/1512 - 1069 eta/216 + 2047 eta^2/1512) + v^5 (-107 pi/21 + 34 pi eta/21)."""
    m1, m2, eta, d = _m(q)
    a = ph.amp_pn_prefactors(eta, 0.0, 0.0, d)
    H = ser((0, 1.0), (2, -107.0 / 42.0 + 55.0 * eta / 42.0), (3, 2.0 * PI),
            (4, -2173.0 / 1512.0 - 1069.0 * eta / 216.0 + 2047.0 * eta * eta / 1512.0),
            (5, -107.0 * PI / 21.0 + 34.0 * PI * eta / 21.0))
    book = smul(H, spow(_pn_R(eta), 0.5))
    # the oracle's coefficients multiply (pi Mf)^(k/3) = v^k: a["2/3"] v^2, a["1"] v^3, ...
    keys = {2: "2/3", 3: "1", 4: "4/3", 5: "5/3"}
    for k, key in keys.items():
        code = a[key] / PI ** (k / 3.0)
        assert abs(code - book[k]) < 1e-12 * max(1.0, abs(book[k])), (k, code, book[k])

"""Pins of the oracle waveform (PhenomD / PhenomHM restatement) and QNM table.

The reference holds no PhenomD/HM implementation, so these tests pin the oracle
against what mathematics and published closed forms fix: textbook PN limits,
continuity at region boundaries, exact scalings, mode symmetries and Kerr QNM
values.  The calibrated fit coefficients themselves stay parity-unpinned.
"""

import numpy as np
import pytest

from oracle import dual as du
from oracle import phenom as ph
from oracle import qnm

PI = np.pi
EG = 0.5772156649015329


def _pd(q=0.6, c1=0.75, c2=0.62):
    M = 1.0
    m1 = M / (1 + q)
    return ph.PhenomD(m1, q * m1, c1, c2)


def test_taylorf2_phase_derivative_matches_textbook_nonspinning():
    """dPsi/dMf of the PN part vs the TaylorF2 phase of Buonanno et al. 2009 eq. (3.18) in v = (pi Mf)^(1/3)."""
    eta = 0.21
    p = ph.tf2_coeffs(eta, 0.0, 0.0)

    def psi_code(Mf):  # the oracle's PN phase (phi_ins without the sigma terms)
        f = du.D(Mf, 1.0)
        f13 = f ** (1 / 3)
        val = -(p[0] / f13 ** 5 + p[2] / f + p[3] / f13 ** 2 + p[4] / f13 + p["5l"] * du.log(f) + p[6] * f13
                + p["6l"] * f13 * du.log(PI * f) + p[7] * f13 ** 2)
        return val.d

    def psi_book(Mf):
        v = (PI * Mf) ** (1 / 3)
        lv = np.log(v)
        a = [1.0, 0.0, 3715 / 756 + 55 * eta / 9, -16 * PI,
             15293365 / 508032 + 27145 * eta / 504 + 3085 * eta ** 2 / 72,
             PI * (38645 / 756 - 65 * eta / 9) * (1 + 3 * lv),
             11583231236531 / 4694215680 - 640 * PI ** 2 / 3 - 6848 * EG / 21 - 6848 / 21 * np.log(4 * v)
             + (-15737765635 / 3048192 + 2255 * PI ** 2 / 12) * eta + 76055 * eta ** 2 / 1728
             - 127825 * eta ** 3 / 1296,
             PI * (77096675 / 254016 + 378515 * eta / 1512 - 74045 * eta ** 2 / 756)]
        return 3 / (128 * eta * v ** 5) * sum(c * v ** k for k, c in enumerate(a))

    Mf = np.geomspace(1e-5, 1e-2, 50)
    h = 1e-7 * Mf
    book = (psi_book(Mf + h) - psi_book(Mf - h)) / (2 * h)  # the v^5 log term makes a closed form messy
    np.testing.assert_allclose(psi_code(Mf), book, rtol=1e-7)


def test_newtonian_limits():
    pd = _pd()
    Mf = np.array([1e-6, 2e-6])
    A = pd.amplitude(du.D(Mf, 1.0))
    # A ~ sqrt(2 eta/3) pi^(-1/6) Mf^(-7/6) (1 + (-323/224 + 451 eta/168) v^2 + O(v^3))
    v2 = (PI * Mf) ** (2 / 3)
    lead = np.sqrt(2 * pd.eta / 3) * PI ** (-1 / 6) * Mf ** (-7 / 6)
    np.testing.assert_allclose(A.v / lead, 1 + (-323 / 224 + 451 * pd.eta / 168) * v2, rtol=2e-5)
    # t(f) = (1/2pi) dPsi/dMf -> -(5/256)(1/eta) v^-8 at leading order (in units of M)
    P = pd.phase(du.D(Mf, 1.0))
    tN = -5 / 256 / pd.eta * (PI * Mf) ** (-8 / 3) * 2 * PI
    np.testing.assert_allclose(P.d / tN, 1 + (4 / 3) * (743 / 336 + 11 * pd.eta / 4) * v2, rtol=2e-4)


@pytest.mark.parametrize("q,c1,c2", [(0.6, 0.75, 0.62), (1.0, 0.0, 0.0), (0.125, -0.9, 0.5), (0.3, 0.9, 0.9)])
def test_phenomd_continuity(q, c1, c2):
    pd = _pd(q, c1, c2)
    eps = 1e-7
    for fj in (ph.AMP_F_INS, pd.fmax):
        a = pd.amplitude(du.D(np.array([fj * (1 - eps), fj * (1 + eps)]), 1.0))
        assert abs(a.v[1] / a.v[0] - 1) < 1e-5
        assert abs(a.d[1] / a.d[0] - 1) < 1e-4
    for fj in (ph.PHI_F_INS, 0.5 * pd.fRD):
        p = pd.phase(du.D(np.array([fj * (1 - eps), fj * (1 + eps)]), 1.0))
        # value continuous: the jump across the join equals the slope times the step
        jump = (p.v[1] - p.v[0]) - 0.5 * (p.d[0] + p.d[1]) * (2 * eps * fj)
        assert abs(jump) < 1e-9 * max(1.0, abs(p.v[0]))
        assert abs(p.d[1] / p.d[0] - 1) < 1e-5


def test_final_spin_and_energy_limits():
    # non-spinning equal mass: a_f ~ 0.686, E_rad ~ 0.048 (Husa et al. 2016 fits)
    assert abs(ph.final_spin(0.25, 0.0, 0.0) - 0.6864) < 2e-3
    assert abs(ph.radiated_energy(0.25, 0.0, 0.0) - 0.0484) < 1e-3
    # test-particle limit eta -> 0: a_f -> chi1, E_rad -> 0
    assert abs(ph.final_spin(1e-6, 0.3, 0.0) - 0.3) < 1e-4


def _params(**kw):
    p = dict(chirp_mass=1e6, mass_ratio=0.6, chi_1=0.3, chi_2=-0.2, luminosity_distance=1e4, reference_phase=0.7,
             inclination=0.9)
    p.update(kw)
    return p


def test_mass_and_distance_scaling():
    f = np.geomspace(1e-4, 3e-2, 200)
    lam = 1.7
    a = ph.waveform_direct(_params(), f * lam, ((2, 2),), 0.0)
    b = ph.waveform_direct(_params(chirp_mass=1e6 * lam), f, ((2, 2),), 0.0)
    np.testing.assert_allclose(b["plus"], lam ** 2 * a["plus"], rtol=1e-9, atol=0)
    np.testing.assert_allclose(b["tf"], lam * a["tf"], rtol=1e-9)
    c = ph.waveform_direct(_params(luminosity_distance=2e4), f, ((2, 2),), 0.0)
    d = ph.waveform_direct(_params(), f, ((2, 2),), 0.0)
    np.testing.assert_allclose(c["cross"], 0.5 * d["cross"], rtol=1e-14)


def test_harmonic_factors_22_closed_form():
    # conventions_difference.ipynb cell 6 (phi_ref = 0)
    for iota in (0.0, 0.4, 1.3, 2.8):
        kp, kc = ph.harmonic_factors((2, 2), iota, 0.0)
        np.testing.assert_allclose(kp, -0.125 * np.sqrt(5 / PI) * (1 + np.cos(iota) ** 2), atol=1e-15)
        np.testing.assert_allclose(kc, 0.125j * np.sqrt(5 / PI) * 2 * np.cos(iota), atol=1e-15)


def test_spin_weighted_harmonics_closed_forms():
    th = np.linspace(0.05, 3.1, 7)
    c, s = np.cos(th / 2), np.sin(th / 2)
    closed = {
        (2, 2): np.sqrt(5 / (64 * PI)) * (1 + np.cos(th)) ** 2,
        (2, 1): np.sqrt(5 / (16 * PI)) * np.sin(th) * (1 + np.cos(th)),
        (2, -2): np.sqrt(5 / (64 * PI)) * (1 - np.cos(th)) ** 2,
        (3, 3): -np.sqrt(21 / (2 * PI)) * c ** 5 * s,
        (3, 2): np.sqrt(7 / PI) * c ** 4 * (3 * np.cos(th) - 2) / 2,
        (4, 4): 3 * np.sqrt(7 / PI) * c ** 6 * s ** 2,
        (4, 3): -3 * np.sqrt(7 / (2 * PI)) * c ** 5 * (2 * np.cos(th) - 1) * s,
        (3, -3): np.sqrt(21 / (2 * PI)) * c * s ** 5,
    }
    for (l, m), v in closed.items():
        np.testing.assert_allclose(ph.sw_harmonic_theta(l, m, th), v, atol=1e-14, err_msg=str((l, m)))


def test_hm_22_is_phenomd_and_odd_modes_vanish_for_equal_bodies():
    f = np.geomspace(1e-4, 3e-2, 300)
    a = ph.waveform_direct(_params(), f, ((2, 2),), 0.0)
    b = ph.waveform_direct(_params(), f, ph.MODES, 0.0)["22"]
    assert np.array_equal(a["plus"], b["plus"]) and np.array_equal(a["tf"], b["tf"])
    eq = ph.waveform_direct(_params(mass_ratio=1.0, chi_1=0.4, chi_2=0.4), f, ph.MODES, 0.0)
    for m in ("21", "33", "43"):
        assert np.max(np.abs(eq[m]["plus"])) == 0.0
    assert np.max(np.abs(eq["32"]["plus"])) > 0 and np.max(np.abs(eq["44"]["plus"])) > 0


def test_hm_inspiral_phase_is_pn_scaled_22():
    hm = ph.PhenomHM(1e6, 0.6, 0.3, -0.2, 0.0, ph.MODES)
    Mf = np.geomspace(1e-5, 4e-3, 50)  # below every mode's f_i = (m/2) 0.014
    x = du.D(Mf, 1.0)
    _, p22 = hm.amp_phase((2, 2), du.D(2 * Mf / 3, 2 / 3))
    _, p33 = hm.amp_phase((3, 3), x)
    np.testing.assert_allclose(p33.v, 1.5 * p22.v, rtol=1e-13)
    np.testing.assert_allclose(p33.d, 1.5 * p22.d, rtol=1e-12)  # t_33(f) = t_22(2f/3)


def test_hm_frequency_map_monotone_and_c1():
    for q, c1, c2 in [(0.6, 0.75, 0.62), (0.125, -0.9, 0.5), (1.0, 0.0, 0.0)]:
        hm = ph.PhenomHM(1e6, q, c1, c2, 0.0, ph.MODES)
        for md in ph.MODES[1:]:
            Mf = np.linspace(1e-4, hm.mode_cut(md), 20001)
            m = hm.fmap(md, du.D(Mf, 1.0))
            assert np.all(np.diff(m.v) > 0) and np.all(m.d > 0)
            # C1: no jump in the derivative between neighbouring samples beyond the curvature
            assert np.max(np.abs(np.diff(m.d))) < 1e-2


def test_qnm_table_pins():
    tab = ph.qnm_table()["modes"]
    a = qnm.TABLE_AMIN + qnm.TABLE_STEP * np.arange(qnm.TABLE_N)
    i0 = int(np.argmin(np.abs(a)))
    for (l, m) in qnm.MODES:
        w = qnm.SCHWARZSCHILD[l]
        assert abs(tab[f"{l}{m}"]["fring"][i0] * 2 * PI - w.real) < 2e-6
        assert abs(tab[f"{l}{m}"]["fdamp"][i0] * 2 * PI + w.imag) < 2e-6
    # Berti-Cardoso-Will fits (PRD 73 064030 table VIII) hold to their quoted few-percent accuracy
    bcw = {(2, 2): (1.5251, -1.1568, 0.1292, 0.7000, 1.4187, -0.4990),
           (3, 3): (1.8956, -1.3043, 0.1818, 0.9000, 2.3430, -0.4810),
           (4, 4): (2.3000, -1.5056, 0.2244, 1.1929, 3.1191, -0.4825),
           (2, 1): (0.6000, -0.2339, 0.4175, -0.3000, 2.3561, -0.2277)}
    for (l, m), (f1, f2, f3, q1, q2, q3) in bcw.items():
        sel = (a >= 0.0) & (a <= 0.95)
        wr = f1 + f2 * (1 - a[sel]) ** f3
        Q = q1 + q2 * (1 - a[sel]) ** q3
        np.testing.assert_allclose(np.array(tab[f"{l}{m}"]["fring"])[sel] * 2 * PI, wr, rtol=0.03)
        np.testing.assert_allclose(np.array(tab[f"{l}{m}"]["fdamp"])[sel] * 2 * PI, wr / (2 * Q), rtol=0.05)
    # known Kerr value: (2,2,0) at a = 0.9 is 0.671614 - 0.064869 i
    i9 = int(np.argmin(np.abs(a - 0.9)))
    assert abs(tab["22"]["fring"][i9] * 2 * PI - 0.671614) < 2e-6
    assert abs(tab["22"]["fdamp"][i9] * 2 * PI - 0.064869) < 2e-6


def test_qnm_prograde_retrograde_symmetry():
    # w_{l,m}(a) = -conj(w_{l,-m}(-a)): the table's a < 0 branch for m > 0 is the m < 0 branch at |a|
    wp = qnm.qnm_branch(2, 2, [-0.1, -0.2, -0.3])
    wn = qnm.qnm_branch(2, -2, [0.1, 0.2, 0.3])
    np.testing.assert_allclose(np.array(wp), np.array(wn), rtol=1e-9)


def test_qnm_interpolation_is_exact_on_grid_and_smooth():
    vals = np.array(ph.qnm_table()["modes"]["33"]["fring"])
    a = qnm.TABLE_AMIN + qnm.TABLE_STEP * np.arange(qnm.TABLE_N)
    np.testing.assert_allclose(qnm.interp_table(vals, a[5:-5]), vals[5:-5], rtol=1e-13)
    mid = 0.5 * (a[10:-10] + a[11:-9])
    w = qnm.qnm_branch(3, 3, [0.3025])[0] / (2 * PI)
    assert abs(qnm.interp_table(vals, 0.3025) - w.real) < 1e-8
    assert np.all(np.diff(qnm.interp_table(vals, mid)) > 0)


def test_hm_c1_construction_changes_nothing_below_the_joints():
    """The product's C1 construction departs from London et al.'s published one (piecewise-linear
    map, ratio frozen at f_i, no taper) only above each mode's inspiral joint f_i = (m/2) 0.014 / M
    (the (2,2) only inside its taper, the last 10% below its cut): below, the two agree to
    rounding; above, the higher modes differ (the deviation is measured, not asserted small)."""
    p = _params(chirp_mass=1.2e6, mass_ratio=0.6, chi_1=0.75, chi_2=0.62)
    f = np.geomspace(1e-5, 5e-2, 4000)
    a = ph.waveform_direct(p, f, ph.MODES, 0.0, construction="c1")
    b = ph.waveform_direct(p, f, ph.MODES, 0.0, construction="published")
    hm = ph.PhenomHM(p["chirp_mass"], p["mass_ratio"], p["chi_1"], p["chi_2"], 0.0, ph.MODES)
    Mf = f * hm.Msec
    for (l, m) in ph.MODES:
        key = f"{l}{m}"
        below = Mf < (0.5 * m * ph.HM_MAP_JOIN if (l, m) != (2, 2) else (1 - ph.TAPER) * ph.MF_CUT)
        np.testing.assert_allclose(a[key]["plus"][below], b[key]["plus"][below], rtol=1e-12, atol=0, err_msg=key)
        np.testing.assert_allclose(a[key]["tf"][below], b[key]["tf"][below], rtol=1e-12, atol=0, err_msg=key)
        if (l, m) != (2, 2):
            above = ~below & (np.abs(b[key]["plus"]) > 0)
            assert np.max(np.abs(a[key]["plus"][above] - b[key]["plus"][above])) > 1e-3 * np.abs(b[key]["plus"]).max()

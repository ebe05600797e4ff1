"""The reference's path methods beyond lnL, through the C-ABI, against the numpy oracle.

* ``get_constellation_vectors`` / ``arm_length_sec`` of both orbit kinds, including a
  geocentric orbit with a rotation rate and reference source other than TianQin's
  (orbit.py:67-152, 260-322, 363-419);
* the six single links of a hip waveform and of a caller-supplied container, and the TDI
  combination of stored links (antenna.py:403-543, tdi.py:1705-1793), through the
  reference's two-stage ``update_detector_response`` and the JAX twins' ``get_response_jit``
  / ``get_combination_jit``;
* a detector on a non-TianQin geocentric orbit in the fused lnL;
* grids above pi f L = 4 (to 0.5 Hz), where the transfer-function sines are large;
* scratch pinning for HIP-graph capture.
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import INJ, add_to_core, inj_row, make_detector, particle_cloud

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)
GEO_CUSTOM = ("geo", 1.7e8, 0.4, 1.1, 2 * np.pi / (5.0 * 86400.0), 0.7, 0.3)


def _core(modes, dets, f_lo=1e-4, f_hi=0.05, n_nodes=1024):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=n_nodes, f_node_min=f_lo, f_node_max=f_hi)
    for d, labels in dets:
        add_to_core(core, d, labels)
    return core


def _relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("orbit", ["LISA_analytic", "Taiji_analytic", "Tianqin_analytic", "custom_geo"])
def test_constellation_vectors(orbit):
    from pespace_amd.detector.orbit import KaplerianHeliocentric, KeplerianGeocentric, available_orbit_models
    if orbit == "custom_geo":
        _, arm, rot0, rev0, om, lr, br = GEO_CUSTOM
        model = KeplerianGeocentric(arm, rot0, rev0, omega_rotation=om, lambda_ref=lr, beta_ref=br)
        spec = GEO_CUSTOM
    else:
        model = available_orbit_models()[orbit]
        spec = odet.ORBITS[orbit]
    assert isinstance(model, (KaplerianHeliocentric, KeplerianGeocentric))
    t = np.linspace(-3e6, 4e7, 997)
    v = model.get_constellation_vectors(t)
    ref = odet.constellation(*spec[:4], t, *spec[4:])
    # the rotation / revolution angle omega t + phi0 is rounded once (ulp ~ 1e-13 at ~600 rad):
    # the link vectors agree to a few ulp of the largest angle
    omega = (spec[4] if len(spec) > 4 else odet.GEO_OMEGA_ROT) if spec[0] == "geo" else 2 * np.pi / 31557600.0
    ang = omega * np.abs(t).max() + 2 * np.pi
    for name, got, r in zip(v._fields, v, ref):
        got = got.cpu().numpy()
        assert got.shape == (t.size, 3)
        if name.startswith("n"):
            atol = 4 * np.finfo(float).eps * max(ang, 1.0)
        else:
            atol = 1e-13 * np.abs(r).max()
        np.testing.assert_allclose(got, r.T, rtol=0, atol=atol, err_msg=name)
    # unit link vectors, and the constant arm length in seconds
    n = np.stack([x.cpu().numpy() for x in v[:3]])
    np.testing.assert_allclose(np.linalg.norm(n, axis=-1), 1.0, atol=1e-14)
    assert model.arm_length_sec(1.0e7) == spec[1] / 299792458.0


@pytest.mark.parametrize("orbit", ["LISA_analytic", "custom_geo"])
def test_links_and_tdi_stages(orbit):
    """update_detector_response in the reference's two stages: the stored links
    (single_link_response_array) and the TDI channels against the oracle, and the JAX-style
    pure functions equal to the stored state."""
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna
    from pespace_amd.detector.orbit import KeplerianGeocentric, available_orbit_models
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    from pespace_amd.waveform.phenom import IMRPhenomHM
    modes = oph.MODES
    if orbit == "custom_geo":
        _, arm, rot0, rev0, om, lr, br = GEO_CUSTOM
        model, spec = KeplerianGeocentric(arm, rot0, rev0, omega_rotation=om, lambda_ref=lr, beta_ref=br), GEO_CUSTOM
        psd_model, fmin = "Tianqin_GWSpace", 1e-3
    else:
        model, spec = available_orbit_models()[orbit], odet.ORBITS[orbit]
        psd_model, fmin = "LISA_SciRDv1", 1e-4
    td = TDIChannelData(backend="hip")
    td.set_fd_data_from_zero(("A", "E", "T"), 655360.0, 5.0, 0.0, fmin, 0.05)
    td.set_fd_noise_power_density_from_model(psd_model, tdi_generation="1.5")
    det = InterferometerAntenna("det", td, model, FDResponseModelMarset2018(), FDMichelsonConstantEqualArm("1.5", True))
    f = np.asarray(td.frequency_samples)
    wf = IMRPhenomHM(f, 0.0, modes=[f"{l}{m}" for l, m in modes], n_nodes=INTERP[2], f_node_min=INTERP[0],
                     f_node_max=INTERP[1])
    wf.update_waveform(INJ)
    lam, beta, psi, tc = (INJ[k] for k in ("ecliptic_longitude", "ecliptic_latitude", "polarization",
                                            "coalescence_time"))
    det.update_detector_response(wf, lam, beta, psi, tc)
    owf = olk.waveform(INJ, f, modes, 0.0, INTERP)
    ref_links = odet.single_links(owf, lam, beta, psi, tc, f, spec)
    got = det.single_link_response_array
    for name in got._fields:
        assert _relnorm(getattr(got, name), ref_links[name]) < 1e-10, name
    ref_tdi = odet.tdi_combination(ref_links, f, spec[1], "1.5", True)
    for c in ("A", "E", "T"):
        tol = 1e-10
        if c == "T":  # (X+Y+Z)/sqrt3 cancels to ~ (pi f L)^2 |X|: both sides carry eps |X| / |T|
            xyz = np.sqrt(np.sum(np.abs(ref_tdi["A"]) ** 2 + np.abs(ref_tdi["E"]) ** 2))
            tol = max(tol, 64 * np.finfo(float).eps * xyz / np.linalg.norm(ref_tdi["T"]))
        assert _relnorm(getattr(det.tdi_response_numpy, c), ref_tdi[c]) < tol, c
    # a reference-style container (the oracle's own waveform) through the container entry points
    links_jit = det.response_model.get_response_jit(owf, lam, beta, psi, tc)
    for name, x in zip(got._fields, links_jit):
        assert _relnorm(x.cpu().numpy(), ref_links[name]) < 1e-12, name
    tdi_jit = det.tdi_combination.get_combination_jit(links_jit)
    for c, x in zip(("A", "E", "T"), tdi_jit):
        tol = 1e-12
        if c == "T":
            xyz = np.sqrt(np.sum(np.abs(ref_tdi["A"]) ** 2 + np.abs(ref_tdi["E"]) ** 2))
            tol = max(tol, 64 * np.finfo(float).eps * xyz / np.linalg.norm(ref_tdi["T"]))
        assert _relnorm(x.cpu().numpy(), ref_tdi[c]) < tol, c


def test_custom_geocentric_orbit_lnl():
    """A detector on a geocentric orbit with its own rotation rate and reference source: response
    and lnL of the fused kernel against the oracle, and the gradient against finite differences
    of the HIP lnL."""
    modes = oph.MODES
    d, labels = make_detector(orbit=GEO_CUSTOM, psd_model="Tianqin_GWSpace", fmin=1e-3, modes=modes)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(3)])
    resp = core.response(0, rows).cpu().numpy()
    for p in range(rows.shape[0]):
        ref = olk.response(olk.as_dict(rows[p]), d, modes, 0.0, INTERP)
        for ci, c in enumerate(labels[:2]):
            assert _relnorm(resp[p, ci], ref[c]) < 1e-10
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), olk.lnl_batch(rows, [d], modes, 0.0, INTERP),
                               rtol=1e-10)
    # the same rows on TianQin's own rotation differ: the orbit constants reach the kernel
    d_tq, _ = make_detector(orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3, modes=modes)
    core_tq = _core(modes, [(d_tq, labels)])
    assert np.max(np.abs(core_tq.response(0, rows[:1]).cpu().numpy() - resp[:1])) > 1e-3 * np.abs(resp[:1]).max()
    # the adjoint through the custom rotation rate (time adjoint omega (p.U d/d(p.V) - ...))
    from tests.test_gpu_grad import _fd_check_core
    _fd_check_core(core, rows[1:3])


def test_links_match_fused_response():
    """pes_links + pes_tdi_from_links reproduce the fused response (A, E to 1e-12; T within the
    cancellation of (X+Y+Z)/sqrt3 that the reference's own formula carries)."""
    modes = oph.MODES
    d, labels = make_detector(modes=modes)
    core = _core(modes, [(d, labels)])
    rows = particle_cloud(4)
    links = core.links(0, rows)
    two_stage = core.tdi_from_links(0, links).cpu().numpy()
    fused = core.response(0, rows).cpu().numpy()
    for p in range(rows.shape[0]):
        assert _relnorm(two_stage[p, 0], fused[p, 0]) < 1e-12
        assert _relnorm(two_stage[p, 1], fused[p, 1]) < 1e-12
        assert _relnorm(two_stage[p, 2], fused[p, 2]) < 1e-8
        ref_links = odet.single_links(olk.waveform(olk.as_dict(rows[p]), d.f, modes, 0.0, INTERP),
                                      *(rows[p, i] for i in (7, 8, 9, 10)), d.f, d.orbit)
        for i, name in enumerate(("link12", "link21", "link23", "link32", "link31", "link13")):
            assert _relnorm(links[p, i].cpu().numpy(), ref_links[name]) < 1e-10


@pytest.mark.parametrize("orbit,psd_model", [("LISA_analytic", "LISA_SciRDv1"), ("Taiji_analytic", "Taiji_TDC")])
def test_high_frequency_grid(orbit, psd_model):
    """Grids to 0.5 Hz (dt = 1 s): pi f L reaches 13.1 (LISA) and 15.7 (Taiji); a light source
    (M ~ 1e5 Msun) keeps its modes up there."""
    modes = oph.MODES
    interp = (1e-4, 0.5, 1024)
    light = dict(chirp_mass=4.0e4)
    d, labels = make_detector(orbit=orbit, psd_model=psd_model, T=16384.0, dt=1.0, fmax=0.5, modes=modes,
                              interp=interp)
    L = odet.ORBITS[orbit][1] / 299792458.0
    assert np.pi * d.f[-1] * L > 4.0
    core = _core(modes, [(d, labels)], f_hi=0.5)
    rows = np.vstack([inj_row(**light), particle_cloud(3, center=inj_row(**light))])
    resp = core.response(0, rows).cpu().numpy()
    hi = d.f > 0.15
    for p in range(rows.shape[0]):
        ref = olk.response(olk.as_dict(rows[p]), d, modes, 0.0, interp)
        for ci, c in enumerate(labels[:2]):
            assert np.linalg.norm(ref[c][hi]) > 1e-3 * np.linalg.norm(ref[c])  # signal up there
            assert _relnorm(resp[p, ci], ref[c]) < 1e-10
            assert _relnorm(resp[p, ci][hi], ref[c][hi]) < 1e-10
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), olk.lnl_batch(rows, [d], modes, 0.0, interp),
                               rtol=1e-10)


def test_scratch_pinning():
    from pespace_amd._lib import PespaceError
    d, labels = make_detector()
    core = _core(((2, 2),), [(d, labels)])
    rows = particle_cloud(16)
    core.reserve(8, grad=False)
    core.pin_scratch(True)
    l8 = core.lnl(rows[:8]).cpu().numpy()
    with pytest.raises(PespaceError, match="pinned"):
        core.lnl(rows)
    with pytest.raises(PespaceError, match="pinned"):
        core.lnl(rows[:8], grad=True)
    core.pin_scratch(False)
    l16, _ = core.lnl(rows, grad=True)
    np.testing.assert_array_equal(l16.cpu().numpy()[:8], l8)

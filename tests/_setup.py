"""Shared small workloads for the parity tests (oracle side and HIP side)."""

import numpy as np

from oracle import detector as odet
from oracle import likelihood as olk

INJ = dict(chirp_mass=1.2e6, mass_ratio=0.6, chi_1=0.75, chi_2=0.62, luminosity_distance=56000.0,
           reference_phase=1.3, inclination=0.4, ecliptic_longitude=1.375, ecliptic_latitude=-1.2108,
           polarization=2.659, coalescence_time=0.8 * 655360.0)

ORBIT_KIND = {"helio": 0, "geo": 1}


def inj_row(**over):
    d = dict(INJ)
    d.update(over)
    return np.array([d[k] for k in olk.PARAM_NAMES], dtype=np.float64)


def particle_cloud(n, seed=2026, rel=1e-3, center=None):
    """Gaussian cloud around the injection; angles wrapped (fisher.py:297-310 semantics)."""
    rng = np.random.default_rng(seed)
    c = inj_row() if center is None else np.asarray(center, dtype=np.float64)
    rows = c + rel * np.abs(c) * rng.standard_normal((n, c.size))
    rows[:, 1] = np.clip(rows[:, 1], 1e-3, 1.0)
    rows[:, 5] = np.mod(rows[:, 5], 2 * np.pi)      # reference phase
    rows[:, 7] = np.mod(rows[:, 7], 2 * np.pi)      # longitude
    rows[:, 9] = np.mod(rows[:, 9], np.pi)          # polarization
    rows[:, 8] = np.clip(rows[:, 8], -np.pi / 2, np.pi / 2)
    return rows


def prior_rows(n, seed=3, tc_range=(0.0, 655360.0)):
    """Rows from the wide prior of the benchmark configurations: M_tot log-uniform 1e5-1e7 Msun,
    q = m1/m2 uniform 1-8, chi uniform [-0.9, 0.9], d_L uniform 5-50 Gpc, sky uniform on the
    sphere, cos(inclination) uniform, polarization and phase uniform, t_c uniform in tc_range."""
    rng = np.random.default_rng(seed)
    M = 10.0 ** rng.uniform(5.0, 7.0, n)
    q = rng.uniform(1.0, 8.0, n)
    eta = q / (1.0 + q) ** 2
    rows = np.empty((n, 11))
    rows[:, 0] = M * eta ** 0.6
    rows[:, 1] = 1.0 / q
    rows[:, 2] = rng.uniform(-0.9, 0.9, n)
    rows[:, 3] = rng.uniform(-0.9, 0.9, n)
    rows[:, 4] = rng.uniform(5000.0, 50000.0, n)
    rows[:, 5] = rng.uniform(0.0, 2 * np.pi, n)
    rows[:, 6] = np.arccos(rng.uniform(-1.0, 1.0, n))
    rows[:, 7] = rng.uniform(0.0, 2 * np.pi, n)
    rows[:, 8] = np.arcsin(rng.uniform(-1.0, 1.0, n))
    rows[:, 9] = rng.uniform(0.0, np.pi, n)
    rows[:, 10] = rng.uniform(*tc_range, n)
    return rows


def make_detector(orbit="LISA_analytic", psd_model="LISA_SciRDv1", gen="1.5", T=655360.0, dt=5.0,
                  fmin=1e-4, fmax=0.05, seed=7, modes=((2, 2),), interp=(1e-4, 0.05, 1024), orthogonal=True,
                  channels=("A", "E", "T"), inject=True, duty=None):
    f, mask, n_t, df, _ = odet.data_grid(T, dt, fmin, fmax)
    labels = ("A", "E", "T") if orthogonal else ("X", "Y", "Z")
    if not orthogonal and channels == ("A", "E", "T"):
        channels = labels
    psd = odet.psd(psd_model, f, labels, False, gen)
    noise = odet.noise_realization(seed, labels, n_t, mask, psd, df)
    orb = odet.ORBITS[orbit] if isinstance(orbit, str) else tuple(orbit)
    d0 = olk.Detector(f, None, psd, df, orb, gen, channels, orthogonal, duty)
    data = dict(noise)
    if inject:
        resp = olk.response(INJ, d0, modes, 0.0, interp)
        data = {c: noise[c] + resp[c] for c in labels}
    return olk.Detector(f, data, psd, df, orb, gen, channels, orthogonal, duty), labels


def add_to_core(core, d, labels):
    kind, arm, rot0, rev0 = d.orbit[:4]
    geo = dict(zip(("omega_rotation", "lambda_ref", "beta_ref"), d.orbit[4:]))
    data = np.stack([d.data[c] for c in labels])
    psd = np.stack([d.psd[c] for c in labels])
    mask = [1 if c in d.channels else 0 for c in labels]
    gen = 15 if d.tdi_generation == "1.5" else 20
    return core.add_detector(ORBIT_KIND[kind], arm, rot0, rev0, gen, d.orthogonal, d.f, data, psd, mask, df=d.df,
                             duty_cycle=getattr(d, "duty", None), **geo)


YEAR_SI = 31557600.0
NETWORK = (("LISA", "LISA_analytic", "LISA_SciRDv1", 1e-4, 0.05),
           ("Taiji", "Taiji_analytic", "Taiji_TDC", 1e-4, 0.06),
           ("TianQin", "Tianqin_analytic", "Tianqin_GWSpace", 1e-3, 0.039))


# TianQin's observation duty cycle in the benchmark network: three months on, three months off,
# windows from 1.5 months into the year, one-day ramps
TIANQIN_DUTY = (YEAR_SI / 2, YEAR_SI / 4, YEAR_SI / 8, 86400.0)


def network_core(modes=("22", "21", "33", "32", "44", "43"), T=YEAR_SI, dt=5.0, seed=7, n_nodes=1024,
                 inj=None, tianqin_duty=None):
    """LISA + Taiji + TianQin on rfft grids of duration T (ragged: 1,574,725 / 1,890,301 /
    1,199,189 bins for one year at dt = 5 s), seeded noise plus the injection, through the
    product's host API (the data of the benchmark configurations), TianQin with the given duty
    cycle (None: always observing).  Returns the core."""
    from pespace_amd.core import LikelihoodCore
    from pespace_amd.detector.antenna import (FDResponseModelMarset2018, InterferometerAntenna,
                                              add_detector_to_core)
    from pespace_amd.detector.orbit import available_orbit_models
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    from pespace_amd.waveform.phenom import IMRPhenomHM
    inj = dict(INJ, coalescence_time=0.8 * T) if inj is None else inj
    fmin = min(n[3] for n in NETWORK)
    fmax = max(n[4] for n in NETWORK)
    dets = []
    for i, (name, orbit, psd, lo, hi) in enumerate(NETWORK):
        td = TDIChannelData(backend="hip")
        td.set_fd_data_from_zero(("A", "E", "T"), T, dt, 0.0, lo, hi)
        td.set_fd_noise_power_density_from_model(psd, tdi_generation="1.5")
        det = InterferometerAntenna(name, td, available_orbit_models()[orbit], FDResponseModelMarset2018(),
                                    FDMichelsonConstantEqualArm("1.5", True),
                                    duty_cycle=tianqin_duty if name == "TianQin" else None)
        dets.append(det)
    wf = IMRPhenomHM(dets[0].tdi_data.frequency_samples, 0.0, modes=modes, n_nodes=n_nodes, f_node_min=fmin,
                     f_node_max=fmax)
    wf.update_waveform(inj)
    for i, det in enumerate(dets):
        det.tdi_data.add_into_fd_data(det.tdi_data.get_fd_noise_realization(seed + i))
        det.inject_signal(wf.waveform_container, inj["ecliptic_longitude"], inj["ecliptic_latitude"],
                          inj["polarization"], inj["coalescence_time"])
    core = LikelihoodCore(modes=wf.modes, n_nodes=n_nodes, f_node_min=fmin, f_node_max=fmax)
    for det in dets:
        add_detector_to_core(core, det, data_channels=("A", "E", "T"))
    return core

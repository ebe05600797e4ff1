"""The benchmark configurations at their stated shapes.

* Config 1 (plain method, FP64): one system drawn from the wide prior with a seed (total mass
  log-uniform 1e5-1e7 Msun, q uniform 1-8, spins uniform [-0.9, 0.9], distance 5-50 Gpc, sky
  uniform, cos(inclination) uniform, polarization and phase uniform), IMRPhenomD (2,2) only,
  LISA analytic orbit, TDI 1.5 A/E/T, 2^14 log-spaced bins over 1e-4-0.05 Hz with per-bin
  weights df_k, data = the injection plus noise coloured by the analytic PSD; the waveform is
  evaluated at every bin (no interpolation).  lnL at the injection and at further prior draws
  against the oracle (rtol 1e-10).
* Config 3 (differentiation): the one-year LISA grid (1,574,725 bins), 64 particles around the
  injection, the adjoint gradient over the 11 parameters against Richardson-extrapolated central
  differences of the HIP lnL (relative step 1e-6; an absolute 0.01 s for t_c, whose relative
  step would be ~25 s, a large fraction of the signal's oscillation period at these frequencies).
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import likelihood as olk
from tests._setup import YEAR_SI, add_to_core, inj_row, make_detector, particle_cloud, prior_rows

pytestmark = pytest.mark.gpu


def _config1_rows(n, seed=2024):
    """Seeded prior draws; t_c inside a band window of 1e6 s."""
    return prior_rows(n, seed=seed, tc_range=(0.0, 1.0e6))


def test_config1_log_grid_seeded_prior_injection():
    from pespace_amd.core import LikelihoodCore
    f = np.geomspace(1e-4, 0.05, 2 ** 14)
    w = np.gradient(f)
    labels = ("A", "E", "T")
    psd = odet.psd("LISA_SciRDv1", f, labels, False, "1.5")
    rng = np.random.default_rng(2024)
    noise = {c: (rng.normal(size=f.size) + 1j * rng.normal(size=f.size)) * 0.5 / np.sqrt(w) * np.sqrt(psd[c])
             for c in labels}
    orb = odet.ORBITS["LISA_analytic"]
    rows = _config1_rows(6)
    inj = olk.as_dict(rows[0])
    d0 = olk.Detector(f, None, psd, w, orb)
    resp = olk.response(inj, d0, ((2, 2),), 0.0, None)
    d = olk.Detector(f, {c: noise[c] + resp[c] for c in labels}, psd, w, orb)
    core = LikelihoodCore(modes=["22"], n_nodes=0)
    core.add_detector(0, orb[1], orb[2], orb[3], 15, True, f, np.stack([d.data[c] for c in labels]),
                      np.stack([psd[c] for c in labels]), [1, 1, 1], weight=w)
    assert core.n_bins == [2 ** 14]
    lnl, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    ref = olk.lnl_batch(rows, [d], ((2, 2),), 0.0, None)
    print("config 1 lnL", lnl, "\noracle", ref)
    np.testing.assert_allclose(lnl, ref, rtol=1e-10)
    assert np.all(np.isfinite(g))
    # the injection row sees noise only: lnL ~ -(number of bins x channels) / 1 per bin
    assert abs(lnl[0] / (3 * 2 ** 14) + 1.0) < 0.05


def _fd_rows(row):
    rows, hs = [], []
    for i in range(row.size):
        h = 1e-2 if i == 10 else (1e-6 * abs(row[i]) if row[i] != 0 else 1e-6)
        for hh in (h, 0.5 * h):
            rp, rm = row.copy(), row.copy()
            rp[i] += hh
            rm[i] -= hh
            rows += [rp, rm]
        hs.append(h)
    return np.array(rows), np.array(hs)


def test_config3_gradient_one_year_64_particles():
    from pespace_amd.core import LikelihoodCore
    d, labels = make_detector(T=YEAR_SI, dt=5.0, modes=((2, 2),), seed=31)
    core = LikelihoodCore(modes=["22"], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    add_to_core(core, d, labels)
    assert core.n_bins == [1574725]
    center = inj_row(coalescence_time=0.8 * 655360.0)
    rows = particle_cloud(64, seed=33, rel=1e-5, center=center)
    lnl, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    fdr, hs = zip(*(_fd_rows(r) for r in rows))
    vals = core.lnl(np.vstack(fdr)).cpu().numpy().reshape(64, 11, 4)
    worst = 0.0
    for p in range(64):
        v = vals[p]
        h = np.asarray(hs[p])
        d1 = (v[:, 0] - v[:, 1]) / (2 * h)
        d2 = (v[:, 2] - v[:, 3]) / h
        fd = (4 * d2 - d1) / 3
        # floor: the rounding of the differences themselves, ~ eps |lnL| / h per component
        floor = 64 * np.finfo(float).eps * abs(lnl[p]) / h
        err = np.abs(g[p] - fd) / np.maximum(np.abs(fd), np.maximum(floor / 2e-6, 1e-9 * np.abs(fd).max()))
        worst = max(worst, err.max())
        assert np.all(err < 2e-6), (p, err)
    print(f"config 3: 64 particles, worst relative error {worst:.2e}")

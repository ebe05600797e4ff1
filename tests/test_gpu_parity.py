"""HIP FP64 path vs the numpy oracle on small workloads (bins x particles).

Covers the waveform container, the TDI response and lnL for: 22-only and six
modes; TDI 1.5 and 2.0; A/E/T and X/Y/Z outputs; heliocentric (LISA, Taiji) and
geocentric (TianQin) orbits; a channel subset; a multi-detector network with
ragged grids; and direct per-bin waveform evaluation on a log-spaced grid.
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import INJ, add_to_core, inj_row, make_detector, particle_cloud, prior_rows

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)


def _core(modes, dets, n_nodes=1024, f_lo=1e-4, f_hi=0.05):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], reference_frequency=0.0, n_nodes=n_nodes,
                          f_node_min=f_lo, f_node_max=f_hi)
    for d, labels in dets:
        add_to_core(core, d, labels)
    return core


def _relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("modes", [((2, 2),), oph.MODES])
def test_waveform_container(modes):
    d, labels = make_detector(modes=modes)
    core = _core(modes, [(d, labels)])
    rows = particle_cloud(3)
    wf = core.waveform(0, rows).cpu().numpy()
    for p in range(rows.shape[0]):
        ref = oph.waveform_interpolated(olk.as_dict(rows[p]), d.f, modes, 0.0, *INTERP)
        for mi, md in enumerate(modes):
            r = ref[f"{md[0]}{md[1]}"]
            hp = wf[p, mi, :, 0] + 1j * wf[p, mi, :, 1]
            hc = wf[p, mi, :, 2] + 1j * wf[p, mi, :, 3]
            assert _relnorm(hp, r["plus"]) < 1e-10
            assert _relnorm(hc, r["cross"]) < 1e-10
            assert _relnorm(wf[p, mi, :, 4], r["tf"]) < 1e-10


CASES = [
    dict(modes=((2, 2),)),
    dict(modes=oph.MODES),
    dict(modes=oph.MODES, gen="2.0"),
    dict(modes=((2, 2),), orthogonal=False),
    dict(modes=oph.MODES, orbit="Taiji_analytic", psd_model="Taiji_TDC"),
    dict(modes=oph.MODES, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3),
    dict(modes=((2, 2),), channels=("A", "E")),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items() if k != "modes")
                         + f"-m{len(c['modes'])}")
def test_response_and_lnl(case):
    case = dict(case)
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(5)])
    resp = core.response(0, rows).cpu().numpy()
    lnl = core.lnl(rows).cpu().numpy()
    ref_lnl = olk.lnl_batch(rows, [d], modes, 0.0, INTERP)
    for p in range(rows.shape[0]):
        ref = olk.response(olk.as_dict(rows[p]), d, modes, 0.0, INTERP)
        for ci, c in enumerate(labels):
            tol = 1e-10
            if c == "T":
                # the oracle forms T = (X+Y+Z)/sqrt3 (tdi.py:1786), which cancels to ~ (pi f L)^2 |X|;
                # its own rounding error is then ~ eps |X| / |T|, larger than 1e-10 for short arms
                xyz = np.sqrt(np.sum(np.abs(ref["A"]) ** 2 + np.abs(ref["E"]) ** 2))
                tol = max(tol, 64 * np.finfo(float).eps * xyz / np.linalg.norm(ref["T"]))
            assert _relnorm(resp[p, ci], ref[c]) < tol, (p, c, _relnorm(resp[p, ci], ref[c]), tol)
    np.testing.assert_allclose(lnl, ref_lnl, rtol=1e-10)


def test_network_ragged():
    """LISA + Taiji + TianQin with different grids; lnL is the sum over detectors."""
    modes = oph.MODES
    dets = [make_detector(modes=modes, seed=7),
            make_detector(modes=modes, orbit="Taiji_analytic", psd_model="Taiji_TDC", fmax=0.03, seed=8),
            make_detector(modes=modes, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3, seed=9)]
    core = _core(modes, dets)
    rows = np.vstack([inj_row(), particle_cloud(3)])
    lnl = core.lnl(rows).cpu().numpy()
    ref = olk.lnl_batch(rows, [d for d, _ in dets], modes, 0.0, INTERP)
    np.testing.assert_allclose(lnl, ref, rtol=1e-10)
    assert [core.n_bins[i] for i in range(3)] == [d.f.size for d, _ in dets]


def test_direct_log_grid():
    """Per-bin waveform evaluation (no interpolation) on a log-spaced grid with weights df_k."""
    f = np.geomspace(1e-4, 0.05, 2 ** 12)
    w = np.gradient(f)
    labels = ("A", "E", "T")
    psd = odet.psd("LISA_SciRDv1", f, labels, False, "1.5")
    rng = np.random.default_rng(11)
    noise = {c: (rng.normal(size=f.size) + 1j * rng.normal(size=f.size)) * 0.5 / np.sqrt(w) * np.sqrt(psd[c])
             for c in labels}
    orb = odet.ORBITS["LISA_analytic"]
    d0 = olk.Detector(f, None, psd, w, orb)
    resp = olk.response(INJ, d0, ((2, 2),), 0.0, None)
    d = olk.Detector(f, {c: noise[c] + resp[c] for c in labels}, psd, w, orb)
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=["22"], n_nodes=0)
    core.add_detector(0, orb[1], orb[2], orb[3], 15, True, f, np.stack([d.data[c] for c in labels]),
                      np.stack([psd[c] for c in labels]), [1, 1, 1], weight=w)
    rows = np.vstack([inj_row(), particle_cloud(3)])
    lnl = core.lnl(rows).cpu().numpy()
    ref = olk.lnl_batch(rows, [d], ((2, 2),), 0.0, None)
    np.testing.assert_allclose(lnl, ref, rtol=1e-10)
    r = core.response(0, rows[:1]).cpu().numpy()[0]
    for ci, c in enumerate(labels):
        assert _relnorm(r[ci], resp[c]) < 1e-10


def test_wide_prior_rows():
    """Rows drawn from the wide prior (masses 1e5-1e7 Msun, t_c anywhere in the data span,
    any sky position and orientation), six modes: signals that end inside, above or below
    the band, and phases up to ~1e7 rad."""
    modes = oph.MODES
    d, labels = make_detector(modes=modes)
    core = _core(modes, [(d, labels)])
    rows = prior_rows(12, seed=21)
    lnl = core.lnl(rows).cpu().numpy()
    ref = olk.lnl_batch(rows, [d], modes, 0.0, INTERP)
    np.testing.assert_allclose(lnl, ref, rtol=1e-10)

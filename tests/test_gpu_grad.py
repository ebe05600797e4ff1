"""Adjoint-kernel gradient vs central finite differences (step 1e-6 relative).

Checked two ways: against finite differences of the HIP lnL itself (isolates the
adjoint), and against finite differences of the numpy oracle (end to end).
"""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud, prior_rows

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)


def _core(modes, dets):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    for d, labels in dets:
        add_to_core(core, d, labels)
    return core


def _fd_rows(row, rel=1e-6):
    """Rows for central differences at steps h and h/2 (Richardson-combined below)."""
    rows, hs = [], []
    for i in range(row.size):
        h = rel * abs(row[i]) if row[i] != 0 else rel
        for hh in (h, 0.5 * h):
            rp, rm = row.copy(), row.copy()
            rp[i] += hh
            rm[i] -= hh
            rows += [rp, rm]
        hs.append(h)
    return np.array(rows), np.array(hs)


def _fd_from(v, hs):
    """Richardson: (4 D(h/2) - D(h)) / 3 removes the O(h^2) truncation term."""
    v = v.reshape(-1, 4)
    d1 = (v[:, 0] - v[:, 1]) / (2 * hs)
    d2 = (v[:, 2] - v[:, 3]) / hs
    return (4 * d2 - d1) / 3


def _check(g, fd, scale, tol):
    err = np.abs(g - fd) / np.maximum(np.abs(fd), scale)
    print("grad", g, "\nfd  ", fd, "\nerr ", err)
    assert np.all(err < tol), err


def _fd_check_core(core, rows, tol=2e-6):
    lnl, g = core.lnl(rows, grad=True)
    lnl0 = core.lnl(rows).cpu().numpy()
    np.testing.assert_allclose(lnl.cpu().numpy(), lnl0, rtol=1e-13)
    g = g.cpu().numpy()
    for p in range(rows.shape[0]):
        fr, hs = _fd_rows(rows[p])
        fd = _fd_from(core.lnl(fr).cpu().numpy(), hs)
        scale = 1e-6 * np.max(np.abs(fd) * np.abs(rows[p]) + 1.0) / np.maximum(np.abs(rows[p]), 1.0)
        # rounding floor of the differences themselves: ~ eps |lnL| / h per component (far from
        # the peak |lnL| reaches 1e6-1e7 and this dominates for the weakly coupled parameters)
        scale = np.maximum(scale, 64 * np.finfo(float).eps * abs(lnl0[p]) / hs / tol)
        _check(g[p], fd, scale, tol)


CASES = [
    dict(modes=((2, 2),), gen="1.5", orbit="LISA_analytic"),
    dict(modes=oph.MODES, gen="2.0", orbit="LISA_analytic"),
    dict(modes=oph.MODES, gen="1.5", orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3),
    dict(modes=oph.MODES, gen="1.5", orbit="Taiji_analytic", psd_model="Taiji_TDC"),
    dict(modes=((2, 2),), gen="2.0", orbit="LISA_analytic", orthogonal=False),
    dict(modes=oph.MODES, gen="1.5", orbit="LISA_analytic", channels=("A", "E")),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items() if k != "modes")
                         + f"-m{len(c['modes'])}")
def test_grad_vs_fd_of_hip_lnl(case):
    case = dict(case)
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(3, rel=1e-5)])
    _fd_check_core(core, rows)


def test_grad_vs_oracle_fd():
    modes = oph.MODES
    d, labels = make_detector(modes=modes, T=131072.0 * 5, dt=5.0)
    core = _core(modes, [(d, labels)])
    row = particle_cloud(1, seed=3, rel=1e-5)[0]
    _, g = core.lnl(row[None, :], grad=True)
    fd = olk.fd_gradient(row, [d], modes, 0.0, INTERP)
    scale = 1e-6 * np.max(np.abs(fd) * np.abs(row) + 1.0) / np.maximum(np.abs(row), 1.0)
    _check(g.cpu().numpy()[0], fd, scale, 2e-4)


def test_grad_network_and_wide_prior():
    """A ragged three-detector network, at rows near the injection and from the wide prior."""
    modes = oph.MODES
    dets = [make_detector(modes=modes, seed=7),
            make_detector(modes=modes, orbit="Taiji_analytic", psd_model="Taiji_TDC", fmax=0.03, seed=8),
            make_detector(modes=modes, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3, seed=9)]
    core = _core(modes, dets)
    _fd_check_core(core, np.vstack([inj_row(), particle_cloud(1, rel=1e-5)]))
    _fd_check_core(core, prior_rows(3, seed=5))


def test_grad_direct_log_grid():
    """Per-bin waveform evaluation (n_nodes = 0) on a log-spaced grid with weights df_k."""
    from oracle import detector as odet
    from pespace_amd.core import LikelihoodCore
    f = np.geomspace(1e-4, 0.05, 2 ** 12)
    w = np.gradient(f)
    labels = ("A", "E", "T")
    psd = odet.psd("LISA_SciRDv1", f, labels, False, "1.5")
    orb = odet.ORBITS["LISA_analytic"]
    d0 = olk.Detector(f, None, psd, w, orb)
    resp = olk.response(olk.as_dict(inj_row()), d0, oph.MODES, 0.0, None)
    rng = np.random.default_rng(12)
    data = np.stack([resp[c] + (rng.normal(size=f.size) + 1j * rng.normal(size=f.size)) * 0.5 / np.sqrt(w)
                     * np.sqrt(psd[c]) for c in labels])
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in oph.MODES], n_nodes=0)
    core.add_detector(0, orb[1], orb[2], orb[3], 15, True, f, data, np.stack([psd[c] for c in labels]), [1, 1, 1],
                      weight=w)
    _fd_check_core(core, np.vstack([inj_row(), particle_cloud(2, rel=1e-5)]))


@pytest.mark.parametrize("orbit,psd_model,fmin", [("LISA_analytic", "LISA_SciRDv1", 1e-4),
                                                  ("Tianqin_analytic", "Tianqin_GWSpace", 1e-3)])
def test_adjoint_matches_forward_mode_gradient_of_the_cpu_restatement(orbit, psd_model, fmin):
    """The HIP adjoint against a second, unrelated differentiation of the same lnL: forward-mode
    dual numbers through the oracle's C++ restatement (node-table derivatives by central
    differences, ~1e-9 relative)."""
    import os

    from oracle import build_cpu
    if not os.path.exists(build_cpu.LIB):
        build_cpu.build()
    from oracle import cpu_baseline as cb
    modes = oph.MODES
    d, labels = make_detector(modes=modes, orbit=orbit, psd_model=psd_model, fmin=fmin)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([particle_cloud(3, seed=41), prior_rows(2, seed=42)])
    lnl, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    net = cb.Network([d])
    for p in range(rows.shape[0]):
        lc, gc = cb.lnl_row(net, cb.RowInputs(rows[p], modes, 0.0, INTERP, grad=True), grad=True, threads=8)
        assert abs(lc - lnl[p]) < 1e-10 * abs(lnl[p])
        scale = np.maximum(np.abs(gc), 1e-6 * np.abs(gc).max())
        err = np.abs(g[p] - gc) / scale
        assert np.all(err < 1e-6), (p, err)

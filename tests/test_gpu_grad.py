"""Adjoint-kernel gradient vs central finite differences (step 1e-6 relative).

Checked two ways: against finite differences of the HIP lnL itself (isolates the
adjoint), and against finite differences of the numpy oracle (end to end).
"""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud

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


@pytest.mark.parametrize("modes,gen,orbit", [(((2, 2),), "1.5", "LISA_analytic"), (oph.MODES, "2.0", "LISA_analytic"),
                                             (oph.MODES, "1.5", "Tianqin_analytic")])
def test_grad_vs_fd_of_hip_lnl(modes, gen, orbit):
    psd = "Tianqin_GWSpace" if orbit.startswith("Tianqin") else "LISA_SciRDv1"
    d, labels = make_detector(modes=modes, gen=gen, orbit=orbit, psd_model=psd)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(3, rel=1e-5)])
    lnl, g = core.lnl(rows, grad=True)
    lnl0 = core.lnl(rows).cpu().numpy()
    np.testing.assert_allclose(lnl.cpu().numpy(), lnl0, rtol=1e-13)
    g = g.cpu().numpy()
    for p in range(rows.shape[0]):
        fr, hs = _fd_rows(rows[p])
        fd = _fd_from(core.lnl(fr).cpu().numpy(), hs)
        scale = 1e-6 * np.max(np.abs(fd) * np.abs(rows[p]) + 1.0) / np.maximum(np.abs(rows[p]), 1.0)
        _check(g[p], fd, scale, 2e-6)


def test_grad_vs_oracle_fd():
    modes = oph.MODES
    d, labels = make_detector(modes=modes, T=131072.0 * 5, dt=5.0)
    core = _core(modes, [(d, labels)])
    row = particle_cloud(1, seed=3, rel=1e-5)[0]
    _, g = core.lnl(row[None, :], grad=True)
    fd = olk.fd_gradient(row, [d], modes, 0.0, INTERP)
    scale = 1e-6 * np.max(np.abs(fd) * np.abs(row) + 1.0) / np.maximum(np.abs(row), 1.0)
    _check(g.cpu().numpy()[0], fd, scale, 2e-4)

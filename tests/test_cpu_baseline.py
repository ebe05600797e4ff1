"""The oracle's C++/OpenMP restatement (the benchmark's CPU baseline) against the numpy oracle.

lnL to 1e-12 relative on small grids (orbit kinds, TDI generations, A/E/T and X/Y/Z, a channel
subset, a two-detector network); the forward-mode gradient against Richardson-extrapolated
central differences of its own lnL.
"""

import os

import numpy as np
import pytest

from oracle import build_cpu
from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import inj_row, make_detector, particle_cloud

INTERP = (1e-4, 0.05, 1024)


@pytest.fixture(scope="module")
def cb():
    if not os.path.exists(build_cpu.LIB):
        build_cpu.build()
    from oracle import cpu_baseline
    return cpu_baseline


CASES = [
    dict(modes=oph.MODES),
    dict(modes=((2, 2),), gen="2.0", orthogonal=False),
    dict(modes=oph.MODES, orbit=("geo", 1.7e8, 0.4, 1.1, 2 * np.pi / (5 * 86400.0), 0.7, 0.3),
         psd_model="Tianqin_GWSpace", fmin=1e-3),
    dict(modes=oph.MODES, channels=("A", "E")),
    # observation duty cycles with an edge at the merger (on the grid's t_f + t_c range)
    dict(modes=oph.MODES, duty=(1.0e8, 5.0e7, 524400.0 - 5.0e7, 2000.0)),
    dict(modes=oph.MODES, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3,
         duty=(1.0e8, 5.0e7, 523000.0, 1500.0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_cpu_restatement_lnl_matches_numpy_oracle(cb, case):
    kw = dict(CASES[case])
    modes = kw.pop("modes")
    d, _ = make_detector(modes=modes, **kw)
    net = cb.Network([d])
    for r in np.vstack([inj_row(), particle_cloud(2)]):
        ri = cb.RowInputs(r, modes, 0.0, INTERP)
        ref = olk.lnl(olk.as_dict(r), [d], modes, 0.0, INTERP)
        got = cb.lnl_row(net, ri, threads=2)
        assert abs(got - ref) < 1e-12 * abs(ref), (got, ref)


def test_cpu_restatement_network_lnl(cb):
    modes = oph.MODES
    dets = [make_detector(modes=modes)[0],
            make_detector(modes=modes, orbit="Taiji_analytic", psd_model="Taiji_TDC", fmax=0.03, seed=8)[0]]
    net = cb.Network(dets)
    r = particle_cloud(1, seed=4)[0]
    ref = olk.lnl(olk.as_dict(r), dets, modes, 0.0, INTERP)
    got = cb.lnl_row(net, cb.RowInputs(r, modes, 0.0, INTERP), threads=3)
    assert abs(got - ref) < 1e-12 * abs(ref)


@pytest.mark.parametrize("duty", [None, (1.0e8, 5.0e7, 524400.0 - 5.0e7, 2000.0)])
def test_cpu_restatement_gradient_matches_richardson_differences(cb, duty):
    modes = oph.MODES
    d, _ = make_detector(modes=modes, T=2 ** 13 * 5.0, duty=duty)
    net = cb.Network([d])
    row = particle_cloud(1, seed=11)[0]
    ri = cb.RowInputs(row, modes, 0.0, INTERP, grad=True)
    lnl, g = cb.lnl_row(net, ri, grad=True, threads=4)
    assert abs(lnl - cb.lnl_row(net, ri, threads=4)) < 1e-12 * abs(lnl)

    def f(r):
        return cb.lnl_row(net, cb.RowInputs(r, modes, 0.0, INTERP), threads=4)

    fd = np.zeros(11)
    for i in range(11):
        h = 1e-6 * abs(row[i]) if row[i] != 0 else 1e-6
        vals = []
        for s in (h, -h, 0.5 * h, -0.5 * h):
            r = row.copy()
            r[i] += s
            vals.append(f(r))
        d1 = (vals[0] - vals[1]) / (2 * h)
        d2 = (vals[2] - vals[3]) / h
        fd[i] = (4 * d2 - d1) / 3
    scale = np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    err = np.abs(g - fd) / scale
    # the node tables' parameter derivatives are central differences (~1e-9 relative); under a
    # duty window their t_f errors are weighed by the window's slope as well (q: ~1e-5)
    assert np.all(err < (1e-5 if duty is None else 3e-5)), err

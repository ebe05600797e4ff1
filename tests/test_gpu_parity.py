"""HIP FP64 path vs the numpy oracle on small workloads (bins x particles)."""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import INJ, add_to_core, inj_row, make_detector, particle_cloud

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
            print(md, _relnorm(hp, r["plus"]), _relnorm(hc, r["cross"]), _relnorm(wf[p, mi, :, 4], r["tf"]))
            assert _relnorm(hp, r["plus"]) < 1e-10
            assert _relnorm(hc, r["cross"]) < 1e-10
            assert _relnorm(wf[p, mi, :, 4], r["tf"]) < 1e-10


@pytest.mark.parametrize("modes", [((2, 2),), oph.MODES])
def test_response_and_lnl(modes):
    d, labels = make_detector(modes=modes)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(5)])
    resp = core.response(0, rows).cpu().numpy()
    lnl = core.lnl(rows).cpu().numpy()
    ref_lnl = olk.lnl_batch(rows, [d], modes, 0.0, INTERP)
    for p in range(rows.shape[0]):
        ref = olk.response(olk.as_dict(rows[p]), d, modes, 0.0, INTERP)
        for ci, c in enumerate(labels):
            assert _relnorm(resp[p, ci], ref[c]) < 1e-10, (p, c, _relnorm(resp[p, ci], ref[c]))
    print(lnl, ref_lnl, np.abs(lnl / ref_lnl - 1))
    np.testing.assert_allclose(lnl, ref_lnl, rtol=1e-10)

"""Detector runs and fused-kernel launches (pes_context_finalize, tile_ranges).

Consecutive detectors of one orbit kind and duty flag form a run; each run is launched in the
fused kernel's instantiation for that kind (a duty-cycled run twice: its plateau work items in the
plain one), each launch writing its own tile groups' partial sums.

* a network ordered heliocentric, geocentric, heliocentric (three runs, three launches) matches
  the oracle, with and without a duty cycle on the geocentric detector;
* 16 runs are accepted, 17 rejected at finalize with a message that names the cause.
"""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)
TQ = dict(orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3)
TQ_DUTY = (1.0e8, 5.0e7, 523000.0, 1500.0)  # window opens near the injection's merger


def _core(modes, dets):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    for d, labels in dets:
        add_to_core(core, d, labels)
    return core


@pytest.mark.parametrize("duty", [False, True])
def test_alternating_orbit_kinds_match_oracle(duty):
    modes = oph.MODES
    dets = [make_detector(modes=modes, seed=7),
            make_detector(modes=modes, seed=9, duty=TQ_DUTY if duty else None, **TQ),
            make_detector(modes=modes, orbit="Taiji_analytic", psd_model="Taiji_TDC", fmax=0.03, seed=8)]
    core = _core(modes, dets)
    rows = np.vstack([inj_row(), particle_cloud(5), inj_row(coalescence_time=5.0e6)])
    print(f"three runs, duty={duty}", flush=True)
    ref = olk.lnl_batch(rows, [d for d, _ in dets], modes, 0.0, INTERP)
    for prec in ("fp64", "mixed"):
        lnl, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision=prec))
        np.testing.assert_allclose(lnl, ref, rtol=1e-10)
        assert np.all(np.isfinite(g))
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), ref, rtol=1e-10)
    # the same detectors in another order (two runs) give the same gradient up to rounding
    other = _core(modes, [dets[0], dets[2], dets[1]])
    g3 = core.lnl(rows, grad=True)[1].cpu().numpy()
    g2 = other.lnl(rows, grad=True)[1].cpu().numpy()
    np.testing.assert_allclose(g3, g2, rtol=1e-9, atol=1e-12 * np.abs(g2).max())


def test_run_limit_at_finalize():
    from pespace_amd._lib import PespaceError
    modes = ((2, 2),)
    lisa = make_detector(modes=modes, T=65536.0, fmin=1e-3, fmax=2e-3, inject=False, seed=1)
    tq = make_detector(modes=modes, T=65536.0, fmin=1e-3, fmax=2e-3, inject=False, seed=2,
                       orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace")
    print("16 and 17 alternating runs", flush=True)
    ok = _core(modes, [lisa, tq] * 8)
    assert np.isfinite(ok.lnl(inj_row()[None]).cpu().numpy()).all()
    bad = _core(modes, [lisa, tq] * 8 + [lisa])
    with pytest.raises(PespaceError, match="too many alternations of orbit kinds"):
        bad.lnl(inj_row()[None])

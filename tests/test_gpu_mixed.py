"""Mixed precision against the FP64 path and the oracle.

lnL is the FP64 forward in both modes, so |dlnL| between them is zero on every row (the bound
of the benchmark metric is |dlnL| < 1e-3); the mixed mode takes FP32 hardware sines, cosines
and reciprocals in the adjoint's recompute only, so its gradient carries FP32-level relative
errors.  Gradient errors are measured relative to the largest |dlnL/dtheta_k| of each
parameter over the batch.

The benchmark configuration with mixed precision (LISA + Taiji + TianQin on one-year grids,
six modes, uneven batches of 256-8192 rows from the wide prior: t_c anywhere in the year, total
mass 1e5-1e7 Msun) runs at full size through size-independent checks (mixed = FP64 lnL,
gradient error, finiteness, batch invariance); a short-grid variant compares both modes with
the oracle's lnL.
"""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import (TIANQIN_DUTY, YEAR_SI, add_to_core, inj_row, make_detector, network_core, particle_cloud,
                          prior_rows)

pytestmark = pytest.mark.gpu

# mixed-mode gradient error relative to the batch's largest |g| per parameter.  Measured: ~3e-7
# near the peak, ~1e-5 for wide-prior rows on a short grid, 1e-4 - 2.5e-4 for wide-prior rows on
# the one-year network (|lnL| up to 2e8: the FP32 phasor and reciprocal errors, ~1e-7 relative per
# term, stand against sums that cancel by orders of magnitude)
GRAD_REL_TOL = 1e-3


def _core(modes, d, labels):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    add_to_core(core, d, labels)
    return core


def _grad_rel(gmx, g64):
    scale = np.maximum(np.abs(g64).max(axis=0), 1e-300)
    return np.abs(gmx - g64) / scale


@pytest.mark.parametrize("modes,gen", [(((2, 2),), "1.5"), (oph.MODES, "1.5"), (oph.MODES, "2.0")])
def test_mixed_accuracy(modes, gen):
    d, labels = make_detector(modes=modes, gen=gen)
    core = _core(modes, d, labels)
    near = np.vstack([inj_row(), particle_cloud(63, seed=5, rel=1e-8)])
    far = particle_cloud(16, seed=6, rel=1e-4)
    rows = np.vstack([near, far])
    l64, g64 = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="fp64"))
    lmx, gmx = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="mixed"))
    rel = _grad_rel(gmx, g64)
    print(f"mixed: |dlnl| max {np.abs(lmx - l64).max():.2e}; grad rel max {rel.max():.2e} "
          f"p99 {np.percentile(rel, 99):.2e}")
    np.testing.assert_array_equal(lmx, l64)  # lnL is the FP64 forward in both modes
    assert rel.max() < GRAD_REL_TOL
    # the forward-only kernel (other work-item size, same arithmetic) agrees to rounding
    np.testing.assert_allclose(core.lnl(rows, precision="mixed").cpu().numpy(), l64, rtol=1e-13)
    r64 = core.response(0, rows[:2], precision="fp64").cpu().numpy()
    rmx = core.response(0, rows[:2], precision="mixed").cpu().numpy()
    np.testing.assert_array_equal(rmx, r64)


@pytest.mark.parametrize("P", [1, 257, 1001])
def test_short_grid_wide_prior_against_oracle(P):
    """Uneven batches from the wide prior on a short grid: both modes' lnL against the oracle
    (rtol 1e-10), the batch against its rows evaluated alone."""
    modes = oph.MODES
    d, labels = make_detector(modes=modes)
    core = _core(modes, d, labels)
    rows = prior_rows(P, seed=P)
    l64, g64 = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="fp64"))
    lmx, gmx = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="mixed"))
    assert np.all(np.isfinite(l64)) and np.all(np.isfinite(g64)) and np.all(np.isfinite(gmx))
    np.testing.assert_array_equal(lmx, l64)
    pick = np.unique([0, P // 2, P - 1])
    ref = olk.lnl_batch(rows[pick], [d], modes, 0.0, (1e-4, 0.05, 1024))
    np.testing.assert_allclose(lmx[pick], ref, rtol=1e-10)
    assert _grad_rel(gmx, g64).max() < GRAD_REL_TOL
    for i in pick:
        l1 = core.lnl(rows[i:i + 1]).cpu().numpy()
        np.testing.assert_allclose(l1, l64[i:i + 1], rtol=1e-13)


def test_network_one_year_uneven_prior_batches():
    """The mixed-precision benchmark configuration at full size: LISA + Taiji + TianQin (its
    3-month on/off duty cycle), 1-yr grids (4.66e6 bins), six modes, batches of 256, 3001 and
    8192 rows from the wide prior."""
    core = network_core(tianqin_duty=TIANQIN_DUTY)
    assert core.n_bins == [1574725, 1890301, 1199189]
    stats = []
    for P in (256, 3001, 8192):
        rows = prior_rows(P, seed=100 + P, tc_range=(0.0, YEAR_SI))
        l64, g64 = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="fp64"))
        lmx, gmx = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="mixed"))
        dl = np.abs(lmx - l64)
        rel = _grad_rel(gmx, g64)
        stats.append((P, dl.max(), np.percentile(dl, 99), rel.max(), np.percentile(rel, 99)))
        print(f"P={P}: |dlnl| max {dl.max():.2e} p99 {np.percentile(dl, 99):.2e}; grad rel max {rel.max():.2e} "
              f"p99 {np.percentile(rel, 99):.2e}; lnl {l64.min():.3g}..{l64.max():.3g}", flush=True)
        assert np.all(np.isfinite(l64)) and np.all(np.isfinite(g64)) and np.all(np.isfinite(gmx))
        assert dl.max() < 1e-3
        assert rel.max() < GRAD_REL_TOL
        # batch invariance: rows evaluated alone (and in another batch) give the same values
        for i in (0, P // 3, P - 1):
            l1, g1 = (x.cpu().numpy() for x in core.lnl(rows[i:i + 1], grad=True, precision="mixed"))
            np.testing.assert_allclose(l1, lmx[i:i + 1], rtol=1e-13)
            np.testing.assert_allclose(g1, gmx[i:i + 1], rtol=1e-12, atol=1e-12 * np.abs(gmx).max())

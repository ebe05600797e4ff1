"""Mixed precision vs the FP64 path.

All arithmetic stays FP64; "mixed" takes the total-phase sine/cosine from the FP32 units.

Near the likelihood peak the error enters lnL through <d - h, dh> ~ SNR eps and |dlnL| stays
below 1e-3.  Far from the peak the FP32 phase error, which changes from bin to bin, adds up
like a random walk over the bins, 2 sum_k w_k Im(d_k* h_k) dphi_k: a few 1e-3 at SNR ~ 2000
(asserted below 5e-3).
"""

import numpy as np
import pytest

from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud, prior_rows

pytestmark = pytest.mark.gpu


def _core(modes, d, labels):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    add_to_core(core, d, labels)
    return core


@pytest.mark.parametrize("modes,gen", [(((2, 2),), "1.5"), (oph.MODES, "1.5"), (oph.MODES, "2.0")])
def test_mixed_accuracy(modes, gen):
    precision = "mixed"
    d, labels = make_detector(modes=modes, gen=gen)
    core = _core(modes, d, labels)
    near = np.vstack([inj_row(), particle_cloud(63, seed=5, rel=1e-8)])
    far = particle_cloud(16, seed=6, rel=1e-4)
    rows = np.vstack([near, far])
    l64, g64 = core.lnl(rows, grad=True, precision="fp64")
    lmx, gmx = core.lnl(rows, grad=True, precision=precision)
    l64, lmx, g64, gmx = (x.cpu().numpy() for x in (l64, lmx, g64, gmx))
    dl = np.abs(lmx - l64)
    n = near.shape[0]
    scale = np.abs(g64[:n]).max(axis=0)
    rel = np.abs(gmx[:n] - g64[:n]) / scale
    print(f"{precision}: near: dlnl max {dl[:n].max():.2e} p99 {np.percentile(dl[:n], 99):.2e}; "
          f"far: dlnl max {dl[n:].max():.2e}; lnl-range near {l64[:n].min():.1f}..{l64[:n].max():.1f}; "
          f"far {l64[n:].min():.3g}..{l64[n:].max():.3g}; grad rel (to max |g|) max {rel.max():.2e}")
    assert dl[:n].max() < 1e-3
    assert dl[n:].max() < 5e-3
    # near the peak the gradient is O(SNR) while its FP32 error is O(eps SNR^2): ~1e-4 relative
    assert rel.max() < 5e-3
    r64 = core.response(0, rows[:2], precision="fp64").cpu().numpy()
    rmx = core.response(0, rows[:2], precision=precision).cpu().numpy()
    err = np.linalg.norm(rmx - r64) / np.linalg.norm(r64)
    print("response rel err", err)
    assert err < 2e-6


@pytest.mark.parametrize("P", [1, 257, 1001])
def test_uneven_batches_wide_prior(P):
    """Batches of uneven size (not multiples of the 40- or 160-row chunks) drawn from the wide prior:
    the mixed path against the FP64 path, and the batch against its rows evaluated alone."""
    modes = oph.MODES
    d, labels = make_detector(modes=modes)
    core = _core(modes, d, labels)
    rows = prior_rows(P, seed=P)
    l64, g64 = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="fp64"))
    lmx = core.lnl(rows, precision="mixed").cpu().numpy()
    rel = np.abs(lmx - l64) / np.abs(l64)
    print(f"P={P}: mixed |dlnl| max {np.abs(lmx - l64).max():.2e}, rel max {rel.max():.2e}")
    assert np.all(np.isfinite(l64)) and np.all(np.isfinite(g64))
    assert rel.max() < 1e-6
    for i in (0, P // 2, P - 1):
        l1 = core.lnl(rows[i:i + 1]).cpu().numpy()
        np.testing.assert_allclose(l1, l64[i:i + 1], rtol=1e-13)

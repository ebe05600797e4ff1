"""pes_lnl_batch_noalloc (the entry point that allocates nothing) against pes_lnl_batch.

Both run the same kernels on the same rows; the no-allocation entry point takes its scratch from
a fixed device array instead of the context's buffers, so results are bitwise equal.  Two calls
in a row, the second with fewer rows (the array keeps the first call's contents), with and
without the gradient, in both precision modes, and on a duty-cycled detector with rows in its
off period (the off-tile flags live in the array too); a batch beyond the array is rejected
before anything is launched.
"""

import ctypes

import numpy as np
import pytest

from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud

pytestmark = pytest.mark.gpu


# TianQin with a duty cycle whose window opens at 523,000 s (1500 s ramp): rows with t_c well
# before it lie in an off period, whose tiles the kernel skips (flags in the scratch too)
TIANQIN_DUTY = dict(orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3,
                    duty=(1.0e8, 5.0e7, 523000.0, 1500.0))


def _core(**det):
    from pespace_amd.core import LikelihoodCore
    modes = oph.MODES
    d, labels = make_detector(modes=modes, **det)
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    add_to_core(core, d, labels)
    core._build()
    return core


def _noalloc(core, rows, grad, precision):
    import torch

    from pespace_amd._lib import check, lib
    from pespace_amd.core import PRECISION
    x = torch.as_tensor(rows, device="cuda").contiguous()
    P = x.shape[0]
    lnl = torch.full((P,), np.nan, dtype=torch.float64, device="cuda")
    g = torch.full((P, 11), np.nan, dtype=torch.float64, device="cuda") if grad else None
    check(lib().pes_lnl_batch_noalloc(core._ctx, x.data_ptr(), P, lnl.data_ptr(),
                                      g.data_ptr() if grad else None, PRECISION[precision],
                                      torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return lnl.cpu().numpy(), (g.cpu().numpy() if grad else None)


@pytest.mark.parametrize("duty", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_pes_lnl_batch_on_two_calls(precision, duty):
    core = _core(**(TIANQIN_DUTY if duty else {}))
    rows = np.vstack([inj_row(), particle_cloud(40, seed=9)])
    if duty:  # every third row in the off period before the window
        rows[::3, 10] = 300000.0 + 1000.0 * np.arange(len(rows[::3]))
    for P in (41, 13):  # the second call is smaller: the array keeps the first call's contents
        for grad in (True, False):
            ref = core.lnl(rows[:P], grad=grad, precision=precision)
            ref_l = (ref[0] if grad else ref).cpu().numpy()
            got_l, got_g = _noalloc(core, rows[:P], grad, precision)
            np.testing.assert_array_equal(got_l, ref_l)
            if grad:
                np.testing.assert_array_equal(got_g, ref[1].cpu().numpy())


def test_noalloc_rejects_batches_beyond_its_array():
    import torch

    from pespace_amd._lib import lib
    core = _core()
    L = lib()
    limit = 64 << 20
    per_row = L.pes_lnl_batch_noalloc_bytes(core._ctx, 1, 1)
    assert 0 < per_row < limit
    P = limit // per_row + 8
    assert L.pes_lnl_batch_noalloc_bytes(core._ctx, P, 1) > limit
    x = torch.as_tensor(np.tile(inj_row(), (P, 1)), device="cuda").contiguous()
    out = torch.empty(P, dtype=torch.float64, device="cuda")
    g = torch.empty(P, 11, dtype=torch.float64, device="cuda")
    rc = L.pes_lnl_batch_noalloc(core._ctx, x.data_ptr(), P, out.data_ptr(), g.data_ptr(), 0,
                                 torch.cuda.current_stream().cuda_stream)
    assert rc == 1 and b"scratch" in L.pes_last_error()
    # the same rows in a batch that fits still run
    Pf = limit // per_row
    lnl, _ = _noalloc(core, np.tile(inj_row(), (Pf, 1)), True, "fp64")
    assert np.all(lnl == lnl[0]) and np.isfinite(lnl[0])

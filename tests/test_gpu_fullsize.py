"""Size-independent properties at the full 1-yr size (1,574,725 bins, six modes).

The oracle would take minutes per row at this size, so these tests check what holds at any
size: with noiseless data equal to the injection, lnL and its gradient vanish at the
injection; lnL is negative elsewhere; evaluations are deterministic, independent of the
batch a row sits in and of the kernel schedule, and the forward-only kernel reproduces the
gradient kernel's lnL; and the mixed path agrees with FP64 at the peak.
"""

import numpy as np
import pytest

from tests._setup import INJ as INJ_SHORT

pytestmark = pytest.mark.gpu

YEAR = 31557600.0


@pytest.fixture(scope="module")
def full():
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    from pespace_amd.inference.likelihood import WhittleLikelihoodHip
    from pespace_amd.waveform.phenom import IMRPhenomHM
    inj = dict(INJ_SHORT)
    inj["coalescence_time"] = 0.8 * YEAR
    td = TDIChannelData(backend="hip")
    td.set_fd_data_from_zero(("A", "E", "T"), YEAR, 5.0, 0.0, 1e-4, 0.05)
    td.set_fd_noise_power_density_from_model("LISA_SciRDv1", tdi_generation="1.5")
    det = InterferometerAntenna("LISA", td, "LISA_analytic", FDResponseModelMarset2018(),
                                FDMichelsonConstantEqualArm("1.5", True), backend="hip")
    wf = IMRPhenomHM(td.frequency_samples, 0.0, n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    wf.update_waveform(inj)
    det.inject_signal(wf.waveform_container, inj["ecliptic_longitude"], inj["ecliptic_latitude"],
                      inj["polarization"], inj["coalescence_time"])
    lk = WhittleLikelihoodHip(wf, det, ("A", "E", "T"))
    row = lk.row(inj)
    return lk, row


def test_noiseless_injection_is_the_maximum(full):
    lk, row = full
    assert lk.core.n_bins == [1574725]
    rng = np.random.default_rng(8)
    rows = np.vstack([row, row + 1e-6 * np.abs(row) * rng.standard_normal((7, 11))])
    lnl, g = (x.cpu().numpy() for x in lk.log_likelihood_and_grad(rows))
    # <h, h> of the injection from the offset rows' scale: lnL(inj) is zero to rounding
    assert abs(lnl[0]) < 1e-9 * np.max(np.abs(lnl[1:]))
    assert np.all(lnl[1:] < 0.0)
    # the gradient vanishes at the maximum relative to its size one step away
    scale = np.max(np.abs(g[1:]), axis=0)
    assert np.all(np.abs(g[0]) < 1e-5 * scale), np.abs(g[0]) / scale


def test_determinism_batch_and_schedule_invariance(full):
    lk, row = full
    core = lk.core
    rows = np.vstack([row + 1e-5 * np.abs(row) * np.random.default_rng(k).standard_normal(11) for k in range(20)])
    l1, g1 = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    l2, g2 = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    np.testing.assert_array_equal(l1, l2)
    np.testing.assert_array_equal(g1, g2)
    # the forward-only kernel (its own compilation unit) gives the gradient kernel's lnL
    l0 = core.lnl(rows).cpu().numpy()
    np.testing.assert_allclose(l0, l1, rtol=1e-10)
    for i in (0, 7, 19):
        la, ga = (x.cpu().numpy() for x in core.lnl(rows[i:i + 1], grad=True))
        np.testing.assert_array_equal(la, l1[i:i + 1])
        np.testing.assert_array_equal(ga, g1[i:i + 1])
    sched = core.schedule()
    try:
        core.set_schedule(5, 3)
        l3, g3 = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    finally:
        core.set_schedule(*sched)
    np.testing.assert_allclose(l3, l1, rtol=1e-12)
    assert np.max(np.abs(g3 - g1) / np.max(np.abs(g1), axis=0)) < 1e-10


def test_mixed_at_the_peak(full):
    lk, row = full
    rows = np.vstack([row, row + 1e-8 * np.abs(row) * np.random.default_rng(3).standard_normal((15, 11))])
    l64 = lk.core.lnl(rows, precision="fp64").cpu().numpy()
    lmx = lk.core.lnl(rows, precision="mixed").cpu().numpy()
    assert np.max(np.abs(lmx - l64)) < 1e-3

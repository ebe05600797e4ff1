"""The reference-style host API on the hip backend, end to end, against the oracle.

Exercises the classes a reference user switches to (as the reference's notebooks drive the
taichi ones): TDIChannelData + InterferometerAntenna with ``inject_signal`` and
``update_detector_response`` (hip waveform model, and a reference-style {plus, cross, tf}
container), IMRPhenomHM.get_waveform, WhittleLikelihoodHip (scalar, batched, autograd),
LikelihoodBilbyInterfaceHip, compute_whittle_likelihood, and HIP-graph capture of the batched
lnL + gradient step.
"""

import numpy as np
import pytest

from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import INJ, inj_row, make_detector, particle_cloud

pytestmark = pytest.mark.gpu

T, DT, FMIN, FMAX = 655360.0, 5.0, 1e-4, 0.05
INTERP = (1e-4, 0.05, 1024)
EXT = dict(lam=INJ["ecliptic_longitude"], beta=INJ["ecliptic_latitude"], psi=INJ["polarization"],
           tc=INJ["coalescence_time"])


def _relnorm(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def _host_detector(seed=7, inject=True):
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    from pespace_amd.waveform.phenom import IMRPhenomHM
    td = TDIChannelData(backend="hip")
    td.set_fd_data_from_zero(("A", "E", "T"), T, DT, 0.0, FMIN, FMAX)
    td.set_fd_noise_power_density_from_model("LISA_SciRDv1", tdi_generation="1.5")
    td.add_into_fd_data(td.get_fd_noise_realization(seed=seed))
    det = InterferometerAntenna("LISA", td, "LISA_analytic", FDResponseModelMarset2018(),
                                FDMichelsonConstantEqualArm("1.5", True), backend="hip")
    wf = IMRPhenomHM(td.frequency_samples, 0.0, n_nodes=INTERP[2], f_node_min=INTERP[0], f_node_max=INTERP[1])
    wf.update_waveform(INJ)
    if inject:
        det.inject_signal(wf.waveform_container, EXT["lam"], EXT["beta"], EXT["psi"], EXT["tc"])
    return det, wf


def test_inject_signal_matches_oracle_data():
    det, _ = _host_detector()
    d, labels = make_detector(modes=oph.MODES, T=T, dt=DT, fmin=FMIN, fmax=FMAX, interp=INTERP)
    for c in labels:
        assert _relnorm(getattr(det.tdi_data.fd_data, c), d.data[c]) < 1e-10, c


def test_update_response_from_container():
    det, wf = _host_detector(inject=False)
    cont = wf.get_waveform(INJ)  # mode-keyed {plus, cross, tf} on the GPU, t_c = 0
    det.update_detector_response(cont, EXT["lam"], EXT["beta"], EXT["psi"], EXT["tc"])
    resp = det.tdi_response_numpy
    d, labels = make_detector(modes=oph.MODES, T=T, dt=DT, fmin=FMIN, fmax=FMAX, interp=INTERP, inject=False)
    ref = olk.response(INJ, d, oph.MODES, 0.0, INTERP)
    for c in ("A", "E"):
        assert _relnorm(getattr(resp, c), ref[c]) < 1e-10, c
    xyz = np.sqrt(np.sum(np.abs(ref["A"]) ** 2 + np.abs(ref["E"]) ** 2))
    assert _relnorm(resp.T, ref["T"]) < max(1e-10, 64 * np.finfo(float).eps * xyz / np.linalg.norm(ref["T"]))


def test_whittle_likelihood_hip_api():
    import torch

    from pespace_amd.inference.likelihood import LikelihoodBilbyInterfaceHip, WhittleLikelihoodHip
    det, wf = _host_detector()
    d, _ = make_detector(modes=oph.MODES, T=T, dt=DT, fmin=FMIN, fmax=FMAX, interp=INTERP)
    lk = WhittleLikelihoodHip(wf, det, ("A", "E", "T"))
    params = lk.likelihood_param_type(**INJ)
    ref0 = olk.lnl(INJ, [d], oph.MODES, 0.0, INTERP)
    assert abs(lk.log_likelihood(params) - ref0) < 1e-10 * abs(ref0)
    rows = np.vstack([inj_row(), particle_cloud(5)])
    ref = olk.lnl_batch(rows, [d], oph.MODES, 0.0, INTERP)
    np.testing.assert_allclose(lk.log_likelihood_batch(rows).cpu().numpy(), ref, rtol=1e-10)
    # autograd: backward of the Function returns the adjoint kernel's gradient
    lnl, g = lk.log_likelihood_and_grad(rows)
    p = torch.as_tensor(rows, device="cuda").requires_grad_(True)
    out = lk(p)
    w = torch.linspace(0.5, 1.5, rows.shape[0], device="cuda", dtype=torch.float64)
    (out * w).sum().backward()
    assert torch.equal(out.detach(), lnl)
    assert torch.allclose(p.grad, g * w[:, None], rtol=0, atol=0)
    # bilby-style interface
    bl = LikelihoodBilbyInterfaceHip(wf, det, ("A", "E", "T"))
    bl.parameters = dict(INJ)
    assert abs(bl.log_likelihood() - ref0) < 1e-10 * abs(ref0)


def test_compute_whittle_likelihood():
    from pespace_amd.inference.likelihood import compute_whittle_likelihood
    det, wf = _host_detector()
    cont = wf.get_waveform(INJ)
    det.update_detector_response(cont, EXT["lam"], EXT["beta"] + 1e-3, EXT["psi"], EXT["tc"])
    td = det.tdi_data
    chans = ("A", "E")
    val = compute_whittle_likelihood(chans, td.fd_data, det.tdi_response, td.fd_noise_power_density,
                                     td.data_info.delta_frequency)
    from oracle import detector as odet
    ref = odet.whittle_lnl(chans, {c: np.asarray(getattr(td.fd_data, c)) for c in chans},
                           {c: getattr(det.tdi_response_numpy, c) for c in chans},
                           {c: np.asarray(getattr(td.fd_noise_power_density, c)) for c in chans},
                           td.data_info.delta_frequency)
    assert isinstance(val, float)
    assert abs(val - ref) < 1e-12 * abs(ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hip_graph_capture_of_lnl_grad_step(precision):
    """Capture one batched lnL + gradient step in a HIP graph, replay it with new parameters
    written into the captured input between replays (with eager work and synchronisation in
    between), and require every replay to equal the eager result bitwise."""
    import torch

    from pespace_amd.inference.likelihood import WhittleLikelihoodHip
    det, wf = _host_detector()
    lk = WhittleLikelihoodHip(wf, det, ("A", "E", "T"))
    core = lk.core
    P = 8
    batches = [np.vstack([inj_row(), particle_cloud(P - 1, seed=100 + k)]) for k in range(4)]
    x = torch.as_tensor(batches[0], device="cuda").contiguous()
    lnl = torch.empty(P, dtype=torch.float64, device="cuda")
    g = torch.empty(P, 11, dtype=torch.float64, device="cuda")
    core.reserve(P, grad=True)  # sizes the scratch, which stays pinned while the graph lives
    core.pin_scratch(True)
    core.lnl(x, grad=True, precision=precision, out=lnl, grad_out=g)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        core.lnl(x, grad=True, precision=precision, out=lnl, grad_out=g)
    for rows in batches[1:] + batches[:1]:
        x.copy_(torch.as_tensor(rows, device="cuda"))
        graph.replay()
        torch.cuda.synchronize()
        l_e, g_e = core.lnl(rows, grad=True, precision=precision)
        torch.cuda.synchronize()
        assert torch.equal(lnl, l_e)
        assert torch.equal(g, g_e)
    # a larger batch would reallocate the scratch under the graph: rejected while pinned
    from pespace_amd._lib import PespaceError
    with pytest.raises(PespaceError, match="pinned"):
        core.lnl(np.vstack(batches), grad=True, precision=precision)
    core.pin_scratch(False)
    del graph


def test_schedules_agree_and_autotune_picks_one():
    """Every schedule of the fused kernel gives the same lnL and gradient (up to the order of
    the partial sums), and the autotuner keeps the fastest candidate it timed."""
    import torch

    from pespace_amd.inference.likelihood import WhittleLikelihoodHip
    det, wf = _host_detector()
    lk = WhittleLikelihoodHip(wf, det, ("A", "E", "T"))
    core = lk.core
    rows = np.vstack([inj_row(), particle_cloud(40, seed=11)])
    ref_l, ref_g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
    scale = np.abs(ref_g).max(axis=0)
    for t, b in ((1, 1), (3, 2), (64, 8), (4096, 1)):
        core.set_schedule(t, b)
        assert core.schedule() == (t, b)
        l, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
        np.testing.assert_allclose(l, ref_l, rtol=1e-12)
        assert np.max(np.abs(g - ref_g) / scale) < 1e-11
    # rows per work item: any value gives the same results (the partial sums are per row)
    for r in (1, 7, 40, 160):
        core.set_schedule(8, 2, r)
        assert core.rows_per_item() == r
        l, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True))
        np.testing.assert_allclose(l, ref_l, rtol=1e-12)
        assert np.max(np.abs(g - ref_g) / scale) < 1e-11
        np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), ref_l, rtol=1e-12)
    best, times = core.autotune(rows, grad=True, candidates=[(8, 2), (32, 4), (8, 2, 10)], reps=1, warmup=0)
    assert best in times and times[best] == min(times.values())
    assert core.schedule() == best[:2] and core.rows_per_item() == (best[2] if len(best) > 2 else 0)
    torch.cuda.synchronize()


def test_invalid_rows_and_empty_batch():
    """Rows outside the model's domain give lnL = -inf and a zero gradient without disturbing
    the valid rows of the same batch; an empty batch is a no-op."""
    import torch

    from pespace_amd.inference.likelihood import WhittleLikelihoodHip
    det, wf = _host_detector()
    lk = WhittleLikelihoodHip(wf, det, ("A", "E", "T"))
    good = np.vstack([inj_row(), particle_cloud(2, seed=4)])
    bad = np.repeat(inj_row()[None, :], 5, axis=0)
    bad[0, 1] = 1.5        # mass ratio above 1
    bad[1, 2] = 1.0        # |chi_1| = 1
    bad[2, 4] = -1.0       # negative distance
    bad[3, 10] = np.nan    # non-finite
    bad[4, 0] = 0.0        # zero chirp mass
    rows = np.vstack([good[:1], bad, good[1:]])
    lnl, g = (x.cpu().numpy() for x in lk.log_likelihood_and_grad(rows))
    ref_l, ref_g = (x.cpu().numpy() for x in lk.log_likelihood_and_grad(good))
    assert np.all(np.isneginf(lnl[1:6])) and np.all(g[1:6] == 0.0)
    np.testing.assert_array_equal(lnl[[0, 6, 7]], ref_l)
    np.testing.assert_array_equal(g[[0, 6, 7]], ref_g)
    empty = lk.log_likelihood_batch(torch.empty(0, 11, dtype=torch.float64, device="cuda"))
    assert empty.shape == (0,)
    # an unknown precision mode is rejected at the C boundary before anything is launched
    from pespace_amd._lib import lib
    x = torch.as_tensor(good, device="cuda")
    out = torch.empty(x.shape[0], dtype=torch.float64, device="cuda")
    rc = lib().pes_lnl_batch(lk.core._ctx, x.data_ptr(), x.shape[0], out.data_ptr(), None, 2,
                             torch.cuda.current_stream().cuda_stream)
    assert rc != 0 and b"precision" in lib().pes_last_error()


@pytest.mark.parametrize("case", ["decreasing", "node_range", "psd", "generation", "orbit", "schedule"])
def test_detector_validation_at_the_boundary(case):
    """Grids and settings the kernel cannot take are rejected by pes_context_add_detector (or
    pes_context_set_schedule) with a message, before any kernel runs; the context is released."""
    from pespace_amd._lib import PespaceError
    from pespace_amd.core import LikelihoodCore
    n = 64
    f = np.linspace(1e-3, 2e-3, n)
    psd = np.ones((3, n))
    kind, gen = 0, 15
    if case == "decreasing":
        f[10] = f[9]
        match = "strictly increasing"
    elif case == "node_range":
        f = np.linspace(0.05, 0.4, n)  # above the node range [1e-4, 0.3] Hz: no extrapolation
        match = "outside the interpolation node range"
    elif case == "psd":
        psd[1, 7] = 0.0
        match = "PSD must be positive"
    elif case == "generation":
        gen = 17
        match = "tdi_generation"
    elif case == "orbit":
        kind = 2
        match = "orbit_kind"
    else:
        match = "schedule"
    core = LikelihoodCore(modes=["22"], n_nodes=64, f_node_min=1e-4, f_node_max=0.3)
    core.add_detector(kind, 2.5e9, 0.0, 0.0, gen, True, f, np.zeros((3, n), complex), psd, (1, 1, 1), df=1e-5)
    with pytest.raises(PespaceError, match=match):
        if case == "schedule":
            core.set_schedule(0, 2)
        else:
            core.lnl(inj_row()[None, :])
    assert case == "schedule" or core._ctx is None

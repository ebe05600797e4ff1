"""Duty-cycled detectors (pes_context_set_duty_cycle) against the oracle.

The benchmark network's TianQin observes three months on, three months off; the reference
models no duty cycle, so the window is this build's (oracle/detector.py duty_window: each
mode's response at a bin weighted by a C1 smoothstep window at t_f + t_c).  Checked here on
short grids whose t_f + t_c spans several windows, so that bins fall on plateaus, ramps and
off periods:

* the TDI response, the links export, the container response and lnL against the oracle
  (rtol 1e-10), heliocentric and geocentric orbits, the forward-only and gradient kernels;
* the adjoint gradient against Richardson central differences of the HIP lnL (the window's
  slope enters t_c and, through t_f, the intrinsic parameters);
* the mixed mode's lnL equals the FP64 lnL (its forward is FP64) and its gradient agrees to 1e-3;
* a three-detector network with only TianQin duty-cycled;
* rows whose signal lies wholly in an off period, whose modes the kernel skips wave by wave:
  lnL is the data-only value (bitwise the same for every such row), the gradient exactly zero,
  and the work count (pes_profile_work) leaves out every (bin, mode) with a zero window,
  checked against the oracle's t_f row by row.
"""

import numpy as np
import pytest

from oracle import detector as odet
from oracle import likelihood as olk
from oracle import phenom as oph
from tests._setup import add_to_core, inj_row, make_detector, particle_cloud, prior_rows

pytestmark = pytest.mark.gpu

INTERP = (1e-4, 0.05, 1024)
# On a linear grid most live bins lie within ~500 s of t_c = 524,288 s (merger and ringdown),
# so the test windows put one edge there, with a ramp of 2000 s (LISA: the window ends at
# 524,400 s; 1.5% of the injection's live (bin, mode) off, 89% on the ramp) or 1500 s (TianQin:
# it starts at 523,000 s; 7% off, 93% on the ramp).  The windows' other edges lie 5e7 s away,
# beyond every bin, so no bin with a large |t_f| sits on a ramp (there the last-digit
# differences of t_f between the kernel and the oracle, ~1e-13 of 1e7 s, would show).  Ramps
# much longer than the finite-difference steps (0.5 s in t_c) keep the gradient check clear of
# the smoothstep's jumps in curvature at the ramp ends.
DUTY_HELIO = (1.0e8, 5.0e7, 524400.0 - 5.0e7, 2000.0)
DUTY_GEO = (1.0e8, 5.0e7, 523000.0, 1500.0)
CASES = {
    "lisa": dict(modes=oph.MODES, duty=DUTY_HELIO),
    "lisa22-2.0-xyz": dict(modes=((2, 2),), gen="2.0", orthogonal=False, duty=DUTY_HELIO),
    "tianqin": dict(modes=oph.MODES, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3,
                    duty=DUTY_GEO),
}


def _core(modes, dets):
    from pespace_amd.core import LikelihoodCore
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in modes], n_nodes=1024, f_node_min=1e-4, f_node_max=0.05)
    for d, labels in dets:
        add_to_core(core, d, labels)
    return core


def _relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _window_coverage(d, row, modes):
    """Fractions of (bin, mode) at w = 0, 0 < w < 1 for this row."""
    p = olk.as_dict(row)
    wf = olk.waveform(p, d.f, modes, 0.0, INTERP)
    w = np.concatenate([odet.duty_window(m["tf"] + p["coalescence_time"], d.duty) for m in wf.values()])
    return np.mean(w == 0.0), np.mean((w > 0.0) & (w < 1.0))


@pytest.mark.parametrize("name", list(CASES))
def test_duty_response_links_and_lnl(name):
    case = dict(CASES[name])
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(3)])
    off, ramp = _window_coverage(d, rows[0], modes)
    print(f"{name}: (bin, mode) off {off:.3f}, on a ramp {ramp:.3f}", flush=True)
    assert off > 0.01 and ramp > 0.05
    resp = core.response(0, rows).cpu().numpy()
    links = core.links(0, rows).cpu().numpy()
    for p in range(rows.shape[0]):
        prm = olk.as_dict(rows[p])
        ref = olk.response(prm, d, modes, 0.0, INTERP)
        for ci, c in enumerate(labels):
            tol = 1e-10
            if c == "T":
                xyz = np.sqrt(np.sum(np.abs(ref[labels[0]]) ** 2 + np.abs(ref[labels[1]]) ** 2))
                tol = max(tol, 64 * np.finfo(float).eps * xyz / np.linalg.norm(ref["T"]))
            assert _relnorm(resp[p, ci], ref[c]) < tol, (p, c)
        wf = olk.waveform(prm, d.f, modes, 0.0, INTERP)
        ref_links = odet.single_links(wf, prm["ecliptic_longitude"], prm["ecliptic_latitude"], prm["polarization"],
                                      prm["coalescence_time"], d.f, d.orbit, d.duty)
        for li, nm in enumerate(("link12", "link21", "link23", "link32", "link31", "link13")):
            assert _relnorm(links[p, li], ref_links[nm]) < 1e-10, (p, nm)
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), olk.lnl_batch(rows, [d], modes, 0.0, INTERP),
                               rtol=1e-10)
    lnl_g, _ = core.lnl(rows, grad=True)
    np.testing.assert_allclose(lnl_g.cpu().numpy(), olk.lnl_batch(rows, [d], modes, 0.0, INTERP), rtol=1e-10)


def test_duty_container_response():
    modes = oph.MODES
    d, labels = make_detector(modes=modes, duty=DUTY_HELIO)
    core = _core(modes, [(d, labels)])
    prm = olk.as_dict(inj_row())
    wf = olk.waveform(prm, d.f, modes, 0.0, INTERP)
    got = core.response_from_container(0, wf, prm["ecliptic_longitude"], prm["ecliptic_latitude"],
                                       prm["polarization"], prm["coalescence_time"]).cpu().numpy()
    ref = olk.response(prm, d, modes, 0.0, INTERP, wf=wf)
    for ci, c in enumerate(labels[:2]):
        assert _relnorm(got[ci], ref[c]) < 1e-10, c


@pytest.mark.parametrize("name", ["lisa", "tianqin"])
def test_duty_gradient_vs_fd(name):
    from tests.test_gpu_grad import _fd_check_core
    case = dict(CASES[name])
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    _fd_check_core(core, np.vstack([inj_row(), particle_cloud(2, rel=1e-5)]))


def test_duty_mixed_mode():
    case = dict(CASES["tianqin"])
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    rows = np.vstack([inj_row(), particle_cloud(3), prior_rows(4, seed=12)])
    l64, g64 = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="fp64"))
    lmx, gmx = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision="mixed"))
    ok = np.isfinite(l64)
    np.testing.assert_allclose(lmx[ok], l64[ok], rtol=1e-13)
    scale = np.maximum(np.abs(g64[ok]).max(axis=0), 1e-300)
    assert np.max(np.abs(gmx[ok] - g64[ok]) / scale) < 1e-3


def test_duty_network_only_tianqin():
    modes = oph.MODES
    dets = [make_detector(modes=modes, seed=7),
            make_detector(modes=modes, orbit="Taiji_analytic", psd_model="Taiji_TDC", fmax=0.03, seed=8),
            make_detector(modes=modes, orbit="Tianqin_analytic", psd_model="Tianqin_GWSpace", fmin=1e-3, seed=9,
                          duty=DUTY_GEO)]
    core = _core(modes, dets)
    rows = np.vstack([inj_row(), particle_cloud(2)])
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(),
                               olk.lnl_batch(rows, [d for d, _ in dets], modes, 0.0, INTERP), rtol=1e-10)
    # the same data with TianQin always observing: the model then has signal in the off
    # periods, which the data lack, so lnL drops (and matches the oracle's without the window)
    import copy
    tq = copy.copy(dets[2][0])
    tq.duty = None
    plain = _core(modes, dets[:2] + [(tq, dets[2][1])])
    l_plain = plain.lnl(rows).cpu().numpy()
    np.testing.assert_allclose(l_plain, olk.lnl_batch(rows, [dets[0][0], dets[1][0], tq], modes, 0.0, INTERP),
                               rtol=1e-10)
    assert np.all(core.lnl(rows).cpu().numpy() - l_plain > 1.0)


def _needed_bin_modes(d, row, modes):
    """(bin, mode) below the mode's cut, and those of them inside an observation window (the
    oracle's t_f + t_c): the work count without and with the duty cycle."""
    p = olk.as_dict(row)
    hm = oph.PhenomHM(p["chirp_mass"], p["mass_ratio"], p["chi_1"], p["chi_2"], 0.0, modes)
    wf = olk.waveform(p, d.f, modes, 0.0, INTERP)
    n_all = n_on = 0
    for md in modes:
        keep = d.f * hm.Msec <= hm.mode_cut(md)
        w = odet.duty_window(wf[f"{md[0]}{md[1]}"]["tf"] + p["coalescence_time"], d.duty)
        n_all += int(keep.sum())
        n_on += int((keep & (w != 0.0)).sum())
    return n_all, n_on


def test_duty_off_period_rows_skip_exactly():
    from pespace_amd._lib import PespaceError
    modes = oph.MODES
    case = dict(CASES["tianqin"])
    case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    # the window opens at 523,000 s: with t_c = 300,000 or 400,000 s every bin's t_f + t_c (at most
    # ~3e4 s before the merger on this grid) lies in the off period before it
    off = np.vstack([inj_row(coalescence_time=300000.0),
                     particle_cloud(2, seed=5, center=inj_row(coalescence_time=400000.0))])
    rows = np.vstack([off, inj_row(), particle_cloud(2)])
    print(f"off-period rows: {len(off)} of {len(rows)}", flush=True)
    ref = olk.lnl_batch(rows, [d], modes, 0.0, INTERP)
    for prec in ("fp64", "mixed"):
        lnl, g = (x.cpu().numpy() for x in core.lnl(rows, grad=True, precision=prec))
        np.testing.assert_allclose(lnl, ref, rtol=1e-10)
        assert np.all(lnl[:len(off)] == lnl[0]) and np.all(g[:len(off)] == 0.0), prec
        assert np.all(np.abs(g[len(off):]).max(axis=1) > 0.0)
        np.testing.assert_array_equal(core.lnl(rows, precision=prec).cpu().numpy(), lnl)
    core.set_profiling(True)
    try:
        for i, row in enumerate(rows):
            core.lnl(row[None], grad=True)
            core.read_profile()
            bm, br = core.profile_work()
            n_all, n_on = _needed_bin_modes(d, row, modes)
            assert bm == n_on, (i, bm, n_on, n_all)
            assert br > 0
            if i < len(off):
                assert n_on == 0 and n_all > 0
            else:
                assert 0 < n_on < n_all  # the injection's window starts near its merger
        # the count is taken from the scratch when asked: another call in between refills it
        core.lnl(rows[:1], grad=True)
        core.read_profile()
        core.response(0, rows[1:2])
        with pytest.raises(PespaceError, match="reused the scratch"):
            core.profile_work()
    finally:
        core.set_profiling(False)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_duty_graph_capture_matches_eager(precision):
    """A batch on a duty-cycled detector (off-period flags, plateau work items, and the range's two
    launches) captured in a HIP graph: every replay with new rows written into the captured input
    equals the eager result bitwise, for rows on the plateau, on the ramps and in the off period."""
    import torch

    from pespace_amd._lib import PespaceError
    case = dict(CASES["tianqin"])
    modes = case.pop("modes")
    d, labels = make_detector(modes=modes, **case)
    core = _core(modes, [(d, labels)])
    P = 12
    batches = []
    for k in range(3):
        rows = np.vstack([inj_row(), particle_cloud(P - 1, seed=40 + k)])
        rows[1:4, 10] = 300000.0 + 1000.0 * np.arange(3) + k  # off period
        rows[8:, 10] = 5.0e6 + 1000.0 * np.arange(4) + k      # on the plateau, far from the ramps
        batches.append(rows)
    # 12 rows give work items of 8 rows: rows 0-7 (ramps and off period) run in the duty
    # instantiation, rows 8-11 (plateau) in the plain one
    x = torch.as_tensor(batches[0], device="cuda").contiguous()
    lnl = torch.empty(P, dtype=torch.float64, device="cuda")
    g = torch.empty(P, 11, dtype=torch.float64, device="cuda")
    core.reserve(P, grad=True)
    core.pin_scratch(True)
    try:
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
            assert torch.equal(lnl, l_e) and torch.equal(g, g_e)
            np.testing.assert_allclose(l_e.cpu().numpy(), olk.lnl_batch(rows, [d], modes, 0.0, INTERP), rtol=1e-10)
        with pytest.raises(PespaceError, match="pinned"):
            core.lnl(np.vstack(batches), grad=True, precision=precision)
        del graph
    finally:
        core.pin_scratch(False)

"""The HIP path against the committed golden fixtures (tests/golden/golden_parity.npz).

Inputs are rebuilt from each fixture's specification with the product's own host API (grid,
PSD and the seeded noise realization, which is bit-identical to the oracle's), so this test
needs no oracle at run time: lnL of six rows, and for the first row the TDI response and the
waveform container of every mode at 24 sample bins, to relative 1e-10.
"""

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_parity.npz"))
NAMES = sorted({k.split("/")[0] for k in FIX.files})
ORBIT_KIND = {"LISA_analytic": 0, "Taiji_analytic": 0, "Tianqin_analytic": 1}


def _core(spec):
    from pespace_amd.core import LikelihoodCore
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna, add_detector_to_core
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    orth = spec.get("orthogonal", True)
    labels = ("A", "E", "T") if orth else ("X", "Y", "Z")
    td = TDIChannelData(backend="hip")
    td.set_fd_data_from_zero(labels, spec["T"], spec["dt"], 0.0, spec["fmin"], spec["fmax"])
    td.set_fd_noise_power_density_from_model(spec["psd_model"], tdi_generation=spec["gen"])
    td.add_into_fd_data(td.get_fd_noise_realization(seed=spec["seed"]))
    det = InterferometerAntenna("det", td, spec["orbit"], FDResponseModelMarset2018(),
                                FDMichelsonConstantEqualArm(spec["gen"], orth), backend="hip")
    fmin_n, fmax_n, nn = spec["interp"]
    core = LikelihoodCore(modes=[f"{l}{m}" for l, m in spec["modes"]], reference_frequency=0.0, n_nodes=nn,
                          f_node_min=fmin_n, f_node_max=fmax_n)
    add_detector_to_core(core, det, data_channels=labels)
    return core


@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(name):
    spec = json.loads(str(FIX[f"{name}/spec"]))
    core = _core(spec)
    rows = FIX[f"{name}/rows"]
    np.testing.assert_allclose(core.lnl(rows).cpu().numpy(), FIX[f"{name}/lnl"], rtol=1e-10)
    idx = FIX[f"{name}/idx"]
    resp = core.response(0, rows[:1]).cpu().numpy()[0][:, idx]
    ref = FIX[f"{name}/resp"]
    for c in range(3):
        tol = 1e-10
        if str(FIX[f"{name}/labels"][c]) == "T":  # the oracle's (X+Y+Z)/sqrt3 cancels at low f
            tol = max(tol, 64 * np.finfo(float).eps * np.max(np.abs(ref[:2])) / np.max(np.abs(ref[2])))
        assert np.max(np.abs(resp[c] - ref[c])) <= tol * np.max(np.abs(ref[c])), (name, c)
    wf = core.waveform(0, rows[:1]).cpu().numpy()[0][:, idx]  # [mode][bin][5]
    refw = FIX[f"{name}/wf"]  # [mode][plus, cross, tf][bin]
    for m in range(refw.shape[0]):
        for k, (re, im) in enumerate(((0, 1), (2, 3))):
            got = wf[m, :, re] + 1j * wf[m, :, im]
            assert np.max(np.abs(got - refw[m, k])) <= 1e-10 * np.max(np.abs(refw[m, k])), (name, m, k)
        assert np.max(np.abs(wf[m, :, 4] - refw[m, 2].real)) <= 1e-10 * np.max(np.abs(refw[m, 2].real)), (name, m)

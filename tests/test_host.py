"""Host-side logic and the C-ABI boundary (no GPU compute here).

* libpespace_amd.so loads and exports every function include/pespace_amd.h declares;
* argument validation of the C entry points (returns status codes, never crashes);
* the product's setup-time numpy code (grid, PSD, noise) is bit-identical to the oracle;
* reference-style errors of the host API (tdi.py, antenna.py, likelihood_jax.py).
"""

import ctypes
import os
import re
from collections import namedtuple

import numpy as np
import pytest

from oracle import detector as odet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pespace_amd.h")
LIB = os.path.join(ROOT, "pespace_amd", "libpespace_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pes_[a-z_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    from pespace_amd import _lib
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert bound == set(names), set(names) ^ bound


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_c_abi_argument_validation():
    from pespace_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    ids = (ctypes.c_int * 1)(0)
    assert L.pes_context_create(None, 0, 1, ids, 0.0, 1024, 1e-4, 0.05) == 1
    assert L.pes_context_create(ctypes.byref(h), 0, 0, ids, 0.0, 1024, 1e-4, 0.05) == 1
    assert L.pes_context_create(ctypes.byref(h), 0, 1, ids, 0.0, 1, 1e-4, 0.05) == 1
    assert L.pes_context_create(ctypes.byref(h), 0, 1, ids, 0.0, 16, 0.05, 1e-4) == 1
    bad = (ctypes.c_int * 2)(0, 0)
    assert L.pes_context_create(ctypes.byref(h), 0, 2, bad, 0.0, 16, 1e-4, 0.05) == 1
    assert b"duplicate" in L.pes_last_error()
    assert L.pes_lnl_batch(None, None, 1, None, None, 0, None) != 0
    assert L.pes_lnl_batch(None, None, 1, None, None, 7, None) == 1
    assert L.pes_whittle_lnl(None, None, None, 1, 1, 1.0, None, None, None) == 1
    assert L.pes_detector_bins(None, 0) == -1
    assert L.pes_context_set_schedule(None, 32, 4) == 1
    t, b = ctypes.c_int(), ctypes.c_int()
    assert L.pes_context_get_schedule(None, ctypes.byref(t), ctypes.byref(b)) == 1
    assert L.pes_links(None, 0, None, 1, None, None) == 1
    assert L.pes_links_from_waveform(None, 0, 1, None, 0.0, 0.0, 0.0, 0.0, None, None) == 3
    assert L.pes_tdi_from_links(None, 0, 1, None, None, None) == 3
    assert L.pes_orbit_vectors(2, 2.5e9, 0.0, 0.0, 0.0, 0.0, 0.0, None, 1, None, None) == 1
    assert L.pes_orbit_vectors(0, 2.5e9, 0.0, 0.0, 0.0, 0.0, 0.0, None, -1, None, None) == 1
    assert L.pes_orbit_vectors(0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, None, 0, None, None) == 1
    assert b"arm length" in L.pes_last_error()
    assert L.pes_context_reserve(None, 8, 1) == 1
    assert L.pes_context_pin_scratch(None, 1) == 1


def test_lnl_output_buffers_are_checked():
    """Caller-supplied lnL / gradient buffers go to the kernels as raw pointers: wrong dtype,
    device, shape or layout is rejected on the host."""
    import torch

    from pespace_amd.core import LikelihoodCore
    dev = torch.device("cuda:0")
    chk = LikelihoodCore._check_out
    good = torch.empty(4, dtype=torch.float64)
    with pytest.raises(ValueError, match="float64"):
        chk(torch.empty(4, dtype=torch.float32), "out", (4,), good.device)
    with pytest.raises(ValueError, match="must be on"):
        chk(good, "out", (4,), dev)
    with pytest.raises(ValueError, match="shape"):
        chk(torch.empty(3, dtype=torch.float64), "out", (4,), good.device)
    with pytest.raises(ValueError, match="shape"):
        chk(torch.empty(4, 10, dtype=torch.float64), "grad_out", (4, 11), good.device)
    with pytest.raises(ValueError, match="contiguous"):
        chk(torch.empty(11, 4, dtype=torch.float64).t(), "grad_out", (4, 11), good.device)
    with pytest.raises(TypeError):
        chk(np.empty(4), "out", (4,), good.device)
    chk(good, "out", (4,), good.device)


def test_product_grid_psd_noise_match_oracle():
    from pespace_amd.detector.noise import available_noise_models
    from pespace_amd.detector.tdi import DataInformation, TDIChannelData
    info = DataInformation(("A", "E", "T"), 2 ** 15 * 5.0, 5.0, 0.0, 1e-4, 0.05)
    f, mask, n_t, df, _ = odet.data_grid(2 ** 15 * 5.0, 5.0, 1e-4, 0.05)
    assert np.array_equal(info.frequency_samples_array, f) and info.delta_frequency == df
    for name in available_noise_models:
        for gen in ("1.5", "2.0"):
            a = available_noise_models[name](f, ("X", "A", "T"), False, gen)
            b = odet.psd(name, f, ("X", "A", "T"), False, gen)
            for c in ("X", "A", "T"):
                assert np.array_equal(a[c], b[c]), (name, gen, c)
    td = TDIChannelData(backend="hip")
    td.set_fd_data_from_zero(("A", "E", "T"), 2 ** 15 * 5.0, 5.0, 0.0, 1e-4, 0.05)
    td.set_fd_noise_power_density_from_model("LISA_SciRDv1", tdi_generation="1.5")
    n = td.get_fd_noise_realization(seed=123)
    ref = odet.noise_realization(123, ("A", "E", "T"), n_t, mask, odet.psd("LISA_SciRDv1", f, "AET", False, "1.5"),
                                 df)
    for c in "AET":
        assert np.array_equal(getattr(n, c), ref[c])


def test_tdi_channel_data_errors():
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    with pytest.raises(ValueError):
        TDIChannelData(backend="taichi")
    td = TDIChannelData()
    with pytest.raises(ValueError):
        td.set_fd_noise_power_density_from_model("LISA_SciRDv1", tdi_generation="1.5")
    with pytest.raises(ValueError):
        td.set_fd_data_from_input(("A", "E"), 655360.0, 5.0, np.zeros((3, 10)))
    td.set_fd_data_from_zero(("A", "E"), 655360.0, 5.0)
    with pytest.raises(ValueError):
        td.set_fd_noise_power_density_from_model("nope", tdi_generation="1.5")
    with pytest.raises(ValueError):
        td.get_fd_noise_realization()
    DC = namedtuple("DC", ("A", "T"))
    n = td.data_info.frequency_series_length
    with pytest.raises(ValueError):
        td.add_into_fd_data(DC(np.zeros(n), np.zeros(n)))
    with pytest.raises(TypeError):
        td.add_into_fd_data({"A": np.zeros(n)})
    with pytest.raises(ValueError):
        FDMichelsonConstantEqualArm(generation="3.0")


def test_antenna_and_orbit_errors():
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna
    from pespace_amd.detector.orbit import KaplerianHeliocentric, KeplerianGeocentric
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    with pytest.raises(ValueError):
        KaplerianHeliocentric(2.5e9, backend="jax")
    with pytest.raises(ValueError):
        KeplerianGeocentric(1.7e8, backend="taichi")
    # orbit_params follow orbit.py:112-139 / 189-228; any geocentric rotation and reference source
    geo = KeplerianGeocentric(1.7e8, 0.4, 1.1, omega_rotation=1e-5, lambda_ref=0.7, beta_ref=0.3)
    op = geo.orbit_params
    assert op.omega_rotation == 1e-5 and op.slam_ref == np.sin(0.7) and op.cbeta_ref == np.cos(0.3)
    assert op.r_det_sec == op.static_arm_sec / np.sqrt(3) and geo.arm_length_sec(5.0) == 1.7e8 / 299792458.0
    hel = KaplerianHeliocentric(2.5e9, 0.3, -0.2)
    assert hel.orbit_params.r_prime == 2.5e9 / 299792458.0 / (2 * np.sqrt(3))
    assert abs(hel.orbit_params.ck3 - np.cos(0.3 + 4 * np.pi / 3)) < 1e-15
    td = TDIChannelData()
    with pytest.raises(ValueError):
        InterferometerAntenna("x", td, "nope", FDResponseModelMarset2018(), FDMichelsonConstantEqualArm())
    with pytest.raises(TypeError):
        InterferometerAntenna("x", td, 3.0, FDResponseModelMarset2018(), FDMichelsonConstantEqualArm())
    with pytest.raises(ValueError):
        InterferometerAntenna("x", td, "LISA_analytic", FDResponseModelMarset2018(), FDMichelsonConstantEqualArm(),
                              backend="taichi")


def test_likelihood_validation_before_gpu():
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna
    from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm, TDIChannelData
    from pespace_amd.inference.likelihood import WhittleLikelihoodHip
    from pespace_amd.waveform.phenom import IMRPhenomD
    td = TDIChannelData()
    td.set_fd_data_from_zero(("A", "E", "T"), 655360.0, 5.0, 0.0, 1e-4, 0.05)
    det = InterferometerAntenna("lisa", td, "LISA_analytic", FDResponseModelMarset2018(),
                                FDMichelsonConstantEqualArm())
    wf = IMRPhenomD(td.frequency_samples)
    with pytest.raises(TypeError):
        WhittleLikelihoodHip(object(), det, ("A", "E"))
    with pytest.raises(ValueError):
        WhittleLikelihoodHip(wf, det, ("X", "E"))


def test_waveform_mass_conventions():
    from pespace_amd.waveform.phenom import _mass_params
    mc, q = _mass_params({"mass_1": 2.0e6, "mass_2": 1.0e6})
    assert abs(q - 0.5) < 1e-15
    mc2, q2 = _mass_params({"chirp_mass": mc, "symmetric_mass_ratio": 2.0 / 9.0})
    assert abs(mc2 - mc) < 1e-9 and abs(q2 - 0.5) < 1e-12
    assert _mass_params({"chirp_mass": 1.0, "mass_ratio": 0.3}) == (1.0, 0.3)

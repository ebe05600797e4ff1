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
    assert L.pes_context_set_duty_cycle(None, 0, 1.0, 0.5, 0.0, 0.1) == 1
    assert L.pes_context_set_rows_per_item(None, 8) == 1
    assert L.pes_context_get_rows_per_item(None, None) == 1
    # time domain: validation precedes any device work
    td_args = [0, 2.5e9, 0.0, 0.0, 0.0, 0.0, 0.0, None, None, 0.0, 1.0]
    assert L.pes_td_links(*td_args, 1, 0.1, 0.2, 0.3, 0.0, 1.0, 4, None, None) == 1  # one waveform sample
    assert L.pes_td_links(*td_args, 8, 0.1, 0.2, 0.3, 0.0, 1.0, 4, None, None) == 1  # NULL arrays
    assert L.pes_td_links(*td_args[:10], -1.0, 8, 0.1, 0.2, 0.3, 0.0, 1.0, 4, None, None) == 1
    assert L.pes_td_links(*td_args, 8, 0.1, 0.2, 0.3, 0.0, 1.0, -1, None, None) == 1
    assert L.pes_td_links(2, *td_args[1:], 8, 0.1, 0.2, 0.3, 0.0, 1.0, 0, None, None) == 1  # orbit kind
    assert L.pes_td_links(*td_args, 8, 0.1, 0.2, 0.3, 0.0, 1.0, 0, None, None) == 0  # nothing to do
    assert L.pes_td_tdi_added_samples(2.5e9, 5.0, 15) == int(3 * (2.5e9 / 299792458.0) // 5.0 + 1)
    assert L.pes_td_tdi_added_samples(2.5e9, 5.0, 16) == -1
    assert L.pes_td_tdi_added_samples(2.5e9, 0.0, 20) == -1
    assert L.pes_td_tdi(2.5e9, 5.0, 15, 1, None, 3, None, None) == 1  # fewer samples than the delays need
    assert b"fewer link samples" in L.pes_last_error()
    assert L.pes_td_tdi(2.5e9, 5.0, 15, 1, None, 10, None, None) == 1  # NULL arrays
    assert L.pes_td_tdi(2.5e9, 5.0, 21, 1, None, 10, None, None) == 1
    assert L.pes_td_tdi(2.5e9, 5.0, 15, 1, None, 6, None, None) == 0  # zero outputs


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


@pytest.mark.parametrize("arm, dt", [(2.5e9, 5.0), (3.0e9, 0.5), (1.7e8, 0.25), (8 * 299792458.0, 2.0),
                                     (2.5e9, 8.339102379953802)])
def test_td_added_samples_follow_the_reference(arm, dt):
    """int(max_num_delay * L // dt + 1) (tdi.py:1445-1450), through the C-ABI, against the oracle."""
    from oracle import td as otd
    from pespace_amd.core import td_added_samples
    L = arm / 299792458.0
    for gen, m in (("1.5", 3), ("2.0", 7)):
        assert td_added_samples(arm, dt, gen) == int(m * L // dt + 1) == otd.added_samples(arm, dt, gen)
    with pytest.raises(ValueError):
        td_added_samples(arm, dt, "3.0")


def test_td_data_container_and_models():
    from pespace_amd.detector.antenna import FDResponseModelMarset2018, InterferometerAntenna, TDResponseModelCornish2003
    from pespace_amd.detector.tdi import TDIChannelData, TDMichelsonConstantEqualArm
    td = TDIChannelData()
    with pytest.raises(ValueError):
        td.set_td_data_from_input(("A", "E", "T"), 100.0, 5.0, np.zeros((3, 20)))  # 21 samples expected
    x = np.arange(63, dtype=np.float64).reshape(3, 21)
    td.set_td_data_from_input(("A", "E", "T"), 100.0, 5.0, x, start_time=10.0)
    assert td.data_info.time_series_length == 21
    np.testing.assert_array_equal(td.time_samples, 10.0 + 5.0 * np.arange(21))
    DC = namedtuple("DC", ("A", "E", "T"))
    td.add_into_td_data(DC(*np.ones((3, 21))))
    np.testing.assert_array_equal(td.td_data.E, x[1] + 1)
    with pytest.raises(ValueError):
        td.add_into_td_data(DC(*np.ones((3, 20))))
    # the reference's file layout holds td_data under tdi_data/td_data/<channel>
    tree = td._file_tree()
    assert set(k for k in tree if k.startswith("tdi_data/")) == {f"tdi_data/td_data/{c}" for c in "AET"}
    back = TDIChannelData._from_file_tree(tree)
    np.testing.assert_array_equal(back.td_data.T, td.td_data.T)
    with pytest.warns(UserWarning):
        td.set_td_data_from_zero(("X", "Y", "Z"), 100.0, 5.0)
    with pytest.raises(ValueError):
        TDMichelsonConstantEqualArm(generation="1.0")
    comb = TDMichelsonConstantEqualArm("2.0", orthogonal=False)
    assert comb.labels == ("X", "Y", "Z") and comb.max_num_delay == 7 and comb.domain == "td"
    det = InterferometerAntenna("LISA", td, "LISA_analytic", TDResponseModelCornish2003(), comb)
    L = 2.5e9 / 299792458.0
    assert comb.added_time_samples_number == int(7 * L // 5.0 + 1)
    assert comb.extended_time_series_length == 21 + comb.added_time_samples_number
    ext = comb.extended_time_samples_array
    assert ext.size == comb.extended_time_series_length and ext[comb.added_time_samples_number] == 0.0
    assert np.all(np.diff(ext) == 5.0)
    assert det.single_link_response is None
    with pytest.raises(ValueError):
        TDResponseModelCornish2003(interpolate_kernel=("lagrange", 5))
    with pytest.raises(ValueError):  # a time-domain response needs a time-domain TDI model
        from pespace_amd.detector.tdi import FDMichelsonConstantEqualArm
        td2 = TDIChannelData()
        td2.set_fd_data_from_zero(("A", "E", "T"), 100.0, 5.0)
        InterferometerAntenna("LISA", td2, "LISA_analytic", TDResponseModelCornish2003(),
                              FDMichelsonConstantEqualArm())
    with pytest.raises(ValueError):  # and a time grid
        InterferometerAntenna("LISA", TDIChannelData(), "LISA_analytic", FDResponseModelMarset2018(),
                              TDMichelsonConstantEqualArm())


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

"""TDIChannelData files (tdi.py:1022-1292): the reference's HDF5 dataset layout and its
consistency checks on recovery.

h5py is not installed in this image, so the layout is tested through the dataset tree that
save_to_file writes and recover_from_file reads (paths -> arrays); the HDF5 round trip itself
runs only where h5py exists, and otherwise the file methods must fail with a clear ImportError.
"""

import importlib.util

import numpy as np
import pytest

from pespace_amd.detector.tdi import TDIChannelData

HAS_H5PY = importlib.util.find_spec("h5py") is not None


def _data():
    td = TDIChannelData()
    n = TDIChannelData()
    n.set_fd_data_from_zero(("A", "E", "T"), 655360.0, 5.0, 0.0, 1e-4, 0.1)
    rng = np.random.default_rng(4)
    arr = rng.standard_normal((3, n.data_info.frequency_series_length)) \
        + 1j * rng.standard_normal((3, n.data_info.frequency_series_length))
    td.set_fd_data_from_input(("A", "E", "T"), 655360.0, 5.0, arr, 0.0, 1e-4, 0.1)
    td.set_fd_noise_power_density_from_model("LISA_SciRDv1", tdi_generation="1.5")
    return td


def test_layout_paths():
    """The dataset paths of tdi.py:1049-1088 for FD data with a PSD."""
    tree = _data()._file_tree()
    info = ["channels", "duration", "delta_time", "start_time", "minimum_frequency", "maximum_frequency",
            "sampling_frequency", "delta_frequency", "time_series_length", "time_samples_array",
            "full_frequency_series_length", "full_frequency_samples_array", "frequency_mask_array",
            "frequency_samples_array", "frequency_series_length"]
    expect = {f"class_attrs/{k}" for k in ("scaling", "backend", "_np_fp", "_np_cp")}
    expect |= {f"data_info/{k}" for k in info}
    expect |= {f"tdi_data/fd_data/{c}" for c in "AET"} | {f"noise_data/fd_noise_power_density/{c}" for c in "AET"}
    assert set(tree) == expect
    assert tree["class_attrs/_np_fp"] == "float64" and tree["class_attrs/_np_cp"] == "complex128"
    assert tree["data_info/time_series_length"] == 131073
    assert tree["data_info/frequency_series_length"] == 65471  # the reference's own grid size
    np.testing.assert_array_equal(tree["data_info/time_samples_array"], np.arange(131073) * 5.0)


def test_tree_round_trip_is_exact():
    td = _data()
    back = TDIChannelData._from_file_tree(td._file_tree())
    for c in "AET":
        np.testing.assert_array_equal(getattr(back.fd_data, c), getattr(td.fd_data, c))
        np.testing.assert_array_equal(getattr(back.fd_noise_power_density, c), getattr(td.fd_noise_power_density, c))
    assert back.data_info.channels == ("A", "E", "T")
    np.testing.assert_array_equal(back.frequency_samples, td.frequency_samples)
    assert back.scaling == td.scaling and back.backend == "hip"


def test_strings_as_bytes_and_reference_backend_names():
    """Files written through h5py return strings as bytes; a file from another backend loads."""
    tree = _data()._file_tree()
    tree["class_attrs/backend"] = np.array(b"taichi", dtype=object)
    tree["class_attrs/_np_fp"] = b"float64"
    tree["data_info/channels"] = np.array([b"A", b"E", b"T"], dtype=object)
    back = TDIChannelData._from_file_tree(tree)
    assert back.data_info.channels == ("A", "E", "T")


@pytest.mark.parametrize("key, value", [("delta_frequency", 1.0), ("time_series_length", 7),
                                        ("frequency_series_length", 65470)])
def test_inconsistent_grid_is_rejected(key, value):
    tree = _data()._file_tree()
    tree[f"data_info/{key}"] = np.array(value)
    with pytest.raises(AssertionError, match=f"the `{key}` is different from the stored value"):
        TDIChannelData._from_file_tree(tree, "bad.h5")


def test_mask_and_arrays_are_checked():
    tree = _data()._file_tree()
    m = tree["data_info/frequency_mask_array"].copy()
    m[np.argmax(m)] = False
    tree["data_info/frequency_mask_array"] = m
    with pytest.raises(AssertionError, match="frequency_mask_array"):
        TDIChannelData._from_file_tree(tree)


def test_precision_mismatch_is_rejected():
    tree = _data()._file_tree()
    tree["class_attrs/_np_fp"] = "float32"
    tree["class_attrs/_np_cp"] = "complex64"
    with pytest.raises(AssertionError, match="numerical precision of stored data"):
        TDIChannelData._from_file_tree(tree)


def test_other_domains():
    tree = _data()._file_tree()
    chans = [k.split("/")[-1] for k in tree if k.startswith("tdi_data/fd_data/")]
    n_t = int(np.asarray(tree["data_info/time_series_length"]))
    for i, c in enumerate(chans):  # time-domain data load beside the FD data
        tree[f"tdi_data/td_data/{c}"] = np.full(n_t, float(i))
    back = TDIChannelData._from_file_tree(tree)
    assert back.fd_data is not None and back.td_data is not None
    np.testing.assert_array_equal(getattr(back.td_data, chans[-1]), np.full(n_t, len(chans) - 1.0))
    bad = dict(tree)
    bad[f"tdi_data/td_data/{chans[0]}"] = np.zeros(4)
    with pytest.raises(ValueError, match="time_series_length"):
        TDIChannelData._from_file_tree(bad)
    tree["tdi_data/wd_data/A"] = np.zeros(4)
    with pytest.warns(UserWarning, match="wd_data"):
        TDIChannelData._from_file_tree(tree)
    tree["noise_data/wd_noise_power_density/A"] = np.zeros(4)
    with pytest.warns(UserWarning), pytest.raises(NotImplementedError):
        TDIChannelData._from_file_tree(tree)


def test_explicit_grid_has_no_file_layout():
    td = TDIChannelData()
    td.set_fd_data_on_grid(("A",), np.geomspace(1e-4, 1e-2, 8), np.full(8, 1e-5))
    with pytest.raises(ValueError, match="rfft grids"):
        td._file_tree()


@pytest.mark.skipif(HAS_H5PY, reason="h5py is installed")
def test_file_methods_need_h5py(tmp_path):
    with pytest.raises(ImportError, match="h5py"):
        _data().save_to_file(str(tmp_path / "d.h5"))
    with pytest.raises(ImportError, match="h5py"):
        TDIChannelData.recover_from_file(str(tmp_path / "d.h5"))


@pytest.mark.skipif(not HAS_H5PY, reason="h5py is not installed")
def test_hdf5_round_trip(tmp_path):
    td = _data()
    path = str(tmp_path / "d.h5")
    td.save_to_file(path)
    back = TDIChannelData.recover_from_file(path)
    for c in "AET":
        np.testing.assert_array_equal(getattr(back.fd_data, c), getattr(td.fd_data, c))
    with pytest.raises(OSError):  # mode "x": an existing file is not overwritten
        td.save_to_file(path)

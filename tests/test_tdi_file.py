"""TDIChannelData files (tdi.py:1022-1292): the reference's HDF5 dataset layout, its
consistency checks on recovery, and the files themselves.

The layout logic is tested on the dataset tree that save_to_file writes and recover_from_file
reads (paths -> arrays).  The files go through pespace_amd.detector.h5file: h5py where it is
importable, else the HDF5 C library through ctypes.  Where an interpreter with h5py is at hand
(PESPACE_H5PY_PYTHON, else a conda installation's), h5py itself checks the files both ways
(tests/h5py_crosscheck.py in that interpreter): it reads ours with the dtypes it writes itself
(bool, <f8, <i8, <c16, UTF-8 variable-length strings), and ours reads a file it writes the way
tdi.py:1049-1088 does; h5dump's datatype, dataspace and storage description of the two files is
the same apart from the data offsets.
"""

import functools
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from pespace_amd.detector import h5file
from pespace_amd.detector.tdi import TDIChannelData

HERE = os.path.dirname(os.path.abspath(__file__))


def _has_hdf5():
    try:
        h5file.backend()
    except ImportError:
        return False
    return True


HAS_HDF5 = _has_hdf5()


@functools.lru_cache(maxsize=None)
def _h5py_python():
    """An interpreter other than this one that imports h5py, or None."""
    cands = [os.environ.get("PESPACE_H5PY_PYTHON"), "/opt/conda/bin/python3.9", "/opt/conda/bin/python3"]
    for exe in cands:
        if not exe or not os.path.exists(exe) or os.path.realpath(exe) == os.path.realpath(sys.executable):
            continue
        r = subprocess.run([exe, "-c", "import h5py, numpy"], env=_clean_env(), capture_output=True, timeout=120)
        if r.returncode == 0:
            return exe
    return None


def _clean_env():
    return {"PATH": os.environ.get("PATH", "/usr/bin:/bin"), "HOME": os.environ.get("HOME", "/tmp"),
            "PYTHONNOUSERSITE": "1"}


def _h5py(*args):
    exe = _h5py_python()
    r = subprocess.run([exe, os.path.join(HERE, "h5py_crosscheck.py"), *map(str, args)], env=_clean_env(),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


need_h5py = pytest.mark.skipif("not _h5py_python()", reason="no interpreter with h5py")
need_hdf5 = pytest.mark.skipif(not HAS_HDF5, reason="neither h5py nor the HDF5 C library is available")


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


def test_file_methods_need_hdf5(tmp_path, monkeypatch):
    """Without h5py and without the C library the file methods fail with a clear ImportError."""
    monkeypatch.setenv("PESPACE_HDF5_LIBRARY", str(tmp_path / "missing" / "libhdf5.so"))
    monkeypatch.setattr(h5file, "_CLIB", None)
    monkeypatch.setattr(h5file, "_h5py", lambda: None)
    with pytest.raises(ImportError, match="PESPACE_HDF5_LIBRARY"):
        _data().save_to_file(str(tmp_path / "d.h5"))
    with pytest.raises(ImportError, match="PESPACE_HDF5_LIBRARY"):
        TDIChannelData.recover_from_file(str(tmp_path / "d.h5"))


@need_hdf5
def test_hdf5_round_trip(tmp_path):
    td = _data()
    n_t = td.data_info.time_series_length
    td._store_td([np.sin(np.arange(n_t) * (0.01 + 0.001 * i)) for i in range(3)])
    path = str(tmp_path / "d.h5")
    td.save_to_file(path)
    back = TDIChannelData.recover_from_file(path)
    for c in "AET":
        np.testing.assert_array_equal(getattr(back.fd_data, c), getattr(td.fd_data, c))
        np.testing.assert_array_equal(getattr(back.td_data, c), getattr(td.td_data, c))
        np.testing.assert_array_equal(getattr(back.fd_noise_power_density, c), getattr(td.fd_noise_power_density, c))
    assert back.data_info.channels == ("A", "E", "T") and back.scaling == td.scaling
    with pytest.raises(OSError):  # mode "x": an existing file is not overwritten
        td.save_to_file(path)
    with pytest.raises(FileNotFoundError):
        TDIChannelData.recover_from_file(str(tmp_path / "none.h5"))
    # what h5py's dataset[()] gives: strings as bytes, numeric scalars as numpy scalars
    tree = h5file.read_tree(path)
    assert tree["class_attrs/backend"] == b"hip" and tree["class_attrs/_np_cp"] == b"complex128"
    assert list(tree["data_info/channels"]) == [b"A", b"E", b"T"]
    assert isinstance(tree["class_attrs/scaling"], np.bool_)
    assert tree["data_info/time_series_length"].dtype == np.int64
    assert tree["tdi_data/fd_data/A"].dtype == np.complex128


def _expected_dtype(value):
    a = np.asarray(value)
    return {"b": "|b1", "f": "<f8", "i": "<i8", "c": "<c16", "U": "|O"}[a.dtype.kind]


@need_hdf5
@need_h5py
def test_h5py_reads_our_file(tmp_path):
    td = _data()
    path = tmp_path / "ours.h5"
    td.save_to_file(str(path))
    _h5py("dump", path, tmp_path / "d.npz", tmp_path / "d.json")
    meta = json.loads((tmp_path / "d.json").read_text())
    vals = np.load(tmp_path / "d.npz", allow_pickle=False)
    tree = td._file_tree()
    assert set(meta) == set(tree)
    for key, value in tree.items():
        m = meta[key]
        assert m["dtype"] == _expected_dtype(value), (key, m)
        assert m["shape"] == list(np.shape(value)), key
        if m["dtype"] == "|O":
            assert m["string"] == "utf-8" and m["class"] == 3, key  # variable-length UTF-8 (H5T_STRING)
            assert m["pytype"] == ("bytes" if np.ndim(value) == 0 else "ndarray"), key
        got = vals[key.replace("/", "|")]
        np.testing.assert_array_equal(got, np.asarray(value), err_msg=key)
    assert meta["class_attrs/scaling"]["class"] == 8  # h5py's bool: an enum (FALSE, TRUE)
    assert meta["tdi_data/fd_data/A"]["class"] == 6  # complex: a compound (r, i)


@need_hdf5
@need_h5py
def test_our_reader_reads_h5py_file(tmp_path):
    td = _data()
    tree = td._file_tree()
    np.savez(tmp_path / "t.npz", **{k.replace("/", "|"): np.asarray(v) for k, v in tree.items()})
    (tmp_path / "t.json").write_text(json.dumps(list(tree)))
    _h5py("write", tmp_path / "t.npz", tmp_path / "t.json", tmp_path / "h5py.h5")
    back = TDIChannelData.recover_from_file(str(tmp_path / "h5py.h5"))
    for c in "AET":
        np.testing.assert_array_equal(getattr(back.fd_data, c), getattr(td.fd_data, c))
        np.testing.assert_array_equal(getattr(back.fd_noise_power_density, c), getattr(td.fd_noise_power_density, c))
    np.testing.assert_array_equal(back.frequency_samples, td.frequency_samples)
    # the same values and Python types as from our own file
    td.save_to_file(str(tmp_path / "ours.h5"))
    a, b = h5file.read_tree(tmp_path / "h5py.h5"), h5file.read_tree(tmp_path / "ours.h5")
    assert list(a) == list(b)
    for k in a:
        assert type(a[k]) is type(b[k]), k
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)
    # h5dump's description of both files: datatypes, dataspaces, layout, filters and fill values
    h5dump = os.path.join(os.path.dirname(_h5py_python()), "h5dump")
    if not os.path.exists(h5dump):
        h5dump = shutil.which("h5dump")
    if h5dump:
        def header(p):
            out = subprocess.run([h5dump, "-p", "-H", str(p)], capture_output=True, text=True, timeout=120,
                                 env=_clean_env()).stdout
            return [ln for ln in out.splitlines()[1:] if "OFFSET" not in ln]
        assert header(tmp_path / "h5py.h5") == header(tmp_path / "ours.h5")


_H5PY_BRANCH = r'''
import importlib.util, sys
import numpy as np
spec = importlib.util.spec_from_file_location("h5file", sys.argv[1])
m = importlib.util.module_from_spec(spec)
spec.loader.exec_module(m)
assert m.backend() == "h5py"
tree = {"a/flag": np.array(True), "a/name": "hip", "a/names": np.array(["A", "E", "T"]),
        "b/n": np.array(131073), "b/x": np.linspace(0.0, 1.0, 7), "c/z": np.arange(5) * (1.0 + 2.0j),
        "c/mask": np.array([True, False, True])}
m.write_tree(sys.argv[2], tree)
back = m.read_tree(sys.argv[2])
assert back["a/name"] == b"hip" and list(back["a/names"]) == [b"A", b"E", b"T"]
assert back["a/flag"] == True and back["b/n"] == 131073
np.testing.assert_array_equal(back["c/z"], tree["c/z"])
print("ok")
'''


@need_hdf5
@need_h5py
def test_h5py_branch_of_h5file(tmp_path):
    """h5file's h5py route (run by the interpreter that has h5py) writes what the C-library route
    reads back with the same values and types."""
    path = tmp_path / "h5py_branch.h5"
    exe = _h5py_python()
    r = subprocess.run([exe, "-c", _H5PY_BRANCH, os.path.join(os.path.dirname(HERE), "pespace_amd", "detector",
                                                                "h5file.py"), str(path)],
                       env=_clean_env(), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]
    back = h5file.read_tree(path)
    assert back["a/name"] == b"hip" and list(back["a/names"]) == [b"A", b"E", b"T"]
    assert isinstance(back["a/flag"], np.bool_) and back["a/flag"]
    assert back["b/n"] == 131073 and back["b/n"].dtype == np.int64
    np.testing.assert_array_equal(back["b/x"], np.linspace(0.0, 1.0, 7))
    np.testing.assert_array_equal(back["c/z"], np.arange(5) * (1.0 + 2.0j))
    np.testing.assert_array_equal(back["c/mask"], [True, False, True])

"""HDF5 files of TDIChannelData (tdi.py:1022-1292): a flat {path: value} tree written to and
read from an HDF5 file, with the datatypes h5py gives the same values.

The reference writes its files with h5py.  Here h5py is used when it is importable; otherwise
the HDF5 C library itself (1.10 or later) is called through ctypes: the library named by
PESPACE_HDF5_LIBRARY, else the one the dynamic loader finds as "hdf5", else that of a conda
installation.  Both routes write the same file: datasets created with their intermediate groups,
contiguous layout, and h5py's datatypes for numpy values

* bool: an enum over a native int8 with FALSE = 0 and TRUE = 1;
* float64, int64: IEEE F64LE, STD I64LE;
* complex128: a compound of two F64LE members "r" (offset 0) and "i" (offset 8);
* str: variable-length UTF-8 strings (h5py.string_dtype(encoding="utf-8")), scalar or 1-D.

Reading returns what h5py's ``dataset[()]`` returns for these types: numpy scalars and arrays,
strings as bytes (a 1-D string dataset as an object array of bytes).
"""

from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np

__all__ = ["write_tree", "read_tree", "backend"]

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_ulonglong
_size = ctypes.c_size_t

# enum values of the HDF5 C API (H5Fpublic.h, H5Spublic.h, H5Tpublic.h, H5Ipublic.h, H5public.h)
_ACC_RDONLY, _ACC_EXCL = 0x0000, 0x0004
_S_SCALAR = 0
_T_INTEGER, _T_FLOAT, _T_STRING, _T_COMPOUND, _T_ENUM = 0, 1, 3, 6, 8
_CSET_UTF8 = 1
_I_GROUP, _I_DATASET = 2, 5
_INDEX_NAME, _ITER_INC = 0, 0
_VARIABLE = ctypes.c_size_t(-1).value
_DEFAULT = 0

_ITERATE = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)


def _library_candidates():
    env = os.environ.get("PESPACE_HDF5_LIBRARY")
    if env:
        return [env]
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda"):
        if prefix:
            out.append(os.path.join(prefix, "lib", "libhdf5.so"))
    out.append("/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so")
    return out


class _CLib:
    """The HDF5 C functions this module calls, with their C signatures."""

    def __init__(self, path):
        L = self.L = ctypes.CDLL(path)
        self.path = path
        sig = {
            "H5open": (_herr, []),
            "H5get_libversion": (_herr, [ctypes.POINTER(ctypes.c_uint)] * 3),
            "H5Eset_auto2": (_herr, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
            "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
            "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Pcreate": (_hid, [_hid]),
            "H5Pset_create_intermediate_group": (_herr, [_hid, ctypes.c_uint]),
            "H5Pclose": (_herr, [_hid]),
            "H5Screate": (_hid, [ctypes.c_int]),
            "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            "H5Sclose": (_herr, [_hid]),
            "H5Tcopy": (_hid, [_hid]),
            "H5Tcreate": (_hid, [ctypes.c_int, _size]),
            "H5Tinsert": (_herr, [_hid, ctypes.c_char_p, _size, _hid]),
            "H5Tenum_create": (_hid, [_hid]),
            "H5Tenum_insert": (_herr, [_hid, ctypes.c_char_p, ctypes.c_void_p]),
            "H5Tset_size": (_herr, [_hid, _size]),
            "H5Tset_cset": (_herr, [_hid, ctypes.c_int]),
            "H5Tget_class": (ctypes.c_int, [_hid]),
            "H5Tget_size": (_size, [_hid]),
            "H5Tget_cset": (ctypes.c_int, [_hid]),
            "H5Tis_variable_str": (ctypes.c_int, [_hid]),
            "H5Tget_nmembers": (ctypes.c_int, [_hid]),
            "H5Tget_member_name": (ctypes.c_void_p, [_hid, ctypes.c_uint]),
            "H5Tget_member_class": (ctypes.c_int, [_hid, ctypes.c_uint]),
            "H5Tget_member_value": (_herr, [_hid, ctypes.c_uint, ctypes.c_void_p]),
            "H5Tget_super": (_hid, [_hid]),
            "H5Tclose": (_herr, [_hid]),
            "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dget_type": (_hid, [_hid]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dclose": (_herr, [_hid]),
            "H5Oopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Oclose": (_herr, [_hid]),
            "H5Iget_type": (ctypes.c_int, [_hid]),
            "H5free_memory": (_herr, [ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        # link iteration: H5Literate2 (1.12 and later) or H5Literate (1.10), the same arguments;
        # reclaiming variable-length read buffers: H5Treclaim (1.12 and later) or H5Dvlen_reclaim
        self.iterate = getattr(L, "H5Literate2", None) or L.H5Literate
        self.iterate.restype = _herr
        self.iterate.argtypes = [_hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _ITERATE, ctypes.c_void_p]
        self.reclaim = getattr(L, "H5Treclaim", None) or L.H5Dvlen_reclaim
        self.reclaim.restype = _herr
        self.reclaim.argtypes = [_hid, _hid, _hid, ctypes.c_void_p]
        if L.H5open() < 0:
            raise OSError(f"HDF5: H5open failed ({path})")
        v = [ctypes.c_uint() for _ in range(3)]
        L.H5get_libversion(*[ctypes.byref(x) for x in v])
        self.version = tuple(x.value for x in v)
        if self.version < (1, 10, 0):
            raise OSError(f"HDF5 {self.version} at {path}: version 1.10 or later is required")
        L.H5Eset_auto2(_DEFAULT, None, None)  # failures are reported from the return codes
        g = lambda sym: _hid.in_dll(L, sym).value
        self.F64LE, self.I64LE = g("H5T_IEEE_F64LE_g"), g("H5T_STD_I64LE_g")
        self.N_DOUBLE, self.N_INT64, self.N_INT8 = g("H5T_NATIVE_DOUBLE_g"), g("H5T_NATIVE_INT64_g"), g("H5T_NATIVE_INT8_g")
        self.C_S1 = g("H5T_C_S1_g")
        self.LINK_CREATE = g("H5P_CLS_LINK_CREATE_ID_g")

    def __getattr__(self, name):
        return getattr(self.L, name)


_CLIB = None


def _clib():
    global _CLIB
    if _CLIB is None:
        errors = []
        for path in _library_candidates():
            try:
                _CLIB = _CLib(path)
                break
            except (OSError, AttributeError) as e:
                errors.append(f"{path}: {e}")
        else:
            raise ImportError("TDIChannelData files are HDF5 (the reference's layout): neither h5py nor the HDF5 C "
                              "library (1.10 or later) was found; set PESPACE_HDF5_LIBRARY to libhdf5.so.  Tried: "
                              + "; ".join(errors))
    return _CLIB


def _h5py():
    try:
        import h5py
    except ImportError:
        return None
    return h5py


def backend():
    """'h5py', or 'libhdf5 <version> (<path>)' when the C library is called directly."""
    if _h5py() is not None:
        return "h5py"
    c = _clib()
    return f"libhdf5 {'.'.join(map(str, c.version))} ({c.path})"


def _check(value, what):
    if value < 0:
        raise OSError(f"HDF5: {what} failed")
    return value


def _as_array(value):
    if isinstance(value, (str, bytes)):
        return np.array(value, dtype=object)
    arr = np.asarray(value)
    if arr.dtype.kind in "US":
        arr = arr.astype(object)
    return arr


# ---- writing -------------------------------------------------------------------------------

def _file_type(c, arr):
    """(file type, memory type, contiguous buffer, keep-alive) for a numpy array."""
    kind = arr.dtype.kind
    if kind == "b":
        t = _check(c.H5Tenum_create(c.N_INT8), "H5Tenum_create")
        for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
            val = ctypes.c_int8(v)
            _check(c.H5Tenum_insert(t, name, ctypes.byref(val)), "H5Tenum_insert")
        buf = np.ascontiguousarray(arr, dtype=np.int8)
        return t, _check(c.H5Tcopy(t), "H5Tcopy"), buf, None
    if kind == "f":
        return (_check(c.H5Tcopy(c.F64LE), "H5Tcopy"), _check(c.H5Tcopy(c.N_DOUBLE), "H5Tcopy"),
                np.ascontiguousarray(arr, dtype=np.float64), None)
    if kind in "iu":
        return (_check(c.H5Tcopy(c.I64LE), "H5Tcopy"), _check(c.H5Tcopy(c.N_INT64), "H5Tcopy"),
                np.ascontiguousarray(arr, dtype=np.int64), None)
    if kind == "c":
        types = []
        for base in (c.F64LE, c.N_DOUBLE):
            t = _check(c.H5Tcreate(_T_COMPOUND, 16), "H5Tcreate")
            _check(c.H5Tinsert(t, b"r", 0, base), "H5Tinsert")
            _check(c.H5Tinsert(t, b"i", 8, base), "H5Tinsert")
            types.append(t)
        return types[0], types[1], np.ascontiguousarray(arr, dtype=np.complex128), None
    if kind == "O":
        encoded = [(x.encode("utf-8") if isinstance(x, str) else bytes(x)) for x in arr.reshape(-1)]
        ptrs = (ctypes.c_char_p * max(len(encoded), 1))(*encoded)
        types = []
        for _ in range(2):
            t = _check(c.H5Tcopy(c.C_S1), "H5Tcopy")
            _check(c.H5Tset_size(t, _VARIABLE), "H5Tset_size")
            _check(c.H5Tset_cset(t, _CSET_UTF8), "H5Tset_cset")
            types.append(t)
        return types[0], types[1], ptrs, encoded
    raise TypeError(f"no HDF5 datatype for numpy dtype {arr.dtype}")


def _space(c, shape):
    if len(shape) == 0:
        return _check(c.H5Screate(_S_SCALAR), "H5Screate")
    dims = (_hsize * len(shape))(*shape)
    return _check(c.H5Screate_simple(len(shape), dims, None), "H5Screate_simple")


def _write_clib(filename, tree):
    c = _clib()
    if os.path.exists(filename):
        raise FileExistsError(f"Unable to create file {filename!r}: the file exists")
    fid = _check(c.H5Fcreate(os.fsencode(filename), _ACC_EXCL, _DEFAULT, _DEFAULT), f"H5Fcreate {filename}")
    lcpl = -1
    try:
        lcpl = _check(c.H5Pcreate(c.LINK_CREATE), "H5Pcreate")
        _check(c.H5Pset_create_intermediate_group(lcpl, 1), "H5Pset_create_intermediate_group")
        for path, value in tree.items():
            arr = _as_array(value)
            ftype, mtype, buf, keep = _file_type(c, arr)
            space = dset = -1
            try:
                space = _space(c, arr.shape)
                dset = _check(c.H5Dcreate2(fid, path.encode(), ftype, space, lcpl, _DEFAULT, _DEFAULT),
                              f"H5Dcreate2 {path}")
                ptr = ctypes.cast(buf, ctypes.c_void_p) if keep is not None else buf.ctypes.data_as(ctypes.c_void_p)
                _check(c.H5Dwrite(dset, mtype, _DEFAULT, _DEFAULT, _DEFAULT, ptr), f"H5Dwrite {path}")
            finally:
                if dset >= 0:
                    c.H5Dclose(dset)
                if space >= 0:
                    c.H5Sclose(space)
                c.H5Tclose(ftype)
                c.H5Tclose(mtype)
    finally:
        if lcpl >= 0:
            c.H5Pclose(lcpl)
        _check(c.H5Fclose(fid), f"H5Fclose {filename}")


def _write_h5py(h5py, filename, tree):
    with h5py.File(filename, "x") as fh:
        for path, value in tree.items():
            arr = _as_array(value)
            if arr.dtype.kind == "O":
                arr = np.array([x.decode() if isinstance(x, bytes) else x for x in arr.reshape(-1)],
                               dtype=h5py.string_dtype(encoding="utf-8")).reshape(arr.shape)
            fh.create_dataset(path, data=arr)


def write_tree(filename, tree):
    """Write {path: value} as datasets of a new HDF5 file (an existing file is not overwritten:
    FileExistsError, as h5py's mode "x")."""
    h5py = _h5py()
    if h5py is not None:
        return _write_h5py(h5py, filename, tree)
    return _write_clib(str(filename), tree)


# ---- reading -------------------------------------------------------------------------------

def _member_name(c, t, i):
    p = c.H5Tget_member_name(t, i)
    if not p:
        raise OSError("HDF5: H5Tget_member_name failed")
    try:
        return ctypes.string_at(p)
    finally:
        c.H5free_memory(p)


def _read_dataset(c, dset, path):
    space = _check(c.H5Dget_space(dset), f"H5Dget_space {path}")
    ftype = _check(c.H5Dget_type(dset), f"H5Dget_type {path}")
    mtype = -1
    try:
        nd = _check(c.H5Sget_simple_extent_ndims(space), f"H5Sget_simple_extent_ndims {path}")
        dims = (_hsize * max(nd, 1))()
        if nd:
            _check(c.H5Sget_simple_extent_dims(space, dims, None), f"H5Sget_simple_extent_dims {path}")
        shape = tuple(int(dims[i]) for i in range(nd))
        n = int(np.prod(shape)) if nd else 1
        cls = c.H5Tget_class(ftype)
        if cls == _T_FLOAT:
            out = np.empty(shape, np.float64)
            mtype = _check(c.H5Tcopy(c.N_DOUBLE), "H5Tcopy")
        elif cls == _T_INTEGER:
            out = np.empty(shape, np.int64)
            mtype = _check(c.H5Tcopy(c.N_INT64), "H5Tcopy")
        elif cls == _T_ENUM:
            names = {_member_name(c, ftype, i) for i in range(c.H5Tget_nmembers(ftype))}
            if names != {b"FALSE", b"TRUE"} or c.H5Tget_size(ftype) != 1:
                raise TypeError(f"{path}: enum datasets other than h5py's bool are not supported")
            mtype = _check(c.H5Tenum_create(c.N_INT8), "H5Tenum_create")
            for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
                val = ctypes.c_int8(v)
                _check(c.H5Tenum_insert(mtype, name, ctypes.byref(val)), "H5Tenum_insert")
            raw = np.empty(shape, np.int8)
            _check(c.H5Dread(dset, mtype, _DEFAULT, _DEFAULT, _DEFAULT, raw.ctypes.data_as(ctypes.c_void_p)),
                   f"H5Dread {path}")
            return raw.astype(bool)[()] if nd == 0 else raw.astype(bool)
        elif cls == _T_COMPOUND:
            names = [_member_name(c, ftype, i) for i in range(c.H5Tget_nmembers(ftype))]
            if names != [b"r", b"i"] or any(c.H5Tget_member_class(ftype, i) != _T_FLOAT for i in range(2)):
                raise TypeError(f"{path}: compound datasets other than complex (r, i) are not supported")
            out = np.empty(shape, np.complex128)
            mtype = _check(c.H5Tcreate(_T_COMPOUND, 16), "H5Tcreate")
            _check(c.H5Tinsert(mtype, b"r", 0, c.N_DOUBLE), "H5Tinsert")
            _check(c.H5Tinsert(mtype, b"i", 8, c.N_DOUBLE), "H5Tinsert")
        elif cls == _T_STRING:
            mtype = _check(c.H5Tcopy(ftype), "H5Tcopy")
            if c.H5Tis_variable_str(ftype) > 0:
                ptrs = (ctypes.c_char_p * n)()
                _check(c.H5Dread(dset, mtype, _DEFAULT, _DEFAULT, _DEFAULT, ptrs), f"H5Dread {path}")
                vals = [ptrs[i] if ptrs[i] is not None else b"" for i in range(n)]
                c.reclaim(mtype, space, _DEFAULT, ptrs)
            else:
                size = c.H5Tget_size(ftype)
                raw = ctypes.create_string_buffer(n * size)
                _check(c.H5Dread(dset, mtype, _DEFAULT, _DEFAULT, _DEFAULT, raw), f"H5Dread {path}")
                vals = [raw.raw[i * size:(i + 1) * size].rstrip(b"\0") for i in range(n)]
            if nd == 0:
                return vals[0]
            out = np.empty(n, dtype=object)
            out[:] = vals
            return out.reshape(shape)
        else:
            raise TypeError(f"{path}: HDF5 datatype class {cls} is not supported")
        _check(c.H5Dread(dset, mtype, _DEFAULT, _DEFAULT, _DEFAULT, out.ctypes.data_as(ctypes.c_void_p)),
               f"H5Dread {path}")
        return out[()] if nd == 0 else out
    finally:
        if mtype >= 0:
            c.H5Tclose(mtype)
        c.H5Tclose(ftype)
        c.H5Sclose(space)


def _link_names(c, gid):
    names = []

    def visit(_group, name, _info, _data):
        names.append(name)
        return 0

    cb = _ITERATE(visit)
    idx = _hsize(0)
    _check(c.iterate(gid, _INDEX_NAME, _ITER_INC, ctypes.byref(idx), cb, None), "H5Literate")
    return names


def _read_group(c, gid, prefix, tree):
    for name in _link_names(c, gid):
        path = prefix + name.decode()
        oid = _check(c.H5Oopen(gid, name, _DEFAULT), f"H5Oopen {path}")
        try:
            kind = c.H5Iget_type(oid)
            if kind == _I_GROUP:
                _read_group(c, oid, path + "/", tree)
            elif kind == _I_DATASET:
                tree[path] = _read_dataset(c, oid, path)
        finally:
            c.H5Oclose(oid)


def read_tree(filename):
    """Every dataset of an HDF5 file as {path: value}, in the groups' name order."""
    h5py = _h5py()
    tree = {}
    if h5py is not None:
        with h5py.File(filename, "r") as fh:
            fh.visititems(lambda name, obj: tree.__setitem__(name, obj[()]) if isinstance(obj, h5py.Dataset)
                          else None)
        return tree
    c = _clib()
    filename = str(filename)
    if not os.path.exists(filename):
        raise FileNotFoundError(f"Unable to open file {filename!r}: no such file")
    fid = _check(c.H5Fopen(os.fsencode(filename), _ACC_RDONLY, _DEFAULT), f"H5Fopen {filename}")
    try:
        _read_group(c, fid, "", tree)
    finally:
        c.H5Fclose(fid)
    return tree

"""Helper of tests/test_tdi_file.py, run by an interpreter that has h5py (a separate process).

  dump <file.h5> <out.npz> <out.json>
      every dataset of the file as h5py reads it: values into out.npz (strings decoded to str),
      and per path the numpy dtype h5py reports, the string encoding (h5py.check_string_dtype),
      the enum mapping (h5py.check_enum_dtype), the HDF5 type class and the Python type of
      ``dataset[()]``.
  write <in.npz> <in.json> <file.h5>
      a new file (mode "x") with the datasets of in.npz in the order in.json lists them, created
      as tdi.py:1049-1088 creates them: ``create_dataset(path, data=np.array(value))``, strings
      with ``h5py.string_dtype(encoding="utf-8")``.

Arrays pass through .npz files without pickles: strings as numpy str arrays.
"""

import json
import sys

import h5py
import numpy as np


def dump(src, out_npz, out_json):
    values, meta = {}, {}

    def visit(name, obj):
        if not isinstance(obj, h5py.Dataset):
            return
        v = obj[()]
        s = h5py.check_string_dtype(obj.dtype)
        e = h5py.check_enum_dtype(obj.dtype)
        meta[name] = {"dtype": obj.dtype.str, "shape": list(obj.shape),
                      "string": s.encoding if s is not None else None,
                      "enum": {k: int(x) for k, x in e.items()} if e is not None else None,
                      "class": int(obj.id.get_type().get_class()), "pytype": type(v).__name__}
        if s is not None:
            v = np.array([x.decode(s.encoding) for x in np.asarray(v, dtype=object).reshape(-1)]).reshape(obj.shape)
        values[name] = np.asarray(v)

    with h5py.File(src, "r") as fh:
        fh.visititems(visit)
    np.savez(out_npz, **{k.replace("/", "|"): v for k, v in values.items()})
    with open(out_json, "w") as f:
        json.dump(meta, f)


def write(in_npz, in_json, dst):
    with open(in_json) as f:
        paths = json.load(f)
    data = np.load(in_npz, allow_pickle=False)
    with h5py.File(dst, "x") as fh:
        for path in paths:
            v = data[path.replace("/", "|")]
            if v.dtype.kind == "U":
                v = np.array(v, dtype=h5py.string_dtype(encoding="utf-8"))
            fh.create_dataset(path, data=np.array(v))


if __name__ == "__main__":
    {"dump": dump, "write": write}[sys.argv[1]](*sys.argv[2:])

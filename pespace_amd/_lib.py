"""ctypes binding of the C-ABI declared in include/pespace_amd.h.

The product path has no CPU fallback: if ``libpespace_amd.so`` is missing or a
GPU is absent, the calls below raise.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpespace_amd.so")
_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)

# (name, restype, argtypes) for every symbol include/pespace_amd.h declares
SIGNATURES = [
    ("pes_last_error", ctypes.c_char_p, []),
    ("pes_context_create", ctypes.c_int,
     [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, c_int_p, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double]),
    ("pes_context_add_detector", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p, c_double_p,
      ctypes.c_double, c_double_p, c_double_p, c_double_p, c_int_p]),
    ("pes_context_finalize", ctypes.c_int, [ctypes.c_void_p]),
    ("pes_context_destroy", None, [ctypes.c_void_p]),
    ("pes_lnl_batch", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p]),
    ("pes_response", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p]),
    ("pes_waveform", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_response_from_waveform", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_whittle_lnl", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_context_set_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("pes_context_get_schedule", ctypes.c_int, [ctypes.c_void_p, c_int_p, c_int_p]),
    ("pes_context_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("pes_profile_read", ctypes.c_int, [ctypes.c_void_p, c_int_p, c_double_p]),
    ("pes_profile_work", ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p]),
    ("pes_detector_bins", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("pes_links", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_links_from_waveform", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_tdi_from_links", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_orbit_vectors", ctypes.c_int,
     [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_context_reserve", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("pes_context_pin_scratch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("pes_lnl_batch_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p]),
    ("pes_lnl_batch_noalloc_bytes", ctypes.c_size_t, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("pes_context_set_rows_per_item", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("pes_context_get_rows_per_item", ctypes.c_int, [ctypes.c_void_p, c_int_p]),
    ("pes_context_set_duty_cycle", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("pes_td_links", ctypes.c_int,
     [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("pes_td_tdi_added_samples", ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    ("pes_td_tdi", ctypes.c_int,
     [ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p]),
]


class PespaceError(RuntimeError):
    pass


def lib():
    """Load libpespace_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PespaceError(
                f"{LIB_PATH} not found: build the HIP extension first "
                "(python -c 'import __graft_entry__; __graft_entry__.build()')")
        handle = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status):
    if status != 0:
        msg = lib().pes_last_error()
        raise PespaceError(f"pespace_amd error {status}: {msg.decode() if msg else ''}")


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(c_double_p)

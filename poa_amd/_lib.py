"""ctypes loading of the package's native libraries (built in-tree by ``poa_amd.build``)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PRODUCT_SO = os.path.join(HERE, "libmpoa.so")
SYNTH_SO = os.path.join(HERE, "libmpoa_synth.so")

_synth = None
_product = None


def synth() -> ctypes.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_SO):
            raise RuntimeError(f"{SYNTH_SO} is not built: run __graft_entry__.build()")
        lib = ctypes.CDLL(SYNTH_SO)
        lib.mpoa_synth_window.restype = ctypes.c_int
        lib.mpoa_synth_window.argtypes = [
            ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_int64,
            ctypes.POINTER(ctypes.c_int64)]
        lib.mpoa_synth_batch.restype = ctypes.c_int
        lib.mpoa_synth_batch.argtypes = [
            ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
            ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        _synth = lib
    return _synth

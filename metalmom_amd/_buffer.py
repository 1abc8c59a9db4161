"""MMBuffer -> NumPy, following the reference ``python/metalmom/_buffer.py:7-34``.

The C buffer is always released with ``mm_buffer_free``, also when the copy or
the reshape raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import lib


def buffer_to_numpy(buf) -> np.ndarray:
    """Copies an ``MMBuffer`` into a new array (float32, or complex64 for
    dtype 1) of shape ``buf.shape[:ndim]`` and frees the C buffer."""
    try:
        shape = tuple(int(buf.shape[i]) for i in range(buf.ndim))
        dt = np.complex64 if buf.dtype == 1 else np.float32
        if not buf.data or buf.count == 0:
            return np.empty(shape, dtype=dt)
        nfloat = int(buf.count) * (2 if buf.dtype == 1 else 1)
        flat = np.ctypeslib.as_array(C.cast(buf.data, C.POINTER(C.c_float)), shape=(nfloat,)).copy()
        return flat.view(dt).reshape(shape)
    finally:
        lib.mm_buffer_free(C.byref(buf))

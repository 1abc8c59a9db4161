"""MMBuffer -> NumPy, following the reference ``python/metalmom/_buffer.py:7-34``.

The array views the C buffer without a copy; the buffer is released with
``mm_buffer_free`` when the last array viewing it is collected, or at once
when the wrap raises or the buffer is empty.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import lib


class _Owner:
    """Holds a copy of an MMBuffer descriptor and frees its data on collection."""
    __slots__ = ("buf",)

    def __init__(self, buf):
        self.buf = buf

    def __del__(self):
        try:
            lib.mm_buffer_free(C.byref(self.buf))
        except Exception:   # interpreter shutdown
            pass


def buffer_to_numpy(buf) -> np.ndarray:
    """Wraps an ``MMBuffer`` as an array (float32, or complex64 for dtype 1) of
    shape ``buf.shape[:ndim]``; the array owns the C buffer from then on."""
    own = None
    try:
        shape = tuple(int(buf.shape[i]) for i in range(buf.ndim))
        dt = np.complex64 if buf.dtype == 1 else np.float32
        if not buf.data or buf.count == 0:
            return np.empty(shape, dtype=dt)
        nfloat = int(buf.count) * (2 if buf.dtype == 1 else 1)
        raw = (C.c_float * nfloat).from_address(C.cast(buf.data, C.c_void_p).value)
        own = _Owner(type(buf).from_buffer_copy(buf))
        raw._owner = own            # the array's base (a memoryview of raw) keeps raw alive
        buf.data = None             # the caller's descriptor no longer owns the data
        return np.ctypeslib.as_array(raw).view(dt).reshape(shape)
    finally:
        if own is None:
            lib.mm_buffer_free(C.byref(buf))

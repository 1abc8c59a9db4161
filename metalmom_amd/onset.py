"""Onset strength, mirroring the reference ``python/metalmom/onset.py:8-71``."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._buffer import buffer_to_numpy
from ._native import MMBuffer, float_ptr, lib


def onset_strength(y=None, sr=22050, S=None, n_fft=2048, hop_length=512, n_mels=128, fmin=0.0, fmax=None,
                   center=True, aggregate=True, **kwargs):
    """Onset strength envelope: spectral flux (lag 1, half-wave rectified) of
    the mel dB spectrogram (``power_to_db(top_db=80)``), averaged over the
    bands. Shape ``(n_frames,)`` with ``aggregate``, else ``(n_mels, n_frames)``.
    ``S`` (a precomputed spectrogram) is not supported, as in the reference.
    """
    if y is None:
        raise ValueError("y must be provided")
    y = np.ascontiguousarray(y, dtype=np.float32)
    c_fmax = float(fmax) if fmax is not None else 0.0
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_onset_strength(ctx, float_ptr(y), len(y), sr, n_fft, hop_length, n_mels, fmin, c_fmax,
                                       1 if center else 0, 1 if aggregate else 0, C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_onset_strength failed with status {status}")
        result = buffer_to_numpy(out)
        return result.ravel() if aggregate else result
    finally:
        lib.mm_destroy(ctx)

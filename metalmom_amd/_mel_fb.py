"""NumPy Slaney mel filterbank for the precomputed-``S`` path of ``melspectrogram``.

The reference builds this matrix in NumPy for that path as well
(``python/metalmom/_mel_fb.py:12-75``, same formula as
``Filters/FilterBank.swift:23-73``). Audio input never uses it: there the
filterbank is generated on the GPU in FP64.
"""
from __future__ import annotations

import numpy as np


def _hz_to_mel(hz):
    f_sp = 200.0 / 3.0
    min_log_hz, min_log_mel, logstep = 1000.0, 1000.0 / f_sp, np.log(6.4) / 27.0
    hz = np.asarray(hz, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(hz >= min_log_hz, min_log_mel + np.log(hz / min_log_hz) / logstep, hz / f_sp)


def _mel_to_hz(mel):
    f_sp = 200.0 / 3.0
    min_log_hz, min_log_mel, logstep = 1000.0, 1000.0 / f_sp, np.log(6.4) / 27.0
    mel = np.asarray(mel, dtype=np.float64)
    return np.where(mel < min_log_mel, mel * f_sp, min_log_hz * np.exp((mel - min_log_mel) * logstep))


def mel_filterbank(sr=22050, n_fft=2048, n_mels=128, fmin=0.0, fmax=None, dtype=np.float32) -> np.ndarray:
    """``[n_mels, n_fft//2+1]`` triangles with Slaney area normalisation,
    computed in float64 and returned as ``dtype`` (float32 by default)."""
    fmax = sr / 2.0 if fmax is None else float(fmax)
    n_freqs = n_fft // 2 + 1
    pts = _mel_to_hz(np.linspace(_hz_to_mel(fmin), _hz_to_mel(fmax), n_mels + 2))
    freqs = np.arange(n_freqs, dtype=np.float64) * sr / n_fft
    lo, ce, hi = pts[:-2, None], pts[1:-1, None], pts[2:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        up = np.where((freqs >= lo) & (freqs <= ce) & (ce != lo), (freqs - lo) / (ce - lo), 0.0)
        down = np.where((freqs > ce) & (freqs <= hi) & (hi != ce), (hi - freqs) / (hi - ce), 0.0)
    w = np.where(up > 0, up, down) * (2.0 / (hi - lo))
    return w.astype(dtype)

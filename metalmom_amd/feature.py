"""Feature entry points, mirroring the reference ``python/metalmom/feature.py``.

``melspectrogram`` (``feature.py:8-91``), ``mfcc`` (``:94-167``),
``amplitude_to_db`` (``:170-220``), ``power_to_db`` (``:223-270``),
``chroma_stft`` (``:273-360``) and the spectral descriptors (``:612-945``)
keep the reference signatures, defaults (hop 512 for mel/MFCC, ``n_mfcc=20``) and error
behaviour (``RuntimeError("mm_... failed with status s")``). Audio input runs
on the GPU through libmetalmom.so. When a precomputed ``S`` is given, the
reference applies the filterbank / DCT in NumPy/SciPy (``feature.py:50-57,
130-134``) and so does this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._buffer import buffer_to_numpy
from ._native import MMBuffer, float_ptr, lib


def melspectrogram(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None,
                   center=True, power=2.0, n_mels=128, fmin=0.0, fmax=None, **kwargs):
    """Mel spectrogram ``(n_mels, n_frames)``: ``FB @ |STFT|^power``."""
    if S is not None:
        from ._mel_fb import mel_filterbank
        S = np.ascontiguousarray(S, dtype=np.float32)
        fb = mel_filterbank(sr=sr, n_fft=n_fft, n_mels=n_mels, fmin=fmin, fmax=fmax)
        return fb @ S
    if y is None:
        raise ValueError("Either y or S must be provided")
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else 512
    win = win_length if win_length is not None else n_fft
    c_fmax = float(fmax) if fmax is not None else 0.0
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_mel_spectrogram(ctx, float_ptr(y), len(y), sr, n_fft, hop, win,
                                        1 if center else 0, power, n_mels, fmin, c_fmax,
                                        C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_mel_spectrogram failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def mfcc(y=None, sr=22050, S=None, n_mfcc=20, n_fft=2048, hop_length=None, win_length=None,
         n_mels=128, fmin=0.0, fmax=None, center=True, **kwargs):
    """MFCCs ``(n_mfcc, n_frames)``: DCT-II (ortho) of ``power_to_db(mel)``."""
    if S is not None:
        from scipy.fftpack import dct
        S = np.ascontiguousarray(S, dtype=np.float32)
        return dct(S, axis=0, type=2, norm="ortho")[:n_mfcc]
    if y is None:
        raise ValueError("Either y or S must be provided")
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else 512
    win = win_length if win_length is not None else n_fft
    c_fmax = float(fmax) if fmax is not None else 0.0
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_mfcc(ctx, float_ptr(y), len(y), sr, n_mfcc, n_fft, hop, win, n_mels, fmin,
                             c_fmax, 1 if center else 0, C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_mfcc failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def _db(fn_name, S, ref, amin, top_db):
    S = np.ascontiguousarray(S, dtype=np.float32)
    shape = S.shape
    if callable(ref):
        ref = float(ref(S))
    flat = S.ravel()
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        c_top_db = float(top_db) if top_db is not None else 0.0
        status = getattr(lib, fn_name)(ctx, float_ptr(flat), len(flat), float(ref), float(amin),
                                       c_top_db, C.byref(out))
        if status != 0:
            raise RuntimeError(f"{fn_name} failed with status {status}")
        return buffer_to_numpy(out).reshape(shape)
    finally:
        lib.mm_destroy(ctx)


def amplitude_to_db(S, ref=1.0, amin=1e-5, top_db=80.0):
    """``20 log10(max(S, amin) / ref)``, clipped to ``max - top_db``."""
    return _db("mm_amplitude_to_db", S, ref, amin, top_db)


def power_to_db(S, ref=1.0, amin=1e-10, top_db=80.0):
    """``10 log10(max(S, amin) / ref)``, clipped to ``max - top_db``."""
    return _db("mm_power_to_db", S, ref, amin, top_db)


def _chroma_filterbank(sr, n_fft, n_chroma=12, tuning=0.0, ctroct=5.0, octwidth=2.0):
    """Chroma filterbank ``(n_chroma, n_fft // 2 + 1)`` for the ``S=`` path,
    as the reference's wrapper builds it in Python (``feature.py:364-420``,
    librosa's ``filters.chroma``): Gaussian pitch-class bumps, L2 column norm,
    octave weighting, rows starting at C."""
    freqs = np.arange(1, n_fft, dtype=np.float64) * sr / n_fft
    a440 = 440.0 * 2.0 ** (tuning / n_chroma)
    frq = n_chroma * np.log2(freqs / (a440 / 16.0))
    frq = np.concatenate(([frq[0] - 1.5 * n_chroma], frq))
    width = np.concatenate((np.maximum(frq[1:] - frq[:-1], 1.0), [1.0]))
    half = np.round(n_chroma / 2.0)
    D = np.subtract.outer(frq, np.arange(n_chroma, dtype=np.float64)).T
    D = np.remainder(D + half + 10 * n_chroma, n_chroma) - half
    w = np.exp(-0.5 * (2.0 * D / width[None, :]) ** 2)
    w = w / np.maximum(np.sqrt((w * w).sum(axis=0)), np.finfo(np.float64).tiny)
    w = w * np.exp(-0.5 * ((frq / n_chroma - ctroct) / octwidth) ** 2)[None, :]
    w = np.roll(w, -3 * (n_chroma // 12), axis=0)
    return np.ascontiguousarray(w[:, :n_fft // 2 + 1], dtype=np.float32)


def _normalize_frames(chroma, norm=2.0):
    """Per-frame normalisation of the reference wrapper (``feature.py:1150-1177``):
    L1, L2, L-inf or a general p-norm; frames with norm <= 1e-10 unchanged."""
    chroma = np.array(chroma, copy=True)
    if norm == 2.0:
        n = np.sqrt(np.sum(chroma ** 2, axis=0))
    elif norm == 1.0:
        n = np.sum(np.abs(chroma), axis=0)
    elif norm == np.inf:
        n = np.max(np.abs(chroma), axis=0) if chroma.shape[0] else np.zeros(chroma.shape[1])
    else:
        n = np.linalg.norm(chroma, ord=norm, axis=0)
    ok = n > 1e-10
    chroma[:, ok] /= n[ok]
    return chroma


def chroma_stft(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, n_chroma=12, center=True,
                norm=None, tuning=0.0, **kwargs):
    """STFT chroma ``(n_chroma, n_frames)`` (the reference's ``feature.py:273-360``):
    the chroma filterbank times the power spectrogram on the GPU
    (``mm_chroma_stft``; default ``hop_length`` 512, ``win_length`` n_fft),
    then the optional per-frame ``norm`` (1, 2, inf). ``S`` (a precomputed
    power spectrogram) is multiplied by the filterbank directly, and only
    that path uses ``tuning``, as in the reference."""
    if S is not None:
        S = np.ascontiguousarray(S, dtype=np.float32)
        fb = _chroma_filterbank(sr, (S.shape[0] - 1) * 2, n_chroma, tuning)
        chroma = fb @ S
        return _normalize_frames(chroma, norm) if norm is not None else chroma
    if y is None:
        raise ValueError("Either y or S must be provided")
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else 512
    win = win_length if win_length is not None else n_fft
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_chroma_stft(ctx, float_ptr(y), len(y), sr, n_fft, hop, win, n_chroma,
                                    1 if center else 0, C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_chroma_stft failed with status {status}")
        result = buffer_to_numpy(out)
        return _normalize_frames(result, norm) if norm is not None else result
    finally:
        lib.mm_destroy(ctx)


def _descriptor(fn_name, y, sr, n_fft, hop_length, win_length, center, *extra):
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else 512
    win = win_length if win_length is not None else n_fft
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = getattr(lib, fn_name)(ctx, float_ptr(y), len(y), sr, n_fft, hop, win, 1 if center else 0,
                                       *extra, C.byref(out))
        if status != 0:
            raise RuntimeError(f"{fn_name} failed with status {status}")
        return buffer_to_numpy(out).reshape(1, -1)
    finally:
        lib.mm_destroy(ctx)


def _freqs(n_freqs, sr, n_fft):
    return np.arange(n_freqs, dtype=np.float32) * sr / n_fft


def spectral_centroid(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, center=True,
                      freq=None, **kwargs):
    """Spectral centroid ``(1, n_frames)`` (the reference's ``feature.py:612-672``):
    ``sum f S / sum S`` per frame of the magnitude STFT, on the GPU for audio
    (default ``hop_length`` 512); ``S`` (a magnitude spectrogram) in NumPy."""
    if S is not None:
        S = np.ascontiguousarray(S, dtype=np.float32)
        f = _freqs(S.shape[0], sr, (S.shape[0] - 1) * 2)
        total = S.sum(axis=0, keepdims=True)
        total = np.where(total > 0, total, 1.0)
        return (f[:, None] * S).sum(axis=0, keepdims=True) / total
    if y is None:
        raise ValueError("Either y or S must be provided")
    return _descriptor("mm_spectral_centroid", y, sr, n_fft, hop_length, win_length, center)


def spectral_bandwidth(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, center=True,
                       freq=None, p=2, **kwargs):
    """Spectral bandwidth ``(1, n_frames)`` (``feature.py:675-740``):
    ``(sum S |f - centroid|^p / sum S)^(1/p)``."""
    if S is not None:
        S = np.ascontiguousarray(S, dtype=np.float32)
        n_fft_s = (S.shape[0] - 1) * 2
        f = _freqs(S.shape[0], sr, n_fft_s)
        c = spectral_centroid(S=S, sr=sr, n_fft=n_fft_s)
        total = S.sum(axis=0, keepdims=True)
        total = np.where(total > 0, total, 1.0)
        return ((S * np.abs(f[:, None] - c) ** p).sum(axis=0, keepdims=True) / total) ** (1.0 / p)
    if y is None:
        raise ValueError("Either y or S must be provided")
    return _descriptor("mm_spectral_bandwidth", y, sr, n_fft, hop_length, win_length, center, float(p))


def spectral_contrast(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, center=True,
                      n_bands=6, fmin=200.0, quantile=0.02, **kwargs):
    """Spectral contrast ``(n_bands + 1, n_frames)`` (the reference's
    ``feature.py:744-806``): per octave band of the magnitude STFT, dB of the
    mean of its top 2 % over the mean of its bottom 2 % (on the GPU; default
    ``hop_length`` 512). As in the reference, ``quantile`` is not passed to
    the C ABI (which always uses 0.02) and ``S`` is not supported.
    """
    if S is not None:
        raise NotImplementedError("spectral_contrast with pre-computed S not yet supported")
    if y is None:
        raise ValueError("Either y or S must be provided")
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else 512
    win = win_length if win_length is not None else n_fft
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_spectral_contrast(ctx, float_ptr(y), len(y), sr, n_fft, hop, win, 1 if center else 0,
                                          int(n_bands), float(fmin), C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_spectral_contrast failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def spectral_rolloff(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, center=True,
                     roll_percent=0.85, **kwargs):
    """Spectral rolloff ``(1, n_frames)`` (``feature.py:808-880``): frequency of
    the first bin where the cumulative magnitude reaches ``roll_percent`` of
    the frame's total."""
    if S is not None:
        S = np.ascontiguousarray(S, dtype=np.float32)
        f = _freqs(S.shape[0], sr, (S.shape[0] - 1) * 2)
        total = S.sum(axis=0, keepdims=True)
        thr = roll_percent * total
        cum = np.cumsum(S, axis=0)
        res = np.zeros((1, S.shape[1]), dtype=np.float32)
        for t in range(S.shape[1]):
            if total[0, t] > 0:
                res[0, t] = f[min(int(np.searchsorted(cum[:, t], thr[0, t])), S.shape[0] - 1)]
        return res
    if y is None:
        raise ValueError("Either y or S must be provided")
    return _descriptor("mm_spectral_rolloff", y, sr, n_fft, hop_length, win_length, center, float(roll_percent))


def spectral_flatness(y=None, sr=22050, S=None, n_fft=2048, hop_length=None, win_length=None, center=True,
                      power=2.0, amin=1e-10, **kwargs):
    """Spectral flatness ``(1, n_frames)`` (``feature.py:883-945``): geometric over
    arithmetic mean of ``max(S^2, 1e-10)``; the audio path uses power 2 and
    amin 1e-10 as the reference's bridge does, ``S`` uses ``amin`` as given."""
    if S is not None:
        S = np.ascontiguousarray(S, dtype=np.float32)
        Sf = np.maximum(S, amin)
        raw = S.sum(axis=0)
        res = np.where(raw > amin, np.exp(np.mean(np.log(Sf), axis=0)) / np.mean(Sf, axis=0), 0.0)
        return res.reshape(1, -1).astype(np.float32)
    if y is None:
        raise ValueError("Either y or S must be provided")
    return _descriptor("mm_spectral_flatness", y, sr, n_fft, hop_length, win_length, center)

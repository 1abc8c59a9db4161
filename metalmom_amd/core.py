"""STFT, iSTFT, audio loading and resampling entry points, mirroring the reference ``python/metalmom/core.py``."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._buffer import buffer_to_numpy
from ._native import MMBuffer, MMSTFTParams, MMSpecParams, float_ptr, lib
from . import _native as N


def _call_stft(fn_name, y, n_fft, hop_length, win_length, center):
    y = np.ascontiguousarray(y, dtype=np.float32)
    hop = hop_length if hop_length is not None else n_fft // 4
    win = win_length if win_length is not None else n_fft
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        params = MMSTFTParams(n_fft, hop, win, 1 if center else 0)
        out = MMBuffer()
        # sr is metadata only for the STFT; the reference passes 22050 (core.py:243)
        status = getattr(lib, fn_name)(ctx, float_ptr(y), len(y), 22050, C.byref(params), C.byref(out))
        if status != 0:
            raise RuntimeError(f"{fn_name} failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def stft(y, n_fft=2048, hop_length=None, win_length=None, center=True):
    """Magnitude STFT, shape ``(n_fft // 2 + 1, n_frames)`` (``core.py:198-250``).

    Defaults: ``hop_length = n_fft // 4``, ``win_length = n_fft``; periodic Hann
    window; ``center`` pads ``n_fft // 2`` zeros on both sides.
    """
    return _call_stft("mm_stft", y, n_fft, hop_length, win_length, center)


def stft_complex(y, n_fft=2048, hop_length=None, win_length=None, center=True):
    """Complex64 STFT with the same framing as :func:`stft` (the reference's
    ``STFT.computeComplex``, ``STFT.swift:512-668``; not exported by its ABI)."""
    return _call_stft("mm_stft_complex", y, n_fft, hop_length, win_length, center)


def istft(stft_matrix, hop_length=None, win_length=None, center=True, length=None):
    """Inverse STFT of a complex spectrogram ``(n_fft // 2 + 1, n_frames)``
    (the reference's ``core.py:253-318``): windowed overlap-add divided by the
    summed squared window. Defaults ``hop_length = n_fft // 4`` and
    ``win_length = n_fft``; ``center`` trims ``n_fft // 2`` samples at both
    ends; ``length`` cuts or zero-pads the result. Real input is taken as
    ``x + 0j``. Returns a 1-D float32 array.
    """
    S = np.asarray(stft_matrix)
    S = S.astype(np.complex64) if not np.iscomplexobj(S) else np.asarray(S, dtype=np.complex64)
    n_freqs, n_frames = S.shape
    n_fft = (n_freqs - 1) * 2
    hop = hop_length if hop_length is not None else n_fft // 4
    win = win_length if win_length is not None else n_fft
    flat = np.ascontiguousarray(S).view(np.float32)
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_istft(ctx, float_ptr(flat), flat.size, n_freqs, n_frames, 22050, hop, win,
                              1 if center else 0, length if length is not None else 0, C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_istft failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def load(path, sr=22050, mono=True, offset=0.0, duration=None, **kwargs):
    """Loads an audio file (the reference's ``core.py:8-58``): returns
    ``(y, sr)`` with ``y`` a 1-D float32 array. ``sr=None`` keeps the file's
    rate; another rate is reached with the GPU resampler. ``mono`` averages
    the channels (otherwise channel 0 is returned); ``offset`` and
    ``duration`` are in seconds. Reads RIFF/WAVE files (integer PCM of 8 to
    32 bits, float32 / float64).
    """
    path_bytes = os.fsencode(path)
    c_sr = int(sr) if sr is not None else 0
    c_dur = float(duration) if duration is not None else 0.0
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        out_sr = C.c_int32()
        status = lib.mm_load(ctx, path_bytes, c_sr, 1 if mono else 0, float(offset), c_dur, C.byref(out),
                             C.byref(out_sr))
        if status != 0:
            raise RuntimeError(f"mm_load failed with status {status}")
        return buffer_to_numpy(out), int(out_sr.value)
    finally:
        lib.mm_destroy(ctx)


def get_duration(path, **kwargs):
    """Duration of an audio file in seconds (the reference's ``core.py:61-85``)."""
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        d = C.c_float()
        status = lib.mm_get_duration(ctx, os.fsencode(path), C.byref(d))
        if status != 0:
            raise RuntimeError(f"mm_get_duration failed with status {status}")
        return float(d.value)
    finally:
        lib.mm_destroy(ctx)


def get_samplerate(path, **kwargs):
    """Sample rate of an audio file in Hz (the reference's ``core.py:88-112``)."""
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        r = C.c_int32()
        status = lib.mm_get_sample_rate(ctx, os.fsencode(path), C.byref(r))
        if status != 0:
            raise RuntimeError(f"mm_get_sample_rate failed with status {status}")
        return int(r.value)
    finally:
        lib.mm_destroy(ctx)


def resample(y, orig_sr, target_sr, **kwargs):
    """Resamples a 1-D signal from ``orig_sr`` to ``target_sr`` (the
    reference's ``core.py:115-160``): Kaiser-windowed sinc interpolation
    (``mm_resample``) on the GPU. Equal rates return the input as float32
    without a copy. Returns a 1-D float32 array of ``ceil(len(y) * target_sr
    / orig_sr)`` samples.
    """
    if orig_sr == target_sr:
        return np.asarray(y, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_resample(ctx, float_ptr(y), len(y), int(orig_sr), int(target_sr), C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_resample failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


def spectral(y, kind="mfcc", sr=22050, n_fft=2048, hop_length=512, win_length=None, center=True,
             pad_mode="constant", power=2.0, n_mels=128, fmin=0.0, fmax=None, n_mfcc=13,
             top_db=80.0, precision="f32"):
    """One clip with every option of ``MMSpecParams`` (additive API).

    ``kind`` is one of ``stft``, ``power``, ``complex``, ``mel``, ``mfcc``,
    ``chroma`` (``n_mels`` chroma bins, tuning 0, no frame normalisation);
    ``precision`` one of ``f32``, ``bf16``, ``f16`` (mel / chroma and DCT
    contraction).
    """
    y = np.ascontiguousarray(y, dtype=np.float32)
    p = make_params(kind, sr, n_fft, hop_length, win_length, center, pad_mode, power, n_mels,
                    fmin, fmax, n_mfcc, top_db, precision, "f32")
    ctx = lib.mm_init()
    if not ctx:
        raise RuntimeError("Failed to initialize MetalMom context")
    try:
        out = MMBuffer()
        status = lib.mm_spectral(ctx, C.byref(p), float_ptr(y), len(y), C.byref(out))
        if status != 0:
            raise RuntimeError(f"mm_spectral failed with status {status}")
        return buffer_to_numpy(out)
    finally:
        lib.mm_destroy(ctx)


_KINDS = {"stft": N.KIND_STFT_MAG, "magnitude": N.KIND_STFT_MAG, "power": N.KIND_STFT_POWER,
          "complex": N.KIND_STFT_COMPLEX, "mel": N.KIND_MEL, "melspectrogram": N.KIND_MEL,
          "mfcc": N.KIND_MFCC, "chroma": N.KIND_CHROMA, "chroma_stft": N.KIND_CHROMA}
_PRECS = {"f32": N.PREC_F32, "fp32": N.PREC_F32, "bf16": N.PREC_BF16, "f16": N.PREC_F16,
          "fp16": N.PREC_F16}
_INPUTS = {"f32": N.INPUT_F32, "float32": N.INPUT_F32, "f16": N.INPUT_F16, "float16": N.INPUT_F16}
_PADS = {"constant": N.PAD_CONSTANT, "reflect": N.PAD_REFLECT}


def make_params(kind, sr, n_fft, hop_length, win_length, center, pad_mode, power, n_mels, fmin,
                fmax, n_mfcc, top_db, precision, input_dtype) -> MMSpecParams:
    """Builds ``MMSpecParams``; ``None`` maps to the ABI sentinels."""
    return MMSpecParams(
        kind=_KINDS[kind] if isinstance(kind, str) else int(kind), sample_rate=int(sr),
        n_fft=int(n_fft), hop_length=int(hop_length or 0), win_length=int(win_length or 0),
        center=1 if center else 0, pad_mode=_PADS[pad_mode], power=float(power),
        n_mels=int(n_mels), f_min=float(fmin), f_max=float(fmax or 0.0), n_mfcc=int(n_mfcc),
        top_db=float(top_db) if top_db is not None else 0.0,
        precision=_PRECS[precision] if isinstance(precision, str) else int(precision),
        input_dtype=_INPUTS[input_dtype] if isinstance(input_dtype, str) else int(input_dtype))

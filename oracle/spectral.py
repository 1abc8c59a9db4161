"""Float64 host oracle for the spectral front end (STFT -> power -> mel -> dB -> DCT-II).

TEST INFRASTRUCTURE ONLY. Nothing in the shipped library imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, and only as the checker / the timed CPU restatement.

Every function restates the reference algorithm (zakkeown/MetalMom) in float64
NumPy and cites the Swift source it follows. How each function is pinned
(golden vectors, closed forms, brute force) is listed in ``PINS`` at the end of
this file and exercised by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "hann_periodic", "prepare_window", "frame_count", "pad_signal",
    "stft_complex", "stft_magnitude", "power_spectrogram", "istft_length", "istft",
    "hz_to_mel", "mel_to_hz", "mel_filterbank", "melspectrogram",
    "linear_to_db", "power_to_db", "amplitude_to_db",
    "dct_ortho_basis", "mfcc", "mfcc_from_logmel", "onset_strength",
    "chroma_filterbank", "chroma_stft", "normalize_frames",
    "spectral_centroid", "spectral_bandwidth", "spectral_rolloff", "spectral_flatness",
    "contrast_bands", "contrast_from_magnitude", "spectral_contrast",
]


# ---------------------------------------------------------------------------
# Window and framing
# ---------------------------------------------------------------------------

def hann_periodic(length: int) -> np.ndarray:
    """Periodic Hann window, ``w[n] = 0.5 - 0.5 cos(2 pi n / length)``.

    Follows ``Sources/MetalMomCore/Filters/Windows.swift:14-48``: a symmetric
    window of length ``length + 1`` with the last sample dropped; length 0 gives
    an empty window and length 1 gives ``[1.0]``.
    """
    if length <= 0:
        return np.zeros(0, dtype=np.float64)
    if length == 1:
        return np.ones(1, dtype=np.float64)
    n = np.arange(length, dtype=np.float64)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * n / length)


def prepare_window(win_length: int, n_fft: int) -> np.ndarray:
    """Full ``n_fft``-sample analysis window.

    Follows ``Sources/MetalMomCore/Spectral/STFT.swift:44-57``: a shorter window
    is zero-padded with ``(n_fft - win)//2`` zeros in front, a longer one is cut
    to its centre ``n_fft`` samples. A non-positive ``win_length`` yields an
    all-zero window (``Windows.hann`` returns ``[]`` there and the padding fills
    the frame with zeros).
    """
    w = hann_periodic(win_length)
    if win_length < n_fft:
        before = (n_fft - win_length) // 2
        out = np.zeros(n_fft, dtype=np.float64)
        if win_length > 0:
            out[before:before + win_length] = w
        return out
    if win_length > n_fft:
        start = (win_length - n_fft) // 2
        return w[start:start + n_fft].copy()
    return w


def frame_count(n_samples: int, n_fft: int, hop: int, center: bool) -> int:
    """Number of STFT frames, ``1 + (padded - n_fft) // hop`` or 0.

    Follows ``STFT.swift:240-254`` (also ``:79-98``): with ``center`` the signal
    is padded by ``n_fft//2`` on both sides; a padded length below ``n_fft``
    gives no frame.
    """
    padded = n_samples + (2 * (n_fft // 2) if center else 0)
    if padded < n_fft:
        return 0
    return 1 + (padded - n_fft) // hop


def pad_signal(y: np.ndarray, n_fft: int, center: bool, pad_mode: str = "constant") -> np.ndarray:
    """Centre padding by ``n_fft//2`` samples per side.

    ``constant`` (zeros) is the reference's behaviour (``STFT.swift:81-85,
    242-246``). ``reflect`` follows ``numpy.pad(mode="reflect")`` and is an
    addition of this build (the reference has no reflect framing).
    """
    y = np.asarray(y, dtype=np.float64)
    if not center:
        return y.copy()
    p = n_fft // 2
    if pad_mode == "constant":
        return np.pad(y, (p, p), mode="constant")
    if pad_mode == "reflect":
        return np.pad(y, (p, p), mode="reflect")
    raise ValueError(f"unknown pad_mode {pad_mode!r}")


# ---------------------------------------------------------------------------
# STFT
# ---------------------------------------------------------------------------

def stft_complex(y, n_fft=2048, hop=None, win_length=None, center=True,
                 pad_mode="constant") -> np.ndarray:
    """Complex STFT, shape ``[n_fft//2+1, n_frames]`` (row-major, bins x frames).

    Restates ``STFT.swift:512-668`` (``computeComplex``): frame ``f`` is
    ``pad(y)[f*hop : f*hop+n_fft] * window`` and bin ``k`` is the plain forward
    DFT ``sum_n x[n] exp(-2 pi i k n / n_fft)`` (vDSP's factor 2 is undone at
    ``STFT.swift:308-310``). Computed here with ``numpy.fft.rfft`` in float64.
    """
    hop = n_fft // 4 if hop is None else hop
    win_length = n_fft if win_length is None else win_length
    y = np.asarray(y, dtype=np.float64)
    n_freqs = n_fft // 2 + 1
    n_frames = frame_count(len(y), n_fft, hop, center)
    if n_frames == 0:
        return np.zeros((n_freqs, 0), dtype=np.complex128)
    x = pad_signal(y, n_fft, center, pad_mode)
    w = prepare_window(win_length, n_fft)
    idx = np.arange(n_frames)[:, None] * hop + np.arange(n_fft)[None, :]
    frames = x[idx] * w[None, :]
    return np.fft.rfft(frames, n=n_fft, axis=1).T.copy()


def istft_length(n_freqs: int, n_frames: int, hop: int, center: bool, length=None) -> int:
    """Sample count of ``istft``'s result (``STFT.swift:703-704, 832-858``).

    The overlap-add buffer holds ``n_fft + (n_frames - 1) * hop`` samples
    (``n_fft = 2 (n_freqs - 1)``). ``length`` (if given) is the result's
    size; otherwise ``center`` trims ``n_fft//2`` from both ends (nothing
    left when the buffer is not longer than ``n_fft``). A non-power-of-two
    ``n_fft`` returns the whole zero buffer untrimmed (``:720-724``).
    """
    n_fft = (n_freqs - 1) * 2
    total = n_fft + (n_frames - 1) * hop
    if not (n_fft > 0 and (n_fft & (n_fft - 1)) == 0):
        return max(total, 0)
    if length is not None:
        return int(length)
    if center:
        return max(total - n_fft, 0) if total - n_fft // 2 > n_fft // 2 else 0
    return total


def istft(stft_matrix, hop=None, win_length=None, center=True, length=None) -> np.ndarray:
    """Inverse STFT by windowed overlap-add, the output of ``mm_istft``.

    Restates ``STFT.swift:686-861`` (``STFT.inverse``): ``n_fft = 2 (F - 1)``;
    every frame is the inverse real DFT of its column with the imaginary parts
    of the DC and Nyquist bins ignored (vDSP's packed format, ``:760-768``),
    scaled to the true inverse (``1/n_fft``); it is multiplied by the
    synthesis window (the analysis window of ``prepare_window``) and added
    at ``f * hop``. Samples whose sum of squared windows exceeds ``1e-10`` are
    divided by it (``:820-826``). Then the centre trim / ``length`` rule of
    ``:832-858``: with ``center`` the result starts at ``n_fft//2`` and is cut
    or zero-padded to ``length``; without it the buffer itself is cut or
    padded. A non-power-of-two ``n_fft`` gives the all-zero buffer, untrimmed.
    Computed here with ``numpy.fft.irfft`` in float64.
    """
    S = np.asarray(stft_matrix)
    n_freqs, n_frames = S.shape
    n_fft = (n_freqs - 1) * 2
    hop = n_fft // 4 if hop is None else hop
    win_length = n_fft if win_length is None else win_length
    total = n_fft + (n_frames - 1) * hop
    if not (n_fft > 0 and (n_fft & (n_fft - 1)) == 0):
        return np.zeros(max(total, 0), dtype=np.float64)
    w = prepare_window(win_length, n_fft)
    out = np.zeros(total, dtype=np.float64)
    wsum = np.zeros(total, dtype=np.float64)
    X = S.astype(np.complex128).copy()
    X[0, :] = X[0, :].real
    X[-1, :] = X[-1, :].real
    frames = np.fft.irfft(X.T, n=n_fft, axis=1) * w[None, :]
    for f in range(n_frames):
        out[f * hop:f * hop + n_fft] += frames[f]
        wsum[f * hop:f * hop + n_fft] += w * w
    nz = wsum > 1e-10
    out[nz] /= wsum[nz]
    n_out = istft_length(n_freqs, n_frames, hop, center, length)
    start = n_fft // 2 if center else 0
    res = np.zeros(n_out, dtype=np.float64)
    avail = max(min(n_out, total - start), 0)
    res[:avail] = out[start:start + avail]
    return res


def stft_magnitude(y, n_fft=2048, hop=None, win_length=None, center=True,
                   pad_mode="constant") -> np.ndarray:
    """``|STFT|``, the output of ``mm_stft`` (``STFT.swift:73-215``)."""
    return np.abs(stft_complex(y, n_fft, hop, win_length, center, pad_mode))


def power_spectrogram(y, n_fft=2048, hop=None, win_length=None, center=True,
                      pad_mode="constant") -> np.ndarray:
    """``re^2 + im^2`` per bin (``STFT.swift:227-340``)."""
    X = stft_complex(y, n_fft, hop, win_length, center, pad_mode)
    return X.real ** 2 + X.imag ** 2


# ---------------------------------------------------------------------------
# Mel filterbank
# ---------------------------------------------------------------------------

_F_SP = 200.0 / 3.0
_MIN_LOG_HZ = 1000.0
_MIN_LOG_MEL = _MIN_LOG_HZ / _F_SP
_LOGSTEP = np.log(6.4) / 27.0


def hz_to_mel(hz):
    """Slaney mel scale, ``Convert/Units.swift:55-67`` (float64)."""
    hz = np.asarray(hz, dtype=np.float64)
    lin = hz / _F_SP
    with np.errstate(divide="ignore", invalid="ignore"):
        log = _MIN_LOG_MEL + np.log(hz / _MIN_LOG_HZ) / _LOGSTEP
    return np.where(hz >= _MIN_LOG_HZ, log, lin)


def mel_to_hz(mel):
    """Inverse Slaney mel scale, ``Convert/Units.swift:74-85`` (float64)."""
    mel = np.asarray(mel, dtype=np.float64)
    lin = mel * _F_SP
    log = _MIN_LOG_HZ * np.exp((mel - _MIN_LOG_MEL) * _LOGSTEP)
    return np.where(mel < _MIN_LOG_MEL, lin, log)


def mel_filterbank(sr, n_fft, n_mels=128, fmin=0.0, fmax=None) -> np.ndarray:
    """Slaney-normalised triangular mel filterbank ``[n_mels, n_fft//2+1]``.

    Follows ``Filters/FilterBank.swift:23-73``: ``n_mels+2`` mel points equally
    spaced between ``hz_to_mel(fmin)`` and ``hz_to_mel(fmax)``; bin ``k`` at
    ``k*sr/n_fft`` Hz gets the rising edge ``(f-fl)/(fc-fl)`` for
    ``fl <= f <= fc`` (when ``fc != fl``), else the falling edge
    ``(fr-f)/(fr-fc)`` for ``fc < f <= fr``; each row is scaled by
    ``2/(fr-fl)``. Everything here is float64 (the reference rounds the mel
    points and bin frequencies to float32 first; librosa, which made the
    goldens, stays in float64).
    """
    fmax = sr / 2.0 if fmax is None else float(fmax)
    n_freqs = n_fft // 2 + 1
    if n_mels <= 0:
        return np.zeros((0, n_freqs), dtype=np.float64)
    mel_lo, mel_hi = hz_to_mel(float(fmin)), hz_to_mel(fmax)
    pts = mel_to_hz(mel_lo + np.arange(n_mels + 2, dtype=np.float64) * (mel_hi - mel_lo) / (n_mels + 1))
    freqs = np.arange(n_freqs, dtype=np.float64) * sr / n_fft
    fb = np.zeros((n_mels, n_freqs), dtype=np.float64)
    for m in range(n_mels):
        fl, fc, fr = pts[m], pts[m + 1], pts[m + 2]
        up = (freqs >= fl) & (freqs <= fc)
        down = (freqs > fc) & (freqs <= fr)
        if fc != fl:
            fb[m, up] = (freqs[up] - fl) / (fc - fl)
        if fr != fc:
            # the reference tests the rising edge first, so a bin on it never
            # takes the falling value (the two masks are disjoint here anyway)
            fb[m, down] = (fr - freqs[down]) / (fr - fc)
        fb[m] *= 2.0 / (fr - fl)
    return fb


def melspectrogram(y, sr=22050, n_fft=2048, hop=512, win_length=None, center=True,
                   power=2.0, n_mels=128, fmin=0.0, fmax=None, pad_mode="constant") -> np.ndarray:
    """``FB @ |STFT|^power``, shape ``[n_mels, n_frames]``.

    Follows ``Spectral/MelSpectrogram.swift:28-142``: power 2 uses
    ``re^2+im^2`` directly, power 1 the magnitude, anything else
    ``magnitude ** power``.
    """
    X = stft_complex(y, n_fft, hop, win_length, center, pad_mode)
    if power == 2.0:
        S = X.real ** 2 + X.imag ** 2
    elif power == 1.0:
        S = np.abs(X)
    else:
        S = np.abs(X) ** power
    return mel_filterbank(sr, n_fft, n_mels, fmin, fmax) @ S


# ---------------------------------------------------------------------------
# dB scaling
# ---------------------------------------------------------------------------

def linear_to_db(S, multiplier, ref=1.0, amin=1e-10, top_db=80.0) -> np.ndarray:
    """``multiplier*log10(max(S, amin)/ref)``, then ``max(x, max(x) - top_db)``.

    Follows ``Spectral/Scaling.swift:73-118``: the clamp floor is the maximum
    over the whole array (one scalar per call); ``top_db=None`` skips it.
    """
    S = np.asarray(S, dtype=np.float64)
    out = multiplier * np.log10(np.maximum(S, amin) / ref)
    if top_db is not None and out.size:
        out = np.maximum(out, out.max() - top_db)
    return out


def power_to_db(S, ref=1.0, amin=1e-10, top_db=80.0) -> np.ndarray:
    """``Scaling.powerToDb`` (``Spectral/Scaling.swift:39-42``)."""
    return linear_to_db(S, 10.0, ref, amin, top_db)


def amplitude_to_db(S, ref=1.0, amin=1e-5, top_db=80.0) -> np.ndarray:
    """``Scaling.amplitudeToDb`` (``Spectral/Scaling.swift:23-26``)."""
    return linear_to_db(S, 20.0, ref, amin, top_db)


# ---------------------------------------------------------------------------
# MFCC
# ---------------------------------------------------------------------------

def dct_ortho_basis(n_mfcc: int, n_mels: int) -> np.ndarray:
    """Orthonormal DCT-II basis ``[n_mfcc, n_mels]``.

    ``basis[k, n] = cos(pi k (2n+1) / (2N)) * (sqrt(1/N) if k == 0 else sqrt(2/N))``
    as in ``Features/MFCC.swift:79-88`` (float64 here).
    """
    N = n_mels
    k = np.arange(n_mfcc, dtype=np.float64)[:, None]
    n = np.arange(N, dtype=np.float64)[None, :]
    basis = np.cos(np.pi * k * (2.0 * n + 1.0) / (2.0 * N))
    norm = np.full((n_mfcc, 1), np.sqrt(2.0 / N))
    if n_mfcc:
        norm[0, 0] = np.sqrt(1.0 / N)
    return basis * norm


def mfcc_from_logmel(logmel: np.ndarray, n_mfcc: int) -> np.ndarray:
    """DCT-II (ortho) along the mel axis, truncated to ``min(n_mfcc, n_mels)``."""
    n_mels = logmel.shape[0]
    k = min(n_mfcc, n_mels)
    return dct_ortho_basis(k, n_mels) @ logmel


def mfcc(y, sr=22050, n_mfcc=20, n_fft=2048, hop=512, win_length=None, n_mels=128,
         fmin=0.0, fmax=None, center=True, top_db=80.0, pad_mode="constant") -> np.ndarray:
    """MFCC ``[min(n_mfcc, n_mels), n_frames]``.

    Follows ``Features/MFCC.swift:28-110``: mel spectrogram with power 2,
    ``powerToDb`` with ``ref=1, amin=1e-10, top_db=80`` over the whole clip,
    then the orthonormal DCT-II basis product.
    """
    M = melspectrogram(y, sr, n_fft, hop, win_length, center, 2.0, n_mels, fmin, fmax, pad_mode)
    logmel = power_to_db(M, 1.0, 1e-10, top_db)
    return mfcc_from_logmel(logmel, n_mfcc)


# How each oracle function is pinned (checked by tests/test_oracle.py).
# ---------------------------------------------------------------------------
# Onset strength (spectral flux of the mel dB spectrogram)
# ---------------------------------------------------------------------------

def onset_strength(y, sr=22050, n_fft=2048, hop=512, n_mels=128, fmin=0.0, fmax=None, center=True,
                   aggregate=True, lag=1) -> np.ndarray:
    """Onset strength envelope, the output of ``mm_onset_strength``.

    Restates ``Rhythm/OnsetDetection.swift:23-128``: power mel spectrogram
    (``win = n_fft``), ``power_to_db(ref 1, amin 1e-10, top_db 80)`` over the
    whole matrix, ``max(0, S[:, t] - S[:, t - lag])``, averaged over the bands
    (``aggregate``) or kept per band, then ``lag + n_fft // (2 hop)`` (centre)
    or ``lag`` zeros in front and, with ``center``, cut to the frame count.
    With ``n_frames <= lag`` the result is a single zero column. Shape
    ``[1, frames]`` or ``[n_mels, frames]``.
    """
    mel = melspectrogram(y, sr, n_fft, hop, n_fft, center, 2.0, n_mels, fmin, fmax)
    n_bands, T = mel.shape
    if T <= lag:
        return np.zeros((1 if aggregate else n_bands, 1))
    db = power_to_db(mel, 1.0, 1e-10, 80.0)
    flux = np.maximum(0.0, db[:, lag:] - db[:, :-lag])
    pad = lag + (n_fft // (2 * hop) if center else 0)
    n_out = min(T - lag + pad, T) if center else T - lag + pad
    env = flux.mean(axis=0, keepdims=True) if aggregate else flux
    out = np.zeros((env.shape[0], n_out))
    m = min(env.shape[1], n_out - pad)
    out[:, pad:pad + m] = env[:, :m]
    return out


# ---------------------------------------------------------------------------
# Chroma (STFT power -> Gaussian pitch-class filterbank)
# ---------------------------------------------------------------------------

def chroma_filterbank(sr, n_fft, n_chroma=12, tuning=0.0, ctroct=5.0, octwidth=2.0) -> np.ndarray:
    """Chroma filterbank ``[n_chroma, n_fft//2 + 1]``.

    Restates ``Features/Chroma.swift:158-275`` (librosa's ``filters.chroma``):
    FFT bin k > 0 sits at ``n_chroma log2(k sr / n_fft / (A440 / 16))``
    fractional chroma bins (``A440 = 440 * 2^(tuning / n_chroma)``), the DC bin
    1.5 octaves below bin 1; bin widths ``max(next - this, 1)`` (1 for the
    last of the n_fft bins); Gaussian bumps ``exp(-0.5 (2 D / width)^2)`` of
    the distance D wrapped to ``[-n_chroma/2, n_chroma/2)``; every column
    L2-normalised; octave weighting ``exp(-0.5 ((octave - ctroct)/octwidth)^2)``;
    rows rolled to start at C (``-3 n_chroma / 12``); cut to n_fft/2 + 1 bins.
    """
    k = np.arange(1, n_fft, dtype=np.float64)
    a440 = 440.0 * 2.0 ** (tuning / n_chroma)
    frq = np.empty(n_fft)
    frq[1:] = n_chroma * np.log2(k * sr / n_fft / (a440 / 16.0))
    frq[0] = frq[1] - 1.5 * n_chroma
    width = np.ones(n_fft)
    width[:-1] = np.maximum(frq[1:] - frq[:-1], 1.0)
    half = np.round(n_chroma / 2.0)
    D = frq[None, :] - np.arange(n_chroma, dtype=np.float64)[:, None]
    D = np.remainder(D + half + 10 * n_chroma, n_chroma) - half
    w = np.exp(-0.5 * (2.0 * D / width[None, :]) ** 2)
    norm = np.sqrt((w * w).sum(axis=0))
    w = w / np.where(norm > 0, norm, 1.0)
    w = w * np.exp(-0.5 * ((frq / n_chroma - ctroct) / octwidth) ** 2)[None, :]
    w = np.roll(w, -3 * (n_chroma // 12), axis=0)
    return w[:, :n_fft // 2 + 1].copy()


def normalize_frames(C, norm):
    """Per-frame (column) normalisation of the reference wrapper
    (``python/metalmom/feature.py:1150-1177``): L1, L2 or L-inf norm, columns
    whose norm is at most 1e-10 left unchanged."""
    C = np.array(C, dtype=np.float64, copy=True)
    if norm == 2.0:
        n = np.sqrt((C * C).sum(axis=0))
    elif norm == 1.0:
        n = np.abs(C).sum(axis=0)
    elif norm == np.inf:
        n = np.abs(C).max(axis=0) if C.shape[0] else np.zeros(C.shape[1])
    else:
        n = (np.abs(C) ** norm).sum(axis=0) ** (1.0 / norm)
    ok = n > 1e-10
    C[:, ok] /= n[ok]
    return C


def chroma_stft(y, sr=22050, n_fft=2048, hop=512, win_length=None, n_chroma=12, center=True, norm=None,
                tuning=0.0) -> np.ndarray:
    """``chroma_filterbank @ |STFT|^2``, shape ``[n_chroma, n_frames]``
    (``Features/Chroma.swift:30-156``), then ``normalize_frames`` when ``norm``
    is given (the reference's Python wrapper, ``feature.py:273-360``)."""
    P = power_spectrogram(y, n_fft, hop, win_length, center)
    C = chroma_filterbank(sr, n_fft, n_chroma, tuning) @ P
    return normalize_frames(C, norm) if norm is not None else C


# ---------------------------------------------------------------------------
# Spectral descriptors (per-frame reductions of |STFT|)
# ---------------------------------------------------------------------------

def _freqs(n_freqs, sr, n_fft):
    return np.arange(n_freqs, dtype=np.float64) * sr / n_fft


def spectral_centroid(y, sr=22050, n_fft=2048, hop=512, win_length=None, center=True) -> np.ndarray:
    """``sum_k f_k S_k / sum_k S_k`` per frame of the magnitude STFT, 0 for a
    frame without energy; shape ``[1, frames]``
    (``Features/SpectralDescriptors.swift:25-56``)."""
    S = stft_magnitude(y, n_fft, hop, win_length, center)
    f = _freqs(S.shape[0], sr, n_fft)
    tot = S.sum(axis=0)
    c = np.where(tot > 0, (f[:, None] * S).sum(axis=0) / np.where(tot > 0, tot, 1.0), 0.0)
    return c[None, :]


def spectral_bandwidth(y, sr=22050, n_fft=2048, hop=512, win_length=None, center=True, p=2.0) -> np.ndarray:
    """``(sum_k S_k |f_k - centroid|^p / sum_k S_k)^(1/p)`` per frame, 0
    without energy (``SpectralDescriptors.swift:70-111``)."""
    S = stft_magnitude(y, n_fft, hop, win_length, center)
    f = _freqs(S.shape[0], sr, n_fft)
    c = spectral_centroid(y, sr, n_fft, hop, win_length, center)[0]
    tot = S.sum(axis=0)
    m = (S * np.abs(f[:, None] - c[None, :]) ** p).sum(axis=0)
    bw = np.where(tot > 0, (m / np.where(tot > 0, tot, 1.0)) ** (1.0 / p), 0.0)
    return bw[None, :]


def spectral_rolloff(y, sr=22050, n_fft=2048, hop=512, win_length=None, center=True,
                     roll_percent=0.85) -> np.ndarray:
    """Frequency of the first bin at which the cumulative magnitude reaches
    ``roll_percent`` of the frame's total, 0 without energy
    (``SpectralDescriptors.swift:228-270``)."""
    S = stft_magnitude(y, n_fft, hop, win_length, center)
    f = _freqs(S.shape[0], sr, n_fft)
    tot = S.sum(axis=0)
    cum = np.cumsum(S, axis=0)
    out = np.zeros(S.shape[1])
    for t in range(S.shape[1]):
        if tot[t] > 0:
            out[t] = f[int(np.argmax(cum[:, t] >= roll_percent * tot[t]))]
    return out[None, :]


def spectral_flatness(y, n_fft=2048, hop=512, win_length=None, center=True, power=2.0, amin=1e-10) -> np.ndarray:
    """Geometric over arithmetic mean of ``max(S^power, amin)`` per frame, 0
    for a frame whose magnitude sum is below ``amin``
    (``SpectralDescriptors.swift:295-340``)."""
    S = stft_magnitude(y, n_fft, hop, win_length, center)
    P = np.maximum(S ** power, amin)
    am = P.mean(axis=0)
    gm = np.exp(np.log(P).mean(axis=0))
    out = np.where((S.sum(axis=0) >= amin) & (am > amin), gm / am, 0.0)
    return out[None, :]


def _round_f32(x):
    """Swift's ``roundf`` (half away from zero) on a float32 value >= 0."""
    return int(np.floor(np.float32(x) + np.float32(0.5)))


def contrast_bands(sr: int, n_fft: int, n_freqs: int, n_bands: int = 6, fmin: float = 200.0,
                   quantile: float = 0.02):
    """Bin edges and per-band quantile counts of spectral contrast, in float32
    as the reference forms them (``SpectralDescriptors.swift:142-158, 180-181``):
    edges 0, fmin 2^(b-1) for b = 1..n_bands, sr/2, each ``min(round(edge /
    (sr / n_fft)), n_freqs)``; a band of ``size`` bins averages
    ``max(1, round(quantile * size))`` values at each end."""
    f32 = np.float32
    res = f32(sr) / f32(n_fft)
    hz = [f32(0)] + [f32(fmin) * f32(2.0) ** f32(b - 1) for b in range(1, n_bands + 1)] + [f32(sr) / f32(2)]
    edges = [min(_round_f32(f32(h) / res), n_freqs) for h in hz]
    q = [max(1, _round_f32(f32(quantile) * f32(e1 - e0))) if e1 > e0 else 0 for e0, e1 in zip(edges[:-1], edges[1:])]
    return edges, q


def contrast_from_magnitude(S, sr: int, n_fft: int, n_bands: int = 6, fmin: float = 200.0,
                            quantile: float = 0.02) -> np.ndarray:
    """Spectral contrast of a magnitude spectrogram ``[F, T]``
    (``SpectralDescriptors.swift:137-204``): per band and frame, the bins
    sorted ascending, valley = mean of the lowest q, peak = mean of the
    highest q, output ``10 log10(max(peak, 1e-10)) - 10 log10(max(valley,
    1e-10))``; an empty band gives 0. ``[n_bands + 1, T]``."""
    S = np.asarray(S, dtype=np.float64)
    F, T = S.shape
    edges, q = contrast_bands(sr, n_fft, F, n_bands, fmin, quantile)
    out = np.zeros((n_bands + 1, T))
    for b in range(n_bands + 1):
        e0, e1 = edges[b], edges[b + 1]
        if e1 <= e0:
            continue
        srt = np.sort(S[e0:e1], axis=0)
        valley = srt[:q[b]].mean(axis=0)
        peak = srt[e1 - e0 - q[b]:].mean(axis=0)
        out[b] = 10 * np.log10(np.maximum(peak, 1e-10)) - 10 * np.log10(np.maximum(valley, 1e-10))
    return out


def spectral_contrast(y, sr=22050, n_fft=2048, hop=512, win_length=None, center=True, n_bands=6,
                      fmin=200.0, quantile=0.02) -> np.ndarray:
    """Spectral contrast of a signal: ``contrast_from_magnitude`` of the
    magnitude STFT (``Bridge.swift:698-737``)."""
    return contrast_from_magnitude(stft_magnitude(y, n_fft, hop, win_length, center), sr, n_fft, n_bands,
                                   fmin, quantile)


PINS = {
    "hann_periodic": "known answer hann(4) = [0, .5, 1, .5] (WindowTests.swift:60-66); closed form",
    "prepare_window": "centre zero-pad / centre truncation indices checked by hand on small cases",
    "frame_count": "22050 -> 44 (STFTTests.swift:8-24), 661500 -> 1292, empty -> 1/0 (EdgeCaseTests.swift:69-81)",
    "stft_complex": "golden stft_440hz_default_magnitude.npy; O(N^2) brute-force DFT for n_fft <= 256; Parseval",
    "istft": "round trip istft(stft_complex(y)) = y (centre, no centre, several hops and window lengths); "
             "O(N^2) brute-force inverse DFT + overlap-add on small cases; length / trim rules on hand cases "
             "(ISTFTTests.swift:66-89); linearity",
    "mel_filterbank": "goldens mel_filterbank_{128_2048,40_1024,64_2048_300_8000}.npy",
    "melspectrogram": "goldens mel_spectrogram_440hz_*.npy",
    "power_to_db": "golden stft_440hz_power_db.npy; 0.5 -> -3.0103 dB, 100 -> 20 dB (ScalingTests.swift:39-56)",
    "amplitude_to_db": "goldens stft_440hz_amplitude_db{,_ref1}.npy",
    "dct_ortho_basis": "scipy.fft.dct(type=2, norm='ortho'); orthonormality",
    "mfcc": "goldens mfcc_440hz_*.npy",
    "chroma_filterbank": "through the chroma goldens; unit column norms before octave weighting; roll puts C first",
    "chroma_stft": "goldens chroma_stft_440hz_{no_norm,default,1024}.npy (librosa.feature.chroma_stft)",
    "normalize_frames": "L1 / L2 / Linf norms of the columns equal 1; zero columns unchanged",
    "spectral_centroid": "golden spectral_centroid_440hz_default.npy; pure tone -> its frequency; silence -> 0",
    "spectral_bandwidth": "golden spectral_bandwidth_440hz_default.npy; a single bin has bandwidth 0",
    "spectral_rolloff": "golden spectral_rolloff_440hz_default.npy; brute-force cumulative search",
    "spectral_flatness": "golden spectral_flatness_440hz_default.npy; white noise near 1, tone near 0, flat -> 1",
    "contrast_bands": "edges and counts worked by hand for sr 22050, n_fft 2048 and 4096 (round half away)",
    "contrast_from_magnitude": "hand cases (a band holding 1..100 gives 10 log10(99.5 / 1.5)); a flat band gives 0; "
                               "bin order within a band does not matter; planted mistakes (peak and valley "
                               "swapped, q rounded down) fail",
    "spectral_contrast": "golden spectral_contrast_440hz_default.npy (librosa's algorithm) at the reference "
                         "test's tolerance (rtol 0.5, atol 15 dB, test_parity_spectral.py:158-172)",
    "onset_strength": "golden onset_strength_440hz_default.npy (librosa.onset.onset_strength); silence -> 0; "
                      "non-negative; shape / padding rules; planted mistakes (no rectification, pad width) fail",
}

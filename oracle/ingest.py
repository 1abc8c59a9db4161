"""Float64 host oracle for audio ingest: sample-rate conversion and PCM decoding.

TEST INFRASTRUCTURE ONLY. Nothing in the shipped library imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use it, as the checker / the timed CPU restatement.

``resample`` restates ``Sources/MetalMomCore/Audio/Resample.swift:12-96`` (the
algorithm behind ``mm_resample``, ``Bridge.swift:1087-1104``). The PCM helpers
restate the sample conversions a WAV reader applies (integer PCM scaled by
2^-(bits-1), unsigned 8-bit centred at 128, IEEE float unchanged); the
reference reads files through AVFoundation (``AudioIO.swift:16-131``), whose
decoder is not part of its sources, so they are pinned by the WAV format
itself (scipy.io.wavfile round trips) instead. How each function is pinned is
listed in ``PINS`` and exercised by ``tests/test_oracle_ingest.py``.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["NUM_ZEROS", "TABLE_PRECISION", "KAISER_BETA", "bessel_i0", "resample_table", "resample_length",
           "resample", "resample_direct", "pcm_to_float", "to_mono"]

NUM_ZEROS = 64            # zero crossings on each side of the sinc (Resample.swift:32)
TABLE_PRECISION = 9       # log2 of the table oversampling (Resample.swift:33)
KAISER_BETA = 6.0         # Resample.swift:41


def bessel_i0(x: float) -> float:
    """Modified Bessel function I0 by its power series, stopping once a term
    falls below 1e-20 of the sum or after 49 terms (Resample.swift:100-112)."""
    s, term, half = 1.0, 1.0, x / 2.0
    for k in range(1, 50):
        term *= half / k
        t2 = term * term
        s += t2
        if t2 < s * 1e-20:
            break
    return s


def resample_table() -> np.ndarray:
    """Kaiser-windowed sinc sampled at x = i / 512 for i in [0, 64 * 512]
    (Resample.swift:35-62); float64 here, rounded to float32 by the reference."""
    over = 1 << TABLE_PRECISION
    n = NUM_ZEROS * over + 1
    x = np.arange(n, dtype=np.float64) / over
    with np.errstate(invalid="ignore", divide="ignore"):
        sinc = np.where(x == 0, 1.0, np.sin(np.pi * x) / (np.pi * x))
    t = x / NUM_ZEROS
    i0b = bessel_i0(KAISER_BETA)
    kaiser = np.array([bessel_i0(KAISER_BETA * math.sqrt(1.0 - v * v)) / i0b if v < 1.0 else 0.0 for v in t])
    return sinc * kaiser


def resample_length(n: int, source_sr: int, target_sr: int) -> int:
    """``ceil(n * target / source)`` in double precision (Resample.swift:21-22);
    n when the rates agree, 0 for an empty signal."""
    if source_sr == target_sr:
        return int(n)
    if n <= 0:
        return 0
    return int(math.ceil(float(n) * (float(target_sr) / float(source_sr))))


def resample(y, source_sr: int, target_sr: int, table: np.ndarray | None = None,
             chunk: int = 4096, _plant: str | None = None) -> np.ndarray:
    """Windowed-sinc sample-rate conversion (Resample.swift:12-96).

    Output sample i sits at input position ``p = i / ratio``; it sums
    ``x[j] * w(|j - p| * scale * 512)`` over ``j`` in
    ``[floor(p) - support + 1, floor(p) + support]`` clipped to the signal,
    where ``w`` interpolates the table linearly and taps whose table index
    reaches the last entry are skipped; ``scale = min(1, ratio)`` stretches
    the filter when downsampling (anti-aliasing), ``support = ceil(64 /
    scale)``, and the sum is multiplied by ``scale``. An output whose weights
    sum to zero or less stays 0. Equal rates return the input unchanged.

    ``_plant`` injects a known mistake for the tests that must catch it:
    "no_scale" (output not multiplied by scale), "unstretched" (filter not
    widened when downsampling).
    """
    y = np.asarray(y, dtype=np.float64)
    if source_sr == target_sr:
        return y.copy()
    n = y.size
    n_out = resample_length(n, source_sr, target_sr)
    if n_out <= 0:
        return np.zeros(0)
    T = resample_table() if table is None else np.asarray(table, dtype=np.float64)
    flen = T.size
    ratio = float(target_sr) / float(source_sr)
    scale = min(1.0, ratio)
    fscale = (1.0 if _plant == "unstretched" else scale) * (1 << TABLE_PRECISION)
    support = int(math.ceil(NUM_ZEROS / scale))
    gain = 1.0 if _plant == "no_scale" else scale
    out = np.zeros(n_out)
    offs = np.arange(-support + 1, support + 1)
    for c0 in range(0, n_out, chunk):
        i = np.arange(c0, min(n_out, c0 + chunk))
        pos = i / ratio
        idx = pos.astype(np.int64)
        j = idx[:, None] + offs[None, :]
        valid = (j >= 0) & (j <= n - 1)
        fp = np.abs(j - pos[:, None]) * fscale
        fi = fp.astype(np.int64)
        use = valid & (fi < flen - 1)
        fi = np.where(use, fi, 0)
        fr = fp - fi
        w = np.where(use, T[fi] * (1.0 - fr) + T[fi + 1] * fr, 0.0)
        xs = y[np.clip(j, 0, n - 1)]
        s = (xs * w).sum(axis=1)
        ws = w.sum(axis=1)
        out[i] = np.where(ws > 0, s * gain, 0.0)
    return out


def resample_direct(y, source_sr: int, target_sr: int) -> np.ndarray:
    """The same sum with the filter evaluated in closed form (sinc times the
    Kaiser window at the exact tap distance) instead of the interpolated
    table: an independent check of ``resample``'s indexing, differing from it
    only by the table's interpolation error. O(n_out * support) Python loops
    over outputs: small inputs only."""
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    n_out = resample_length(n, source_sr, target_sr)
    ratio = float(target_sr) / float(source_sr)
    scale = min(1.0, ratio)
    support = int(math.ceil(NUM_ZEROS / scale))
    i0b = bessel_i0(KAISER_BETA)
    out = np.zeros(n_out)
    for i in range(n_out):
        p = i / ratio
        k = int(p)
        j = np.arange(max(0, k - support + 1), min(n - 1, k + support) + 1)
        x = np.abs(j - p) * scale
        t = x / NUM_ZEROS
        keep = x * (1 << TABLE_PRECISION) < NUM_ZEROS * (1 << TABLE_PRECISION)
        with np.errstate(invalid="ignore", divide="ignore"):
            sinc = np.where(x == 0, 1.0, np.sin(np.pi * x) / (np.pi * x))
        kai = np.array([bessel_i0(KAISER_BETA * math.sqrt(max(0.0, 1.0 - v * v))) / i0b for v in t])
        w = np.where(keep, sinc * kai, 0.0)
        if w.sum() > 0:
            out[i] = (y[j] * w).sum() * scale
    return out


def pcm_to_float(raw: bytes, bits: int, is_float: bool) -> np.ndarray:
    """Little-endian interleaved PCM samples -> float64: signed integers of
    16/24/32 bits divided by 2^(bits-1), unsigned 8-bit as (u - 128) / 128,
    IEEE float32 / float64 unchanged."""
    b = np.frombuffer(raw, dtype=np.uint8)
    if is_float:
        return np.frombuffer(raw, dtype="<f4" if bits == 32 else "<f8").astype(np.float64)
    if bits == 8:
        return (b.astype(np.float64) - 128.0) / 128.0
    if bits == 16:
        return np.frombuffer(raw, dtype="<i2").astype(np.float64) / 32768.0
    if bits == 24:
        t = b.reshape(-1, 3).astype(np.int32)
        v = t[:, 0] | (t[:, 1] << 8) | (t[:, 2] << 16)
        v = np.where(v >= 1 << 23, v - (1 << 24), v)
        return v.astype(np.float64) / 8388608.0
    if bits == 32:
        return np.frombuffer(raw, dtype="<i4").astype(np.float64) / 2147483648.0
    raise ValueError(f"unsupported PCM width {bits}")


def to_mono(x: np.ndarray, channels: int, mono: bool) -> np.ndarray:
    """Interleaved frames -> one channel: the mean of the channels when
    ``mono``, else the first channel (the reference returns channel 0 of a
    multichannel read, AudioIO.swift:115-117, 126-128)."""
    f = np.asarray(x, dtype=np.float64).reshape(-1, channels)
    return f.mean(axis=1) if mono else f[:, 0].copy()


PINS = {
    "bessel_i0": "scipy.special.i0 at several points (1e-13)",
    "resample_table": "closed form sinc * Kaiser(beta 6) at the table points; ends 1 and 0",
    "resample_length": "ResampleTests.swift:12-35 / test_resample.py lengths (22050 <-> 44100, +-1)",
    "resample": "identity, lengths, round-trip SNR > 40 dB, RMS within 0.05 (ResampleTests.swift, test_resample.py); "
                "resample_direct (closed-form filter) on small inputs; DC gain 1 in the interior; a tone above "
                "the new Nyquist attenuated > 40 dB; planted mistakes (no scale factor, unstretched filter) fail",
    "pcm_to_float": "scipy.io.wavfile round trips for u8 / s16 / s24 (hand-packed) / s32 / f32 / f64",
    "to_mono": "mean of channels for mono, channel 0 otherwise, on hand cases",
}

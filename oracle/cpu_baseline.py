"""Float32 NumPy restatement of the reference's CPU path, used as the CPU baseline.

TEST / BENCH INFRASTRUCTURE ONLY (the ``cpu_baseline`` leg of bench.py). The
reference's own CPU arithmetic is Apple Accelerate (macOS only), so this is a
port of its semantics in float32:

* zero centre padding and ``1 + (len_padded - n_fft)//hop`` frames
  (``Spectral/STFT.swift:240-254``);
* periodic Hann (``Filters/Windows.swift:14-48``);
* per-frame real FFT and ``re^2 + im^2`` (``STFT.swift:287-329``), here
  ``numpy.fft.rfft`` on float32 frames (numpy >= 2 keeps float32);
* dense ``FB @ P`` as ``vDSP_mmul`` does (``Spectral/MelSpectrogram.swift:129-136``);
* ``10 log10(max(x, 1e-10))`` and the whole-clip ``top_db`` floor
  (``Spectral/Scaling.swift:73-114``);
* DCT-II ortho basis product (``Features/MFCC.swift:79-106``);
* inverse STFT (``STFT.swift:686-861``): per-frame inverse real FFT
  (``numpy.fft.irfft`` on complex64 columns), synthesis window, overlap-add
  and division by the summed squared window where it exceeds 1e-10, centre
  trim and length.
"""
from __future__ import annotations

import numpy as np

from . import spectral as o


class MfccF32:
    def __init__(self, sr=22050, n_fft=2048, hop=512, n_mels=128, n_mfcc=13, top_db=80.0):
        self.sr, self.n_fft, self.hop, self.top_db = sr, n_fft, hop, top_db
        self.win = o.hann_periodic(n_fft).astype(np.float32)
        self.fb = o.mel_filterbank(sr, n_fft, n_mels).astype(np.float32)
        self.dct = o.dct_ortho_basis(min(n_mfcc, n_mels), n_mels).astype(np.float32)

    def __call__(self, y: np.ndarray) -> np.ndarray:
        y = np.asarray(y, dtype=np.float32)
        p = self.n_fft // 2
        x = np.pad(y, (p, p))
        T = 1 + (len(x) - self.n_fft) // self.hop
        frames = np.lib.stride_tricks.as_strided(
            x, shape=(T, self.n_fft), strides=(x.strides[0] * self.hop, x.strides[0]))
        X = np.fft.rfft(frames * self.win, axis=1)
        P = (X.real * X.real + X.imag * X.imag).astype(np.float32)
        mel = self.fb @ P.T
        db = 10.0 * np.log10(np.maximum(mel, np.float32(1e-10)))
        if self.top_db is not None:
            db = np.maximum(db, db.max() - np.float32(self.top_db))
        return self.dct @ db


class IstftF32:
    """Float32 inverse STFT with a hop that divides n_fft (the overlap-add runs
    as n_fft/hop shifted block sums)."""

    def __init__(self, n_fft=2048, hop=512):
        assert n_fft % hop == 0
        self.n_fft, self.hop = n_fft, hop
        self.win = o.hann_periodic(n_fft).astype(np.float32)

    def __call__(self, S: np.ndarray, length=None) -> np.ndarray:
        N, hop = self.n_fft, self.hop
        T = S.shape[1]
        R = N // hop
        frames = np.fft.irfft(np.ascontiguousarray(S.T), n=N, axis=1).astype(np.float32) * self.win
        total = N + (T - 1) * hop
        out = np.zeros((total // hop, hop), dtype=np.float32)
        ws = np.zeros_like(out)
        w2 = (self.win * self.win).reshape(R, hop)
        fb = frames.reshape(T, R, hop)
        for b in range(R):
            out[b:b + T] += fb[:, b, :]
            ws[b:b + T] += w2[b]
        out, ws = out.reshape(-1), ws.reshape(-1)
        nz = ws > np.float32(1e-10)
        out[nz] /= ws[nz]
        start = N // 2
        n = length if length is not None else total - N
        res = np.zeros(n, dtype=np.float32)
        m = max(min(n, total - start), 0)
        res[:m] = out[start:start + m]
        return res


class OnsetF32:
    """Float32 onset strength (``Rhythm/OnsetDetection.swift:23-128``): power
    mel spectrogram, whole-matrix ``power_to_db(top_db 80)``, lag-1 half-wave
    rectified difference averaged over the bands, centre padding."""

    def __init__(self, sr=22050, n_fft=2048, hop=512, n_mels=128):
        self.n_fft, self.hop = n_fft, hop
        self.win = o.hann_periodic(n_fft).astype(np.float32)
        self.fb = o.mel_filterbank(sr, n_fft, n_mels).astype(np.float32)

    def __call__(self, y: np.ndarray) -> np.ndarray:
        y = np.asarray(y, dtype=np.float32)
        p = self.n_fft // 2
        x = np.pad(y, (p, p))
        T = 1 + (len(x) - self.n_fft) // self.hop
        frames = np.lib.stride_tricks.as_strided(
            x, shape=(T, self.n_fft), strides=(x.strides[0] * self.hop, x.strides[0]))
        X = np.fft.rfft(frames * self.win, axis=1)
        P = (X.real * X.real + X.imag * X.imag).astype(np.float32)
        db = 10.0 * np.log10(np.maximum(self.fb @ P.T, np.float32(1e-10)))
        db = np.maximum(db, db.max() - np.float32(80.0))
        flux = np.maximum(np.float32(0), db[:, 1:] - db[:, :-1]).mean(axis=0)
        pad = 1 + self.n_fft // (2 * self.hop)
        return np.concatenate([np.zeros(pad, np.float32), flux])[:T]


def _magnitude_f32(y, n_fft, hop, win):
    y = np.asarray(y, dtype=np.float32)
    p = n_fft // 2
    x = np.pad(y, (p, p))
    T = 1 + (len(x) - n_fft) // hop
    frames = np.lib.stride_tricks.as_strided(x, shape=(T, n_fft), strides=(x.strides[0] * hop, x.strides[0]))
    X = np.fft.rfft(frames * win, axis=1)
    return np.abs(X).astype(np.float32).T            # [F, T]


class ChromaF32:
    """Float32 STFT chroma (``Features/Chroma.swift:30-156``): power spectrogram
    times the chroma filterbank (``vDSP_mmul``), no frame normalisation."""

    def __init__(self, sr=22050, n_fft=2048, hop=512, n_chroma=12):
        self.n_fft, self.hop = n_fft, hop
        self.win = o.hann_periodic(n_fft).astype(np.float32)
        self.fb = o.chroma_filterbank(sr, n_fft, n_chroma).astype(np.float32)

    def __call__(self, y):
        S = _magnitude_f32(y, self.n_fft, self.hop, self.win)
        return self.fb @ (S * S)


class ResampleF32:
    """Float32 restatement of the windowed-sinc converter
    (``Audio/Resample.swift:12-96``): tap positions in float64 as the
    reference forms them, float32 table, weights and sums, vectorised over
    chunks of outputs."""

    def __init__(self, source_sr=44100, target_sr=22050, chunk=8192):
        from . import ingest as g
        self.src, self.dst, self.chunk = source_sr, target_sr, chunk
        self.T = g.resample_table().astype(np.float32)
        self.ratio = target_sr / source_sr
        self.scale = min(1.0, self.ratio)
        self.support = int(np.ceil(g.NUM_ZEROS / self.scale))
        self.n_out = g.resample_length

    def __call__(self, y):
        y = np.asarray(y, dtype=np.float32)
        n = y.size
        m = self.n_out(n, self.src, self.dst)
        out = np.zeros(m, np.float32)
        offs = np.arange(-self.support + 1, self.support + 1)
        fs = self.scale * 512.0
        for c0 in range(0, m, self.chunk):
            i = np.arange(c0, min(m, c0 + self.chunk))
            pos = i / self.ratio
            j = pos.astype(np.int64)[:, None] + offs[None, :]
            fp = np.abs(j - pos[:, None]) * fs
            fi = fp.astype(np.int64)
            use = (j >= 0) & (j < n) & (fi < self.T.size - 1)
            fi = np.where(use, fi, 0)
            fr = (fp - fi).astype(np.float32)
            w = np.where(use, self.T[fi] * (np.float32(1) - fr) + self.T[fi + 1] * fr, np.float32(0))
            s = (y[np.clip(j, 0, n - 1)] * w).sum(axis=1, dtype=np.float32)
            ws = w.sum(axis=1, dtype=np.float32)
            out[i] = np.where(ws > 0, s * np.float32(self.scale), np.float32(0))
        return out


class ContrastF32:
    """Float32 spectral contrast (``SpectralDescriptors.swift:137-204``) of the
    magnitude STFT: per band, sort, mean of the q lowest and highest, dB
    difference."""

    def __init__(self, sr=22050, n_fft=2048, hop=512, n_bands=6, fmin=200.0):
        self.n_fft, self.hop = n_fft, hop
        self.win = o.hann_periodic(n_fft).astype(np.float32)
        self.edges, self.q = o.contrast_bands(sr, n_fft, n_fft // 2 + 1, n_bands, fmin)

    def __call__(self, y):
        S = _magnitude_f32(y, self.n_fft, self.hop, self.win)
        out = np.zeros((len(self.q), S.shape[1]), np.float32)
        for b, q in enumerate(self.q):
            e0, e1 = self.edges[b], self.edges[b + 1]
            if e1 <= e0:
                continue
            srt = np.sort(S[e0:e1], axis=0)
            valley = srt[:q].mean(axis=0)
            peak = srt[e1 - e0 - q:].mean(axis=0)
            out[b] = 10 * np.log10(np.maximum(peak, 1e-10)) - 10 * np.log10(np.maximum(valley, 1e-10))
        return out


class CentroidF32:
    """Float32 spectral centroid (``SpectralDescriptors.swift:25-56``) of the
    magnitude STFT."""

    def __init__(self, sr=22050, n_fft=2048, hop=512):
        self.n_fft, self.hop = n_fft, hop
        self.win = o.hann_periodic(n_fft).astype(np.float32)
        self.f = (np.arange(n_fft // 2 + 1, dtype=np.float32) * np.float32(sr) / np.float32(n_fft))

    def __call__(self, y):
        S = _magnitude_f32(y, self.n_fft, self.hop, self.win)
        tot = S.sum(axis=0)
        return np.where(tot > 0, (self.f[:, None] * S).sum(axis=0) / np.where(tot > 0, tot, 1), 0)

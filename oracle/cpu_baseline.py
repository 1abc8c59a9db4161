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
* DCT-II ortho basis product (``Features/MFCC.swift:79-106``).
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

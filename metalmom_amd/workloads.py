"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

plain_clip      30 s mono at 22.05 kHz: sines 220/440/880 Hz (0.5/0.3/0.2) plus
                Gaussian noise sigma 0.01, default_rng(0).
uneven_batch    256 clips, lengths uniform in [1 s, 60 s] from default_rng(1);
                per clip, in clip order from the same generator: white noise
                sigma 0.1 plus a linear chirp 100 Hz -> 8 kHz over the clip,
                amplitude 0.5, phase 2 pi (f0 t + (f1 - f0) t^2 / (2 T))
                (the chirp law of the reference's SignalGen.swift:70-78).
large_fft_batch 1024 clips of 10 s at 44.1 kHz, float16: 0.2 * U(-1, 1) plus 3
                sines of amplitude 0.25 with frequencies U(50 Hz, 10 kHz),
                clipped to [-1, 1], default_rng(3).
"""
from __future__ import annotations

import numpy as np


def plain_clip(n: int = 661500, sr: int = 22050, seed: int = 0) -> np.ndarray:
    t = np.arange(n) / sr
    rng = np.random.default_rng(seed)
    x = (0.5 * np.sin(2 * np.pi * 220 * t) + 0.3 * np.sin(2 * np.pi * 440 * t)
         + 0.2 * np.sin(2 * np.pi * 880 * t) + rng.normal(0.0, 0.01, n))
    return x.astype(np.float32)


def chirp_noise_clip(n: int, sr: int, rng, f0: float = 100.0, f1: float = 8000.0) -> np.ndarray:
    t = np.arange(n, dtype=np.float64) / sr
    dur = max(n / sr, 1e-9)
    phase = 2 * np.pi * (f0 * t + 0.5 * (f1 - f0) / dur * t * t)
    return (rng.normal(0.0, 0.1, n) + 0.5 * np.sin(phase)).astype(np.float32)


def uneven_lengths(n_clips: int = 256, sr: int = 22050, seed: int = 1, lo: float = 1.0, hi: float = 60.0):
    rng = np.random.default_rng(seed)
    return rng, np.round(rng.uniform(lo, hi, n_clips) * sr).astype(np.int64)


def uneven_batch(n_clips: int = 256, sr: int = 22050, seed: int = 1):
    rng, lens = uneven_lengths(n_clips, sr, seed)
    return [chirp_noise_clip(int(n), sr, rng) for n in lens]


def large_fft_batch(n_clips: int = 1024, sr: int = 44100, seconds: float = 10.0, seed: int = 3):
    rng = np.random.default_rng(seed)
    n = int(round(seconds * sr))
    t = np.arange(n) / sr
    out = []
    for _ in range(n_clips):
        x = 0.2 * rng.uniform(-1.0, 1.0, n)
        for f in rng.uniform(50.0, 10000.0, 3):
            x += 0.25 * np.sin(2 * np.pi * f * t)
        out.append(np.clip(x, -1.0, 1.0).astype(np.float16))
    return out

"""GPU parity of STFT chroma (mm_chroma_stft, Chroma.swift:30-275) against the
float64 oracle at tuning 0 (what the reference computes), through the
drop-in Python API.

Tolerance: the contraction of the FP32 power spectrum with the filterbank
accumulates in FP32 over the n_fft/2 + 1 bins: max error / max |chroma|
<= 1e-5; normalised frames the same. Shapes are exact.
"""
import numpy as np
import pytest

from oracle import spectral as o

pytestmark = pytest.mark.gpu
RTOL = 1e-5


def _tones(n, sr, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / sr
    x = 0.05 * rng.standard_normal(n)
    for f in (261.63, 329.63, 392.0, 440.0, 880.0):
        x += 0.2 * np.sin(2 * np.pi * f * t + rng.uniform(0, 6.28))
    return x.astype(np.float32)


def test_chroma_golden_signal(signal_440):
    import metalmom_amd as mm
    y = signal_440.astype(np.float32)
    C = mm.chroma_stft(y=y, sr=22050)
    ref = o.chroma_stft(y.astype(np.float64), 22050)
    assert C.shape == (12, 44) and C.dtype == np.float32
    assert np.abs(C - ref).max() <= RTOL * np.abs(ref).max()
    assert (np.argmax(C[:, 5:-5], axis=0) == 9).all()      # 440 Hz is A


@pytest.mark.parametrize("sr,n,kw", [
    (22050, 60000, {}),
    (22050, 60000, {"n_fft": 1024}),
    (22050, 60000, {"hop_length": 300, "win_length": 1500}),
    (44100, 90000, {"n_chroma": 24}),
    (22050, 60000, {"center": False, "norm": np.inf}),
    (22050, 60000, {"norm": 2.0}),
    (22050, 60000, {"n_fft": 4096, "hop_length": 1024, "norm": 1.0}),
    (22050, 1_500_000, {}),
], ids=["default", "nfft1024", "hop300-win1500", "44k-24bins", "uncentred-linf", "l2", "nfft4096-l1", "long"])
def test_chroma_matches_oracle(sr, n, kw):
    import metalmom_amd as mm
    print(f"chroma sr={sr} n={n} {kw}", flush=True)
    y = _tones(n, sr, n % 89)
    C = mm.chroma_stft(y=y, sr=sr, **kw)
    ref = o.chroma_stft(y.astype(np.float64), sr, n_fft=kw.get("n_fft", 2048), hop=kw.get("hop_length", 512),
                        win_length=kw.get("win_length"), n_chroma=kw.get("n_chroma", 12),
                        center=kw.get("center", True), norm=kw.get("norm"))
    assert C.shape == ref.shape
    assert np.abs(C - ref).max() <= RTOL * np.abs(ref).max()


def test_chroma_non_power_of_two_is_empty():
    import metalmom_amd as mm
    C = mm.chroma_stft(y=np.ones(5000, np.float32), n_fft=1000)
    assert C.shape == (12, 0)


def test_chroma_batch_equals_single_clips_and_bf16_error():
    """kind='chroma' in a device-resident batch: every clip bitwise equal to
    mm_chroma_stft alone (FP32); the BF16 contraction within 1e-2 of the
    frame's largest chroma value (bf16 weights and power, FP32 accumulation)."""
    import torch
    import metalmom_amd as mm
    from metalmom_amd.batch import SpectralBatch
    clips = [_tones(n, 22050, n) for n in (30000, 77777, 5000, 150000)]
    lens = [len(c) for c in clips]
    for prec in ("f32", "bf16"):
        B = SpectralBatch(lens, kind="chroma", sr=22050, n_fft=2048, hop_length=512, n_mels=12, precision=prec)
        d_in = B.pack(clips)
        out = B.empty_output()
        B.run(d_in, out)
        torch.cuda.synchronize()
        for c, r in zip(clips, B.unpack(out)):
            single = mm.chroma_stft(y=c, sr=22050, hop_length=512)
            assert r.shape == single.shape
            if prec == "f32":
                assert np.array_equal(r, single)
            else:
                assert (np.abs(r - single) <= 1e-2 * single.max(axis=0, keepdims=True) + 1e-30).all()
        B.close()


def _sine(f, sr=22050, seconds=1.0):
    i = np.arange(int(sr * seconds), dtype=np.float32)
    return np.sin(np.float32(2 * np.pi * f) * i / np.float32(sr)).astype(np.float32)


@pytest.mark.parametrize("f,pc", [(440.0, 9), (261.63, 0)], ids=["A4", "C4"])
def test_pure_tone_peaks_at_its_pitch_class(f, pc):
    """Known answers of ChromaTests.swift:134-176: one second of a sine, the
    frame-averaged chroma peaks at its pitch class (C = 0, A = 9)."""
    import metalmom_amd as mm
    C = mm.chroma_stft(y=_sine(f), sr=22050)
    assert int(np.argmax(C.mean(axis=1))) == pc


def test_silence_gives_zero_chroma():
    """ChromaTests.swift:180-188: a second of silence, every value 0, with
    and without frame normalisation (a silent frame stays zero)."""
    import metalmom_amd as mm
    y = np.zeros(22050, np.float32)
    for norm in (None, np.inf, 2.0):
        C = mm.chroma_stft(y=y, sr=22050, norm=norm)
        assert C.shape == (12, 44) and not C.any(), norm


def test_l2_normalised_frames_have_unit_norm():
    """ChromaTests.swift:192-209: norm=2 leaves every non-silent frame with
    L2 norm 1 within 1e-5."""
    import metalmom_amd as mm
    C = mm.chroma_stft(y=_sine(440.0), sr=22050, norm=2.0)
    l2 = np.sqrt((C.astype(np.float64) ** 2).sum(axis=0))
    live = l2 > 1e-10
    assert live.all()
    assert np.abs(l2[live] - 1.0).max() <= 1e-5

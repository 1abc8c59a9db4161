"""GPU ingest: mm_load at another rate (raw PCM converted and resampled on the
device) and the server's raw-PCM requests (mm_server_submit_pcm).

mm_load(sr) must equal mm_resample of the native-rate load bit for bit (the
device PCM conversion rounds exactly as the host one, and the same converter
runs), and the oracle within the resampler's tolerance; a PCM request to the
server must equal the same clip converted, resampled and analysed alone.
"""
import math

import numpy as np
import pytest
import scipy.io.wavfile as wavfile

from oracle import ingest as g

pytestmark = pytest.mark.gpu


def test_load_with_sr(tmp_path):
    """test_audio_io.py:35-45."""
    import metalmom_amd as mm
    n = 44100
    t = np.arange(n, dtype=np.float32) / 44100
    wavfile.write(str(tmp_path / "a.wav"), 44100, np.sin(2 * np.pi * 440 * t).astype(np.float32))
    y, sr = mm.load(str(tmp_path / "a.wav"), sr=22050)
    assert sr == 22050 and abs(len(y) - 22050) < 1024


@pytest.mark.parametrize("dtype,channels,src,dst", [(np.int16, 2, 44100, 22050), (np.int32, 1, 48000, 22050),
                                                    (np.float32, 3, 16000, 22050), (np.uint8, 2, 8000, 44100)])
def test_load_resampled_equals_resample_of_native(tmp_path, dtype, channels, src, dst):
    import metalmom_amd as mm
    print(f"load {dtype.__name__} x{channels} {src}->{dst}", flush=True)
    rng = np.random.default_rng(channels + src)
    n = 30000
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        x = rng.integers(info.min // 4, info.max // 4, size=(n, channels)).astype(dtype)
    else:
        x = (0.3 * rng.standard_normal((n, channels))).astype(dtype)
    wavfile.write(str(tmp_path / "f.wav"), src, x if channels > 1 else x[:, 0])
    for mono in (True, False):
        native, sr0 = mm.load(str(tmp_path / "f.wav"), sr=None, mono=mono)
        y, sr = mm.load(str(tmp_path / "f.wav"), sr=dst, mono=mono)
        assert sr0 == src and sr == dst
        assert np.array_equal(y, mm.resample(native, src, dst))
        bits = 8 * np.dtype(dtype).itemsize
        ref = g.resample(g.to_mono(g.pcm_to_float(x.tobytes(), bits, dtype == np.float32), channels, mono)
                         .astype(np.float32).astype(np.float64), src, dst)
        tol = 2e-6 * math.ceil(64 / min(1.0, dst / src)) * float(np.abs(native).max())
        assert np.abs(y - ref).max() <= tol


def _pcm(n, channels, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / 44100.0
    base = 0.4 * np.sin(2 * np.pi * (200 + 3000 * t / t[-1]) * t)
    x = base[:, None] + 0.1 * rng.standard_normal((n, channels))
    return np.clip(x * 32767, -32768, 32767).astype("<i2")


def test_server_pcm_requests_equal_alone():
    import metalmom_amd as mm
    from metalmom_amd.serve import SpectralServer
    from metalmom_amd import _native as N
    srv = SpectralServer(hbm_budget_bytes=2 << 30)
    try:
        cases = []
        for i, (kind, src, dst, ch, n) in enumerate([("mfcc", 44100, 22050, 2, 90000), ("mel", 48000, 22050, 1, 50000),
                                                     ("stft", 22050, 22050, 2, 40000), ("mfcc", 16000, 22050, 2, 30000),
                                                     ("mel", 44100, 44100, 1, 70000)]):
            x = _pcm(n, ch, i)
            p = srv.params(kind, sr=dst)
            rid = srv.submit_pcm(x, ch, N.PCM_S16, src, p, mono=True)
            # the same clip alone: host conversion, GPU resampler, one-clip analysis
            clip = g.to_mono(g.pcm_to_float(x.tobytes(), 16, False), ch, True).astype(np.float32)
            clip = mm.resample(clip, src, dst)
            alone = mm.spectral(clip, kind=kind, sr=dst, n_fft=2048, hop_length=512)
            cases.append((rid, alone, kind, src, dst))
        for rid, alone, kind, src, dst in cases:
            got, _ = srv.wait(rid)
            print(f"server pcm {kind} {src}->{dst}", flush=True)
            assert got.shape == alone.shape and np.array_equal(got, alone)
    finally:
        srv.close()


@pytest.mark.parametrize("fmt,mono", [("u8", True), ("s24", False), ("s32", True), ("f32", False), ("f64", True)])
def test_server_pcm_formats(fmt, mono):
    """Every PCM sample format through the server, mixed down or channel 0,
    against the same clip converted on the host (oracle), resampled and
    analysed alone."""
    import metalmom_amd as mm
    from metalmom_amd.serve import SpectralServer
    from metalmom_amd import _native as N
    rng = np.random.default_rng(len(fmt) + mono)
    n, ch, src, dst = 40000, 3, 44100, 22050
    x = np.clip(0.4 * rng.standard_normal((n, ch)), -0.99, 0.99)
    if fmt == "u8":
        raw, bits, flt, code = np.round(x * 127 + 128).astype(np.uint8).tobytes(), 8, False, N.PCM_U8
    elif fmt == "s24":
        v = np.round(x * 8388607).astype(np.int64).ravel()
        raw = b"".join(int(a & 0xFFFFFF).to_bytes(3, "little") for a in v)
        bits, flt, code = 24, False, N.PCM_S24
    elif fmt == "s32":
        raw, bits, flt, code = np.round(x * 2147483647).astype("<i4").tobytes(), 32, False, N.PCM_S32
    elif fmt == "f32":
        raw, bits, flt, code = x.astype("<f4").tobytes(), 32, True, N.PCM_F32
    else:
        raw, bits, flt, code = x.astype("<f8").tobytes(), 64, True, N.PCM_F64
    srv = SpectralServer(hbm_budget_bytes=1 << 30)
    try:
        p = srv.params("mel", sr=dst)
        rid = srv.submit_pcm(np.frombuffer(raw, np.uint8), ch, code, src, p, mono=mono)
        got, _ = srv.wait(rid)
    finally:
        srv.close()
    clip = g.to_mono(g.pcm_to_float(raw, bits, flt), ch, mono).astype(np.float32)
    alone = mm.spectral(mm.resample(clip, src, dst), kind="mel", sr=dst, n_fft=2048, hop_length=512)
    assert np.array_equal(got, alone)


def test_load_slice_then_resample(tmp_path):
    """offset / duration pick the frames first (AudioIO.swift:31-45); the
    slice is then converted and resampled on the GPU."""
    import metalmom_amd as mm
    rng = np.random.default_rng(12)
    x = rng.integers(-20000, 20000, size=(96000, 2)).astype(np.int16)
    wavfile.write(str(tmp_path / "s.wav"), 48000, x)
    native, _ = mm.load(str(tmp_path / "s.wav"), sr=None, offset=0.3, duration=0.75)
    assert native.size == int(0.75 * 48000)
    y, sr = mm.load(str(tmp_path / "s.wav"), sr=22050, offset=0.3, duration=0.75)
    assert sr == 22050 and np.array_equal(y, mm.resample(native, 48000, 22050))

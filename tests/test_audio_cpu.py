"""WAV ingest without a GPU: mm_load at the file's own rate, mm_get_duration
and mm_get_sample_rate (Bridge.swift:1008-1085, AudioIO.swift:16-147).

The reference decodes through AVFoundation, whose decoder is not in its
sources; the sample values are pinned by the WAV format (oracle/ingest.py,
checked against scipy.io.wavfile) and must match bit for bit. The
reference's own load tests (test_audio_io.py) run here for every case that
keeps the native rate.
"""
import os
import struct

import numpy as np
import pytest
import scipy.io.wavfile as wavfile

from oracle import ingest as g


def _write_wav(path, sr, channels, tag, bits, container, data, extensible=False, extra=b"", data_size=None):
    """Hand-written RIFF/WAVE file (formats scipy cannot write)."""
    block = channels * container
    if extensible:
        fmt = struct.pack("<HHIIHHHHI", 0xFFFE, channels, sr, sr * block, block, 8 * container, 22, bits, 0)
        fmt += struct.pack("<H", tag) + b"\x00\x00\x00\x00\x10\x00\x80\x00\x00\xaa\x00\x38\x9b\x71"
    else:
        fmt = struct.pack("<HHIIHH", tag, channels, sr, sr * block, block, 8 * container)
    body = b"WAVE" + b"fmt " + struct.pack("<I", len(fmt)) + fmt + extra
    n = len(data) if data_size is None else data_size
    body += b"data" + struct.pack("<I", n) + data
    with open(path, "wb") as f:
        f.write(b"RIFF" + struct.pack("<I", len(body)) + body)


def _expect(raw, bits, flt, channels, mono):
    return g.to_mono(g.pcm_to_float(raw, bits, flt), channels, mono).astype(np.float32)


def _load(path, **kw):
    import metalmom_amd as mm
    return mm.load(str(path), **kw)


# -- the reference's test_audio_io.py, native-rate cases ----------------------

def _tone_wav(path, sr=22050, duration=1.0, freq=440.0):
    n = int(sr * duration)
    t = np.arange(n, dtype=np.float32) / sr
    audio = np.sin(2 * np.pi * freq * t).astype(np.float32)
    wavfile.write(str(path), sr, audio)
    return audio


def test_load_wav(tmp_path):
    ref = _tone_wav(tmp_path / "a.wav")
    y, sr = _load(tmp_path / "a.wav", sr=None)
    assert sr == 22050 and y.dtype == np.float32
    assert np.array_equal(y, ref)


def test_load_with_offset_and_duration(tmp_path):
    ref = _tone_wav(tmp_path / "a.wav", duration=3.0)
    y, sr = _load(tmp_path / "a.wav", sr=None, offset=1.0)
    assert sr == 22050 and np.array_equal(y, ref[22050:])
    y, sr = _load(tmp_path / "a.wav", sr=None, duration=1.0)
    assert np.array_equal(y, ref[:22050])
    y, _ = _load(tmp_path / "a.wav", sr=None, offset=0.5, duration=0.25)
    assert np.array_equal(y, ref[11025:11025 + 5512])      # trunc(0.25 * 22050)
    y, sr = _load(tmp_path / "a.wav", sr=None, offset=5.0)  # past the end: empty (AudioIO.swift:34-36)
    assert y.shape == (0,) and sr == 22050


def test_get_duration_and_samplerate(tmp_path):
    import metalmom_amd as mm
    _tone_wav(tmp_path / "a.wav", sr=22050, duration=2.0)
    _tone_wav(tmp_path / "b.wav", sr=44100, duration=1.0)
    assert abs(mm.get_duration(str(tmp_path / "a.wav")) - 2.0) < 0.01
    assert mm.get_samplerate(str(tmp_path / "b.wav")) == 44100


def test_missing_and_invalid_files(tmp_path):
    import metalmom_amd as mm
    for fn in (lambda p: mm.load(p), mm.get_duration, mm.get_samplerate):
        with pytest.raises(RuntimeError):
            fn("/nonexistent/file.wav")
    (tmp_path / "x.wav").write_bytes(b"not a wav file at all")
    with pytest.raises(RuntimeError):
        mm.load(str(tmp_path / "x.wav"), sr=None)
    _tone_wav(tmp_path / "a.wav")
    with pytest.raises(RuntimeError):
        mm.load(str(tmp_path / "a.wav"), sr=None, offset=-1.0)
    # a-law (format tag 6) is not decoded
    _write_wav(tmp_path / "alaw.wav", 8000, 1, 6, 8, 1, bytes(100))
    with pytest.raises(RuntimeError):
        mm.load(str(tmp_path / "alaw.wav"), sr=None)


# -- sample formats, channels, chunk layout -----------------------------------

@pytest.mark.parametrize("dtype,bits,flt", [(np.uint8, 8, False), (np.int16, 16, False), (np.int32, 32, False),
                                            (np.float32, 32, True), (np.float64, 64, True)])
@pytest.mark.parametrize("channels", [1, 2, 3])
def test_formats_and_channels(tmp_path, dtype, bits, flt, channels):
    rng = np.random.default_rng(bits * 10 + channels)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        x = rng.integers(info.min, info.max, size=(1000, channels), endpoint=True).astype(dtype)
    else:
        x = rng.uniform(-1, 1, size=(1000, channels)).astype(dtype)
    wavfile.write(str(tmp_path / "f.wav"), 16000, x if channels > 1 else x[:, 0])
    raw = x.tobytes()
    for mono in (True, False):
        y, sr = _load(tmp_path / "f.wav", sr=None, mono=mono)
        assert sr == 16000
        assert np.array_equal(y, _expect(raw, bits, flt, channels, mono))


def test_pcm24_and_extensible(tmp_path):
    rng = np.random.default_rng(24)
    v = rng.integers(-(1 << 23), 1 << 23, size=600)
    raw = b"".join(int(a & 0xFFFFFF).to_bytes(3, "little") for a in v)
    _write_wav(tmp_path / "s24.wav", 48000, 2, 1, 24, 3, raw)
    y, sr = _load(tmp_path / "s24.wav", sr=None)
    assert sr == 48000 and np.array_equal(y, _expect(raw, 24, False, 2, True))
    # 24 valid bits left-justified in 32-bit containers, WAVE_FORMAT_EXTENSIBLE
    raw32 = (v.astype(np.int64) << 8).astype("<i4").tobytes()
    _write_wav(tmp_path / "ext.wav", 48000, 1, 1, 24, 4, raw32, extensible=True)
    y, _ = _load(tmp_path / "ext.wav", sr=None)
    assert np.array_equal(y, (v / 8388608.0).astype(np.float32))
    rawf = rng.uniform(-1, 1, 500).astype("<f4").tobytes()
    _write_wav(tmp_path / "extf.wav", 8000, 1, 3, 32, 4, rawf, extensible=True)
    y, _ = _load(tmp_path / "extf.wav", sr=None)
    assert np.array_equal(y, np.frombuffer(rawf, "<f4"))


def test_odd_chunks_and_streamed_size(tmp_path):
    """An odd-sized chunk before the data is padded to even; a data size of
    0xFFFFFFFF (streamed writers) means 'to the end of the file'."""
    raw = np.arange(-500, 500, dtype="<i2").tobytes()
    _write_wav(tmp_path / "odd.wav", 22050, 1, 1, 16, 2, raw, extra=b"LIST" + struct.pack("<I", 3) + b"abc\x00")
    y, _ = _load(tmp_path / "odd.wav", sr=None)
    assert np.array_equal(y, (np.arange(-500, 500) / 32768.0).astype(np.float32))
    _write_wav(tmp_path / "stream.wav", 22050, 1, 1, 16, 2, raw, data_size=0xFFFFFFFF)
    y, _ = _load(tmp_path / "stream.wav", sr=None)
    assert y.size == 1000
    import metalmom_amd as mm
    assert abs(mm.get_duration(str(tmp_path / "stream.wav")) - 1000 / 22050) < 1e-6

"""Pins of the ingest oracle (oracle/ingest.py): the sample-rate converter
behind mm_resample (Resample.swift) and the PCM sample conversions.

The resampler is checked against the reference's own tests
(ResampleTests.swift, test_resample.py), against the same sum with the filter
in closed form, and against what the mathematics fixes (unit DC gain, a
stop band above the new Nyquist); planted mistakes fail those checks.
"""
import io
import math

import numpy as np
import pytest
import scipy.io.wavfile as wavfile
import scipy.special

from oracle import ingest as g


def _tone(n, sr, f=440.0):
    return np.sin(np.arange(n, dtype=np.float32) * f * 2 * np.pi / sr).astype(np.float32)


def test_bessel_i0_matches_scipy():
    for x in (0.0, 0.3, 1.0, 3.0, 6.0, 12.0):
        assert abs(g.bessel_i0(x) - scipy.special.i0(x)) <= 1e-13 * scipy.special.i0(x)


def test_table_closed_form():
    T = g.resample_table()
    assert T.size == 64 * 512 + 1
    assert T[0] == 1.0 and T[-1] == 0.0
    k = np.array([1, 256, 511, 512, 700, 5000, 20000, 32767])
    x = k / 512.0
    ref = np.sin(np.pi * x) / (np.pi * x) * scipy.special.i0(6.0 * np.sqrt(1 - (x / 64) ** 2)) / scipy.special.i0(6.0)
    assert np.abs(T[k] - ref).max() < 1e-12
    assert abs(T[512]) < 1e-15 and abs(T[1024]) < 1e-15     # zero crossings of the sinc


def test_lengths():
    """ResampleTests.swift:12-35, test_resample.py:11-20."""
    assert abs(len(g.resample(_tone(44100, 44100), 44100, 22050)) - 22050) <= 1
    assert abs(len(g.resample(_tone(22050, 22050), 22050, 44100)) - 44100) <= 1
    assert g.resample_length(1000, 48000, 22050) == math.ceil(1000 * 22050 / 48000)
    assert g.resample_length(0, 44100, 22050) == 0 and g.resample(np.zeros(0), 44100, 22050).size == 0
    assert g.resample_length(7, 16000, 16000) == 7


def test_identity():
    y = _tone(22050, 22050)
    assert np.array_equal(g.resample(y, 22050, 22050), y)


def test_round_trip_snr_and_energy():
    """test_resample.py:23-44 (> 40 dB round trip, RMS within 0.05)."""
    y = _tone(22050, 22050)
    up = g.resample(y, 22050, 44100)
    rt = g.resample(up, 44100, 22050)
    m = min(len(y), len(rt))
    s = y[1000:m - 1000].astype(np.float64)
    noise = s - rt[1000:m - 1000]
    assert 10 * np.log10(np.sum(s ** 2) / max(np.sum(noise ** 2), 1e-20)) > 40
    assert abs(np.sqrt(np.mean(y.astype(np.float64) ** 2)) - np.sqrt(np.mean(up ** 2))) < 0.05


@pytest.mark.parametrize("src,dst", [(44100, 22050), (22050, 44100), (48000, 22050), (16000, 22050), (44100, 8000)])
def test_table_sum_matches_closed_form_filter(src, dst):
    """The interpolated table differs from the exact filter by at most its
    linear-interpolation error (h^2/8 max|f''| ~ 5e-6 per tap)."""
    rng = np.random.default_rng(src + dst)
    y = rng.standard_normal(400)
    a = g.resample(y, src, dst)
    b = g.resample_direct(y, src, dst)
    assert a.shape == b.shape
    assert np.abs(a - b).max() < 2e-4 * np.abs(y).max()


@pytest.mark.parametrize("src,dst", [(44100, 22050), (22050, 44100), (48000, 22050)])
def test_dc_gain_is_one(src, dst):
    """A constant maps to the same constant away from the edges: the
    filter's taps sum to 1 / scale and the output is multiplied by scale."""
    y = np.ones(6000)
    r = g.resample(y, src, dst)
    inner = r[400:-400]
    assert np.abs(inner - 1.0).max() < 2e-3


def test_stop_band_above_new_nyquist():
    """A 15 kHz tone resampled 44.1 -> 22.05 kHz (new Nyquist 11.025 kHz)
    is attenuated by more than 40 dB; a 2 kHz tone keeps its level."""
    hi = g.resample(_tone(44100, 44100, 15000.0), 44100, 22050)[500:-500]
    lo = g.resample(_tone(44100, 44100, 2000.0), 44100, 22050)[500:-500]
    assert 20 * np.log10(np.sqrt(np.mean(hi ** 2)) / np.sqrt(0.5)) < -40
    assert abs(np.sqrt(np.mean(lo ** 2)) - np.sqrt(0.5)) < 0.01


def test_planted_mistakes_are_caught():
    y = np.ones(6000)
    assert np.abs(g.resample(y, 44100, 22050, _plant="no_scale")[400:-400] - 1.0).max() > 0.5
    hi = g.resample(_tone(44100, 44100, 15000.0), 44100, 22050, _plant="unstretched")[500:-500]
    assert 20 * np.log10(np.sqrt(np.mean(hi ** 2)) / np.sqrt(0.5)) > -40


def _wav_bytes(sr, data):
    b = io.BytesIO()
    wavfile.write(b, sr, data)
    return b.getvalue()


def _data_chunk(raw):
    i = raw.index(b"data")
    n = int.from_bytes(raw[i + 4:i + 8], "little")
    return raw[i + 8:i + 8 + n]


@pytest.mark.parametrize("dtype,bits,flt", [(np.uint8, 8, False), (np.int16, 16, False), (np.int32, 32, False),
                                            (np.float32, 32, True), (np.float64, 64, True)])
def test_pcm_to_float_matches_wav_files(dtype, bits, flt):
    rng = np.random.default_rng(bits)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        x = rng.integers(info.min, info.max, size=(300, 2), endpoint=True).astype(dtype)
    else:
        x = rng.uniform(-1, 1, size=(300, 2)).astype(dtype)
    raw = _data_chunk(_wav_bytes(22050, x))
    got = g.pcm_to_float(raw, bits, flt).reshape(-1, 2)
    if dtype == np.uint8:
        ref = (x.astype(np.float64) - 128) / 128
    elif flt:
        ref = x.astype(np.float64)
    else:
        ref = x.astype(np.float64) / float(2 ** (bits - 1))
    assert np.array_equal(got, ref)
    assert got.min() >= -1.0 and got.max() <= 1.0


def test_pcm24_hand_packed():
    v = np.array([0, 1, -1, 8388607, -8388608, 4660, -4660], dtype=np.int64)
    raw = b"".join(int(x & 0xFFFFFF).to_bytes(3, "little") for x in v)
    assert np.array_equal(g.pcm_to_float(raw, 24, False), v / 8388608.0)


def test_to_mono():
    x = np.array([[1.0, 3.0], [-2.0, 2.0], [0.5, 0.5]]).ravel()
    assert np.array_equal(g.to_mono(x, 2, True), [2.0, 0.0, 0.5])
    assert np.array_equal(g.to_mono(x, 2, False), [1.0, -2.0, 0.5])
    assert np.array_equal(g.to_mono(np.arange(4.0), 1, True), np.arange(4.0))


def test_cpu_baseline_port_matches_oracle():
    """The float32 restatement timed as the CPU baseline computes the same
    conversion (FP32 rounding only)."""
    from oracle.cpu_baseline import ResampleF32
    y = np.random.default_rng(0).standard_normal(20000).astype(np.float32)
    a = ResampleF32(44100, 22050)(y)
    b = g.resample(y.astype(np.float64), 44100, 22050)
    assert a.shape == b.shape and np.abs(a - b).max() < 1e-5

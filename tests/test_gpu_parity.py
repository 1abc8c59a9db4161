"""GPU parity: libmetalmom.so (HIP, gfx950) against the reference's librosa
goldens and the float64 oracle, through the drop-in Python API and the C ABI.

Tolerances: the reference tiers for the goldens (test_parity_*.py: 1e-4, 5e-4
for the fmin/fmax MFCC); against the FP64 oracle the FP32 targets are STFT
magnitude max-abs error / max magnitude <= 1e-5, mel dB max-abs <= 1e-3 dB and
MFCC max-abs error / max |MFCC| <= 1e-4; 16-bit mel contractions <= 0.05 dB.
Integer semantics (frame counts, shapes) are exact.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import spectral as o

pytestmark = pytest.mark.gpu


def mm():
    import metalmom_amd
    return metalmom_amd


def config0_signal(n=661500, sr=22050, seed=0):
    t = np.arange(n) / sr
    rng = np.random.default_rng(seed)
    x = (0.5 * np.sin(2 * np.pi * 220 * t) + 0.3 * np.sin(2 * np.pi * 440 * t)
         + 0.2 * np.sin(2 * np.pi * 880 * t) + rng.normal(0, 0.01, n))
    return x.astype(np.float32)


def rel_err(a, b):
    a = np.asarray(a)
    a = a.astype(np.complex128) if np.iscomplexobj(a) else a.astype(np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


# -- goldens (reference test tiers) ------------------------------------------

def test_stft_golden(signal_440):
    G = load_golden("stft_440hz_default_magnitude")
    R = mm().stft(signal_440, n_fft=2048, hop_length=512, win_length=2048, center=True)
    assert R.shape == G.shape and R.dtype == np.float32
    np.testing.assert_allclose(R, G, rtol=1e-4, atol=1e-4)
    mid = R.shape[1] // 2
    assert np.argmax(R[:, mid]) == np.argmax(G[:, mid])
    assert 0.99 < np.sum(R ** 2) / np.sum(G ** 2) < 1.01


@pytest.mark.parametrize("name,kw", [
    ("mel_spectrogram_440hz_default", {}),
    ("mel_spectrogram_440hz_40_1024", dict(n_fft=1024, n_mels=40)),
    ("mel_spectrogram_440hz_64_fmin300_fmax8000", dict(n_mels=64, fmin=300, fmax=8000)),
])
def test_melspectrogram_golden(signal_440, name, kw):
    G = load_golden(name)
    R = mm().melspectrogram(y=signal_440, sr=22050, **kw)
    assert R.shape == G.shape
    np.testing.assert_allclose(R, G, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("name,kw,tol", [
    ("mfcc_440hz_default", dict(n_mfcc=20), 1e-4),
    ("mfcc_440hz_13_1024", dict(n_mfcc=13, n_fft=1024, n_mels=40), 1e-4),
    ("mfcc_440hz_20_64_fmin300_fmax8000", dict(n_mfcc=20, n_mels=64, fmin=300, fmax=8000), 5e-4),
])
def test_mfcc_golden(signal_440, name, kw, tol):
    G = load_golden(name)
    R = mm().mfcc(y=signal_440, sr=22050, **kw)
    assert R.shape == G.shape
    np.testing.assert_allclose(R, G, rtol=tol, atol=tol)


def test_db_goldens(signal_440):
    M = load_golden("stft_440hz_default_magnitude")
    P = M ** 2
    np.testing.assert_allclose(mm().power_to_db(P, ref=np.max), load_golden("stft_440hz_power_db"),
                               rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(mm().amplitude_to_db(M, ref=np.max), load_golden("stft_440hz_amplitude_db"),
                               rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(mm().amplitude_to_db(M, ref=1.0),
                               load_golden("stft_440hz_amplitude_db_ref1"), rtol=1e-4, atol=1e-4)


# -- FP64 oracle at the FP32 targets (30 s clip of the plain configuration) --

@pytest.fixture(scope="module")
def clip30():
    return config0_signal()


def test_config0_stft_fp32(clip30):
    R = mm().stft(clip30, n_fft=2048, hop_length=512)
    O = o.stft_magnitude(clip30, 2048, 512)
    assert R.shape == O.shape == (1025, 1292)
    assert rel_err(R, O) <= 1e-5


def test_config0_complex_stft(clip30):
    R = mm().stft_complex(clip30, n_fft=2048, hop_length=512)
    O = o.stft_complex(clip30, 2048, 512)
    assert R.dtype == np.complex64 and R.shape == O.shape
    assert rel_err(R, O) <= 1e-5
    assert np.all(R[0].imag == 0) and np.all(R[-1].imag == 0)


def test_config0_mel_db_fp32(clip30):
    R = mm().melspectrogram(y=clip30, sr=22050, n_fft=2048, hop_length=512, n_mels=128)
    O = o.melspectrogram(clip30, 22050, 2048, 512, n_mels=128)
    assert rel_err(R, O) <= 1e-5
    dR = o.power_to_db(R.astype(np.float64))
    dO = o.power_to_db(O)
    assert np.abs(dR - dO).max() <= 1e-3


def test_config0_mfcc_fp32(clip30):
    R = mm().mfcc(y=clip30, sr=22050, n_mfcc=13, n_fft=2048, hop_length=512, n_mels=128)
    O = o.mfcc(clip30, 22050, 13, 2048, 512, n_mels=128)
    assert R.shape == O.shape == (13, 1292)
    assert rel_err(R, O) <= 1e-4


# -- parameter sweeps and edge cases (EdgeCaseTests.swift, ParameterSweepTests.swift) --

@pytest.mark.parametrize("n_fft", [2, 4, 16, 64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_stft_sizes(n_fft):
    rng = np.random.default_rng(n_fft)
    y = rng.standard_normal(9000).astype(np.float32)
    hop = max(1, n_fft // 4)
    R = mm().stft(y, n_fft=n_fft, hop_length=hop)
    O = o.stft_magnitude(y, n_fft, hop)
    assert R.shape == O.shape
    assert rel_err(R, O) <= 1e-5


@pytest.mark.parametrize("hop,win,center", [(333, 2048, True), (512, 1500, True), (512, 3000, True),
                                            (700, 2048, False), (1, 2048, True), (4096, 2048, True)])
def test_stft_framing_variants(hop, win, center):
    y = np.random.default_rng(hop).standard_normal(30000).astype(np.float32)
    R = mm().stft(y, n_fft=2048, hop_length=hop, win_length=win, center=center)
    O = o.stft_magnitude(y, 2048, hop, win, center)
    assert R.shape == O.shape
    assert rel_err(R, O) <= 1e-5


def test_short_and_special_inputs():
    r = mm().stft(np.sin(np.arange(10, dtype=np.float32)))
    assert r.shape == (1025, 1)
    np.testing.assert_allclose(r, o.stft_magnitude(np.sin(np.arange(10, dtype=np.float32)), 2048, 512),
                               atol=1e-5)
    z = mm().stft(np.zeros(4096, np.float32))
    assert not z.any()
    nan = mm().stft(np.full(4096, np.nan, np.float32))
    assert nan.shape == (1025, 9)
    r = mm().stft(np.ones(2048, np.int32))
    assert r.dtype == np.float32
    with pytest.raises(RuntimeError, match="failed with status"):
        mm().stft(np.array([], np.float32))


def test_uncentred_short_signal_is_empty():
    r = mm().stft(np.ones(100, np.float32), n_fft=2048, center=False)
    assert r.shape == (1025, 0)


@pytest.mark.parametrize("n_mels,n_mfcc", [(128, 13), (40, 20), (20, 50), (256, 20), (1, 1)])
def test_mfcc_shapes_and_values(n_mels, n_mfcc):
    y = config0_signal(30000, seed=n_mels)
    R = mm().mfcc(y=y, sr=22050, n_mfcc=n_mfcc, n_mels=n_mels)
    O = o.mfcc(y, 22050, n_mfcc, 2048, 512, n_mels=n_mels)
    assert R.shape == O.shape == (min(n_mfcc, n_mels), 59)
    assert rel_err(R, O) <= 1e-4


@pytest.mark.parametrize("power", [1.0, 2.0, 1.5])
def test_mel_power(power):
    y = config0_signal(40000, seed=3)
    R = mm().melspectrogram(y=y, sr=22050, power=power)
    O = o.melspectrogram(y, 22050, 2048, 512, power=power)
    assert rel_err(R, O) <= 1e-5


def test_mfcc_top_db_clamp_active():
    """A loud tone then digital silence: most frames sit at the top_db floor,
    so the second (clamped) pass rewrites their tiles."""
    y = config0_signal(60000, seed=4)
    y[20000:] = 0.0
    R = mm().mfcc(y=y, sr=22050, n_mfcc=13)
    O = o.mfcc(y, 22050, 13, 2048, 512)
    assert rel_err(R, O) <= 1e-4


def test_mel_nonzero_fmin_fmax_and_sr():
    y = config0_signal(50000, 44100, seed=5)
    R = mm().melspectrogram(y=y, sr=44100, n_fft=4096, hop_length=1024, n_mels=256, fmin=50, fmax=16000)
    O = o.melspectrogram(y, 44100, 4096, 1024, n_mels=256, fmin=50, fmax=16000)
    assert rel_err(R, O) <= 1e-5


def test_reflect_padding():
    y = config0_signal(20000, seed=6)
    R = mm().spectral(y, kind="stft", n_fft=2048, hop_length=512, pad_mode="reflect")
    O = o.stft_magnitude(y, 2048, 512, pad_mode="reflect")
    assert rel_err(R, O) <= 1e-5
    y2 = y[:700]    # pad longer than the signal: periodic mirror
    R = mm().spectral(y2, kind="stft", n_fft=2048, hop_length=256, pad_mode="reflect")
    O = o.stft_magnitude(y2, 2048, 256, pad_mode="reflect")
    assert rel_err(R, O) <= 1e-5


def test_repeat_calls_are_bitwise_identical(clip30):
    a = mm().mfcc(y=clip30[:100000], n_mfcc=13)
    b = mm().mfcc(y=clip30[:100000], n_mfcc=13)
    assert np.array_equal(a, b)


def test_concurrent_contexts(clip30):
    from concurrent.futures import ThreadPoolExecutor
    y = clip30[:200000]
    ref = mm().mfcc(y=y, n_mfcc=13)
    with ThreadPoolExecutor(4) as ex:
        outs = list(ex.map(lambda _: mm().mfcc(y=y, n_mfcc=13), range(8)))
    for r in outs:
        assert np.array_equal(r, ref)


# -- batches -------------------------------------------------------------------

def _uneven_clips(n, seed, sr=22050, lo=0.2, hi=3.0):
    rng = np.random.default_rng(seed)
    lens = (rng.uniform(lo, hi, n) * sr).astype(np.int64)
    return [config0_signal(int(L), sr, seed + i) for i, L in enumerate(lens)]


@pytest.mark.parametrize("kind", ["mfcc", "mel", "stft", "complex"])
def test_batch_equals_single_clip_bitwise(kind):
    from metalmom_amd.batch import spectral_batch
    clips = _uneven_clips(9, 11)
    outs = spectral_batch(clips, kind=kind, n_mfcc=13)
    for c, b in zip(clips, outs):
        s = mm().spectral(c, kind=kind, n_mfcc=13)
        assert s.shape == b.shape
        assert np.array_equal(s, b)


def test_device_batch_graph_replay_matches_eager():
    import torch
    from metalmom_amd.batch import SpectralBatch
    clips = _uneven_clips(17, 21)
    B = SpectralBatch([len(c) for c in clips], kind="mfcc", n_mfcc=13)
    d_in = B.pack(clips)
    out_e = B.empty_output()
    out_g = B.empty_output()
    B.run(d_in, out_e)
    torch.cuda.synchronize()
    eager = out_e.cpu().numpy().copy()
    for _ in range(3):
        out_g.zero_()
        B.run(d_in, out_g, graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(out_g.cpu().numpy(), eager)
    for c, r in zip(clips, B.unpack(out_e)):
        assert rel_err(r, o.mfcc(c, 22050, 13, 2048, 512)) <= 1e-4


@pytest.mark.parametrize("prec,tol_db", [("bf16", 0.05), ("f16", 0.05)])
def test_low_precision_mel_error(prec, tol_db):
    clips = _uneven_clips(6, 31, lo=1.0, hi=4.0)
    from metalmom_amd.batch import spectral_batch
    mel = spectral_batch(clips, kind="mel", precision=prec)
    mf = spectral_batch(clips, kind="mfcc", n_mfcc=13, precision=prec)
    for c, m, f in zip(clips, mel, mf):
        Om = o.melspectrogram(c, 22050, 2048, 512)
        dO = o.power_to_db(Om)
        dR = o.power_to_db(m.astype(np.float64))
        assert np.abs(dR - dO).max() <= tol_db
        Of = o.mfcc(c, 22050, 13, 2048, 512)
        assert rel_err(f, Of) <= 2e-3


def test_f16_input_large_fft():
    """Large-FFT configuration: FP16 input, n_fft 4096, hop 1024, 256 mels, 20 MFCC."""
    import torch
    from metalmom_amd.batch import SpectralBatch
    rng = np.random.default_rng(3)
    sr, n = 44100, 3 * 44100
    clips = []
    for i in range(4):
        t = np.arange(n) / sr
        x = 0.2 * rng.uniform(-1, 1, n)
        for f in rng.uniform(50, 10000, 3):
            x += 0.25 * np.sin(2 * np.pi * f * t)
        clips.append(np.clip(x, -1, 1).astype(np.float16))
    B = SpectralBatch([n] * 4, kind="mfcc", sr=sr, n_fft=4096, hop_length=1024, n_mels=256, n_mfcc=20,
                      precision="f16", input_dtype="f16")
    d_in = B.pack(clips)
    out = B.empty_output()
    B.run(d_in, out)
    torch.cuda.synchronize()
    for c, r in zip(clips, B.unpack(out)):
        O = o.mfcc(c.astype(np.float64), sr, 20, 4096, 1024, n_mels=256)
        assert r.shape == O.shape == (20, 1 + n // 1024)
        assert rel_err(r, O) <= 2e-3


@pytest.mark.parametrize("prec,idt", [("bf16", "f32"), ("bf16", "f16"), ("f16", "f32"), ("f16", "f16")])
def test_multi_round_mel_blocks(prec, idt):
    """n_fft 4096 with 256 mels at 44.1 kHz: the top block of 8 mels spans 9
    chunks of 32 bins, so its contraction runs in two rounds with the
    accumulator carried between them. A round whose MFMA result is read too
    early loses one chunk: the top bands come out several dB low in whole
    tiles, depending on timing, so the launches are repeated. Every band is
    checked in dB (16-bit tier, 0.05 dB) and the MFCC at the 16-bit tier, for
    the default workgroup shape and then each 8-wave shape (8 frames per tile
    and 16, two per wave), all of which exist at n_fft 4096; the two 8-wave
    shapes give the same bits."""
    import torch
    from metalmom_amd.batch import SpectralBatch
    sr, n = 44100, 3 * 44100
    fb = o.mel_filterbank(sr, 4096, 256)
    nz = np.nonzero(fb[248:256].any(axis=0))[0]
    s16 = int(nz[0]) & ~7
    assert (int(nz[-1]) + 1 - s16 + 31) // 32 >= 9   # the case this test is for
    rng = np.random.default_rng(11)
    t = np.arange(n) / sr
    clips = []
    for i in range(4):
        x = 0.2 * rng.uniform(-1, 1, n)
        for f in rng.uniform(50, 20000, 4):
            x += 0.2 * np.sin(2 * np.pi * f * t)
        clips.append(np.clip(x, -1, 1).astype(np.float16 if idt == "f16" else np.float32))
    ref = [c.astype(np.float64) for c in clips]
    Om = [o.power_to_db(o.melspectrogram(c, sr, 4096, 1024, n_mels=256)) for c in ref]
    Of = [o.mfcc(c, sr, 20, 4096, 1024, n_mels=256) for c in ref]
    Bm = SpectralBatch([n] * 4, kind="mel", sr=sr, n_fft=4096, hop_length=1024, n_mels=256,
                       precision=prec, input_dtype=idt)
    Bf = SpectralBatch([n] * 4, kind="mfcc", sr=sr, n_fft=4096, hop_length=1024, n_mels=256, n_mfcc=20,
                       precision=prec, input_dtype=idt)
    dm, df = Bm.pack(clips), Bf.pack(clips)
    last = {}
    for shape in (None, 2, 3):
        if shape is not None:
            for B in (Bm, Bf):
                B.set_shape(shape)
                sh = B.launch_shape()
                assert (sh["shape"], sh["waves_per_wg"], sh["tile_frames"]) == (shape, 8, 8 * (shape - 1)), sh
        for rep in range(3):
            om, of = Bm.empty_output(), Bf.empty_output()
            Bm.run(dm, om)
            Bf.run(df, of)
            torch.cuda.synchronize()
            for i, (m, f) in enumerate(zip(Bm.unpack(om), Bf.unpack(of))):
                d = np.abs(o.power_to_db(m.astype(np.float64)) - Om[i])
                assert d.max() <= 0.05, (shape, rep, i, int(np.argmax(d.max(axis=1))), float(d.max()))
                assert rel_err(f, Of[i]) <= 2e-3, (shape, rep, i)
            if shape is not None:
                last[shape] = (Bm.unpack(om), Bf.unpack(of))
    # the two 8-wave shapes give the same bits
    for k in (0, 1):
        for a, b in zip(last[2][k], last[3][k]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("n_fft,kind,prec,shape", [
    (1024, "stft", "f32", 3), (1024, "mel", "f32", 1), (1024, "mfcc", "bf16", 1), (2048, "stft", "f32", 3), (2048, "complex", "f32", 1), (2048, "mel", "f32", 1), (2048, "mel", "bf16", 1),
    (2048, "mfcc", "f32", 1), (2048, "mfcc", "bf16", 4), (2048, "mfcc", "f16", 4),
    (4096, "stft", "f32", 3), (4096, "complex", "f32", 2), (4096, "mel", "bf16", 3), (4096, "mfcc", "f32", 3),
    (4096, "mfcc", "f16", 3)])
def test_default_shapes(n_fft, kind, prec, shape):
    """The untuned workgroup shape (single-clip calls, the server, plans
    never autotuned): the measured-fastest one per configuration, 8-frame
    tiles at n_fft 4096 where the two-frame shape's LDS does not fit."""
    from metalmom_amd.batch import SpectralBatch
    B = SpectralBatch([3 * 44100, 44100], kind=kind, sr=44100, n_fft=n_fft, hop_length=n_fft // 4, n_mels=128,
                      n_mfcc=13, precision=prec)
    B.set_shape(0)
    assert B.launch_shape()["shape"] == shape
    B.close()

@pytest.mark.parametrize("kind", ["stft", "mel", "mfcc"])
def test_n4096_eight_wave_shapes_f32(kind):
    """n_fft 4096 in FP32: the 8-frame and the 16-frame (two per wave) tiles
    give the same bits, and both meet the FP32 targets against the oracle."""
    import torch
    from metalmom_amd.batch import SpectralBatch
    sr = 44100
    clips = _uneven_clips(7, 5, sr=sr)
    B = SpectralBatch([len(c) for c in clips], kind=kind, sr=sr, n_fft=4096, hop_length=1024, n_mfcc=13)
    d_in = B.pack(clips)
    res = {}
    for shape in (2, 3):
        B.set_shape(shape)
        sh = B.launch_shape()
        assert (sh["shape"], sh["waves_per_wg"], sh["tile_frames"]) == (shape, 8, 8 * (shape - 1)), sh
        out = B.empty_output()
        B.run(d_in, out)
        torch.cuda.synchronize()
        res[shape] = B.unpack(out)
    B.close()
    for c, a, b in zip(clips, res[2], res[3]):
        assert np.array_equal(a, b)
        if kind == "stft":
            assert rel_err(a, o.stft_magnitude(c, 4096, 1024)) <= 1e-5
        elif kind == "mel":
            assert rel_err(a, o.melspectrogram(c, sr, 4096, 1024, n_mels=128)) <= 1e-5
        else:
            assert rel_err(a, o.mfcc(c, sr, 13, 4096, 1024, n_mels=128)) <= 1e-4


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 9, 1003, 40001])
def test_db_any_length(n):
    """power_to_db / amplitude_to_db over lengths that leave 1-3 values after
    the kernel's 16-byte groups: same values as the oracle, top_db from the
    maximum over all of them (the maximum may sit in the tail)."""
    rng = np.random.default_rng(n)
    x = (rng.random(n) * 10.0).astype(np.float32)
    x[-1] = 50.0   # the maximum in the last position
    got = mm().power_to_db(x, top_db=80.0).astype(np.float64).ravel()
    ref = o.power_to_db(x.astype(np.float64), top_db=80.0).ravel()
    assert np.abs(got - ref).max() < 1e-3
    got = mm().amplitude_to_db(x, top_db=40.0).astype(np.float64).ravel()
    ref = o.amplitude_to_db(x.astype(np.float64), top_db=40.0).ravel()
    assert np.abs(got - ref).max() < 1e-3

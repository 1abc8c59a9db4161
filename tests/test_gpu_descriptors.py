"""GPU parity of the spectral descriptors (mm_spectral_centroid / _bandwidth /
_rolloff / _flatness, SpectralDescriptors.swift) against the reference's
librosa goldens and the float64 oracle.

Tolerances: centroid and bandwidth are ratios of FP32 sums over the FP32
magnitude spectrum: 1e-5 relative to the largest value. Rolloff is a bin
frequency: equal to the oracle on at least 97 % of frames and never more
than one bin (sr / n_fft) away (a cumulative sum within rounding of the
threshold may cross one bin later or earlier), compared as bin indices
since the GPU writes the frequency in FP32 as the reference does. Flatness of noise-like
frames: 1e-4 relative. Flatness of a pure tone is set by the FP32 FFT's
noise floor against the 1e-10 power floor (most bins hold only rounding
noise), so the 440 Hz golden is checked to within a factor of 2; the same
noise, weighted by the squared distance from the centroid, puts the pure
tone's bandwidth at 1e-3 relative (noise-like frames stay at 1e-5).
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import spectral as o

pytestmark = pytest.mark.gpu


def _mix(n, sr, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / sr
    x = 0.3 * rng.standard_normal(n)
    for f in (180.0, 700.0, 2500.0):
        x += 0.3 * np.sin(2 * np.pi * f * t * (1 + 0.1 * t))
    return x.astype(np.float32)


def _check(name, got, ref, sr, n_fft):
    assert got.shape == ref.shape
    if name == "spectral_rolloff":   # compared as bin indices (the GPU writes Float(k) * sr / n_fft in FP32)
        d = np.abs(np.rint(got * n_fft / sr) - np.rint(ref * n_fft / sr))
        assert (d <= 1).all() and (d == 0).mean() >= 0.97, (d.max(), (d == 0).mean())
    elif name == "spectral_flatness":
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
    else:
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["spectral_centroid", "spectral_bandwidth", "spectral_rolloff", "spectral_flatness"])
def test_descriptor_goldens(signal_440, name):
    import metalmom_amd as mm
    y = signal_440.astype(np.float32)
    got = getattr(mm, name)(y=y, sr=22050)
    G = load_golden(f"{name}_440hz_default")
    assert got.shape == G.shape == (1, 44) and got.dtype == np.float32
    if name == "spectral_flatness":
        assert (got > 0).all() and (got < 2 * G).all() and (got > 0.5 * G).all()
    elif name == "spectral_bandwidth":
        assert np.abs(got - G).max() <= 1e-3 * np.abs(G).max()
    else:
        _check(name, got.astype(np.float64), G.astype(np.float64), 22050, 2048)


@pytest.mark.parametrize("name", ["spectral_centroid", "spectral_bandwidth", "spectral_rolloff", "spectral_flatness"])
@pytest.mark.parametrize("sr,n,kw", [
    (22050, 80000, {}),
    (44100, 90000, {"n_fft": 4096, "hop_length": 1024}),
    (22050, 50000, {"n_fft": 1024, "hop_length": 300, "win_length": 800, "center": False}),
    (22050, 900000, {}),
    (22050, 40000, {"n_fft": 512, "hop_length": 128}),
], ids=["default", "44k-4096", "1024-hop300-uncentred", "long", "512-two-pass"])
def test_descriptor_matches_oracle(name, sr, n, kw):
    import metalmom_amd as mm
    print(f"{name} sr={sr} n={n} {kw}", flush=True)
    y = _mix(n, sr, n % 83)
    got = getattr(mm, name)(y=y, sr=sr, **kw).astype(np.float64)
    okw = dict(n_fft=kw.get("n_fft", 2048), hop=kw.get("hop_length", 512), win_length=kw.get("win_length"),
               center=kw.get("center", True))
    ref = getattr(o, name)(y.astype(np.float64), **okw) if name == "spectral_flatness" else \
        getattr(o, name)(y.astype(np.float64), sr, **okw)
    _check(name, got, ref, sr, okw["n_fft"])


def test_descriptor_params_and_edges():
    import metalmom_amd as mm
    y = _mix(60000, 22050, 3)
    for p in (1.0, 3.0):
        got = mm.spectral_bandwidth(y=y, p=p).astype(np.float64)
        ref = o.spectral_bandwidth(y.astype(np.float64), p=p)
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    got = mm.spectral_rolloff(y=y, roll_percent=0.5).astype(np.float64)
    ref = o.spectral_rolloff(y.astype(np.float64), roll_percent=0.5)
    _check("spectral_rolloff", got, ref, 22050, 2048)
    z = np.zeros(30000, np.float32)
    for name in ("spectral_centroid", "spectral_bandwidth", "spectral_rolloff", "spectral_flatness"):
        assert not getattr(mm, name)(y=z).any()                   # silence: 0 (no energy)
        assert getattr(mm, name)(y=z, n_fft=1000).shape == (1, 0)  # no STFT frames


def test_descriptor_device_entry_matches_host_entry():
    import ctypes as C
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    y = _mix(250000, 22050, 9)
    d_y = torch.from_numpy(y).cuda()
    T = 1 + len(y) // 512
    d_o = torch.full((T,), float("nan"), device="cuda")
    n = C.c_int64(0)
    ctx = lib.mm_init()
    try:
        for kind, name, param, kw in [(0, "spectral_centroid", 0.0, {}), (1, "spectral_bandwidth", 2.0, {}),
                                      (2, "spectral_rolloff", 0.85, {}), (3, "spectral_flatness", 0.0, {}),
                                      (0, "spectral_centroid", 0.0, {})]:
            s = lib.mm_spectral_descriptor_device(ctx, C.c_void_p(d_y.data_ptr()), len(y), 22050, 2048, 512, 2048, 1,
                                                  kind, param, C.c_void_p(d_o.data_ptr()), C.byref(n),
                                                  C.c_void_p(torch.cuda.current_stream().cuda_stream))
            assert s == 0 and n.value == T
            torch.cuda.synchronize()
            assert np.array_equal(d_o.cpu().numpy()[None, :], getattr(mm, name)(y=y, **kw))
    finally:
        lib.mm_destroy(ctx)


# -- spectral contrast (mm_spectral_contrast, SpectralDescriptors.swift:137-204)

def test_contrast_golden_at_reference_tolerance(signal_440):
    """The librosa golden at the reference's own tolerance (rtol 0.5, atol
    15 dB; test_parity_spectral.py:158-172), and the oracle closely."""
    import metalmom_amd as mm
    G = load_golden("spectral_contrast_440hz_default")
    got = mm.feature.spectral_contrast(y=signal_440, sr=22050)
    assert got.shape == G.shape and np.all(np.isfinite(got))
    np.testing.assert_allclose(got, G, rtol=0.5, atol=15.0)


@pytest.mark.parametrize("sr,n,kw", [
    (22050, 80000, {}),
    (44100, 120000, {"n_fft": 4096, "hop_length": 1024}),                          # top band q = 29
    (44100, 150000, {"n_fft": 8192, "hop_length": 2048}),                          # top band q = 58: bisection
    (22050, 50000, {"n_fft": 1024, "hop_length": 300, "center": False, "n_bands": 4, "fmin": 100.0}),
], ids=["default", "4096", "8192-bisection", "1024-4bands-uncentred"])
def test_contrast_matches_oracle(sr, n, kw):
    """Noise-rich signals, so every band's lowest magnitudes sit far above
    the FP32 FFT's rounding floor: the dB difference of two FP32 means
    agrees with the float64 oracle to 0.01 dB."""
    import metalmom_amd as mm
    print(f"contrast sr={sr} n={n} {kw}", flush=True)
    y = _mix(n, sr, n % 89)
    got = mm.feature.spectral_contrast(y=y, sr=sr, **kw).astype(np.float64)
    ref = o.spectral_contrast(y.astype(np.float64), sr, n_fft=kw.get("n_fft", 2048), hop=kw.get("hop_length", 512),
                              center=kw.get("center", True), n_bands=kw.get("n_bands", 6),
                              fmin=kw.get("fmin", 200.0))
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 0.01, float(np.abs(got - ref).max())


def test_contrast_edges():
    import metalmom_amd as mm
    z = np.zeros(30000, np.float32)
    assert not mm.feature.spectral_contrast(y=z).any()                        # silence: 0 dB in every band
    assert mm.feature.spectral_contrast(y=z, n_fft=1000).shape == (7, 0)      # no STFT frames
    c = mm.feature.spectral_contrast(y=_mix(20000, 8000, 1), sr=8000, n_fft=512, n_bands=7)
    assert c.shape[0] == 8 and not c[-1].any()                                 # octaves past Nyquist: empty


def test_contrast_device_entry_matches_host_entry():
    import ctypes as C
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    y = _mix(70000, 22050, 4)
    host = mm.feature.spectral_contrast(y=y, sr=22050, n_bands=5, fmin=150.0)
    X = torch.from_numpy(y.copy()).cuda()
    T = host.shape[1]
    Y = torch.full((6 * T,), float("nan"), device="cuda")
    fr = C.c_int64()
    ctx = lib.mm_init()
    try:
        for _ in range(2):   # the second call reuses the uploaded band table
            s = lib.mm_spectral_contrast_device(ctx, C.c_void_p(X.data_ptr()), len(y), 22050, 2048, 512, 2048, 1, 5,
                                                150.0, C.c_void_p(Y.data_ptr()), C.byref(fr),
                                                C.c_void_p(torch.cuda.current_stream().cuda_stream))
            assert s == 0 and fr.value == T
            torch.cuda.synchronize()
            assert np.array_equal(Y.cpu().numpy().reshape(6, T), host)
    finally:
        lib.mm_destroy(ctx)


@pytest.mark.parametrize("n", [1, 100, 2047, 2049])
def test_descriptors_on_tiny_clips(n):
    """Clips of one or a few frames, mostly centre padding: every descriptor
    and contrast against the oracle (rolloff by bin, as above)."""
    import metalmom_amd as mm
    y = _mix(n, 22050, n)
    for name in ("spectral_centroid", "spectral_bandwidth", "spectral_rolloff", "spectral_flatness"):
        got = getattr(mm, name)(y=y, sr=22050).astype(np.float64)
        ref = getattr(o, name)(y.astype(np.float64)) if name == "spectral_flatness" else \
            getattr(o, name)(y.astype(np.float64), 22050)
        _check(name, got, ref, 22050, 2048)
    got = mm.feature.spectral_contrast(y=y, sr=22050).astype(np.float64)
    ref = o.spectral_contrast(y.astype(np.float64), 22050)
    assert got.shape == ref.shape and np.abs(got - ref).max() <= 0.05


def _sine(f, sr=22050, seconds=1.0):
    i = np.arange(int(sr * seconds), dtype=np.float32)
    return np.sin(np.float32(2 * np.pi * f) * i / np.float32(sr)).astype(np.float32)


def _lcg_noise(n=22050, seed=42):
    """The white noise of SpectralDescriptorsTests.swift:289-300: a 64-bit
    LCG, each state read as a signed integer over Int64.max."""
    s, out = seed, np.empty(n, np.float32)
    m = (1 << 64) - 1
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & m
        out[i] = np.float32(np.float64(s - (1 << 64) if s >> 63 else s) / np.float64((1 << 63) - 1))
    return out


def test_descriptor_known_answers_of_the_reference_tests():
    """The tone and noise known answers of SpectralDescriptorsTests.swift
    (58-77, 113-131, 212-230, 269-314, 329-343), through the drop-in entries
    at the defaults: a second of 440 Hz has centroid 440 +- 50 Hz, bandwidth
    under 500 Hz, 85% rolloff 440 +- 200 Hz and flatness under 0.1 (frames
    2..n-3 averaged); LCG white noise has mean flatness above 0.3; a 1 kHz
    tone's middle-frame centroid is above 440 Hz's."""
    import metalmom_amd as mm
    y = _sine(440.0)
    mid = slice(2, -2)
    cen = mm.spectral_centroid(y=y, sr=22050)[0]
    assert abs(cen[mid].mean() - 440.0) <= 50.0
    assert mm.spectral_bandwidth(y=y, sr=22050)[0][mid].mean() < 500.0
    assert abs(mm.spectral_rolloff(y=y, sr=22050)[0][mid].mean() - 440.0) <= 200.0
    assert mm.spectral_flatness(y=y)[0][mid].mean() < 0.1
    assert mm.spectral_flatness(y=_lcg_noise())[0].mean() > 0.3
    c1k = mm.spectral_centroid(y=_sine(1000.0), sr=22050)[0]
    assert c1k[len(c1k) // 2] > cen[len(cen) // 2]

"""GPU parity of onset strength (mm_onset_strength, OnsetDetection.swift:23-128)
against the reference's librosa golden and the float64 oracle.

Tolerance: every value is a difference of two mel dB values (FP32 path,
<= 1e-3 dB each against FP64) averaged over bands, so |error| <= 2e-3 (dB)
absolute, aggregated or per band. Silence is exactly zero; shapes, padding
and the zero-column cases are exact.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import spectral as o

pytestmark = pytest.mark.gpu
TOL = 2e-3


def test_onset_strength_golden(signal_440):
    import metalmom_amd as mm
    G = load_golden("onset_strength_440hz_default")
    env = mm.onset_strength(y=signal_440.astype(np.float32), sr=22050)
    assert env.shape == (44,) and env.dtype == np.float32
    assert np.abs(env - G).max() <= TOL
    assert np.abs(env - o.onset_strength(signal_440.astype(np.float32).astype(np.float64))[0]).max() <= TOL


def _clip(n, sr, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / sr
    x = 0.1 * rng.standard_normal(n)
    for k in range(6):   # note onsets: bursts of a tone every ~0.25 s
        a = int((0.1 + 0.25 * k) * sr)
        x[a:a + sr // 10] += np.sin(2 * np.pi * (300 + 150 * k) * t[:sr // 10][:len(x[a:a + sr // 10])])
    return x.astype(np.float32)


@pytest.mark.parametrize("sr,n,kw", [
    (22050, 60000, {}),
    (22050, 60000, {"aggregate": False}),
    (22050, 60000, {"center": False}),
    (22050, 50000, {"hop_length": 300}),
    (44100, 90000, {"n_mels": 40, "fmin": 100.0, "fmax": 8000.0}),
    (22050, 50000, {"n_fft": 1024, "hop_length": 256}),
    (22050, 3000, {}),
    (22050, 2_000_000, {}),
], ids=["default", "per-band", "uncentred", "hop300", "44k-40mels-band", "nfft1024", "short", "long"])
def test_onset_strength_matches_oracle(sr, n, kw):
    import metalmom_amd as mm
    print(f"onset sr={sr} n={n} {kw}", flush=True)
    y = _clip(n, sr, n % 97)
    got = mm.onset_strength(y=y, sr=sr, **kw)
    okw = dict(n_fft=kw.get("n_fft", 2048), hop=kw.get("hop_length", 512), n_mels=kw.get("n_mels", 128),
               fmin=kw.get("fmin", 0.0), fmax=kw.get("fmax"), center=kw.get("center", True),
               aggregate=kw.get("aggregate", True))
    ref = o.onset_strength(y.astype(np.float64), sr, **okw)
    if okw["aggregate"]:
        ref = ref[0]
    assert got.shape == ref.shape
    assert (got >= 0).all()
    assert np.abs(got - ref).max() <= TOL


def test_onset_strength_silence_single_frame_and_repeatability():
    import metalmom_amd as mm
    assert not mm.onset_strength(y=np.zeros(22050, np.float32)).any()          # test_onset_strength.py:23-26
    one = mm.onset_strength(y=np.ones(1000, np.float32), center=False)          # a single frame
    assert one.shape == (1,) and one[0] == 0
    y = _clip(200000, 22050, 5)
    assert np.array_equal(mm.onset_strength(y=y), mm.onset_strength(y=y))


def test_onset_device_entry_matches_host_entry():
    """mm_onset_strength_device on a torch tensor equals the host entry
    bitwise, also after another call on the same context replaced the
    clip work list it caches."""
    import ctypes as C
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import MMBuffer, MMSTFTParams, lib
    y = _clip(300000, 22050, 11)
    host = mm.onset_strength(y=y)
    d_y = torch.from_numpy(y).cuda()
    d_o = torch.full((128 * 600,), float("nan"), device="cuda")
    rows, cols = C.c_int64(0), C.c_int64(0)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ctx = lib.mm_init()
    try:
        def run():
            s = lib.mm_onset_strength_device(ctx, C.c_void_p(d_y.data_ptr()), len(y), 22050, 2048, 512, 128, 0.0, 0.0,
                                             1, 1, C.c_void_p(d_o.data_ptr()), C.byref(rows), C.byref(cols), st)
            assert s == 0
            torch.cuda.synchronize()
            return d_o[:rows.value * cols.value].cpu().numpy().copy()
        a = run()
        assert (rows.value, cols.value) == (1, host.size) and np.array_equal(a, host)
        out = MMBuffer()   # a host STFT on the same context rewrites its work-list buffer
        assert lib.mm_stft(ctx, y.ctypes.data_as(C.POINTER(C.c_float)), 5000, 22050,
                           C.byref(MMSTFTParams(2048, 512, 2048, 1)), C.byref(out)) == 0
        lib.mm_buffer_free(C.byref(out))
        assert np.array_equal(run(), host)
    finally:
        lib.mm_destroy(ctx)

"""GPU parity of the inverse STFT (mm_istft, STFT.swift:686-861) against the
float64 oracle, through the drop-in Python API and the C ABI.

Tolerance: output sample s is the window-weighted sum of FP32 inverse FFT
samples divided by the summed squared window W2(s), so its error is bounded
by rtol * A * W1(s) / W2(s), with A the largest inverse-FFT sample of the
input, W1 the summed |window| and rtol = 2e-6 (a few FP32 ulps of the FFT).
Where the division is skipped (W2 <= 1e-10) the bound is rtol * A * W1(s).
With centre padding and a full-length Hann window W1/W2 <= 1.34, so the
bound is ~3e-6 A; only the uncentred edges, where W2 runs to zero, are
looser. Samples whose W2 lies within 10 % of the 1e-10 threshold (where
FP32 and FP64 could decide differently) are not compared.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import spectral as o

pytestmark = pytest.mark.gpu

RTOL = 2e-6


def _weights(F, T, hop, win, center, length):
    N = 2 * (F - 1)
    w = o.prepare_window(win, N)
    total = N + (T - 1) * hop
    w2, w1 = np.zeros(total), np.zeros(total)
    for f in range(T):
        w2[f * hop:f * hop + N] += w * w
        w1[f * hop:f * hop + N] += np.abs(w)
    n = o.istft_length(F, T, hop, center, length)
    start = N // 2 if center else 0
    m = max(min(n, total - start), 0)
    W1, W2 = np.zeros(n), np.zeros(n)
    W1[:m], W2[:m] = w1[start:start + m], w2[start:start + m]
    return W1, W2


def check_istft(got, S, hop, win, center, length, rtol=RTOL):
    """got: GPU result for the complex64 spectrogram S."""
    S = np.asarray(S, dtype=np.complex64).astype(np.complex128)
    ref = o.istft(S, hop, win, center, length)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    if ref.size == 0:
        return
    F, T = S.shape
    N = 2 * (F - 1)
    A = max(np.abs(np.fft.irfft(S.T, n=N, axis=1)).max(), 1e-30)
    W1, W2 = _weights(F, T, hop, win, center, length)
    div = W2 > 1e-10
    bound = rtol * A * W1 / np.where(div, W2, 1.0) + 1e-30
    near = np.abs(W2 - 1e-10) < 1e-11
    err = np.abs(got.astype(np.float64) - ref)
    bad = (err > bound) & ~near
    assert not bad.any(), (int(np.argmax(bad)), float(err[bad].max()), float(bound[bad].min()))


def _signal(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / 22050.0
    return (0.3 * rng.standard_normal(n) + 0.5 * np.sin(2 * np.pi * 440 * t)).astype(np.float32)


# (name, n_fft, hop, win, center, length, samples); fused = n_fft 1024/2048/4096
# with at most half a tile of halo frames, generic otherwise
CASES = [
    ("fused-2048-h512-centre", 2048, 512, 2048, True, None, 50000),
    ("fused-2048-h512-centre-short", 2048, 512, 2048, True, 30000, 50000),
    ("fused-1024-h256-uncentred-long", 1024, 256, 1024, False, 52000, 50000),
    ("fused-4096-h1024-centre", 4096, 1024, 4096, True, None, 70000),
    ("fused-2048-h256-centre", 2048, 256, 2048, True, None, 30000),
    ("fused-2048-win1500", 2048, 512, 1500, True, None, 40000),
    ("fused-2048-h500", 2048, 500, 2048, True, None, 40000),
    ("fused-4096-h4096-no-overlap", 4096, 4096, 4096, False, None, 40000),
    ("generic-2048-h128", 2048, 128, 2048, True, None, 20000),
    ("generic-256-h64", 256, 64, 256, True, None, 20000),
    ("generic-512-h100-uncentred", 512, 100, 512, False, None, 20000),
    ("generic-16384-h4096", 16384, 4096, 16384, True, None, 100000),
    ("generic-8-h2", 8, 2, 8, True, None, 300),
]


@pytest.mark.parametrize("name,n_fft,hop,win,center,length,n", CASES, ids=[c[0] for c in CASES])
def test_istft_matches_oracle(name, n_fft, hop, win, center, length, n):
    import metalmom_amd as mm
    print(f"istft case {name}", flush=True)
    y = _signal(n, n_fft + hop)
    S = o.stft_complex(y.astype(np.float64), n_fft, hop, win, center).astype(np.complex64)
    got = mm.istft(S, hop_length=hop, win_length=win, center=center, length=length)
    assert got.dtype == np.float32
    check_istft(got, S, hop, win, center, length)


@pytest.mark.parametrize("F,T,hop,center", [(1025, 30, 512, True), (1025, 30, 512, False), (129, 40, 32, True),
                                            (5, 9, 2, False), (2, 3, 0, False)])
def test_istft_arbitrary_spectra(F, T, hop, center):
    """Spectra that are no STFT of anything, with DC and Nyquist imaginary
    parts set: those parts are ignored (STFT.swift:760-768)."""
    import metalmom_amd as mm
    print(f"istft arbitrary F={F} T={T} hop={hop} center={center}", flush=True)
    rng = np.random.default_rng(F + T)
    S = (rng.standard_normal((F, T)) + 1j * rng.standard_normal((F, T))).astype(np.complex64)
    N = 2 * (F - 1)
    got = mm.istft(S, hop_length=hop, center=center)
    check_istft(got, S, hop, N, center, None)


def test_istft_inverts_gpu_stft():
    """y -> stft_complex -> istft on the GPU gives y back (ISTFTTests.swift:8-37,
    test_parity_istft.py round trips)."""
    import metalmom_amd as mm
    for n_fft, n in [(2048, 44 * 512), (1024, 8192), (256, 5000)]:
        y = _signal(n, n_fft)
        S = mm.stft_complex(y, n_fft=n_fft)
        r = mm.istft(S, hop_length=n_fft // 4, length=len(y))
        assert r.shape == y.shape
        assert np.abs(r - y).max() <= 1e-5 * np.abs(y).max()
        e = float(np.sum(r.astype(np.float64) ** 2) / np.sum(y.astype(np.float64) ** 2))
        assert 0.99 < e < 1.01   # test_parity_istft.py:108-125


def test_istft_long_clip_and_repeatability():
    """A clip long enough for every workgroup of the persistent grid to take
    several tiles; two calls are bitwise identical (fixed summation order)."""
    import metalmom_amd as mm
    y = _signal(2_000_000, 7)
    S = o.stft_complex(y.astype(np.float64), 2048, 512).astype(np.complex64)
    a = mm.istft(S, hop_length=512, length=len(y))
    b = mm.istft(S, hop_length=512, length=len(y))
    assert np.array_equal(a, b)
    check_istft(a, S, 512, 2048, True, len(y))


def test_istft_device_entry_matches_host_entry():
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    for n_fft, hop in [(2048, 512), (512, 100)]:
        y = _signal(40000, 3)
        S = o.stft_complex(y.astype(np.float64), n_fft, hop).astype(np.complex64)
        host = mm.istft(S, hop_length=hop)
        flat = np.ascontiguousarray(S).view(np.float32)
        F, T = S.shape
        n = lib.mm_istft_length(F, T, hop, 1, 0)
        assert n == host.size
        dX = torch.from_numpy(flat.copy()).cuda()
        dY = torch.full((n,), float("nan"), device="cuda")
        ctx = lib.mm_init()
        try:
            s = lib.mm_istft_device(ctx, C.c_void_p(dX.data_ptr()), 0, F, T, 22050, hop, n_fft, 1, 0,
                                    C.c_void_p(dY.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
            assert s == 0
            torch.cuda.synchronize()
        finally:
            lib.mm_destroy(ctx)
        assert np.array_equal(dY.cpu().numpy(), host)
        # rows padded to a multiple of 16 frames (ld_frames), as the host entry lays them out
        ld = (T + 15) // 16 * 16
        padded = np.zeros((F, ld), np.complex64)
        padded[:, :T] = S
        dP = torch.from_numpy(np.ascontiguousarray(padded).view(np.float32).copy()).cuda()
        dZ = torch.full((n,), float("nan"), device="cuda")
        ctx = lib.mm_init()
        try:
            assert lib.mm_istft_device(ctx, C.c_void_p(dP.data_ptr()), T - 1, F, T, 22050, hop, n_fft, 1, 0,
                                       C.c_void_p(dZ.data_ptr()), None) == -1   # ld below n_frames
            s = lib.mm_istft_device(ctx, C.c_void_p(dP.data_ptr()), ld, F, T, 22050, hop, n_fft, 1, 0,
                                    C.c_void_p(dZ.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
            assert s == 0
            torch.cuda.synchronize()
        finally:
            lib.mm_destroy(ctx)
        assert np.array_equal(dZ.cpu().numpy(), host)

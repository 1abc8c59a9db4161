"""GPU parity of the sample-rate converter (mm_resample, Resample.swift:12-96)
against the float64 oracle, and the reference's own resample tests
(ResampleTests.swift, test_resample.py) on the GPU path.

Tolerance: each output is a sum of 2 * support FP32 products x_j w_j
(support = ceil(64 / scale)), so it differs from the float64 sum by a few
FP32 rounding steps of the running sum: |got - ref| <= 2e-6 * support *
max|x| (3e-4 of max|x| would already be a wrong tap). The oracle uses the
FP32-rounded table the kernel reads, so the comparison sees the arithmetic
only.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import ingest as g

pytestmark = pytest.mark.gpu

_T32 = None


def _table32():
    global _T32
    if _T32 is None:
        _T32 = g.resample_table().astype(np.float32).astype(np.float64)
    return _T32


def _signal(n, sr, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n) / sr
    x = 0.2 * rng.standard_normal(n) + 0.5 * np.sin(2 * np.pi * 440 * t) + 0.2 * np.sin(2 * np.pi * 0.3 * sr * t)
    return x.astype(np.float32)


def _check(got, y, src, dst):
    ref = g.resample(y.astype(np.float64), src, dst, table=_table32())
    assert got.dtype == np.float32 and got.shape == ref.shape, (got.shape, ref.shape)
    scale = min(1.0, dst / src)
    support = math.ceil(64 / scale)
    tol = 2e-6 * support * max(float(np.abs(y).max()), 1e-30)
    err = np.abs(got.astype(np.float64) - ref)
    assert err.max() <= tol, (float(err.max()), tol, int(np.argmax(err)))


# (source, target, samples): one-phase pairs take the polyphase kernel with
# the row in scalar registers (decimation by 2: two outputs per lane); pairs
# with 2 .. 4096 phases take the phase-major kernel (two phases per wave),
# whose workgroups cover 64 outputs of
# every phase (64 x 320 input samples at 48 -> 22.05 kHz): the short clips
# reach only the clipped first and last workgroups, the 150,000-300,000-sample
# ones the interior workgroups (transposed LDS window) of every layout, input
# step 1 (8 -> 48 kHz), 147, 320 and 441 included; 44.1 -> 32.001 kHz has
# too many phases and takes the per-tap kernel
CASES = [(44100, 22050, 50000), (22050, 44100, 30000), (48000, 22050, 40000), (16000, 22050, 20000),
         (22050, 16000, 20000), (44100, 8000, 30000), (8000, 48000, 5000), (44100, 48000, 30000),
         (44100, 22050, 1), (44100, 22050, 2), (22050, 44100, 3), (48000, 22050, 129),
         (48000, 22050, 300000), (16000, 22050, 150000), (22050, 16000, 150000), (44100, 8000, 300000),
         # one phase: decimation by 2 (two outputs per lane), 3 (scalar-index
         # LDS reads) and 4 (float4 reads), long enough for interior chunks
         (48000, 24000, 100000), (32000, 16000, 70000), (48000, 16000, 90000), (44100, 11025, 90000),
         (44100, 32001, 20000)]   # 10,667 phases: too many rows for LDS, the per-tap kernel


@pytest.mark.parametrize("src,dst,n", CASES, ids=[f"{a}-{b}-{n}" for a, b, n in CASES])
def test_resample_matches_oracle(src, dst, n):
    import metalmom_amd as mm
    print(f"resample {src}->{dst} n={n}", flush=True)
    y = _signal(n, src, src + n)
    _check(mm.resample(y, src, dst), y, src, dst)


def test_reference_resample_tests_on_gpu():
    """test_resample.py:11-44 and ResampleTests.swift:12-90."""
    import metalmom_amd as mm
    y = np.sin(np.arange(44100, dtype=np.float32) * 440.0 * 2 * np.pi / 44100)
    assert abs(len(mm.resample(y, 44100, 22050)) - 22050) <= 1
    y = np.sin(np.arange(22050, dtype=np.float32) * 440.0 * 2 * np.pi / 22050)
    up = mm.resample(y, 22050, 44100)
    assert abs(len(up) - 44100) <= 1
    rt = mm.resample(up, 44100, 22050)
    m = min(len(y), len(rt))
    s = y[1000:m - 1000]
    noise = s - rt[1000:m - 1000]
    assert 10 * np.log10(np.sum(s ** 2) / max(np.sum(noise ** 2), 1e-20)) > 40
    assert abs(np.sqrt(np.mean(y ** 2)) - np.sqrt(np.mean(up ** 2))) < 0.05


def test_long_signal_and_repeatability():
    """Enough chunks for every persistent workgroup to take several; two
    calls are bitwise identical."""
    import metalmom_amd as mm
    y = _signal(3_000_000, 44100, 5)
    a = mm.resample(y, 44100, 22050)
    b = mm.resample(y, 44100, 22050)
    assert np.array_equal(a, b)
    i = np.random.default_rng(1).integers(0, len(a) - 2000)
    # the oracle on a window (outputs depend on inputs within +-support)
    lo = max(0, 2 * i - 200)
    seg = y[lo:2 * i + 4200]
    ref = g.resample(seg.astype(np.float64), 44100, 22050, table=_table32())
    off = (2 * i - lo) // 2
    assert np.abs(a[i:i + 2000] - ref[off:off + 2000]).max() <= 2e-6 * 128 * np.abs(y).max()


def test_long_many_phase_signal_and_repeatability():
    """48 -> 22.05 kHz (147 phases, 320 input samples per 147 outputs) over
    3 M samples: two calls bitwise identical; 1,000 outputs from the middle
    against the oracle run on the inputs they reach (a segment starting on a
    multiple of 320 samples, so its outputs keep their phases)."""
    import metalmom_amd as mm
    y = _signal(3_000_000, 48000, 6)
    a = mm.resample(y, 48000, 22050)
    assert np.array_equal(a, mm.resample(y, 48000, 22050))
    k = int(np.random.default_rng(2).integers(10, len(a) // 147 - 20))
    lo = 320 * (k - 4)                  # 1,280 samples before output 147 k's window
    ref = g.resample(y[lo:320 * k + 3000].astype(np.float64), 48000, 22050, table=_table32())
    got = a[147 * k:147 * k + 1000]
    assert np.abs(got - ref[588:1588]).max() <= 2e-6 * 140 * np.abs(y).max()


def test_device_entry_matches_host_entry():
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    for src, dst in [(44100, 22050), (16000, 22050), (22050, 22050)]:
        y = _signal(40000, src, 3)
        host = mm.resample(y, src, dst)
        n = lib.mm_resample_length(len(y), src, dst)
        assert n == host.size
        dX = torch.from_numpy(y.copy()).cuda()
        dY = torch.full((n,), float("nan"), device="cuda")
        s = lib.mm_resample_device(None, C.c_void_p(dX.data_ptr()), len(y), src, dst, C.c_void_p(dY.data_ptr()),
                                   C.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert s == 0
        torch.cuda.synchronize()
        assert np.array_equal(dY.cpu().numpy(), host)

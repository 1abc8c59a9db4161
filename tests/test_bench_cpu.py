"""bench.py's committed HBM-traffic summaries (rocprofv3 FETCH_SIZE and
WRITE_SIZE, copy-calibrated) belong to the workloads bench.py runs, so the
`traffic` fields of its JSON line are filled: a summary taken on another
workload (other algorithmic bytes) reads as None. Each measured figure is at
least the algorithmic bytes and within 1.2 x of them."""
import json
import math

import numpy as np
import pytest


def _lens():
    from metalmom_amd.workloads import uneven_lengths
    _, lens = uneven_lengths(256, 22050, 1)
    return [int(v) for v in lens]


@pytest.mark.parametrize("src,dst,name", [(44100, 22050, "RESAMPLE_TRAFFIC_FILE"),
                                          (48000, 22050, "RESAMPLE_PHASE_TRAFFIC_FILE")])
def test_resample_traffic_summaries_match_the_bench_workload(src, dst, name):
    import bench
    n = sum(_lens())                                  # the uneven batch joined into one clip
    m = math.ceil(n * (dst / src))                    # Resample.swift's output length
    path = getattr(bench, name)
    alg = 4 * (n + m)
    t = bench.measured_traffic(alg, path)
    assert t is not None, (path, alg, json.load(open(path))["algorithmic_bytes_per_launch"])
    assert alg <= t <= 1.2 * alg


def test_fma_roofline_prices_two_flops_per_tap():
    import bench
    r = bench._fma_roofline(10 ** 12, 1000.0)         # 1e12 taps in one second
    assert r["achieved"] == pytest.approx(2.0) and r["unit"] == "TFLOP/s"
    assert r["frac"] == pytest.approx(2.0 / 157.3) and r["frac_of_unpacked_peak"] == pytest.approx(4.0 / 157.3)


def test_contraction_flop_counts():
    """Executed MFMA FLOPs of one n_fft 2048 / 128-mel / 13-MFCC BF16 tile:
    42 chunks of 32 bins (16x16x32 MFMAs), 16 blocks x 3 DCT MFMAs
    (16x16x16) and the matrix-core FFT's 48 16x16x32 MFMAs per frame for 16
    frames (the FP32 kernel and n_fft 4096 keep the FFT on the VALU); the
    FFT's 2.5 N log2 N; useful FLOPs from the filterbank's 2,018 nonzeros
    and the 128 x 13 DCT."""
    import bench
    fft_tile = 16 * 48 * 16384
    assert bench.mfma_flops_per_tile(22050, 2048, 128, 13, "bf16") == 42 * 16384 + 16 * 3 * 8192 + fft_tile
    assert bench.mfma_flops_per_tile(22050, 2048, 128, 13, "f16") == 42 * 16384 + 16 * 3 * 8192 + fft_tile
    assert bench.mfma_flops_per_tile(22050, 2048, 128, 13, "f32") == 80 * 8192 + 16 * 4 * 2048
    assert bench.mfma_flops_per_tile(44100, 4096, 256, 20, "f16") < fft_tile
    assert bench.fft_flops(2048) == 56320
    assert bench.useful_mfma_flops_per_frame(22050, 2048, 128, 13) == 2 * 2018 + 2 * 128 * 13
    r = bench.kernel_roofline(1000, 2_100_000, 2048, 63, 1_081_344, 7364, "bf16", 1.0)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s"
    assert r["achieved"] == pytest.approx(1000 * 56320 / 1e-3 / 1e12)
    assert r["hbm"]["achieved"] == pytest.approx(2.1) and r["mfma"]["peak"] == 2500.0


def test_bench_fp64_reference_matches_oracle():
    """The bench's float64 error reference equals the oracle's STFT
    magnitude, mel and MFCC."""
    import bench
    from oracle import spectral as o
    y = np.random.default_rng(3).standard_normal(30000).astype(np.float32)
    mag, mel, mf = bench.fp64_reference(y)
    np.testing.assert_allclose(mag, o.stft_magnitude(y, 2048, 512), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(mel, o.melspectrogram(y, 22050, 2048, 512), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(mf, o.mfcc(y, 22050, 13, 2048, 512), rtol=1e-9, atol=1e-9)

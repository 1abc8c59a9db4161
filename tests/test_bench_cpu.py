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

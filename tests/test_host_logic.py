"""Host-side work description of the fused kernel, checked without a GPU.

Builds tests/host/meta_check.cpp against the built libmetalmom.so and runs it:
the per-tile prefetch records (which frames may be fetched by the plain
interior load, and from which packed-signal address) and the assignment of
8-mel filterbank blocks to wavefronts (every block once, slot bounds, busy
waves form a prefix, packed fields decode). Both describe device memory
accesses, so a mistake here is an out-of-bounds read on the GPU.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

LIB_DIR = os.path.join(ROOT, "metalmom_amd")


def test_tile_records_and_block_assignment(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    if not os.path.exists(os.path.join(LIB_DIR, "libmetalmom.so")):
        pytest.skip("libmetalmom.so not built")
    exe = str(tmp_path / "meta_check")
    src = os.path.join(ROOT, "tests", "host", "meta_check.cpp")
    subprocess.run([hipcc, "-std=c++17", "-O1", "-I" + ROOT, "-o", exe, src, "-L" + LIB_DIR, "-lmetalmom",
                    "-Wl,-rpath," + LIB_DIR], check=True, capture_output=True, cwd=str(tmp_path))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip() == "ok"


def test_serving_schedule_backlog_and_types():
    """make_schedule: Poisson arrivals at a rate, or a backlog of n requests
    all at t = 0 (the saturation point), restricted to the given types."""
    import numpy as np
    from metalmom_amd.serve import make_schedule
    s = make_schedule(float("inf"), 1.0, seed=4, kinds=("mfcc",), n_requests=50)
    assert len(s) == 50 and all(t == 0.0 and k == "mfcc" for t, k, _, _, _ in s)
    p = make_schedule(1000.0, 0.5, seed=4)
    ts = np.array([t for t, *_ in p])
    assert len(p) == 500 and np.all(np.diff(ts) > 0) and {k for _, k, *_ in p} == {"stft", "mel", "mfcc"}
    for _, _, sr, n, off in p:
        assert 0.5 * sr - 1 <= n <= 300 * sr + 1 and 0 <= off and off + n <= 330 * sr

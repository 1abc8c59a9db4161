"""C-ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, keeps the reference struct layouts, and rejects bad
arguments with the reference status codes before touching the device
(Bridge.swift:53-68, 277-279, 326-331, 418-423; BridgeErrorTests.swift)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "metalmom.h")


def _lib():
    from metalmom_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libmetalmom.so is not built (python -m metalmom_amd.build)")
    return _native.load()


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_drop_in_set():
    names = header_functions()
    for n in ["mm_init", "mm_destroy", "mm_stft", "mm_power_to_db", "mm_amplitude_to_db",
              "mm_mel_spectrogram", "mm_mfcc", "mm_buffer_free"]:
        assert n in names


def test_library_exports_every_header_symbol():
    lib = _lib()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    from metalmom_amd._native import SIGNATURES
    assert set(header_functions()) == set(SIGNATURES), "ctypes declarations out of sync with the header"


def test_struct_layouts_match_reference():
    from metalmom_amd._native import MMBuffer, MMSTFTParams
    # metalmom.h:14-28: float* + int64[8] + int32 + int32 + int64; four int32
    assert C.sizeof(MMBuffer) == 8 + 64 + 4 + 4 + 8
    assert MMBuffer.count.offset == 80
    assert C.sizeof(MMSTFTParams) == 16


def test_status_codes_and_double_destroy():
    lib = _lib()
    from metalmom_amd._native import MMBuffer, MMSTFTParams
    ctx = lib.mm_init()
    assert ctx
    sig = np.zeros(4096, np.float32)
    p = MMSTFTParams(2048, 512, 2048, 1)
    out = MMBuffer()
    fp = sig.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.mm_stft(None, fp, 4096, 22050, C.byref(p), C.byref(out)) == -1      # null ctx
    assert lib.mm_stft(ctx, None, 4096, 22050, C.byref(p), C.byref(out)) == -1     # null signal
    assert lib.mm_stft(ctx, fp, 0, 22050, C.byref(p), C.byref(out)) == -1          # empty
    assert lib.mm_stft(ctx, fp, 4096, 22050, None, C.byref(out)) == -1             # null params
    assert lib.mm_stft(ctx, fp, 4096, 22050, C.byref(p), None) == -1               # null out
    bad = MMSTFTParams(1000, 512, 1000, 1)
    assert lib.mm_stft(ctx, fp, 4096, 22050, C.byref(bad), C.byref(out)) == -1     # n_fft not 2^k
    assert not out.data
    assert lib.mm_power_to_db(None, None, 10, 1.0, 1e-10, 80.0, C.byref(out)) == -1
    assert lib.mm_power_to_db(None, fp, 0, 1.0, 1e-10, 80.0, C.byref(out)) == -1
    assert lib.mm_mel_spectrogram(None, fp, 4096, 22050, 2048, 512, 2048, 1, 2.0, 128, 0.0, 0.0,
                                  C.byref(out)) == -1
    assert lib.mm_mfcc(ctx, fp, -5, 22050, 13, 2048, 512, 2048, 128, 0.0, 0.0, 1, C.byref(out)) == -1
    lib.mm_destroy(ctx)
    lib.mm_destroy(ctx)        # second destroy is a no-op (BridgeErrorTests.swift:71-84)
    lib.mm_buffer_free(None)   # null-safe


def test_non_power_of_two_mel_returns_empty_result():
    """The reference core returns [n_mels, 0] for a non-power-of-two n_fft."""
    lib = _lib()
    from metalmom_amd._native import MMBuffer
    ctx = lib.mm_init()
    sig = np.zeros(4096, np.float32)
    out = MMBuffer()
    s = lib.mm_mel_spectrogram(ctx, sig.ctypes.data_as(C.POINTER(C.c_float)), 4096, 22050, 1000, 256,
                               1000, 1, 2.0, 40, 0.0, 0.0, C.byref(out))
    assert s == 0 and out.ndim == 2 and out.shape[0] == 40 and out.shape[1] == 0 and out.count == 0
    assert not out.data
    lib.mm_destroy(ctx)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU status code")
def test_compute_without_gpu_reports_unavailable():
    lib = _lib()
    from metalmom_amd._native import MMBuffer, MMSTFTParams
    ctx = lib.mm_init()
    sig = np.zeros(4096, np.float32)
    out = MMBuffer()
    s = lib.mm_stft(ctx, sig.ctypes.data_as(C.POINTER(C.c_float)), 4096, 22050,
                    C.byref(MMSTFTParams(2048, 512, 2048, 1)), C.byref(out))
    assert s == -2
    lib.mm_destroy(ctx)

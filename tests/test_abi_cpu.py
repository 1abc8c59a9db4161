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


def test_istft_shapes_match_the_oracle():
    """mm_istft_length follows STFT.swift:703-704, 720-724, 832-858 (centre trim,
    length, untrimmed zeros for a non-power-of-two n_fft) and rejects what
    mm_istft rejects."""
    from oracle import spectral as o
    lib = _lib()
    for F, T, hop, center, length in [(1025, 45, 512, 1, 0), (1025, 45, 512, 1, 22050), (1025, 45, 512, 0, 0),
                                      (1025, 1, 512, 1, 0), (513, 10, 100, 0, 7), (4, 3, 2, 1, 0),
                                      (2, 3, 0, 1, 0), (2, 3, 0, 0, 0), (8193, 2, 4096, 1, 0), (3, 5, 1, 1, 3)]:
        want = o.istft_length(F, T, hop, bool(center), length if length > 0 else None)
        assert lib.mm_istft_length(F, T, hop, center, length) == want, (F, T, hop, center, length)
    assert lib.mm_istft_length(0, 3, 1, 1, 0) == -1          # no bins
    assert lib.mm_istft_length(5, 0, 1, 1, 0) == -1          # no frames
    assert lib.mm_istft_length(5, 3, -1, 1, 0) == -1         # negative hop
    assert lib.mm_istft_length(16385, 2, 4096, 1, 0) == -1   # n_fft 32768: beyond the kernels


def test_istft_validation_and_host_only_results():
    """Bridge.swift:157-162 argument checks, the stft_count / shape check, the
    empty result (data NULL, shape all zero, Bridge.swift:190-201) and the
    zero buffer of a non-power-of-two n_fft, none of which needs the GPU."""
    lib = _lib()
    from metalmom_amd._native import MMBuffer
    ctx = lib.mm_init()
    fp = C.POINTER(C.c_float)
    S = np.zeros((4, 3), np.complex64)
    flat = np.ascontiguousarray(S).view(np.float32)
    out = MMBuffer()
    assert lib.mm_istft(None, flat.ctypes.data_as(fp), flat.size, 4, 3, 22050, 2, 6, 1, 0, C.byref(out)) == -1
    assert lib.mm_istft(ctx, None, flat.size, 4, 3, 22050, 2, 6, 1, 0, C.byref(out)) == -1
    assert lib.mm_istft(ctx, flat.ctypes.data_as(fp), 0, 4, 3, 22050, 2, 6, 1, 0, C.byref(out)) == -1
    assert lib.mm_istft(ctx, flat.ctypes.data_as(fp), flat.size - 2, 4, 3, 22050, 2, 6, 1, 0, C.byref(out)) == -1
    assert lib.mm_istft(ctx, flat.ctypes.data_as(fp), flat.size, 4, 3, 22050, -1, 6, 1, 0, C.byref(out)) == -1
    assert lib.mm_istft(ctx, flat.ctypes.data_as(fp), flat.size, 4, 3, 22050, 2, 6, 1, 0, None) == -1
    # n_fft 6 (not a power of two): 6 + 2 * 2 zeros, untrimmed
    s = lib.mm_istft(ctx, flat.ctypes.data_as(fp), flat.size, 4, 3, 22050, 2, 6, 1, 0, C.byref(out))
    assert s == 0 and out.ndim == 1 and out.shape[0] == 10 and out.count == 10
    assert not np.ctypeslib.as_array(out.data, shape=(10,)).any()
    lib.mm_buffer_free(C.byref(out))
    # one 2048-point frame with centre trim: nothing left
    S1 = np.ones((1025, 1), np.complex64)
    f1 = np.ascontiguousarray(S1).view(np.float32)
    s = lib.mm_istft(ctx, f1.ctypes.data_as(fp), f1.size, 1025, 1, 22050, 512, 2048, 1, 0, C.byref(out))
    assert s == 0 and out.ndim == 1 and out.count == 0 and not out.data and out.shape[0] == 0
    lib.mm_destroy(ctx)


def test_onset_strength_validation_and_zero_columns():
    """Bridge.swift:1334-1336 checks the signal and out only (ctx may be NULL);
    sr <= 0 and n_mels < 1 are rejected; no STFT frame or a single frame gives
    one zero column (OnsetDetection.swift:63-71) without any GPU work."""
    lib = _lib()
    from metalmom_amd._native import MMBuffer
    fp = C.POINTER(C.c_float)
    y = np.ones(4096, np.float32)
    out = MMBuffer()
    assert lib.mm_onset_strength(None, None, 4096, 22050, 2048, 512, 128, 0.0, 0.0, 1, 1, C.byref(out)) == -1
    assert lib.mm_onset_strength(None, y.ctypes.data_as(fp), 0, 22050, 2048, 512, 128, 0.0, 0.0, 1, 1,
                                 C.byref(out)) == -1
    assert lib.mm_onset_strength(None, y.ctypes.data_as(fp), 4096, 22050, 2048, 512, 128, 0.0, 0.0, 1, 1, None) == -1
    assert lib.mm_onset_strength(None, y.ctypes.data_as(fp), 4096, 0, 2048, 512, 128, 0.0, 0.0, 1, 1,
                                 C.byref(out)) == -1
    assert lib.mm_onset_strength(None, y.ctypes.data_as(fp), 4096, 22050, 2048, 512, 0, 0.0, 0.0, 1, 1,
                                 C.byref(out)) == -1
    for n_fft, n, center, agg, rows in [(1000, 4096, 1, 1, 1), (1000, 4096, 1, 0, 128), (2048, 2048, 0, 1, 1),
                                        (2048, 100, 0, 0, 128)]:
        s = lib.mm_onset_strength(None, y.ctypes.data_as(fp), n, 22050, n_fft, 512, 128, 0.0, 0.0, center, agg,
                                  C.byref(out))
        assert s == 0 and out.ndim == 2 and out.shape[0] == rows and out.shape[1] == 1 and out.count == rows
        assert not np.ctypeslib.as_array(out.data, shape=(rows,)).any()
        lib.mm_buffer_free(C.byref(out))


def test_chroma_precomputed_spectrogram_path_matches_goldens(signal_440):
    """The S= path multiplies a precomputed power spectrogram by the chroma
    filterbank in NumPy, with tuning, as the reference wrapper does
    (feature.py:313-323); with librosa's estimated tunings of the golden
    signal (0.02 at n_fft 2048, -0.07 at 1024) it reproduces the goldens."""
    from conftest import load_golden
    from oracle import spectral as o
    import metalmom_amd as mm
    for name, n_fft, tuning, norm in [("chroma_stft_440hz_no_norm", 2048, 0.02, None),
                                      ("chroma_stft_440hz_default", 2048, 0.02, np.inf),
                                      ("chroma_stft_440hz_1024", 1024, -0.07, None)]:
        G = load_golden(name)
        P = o.power_spectrogram(signal_440, n_fft, 512).astype(np.float32)
        C = mm.feature.chroma_stft(S=P, sr=22050, tuning=tuning, norm=norm)
        assert C.shape == G.shape and np.abs(C - G).max() <= 1e-5 * np.abs(G).max()


# -- mm_resample (Bridge.swift:1087-1104, Resample.swift:12-96) ----------------

def test_resample_length_and_validation():
    lib = _lib()
    from oracle import ingest as g
    for n, a, b in [(44100, 44100, 22050), (22050, 22050, 44100), (1000, 48000, 22050), (7, 16000, 16000),
                    (1, 44100, 8000), (0, 44100, 22050)]:
        assert lib.mm_resample_length(n, a, b) == g.resample_length(n, a, b)
    assert lib.mm_resample_length(10, 0, 22050) == -1 and lib.mm_resample_length(-1, 1, 2) == -1
    from metalmom_amd._native import MMBuffer, float_ptr
    y = np.ones(100, np.float32)
    out = MMBuffer()
    assert lib.mm_resample(None, None, 100, 44100, 22050, C.byref(out)) == -1
    assert lib.mm_resample(None, float_ptr(y), 0, 44100, 22050, C.byref(out)) == -1
    assert lib.mm_resample(None, float_ptr(y), 100, 0, 22050, C.byref(out)) == -1
    assert lib.mm_resample(None, float_ptr(y), 100, 44100, -5, C.byref(out)) == -1
    assert lib.mm_resample(None, float_ptr(y), 100, 44100, 22050, None) == -1


def test_resample_identity_needs_no_gpu():
    """test_resample.py:5-8 / ResampleTests.swift:5-10: equal rates return the
    signal itself."""
    import metalmom_amd as mm
    from metalmom_amd._native import MMBuffer, float_ptr
    y = np.sin(np.arange(22050, dtype=np.float32) * 440.0 * 2 * np.pi / 22050)
    assert np.array_equal(mm.resample(y, 22050, 22050), y)
    lib = _lib()
    out = MMBuffer()
    assert lib.mm_resample(None, float_ptr(y), len(y), 22050, 22050, C.byref(out)) == 0
    from metalmom_amd._buffer import buffer_to_numpy
    r = buffer_to_numpy(out)
    assert r.shape == y.shape and np.array_equal(r, y)


def test_contrast_validation_needs_no_gpu():
    """Bridge.swift:711-716 (ctx, signal, length, out); n_bands < 1 traps in
    the reference and sr <= 0 has no bin width: -1 here. A non-power-of-two
    n_fft has no STFT frames: [n_bands + 1, 0]."""
    lib = _lib()
    from metalmom_amd._native import MMBuffer, float_ptr
    y = np.ones(4096, np.float32)
    out = MMBuffer()
    assert lib.mm_spectral_contrast(None, float_ptr(y), 4096, 22050, 2048, 512, 2048, 1, 6, 200.0, C.byref(out)) == -1
    ctx = lib.mm_init()
    try:
        assert lib.mm_spectral_contrast(ctx, float_ptr(y), 4096, 22050, 2048, 512, 2048, 1, 0, 200.0,
                                        C.byref(out)) == -1
        assert lib.mm_spectral_contrast(ctx, float_ptr(y), 4096, 0, 2048, 512, 2048, 1, 6, 200.0, C.byref(out)) == -1
        assert lib.mm_spectral_contrast(ctx, float_ptr(y), 4096, 22050, 1000, 512, 1000, 1, 6, 200.0,
                                        C.byref(out)) == 0
        assert out.ndim == 2 and out.shape[0] == 7 and out.shape[1] == 0 and out.count == 0
        lib.mm_buffer_free(C.byref(out))
    finally:
        lib.mm_destroy(ctx)


def test_device_entries_reject_bad_arguments_without_gpu():
    """The additive device entries validate before touching the device."""
    lib = _lib()
    buf = C.c_void_p(1 << 20)   # never dereferenced: every call below fails validation first
    assert lib.mm_resample_device(None, None, 100, 44100, 22050, buf, None) == -1
    assert lib.mm_resample_device(None, buf, 0, 44100, 22050, buf, None) == -1
    assert lib.mm_resample_device(None, buf, 100, 0, 22050, buf, None) == -1
    assert lib.mm_resample_device(None, buf, 100, 44100, 22050, None, None) == -1
    assert lib.mm_spectral_contrast_device(None, buf, 4096, 22050, 2048, 512, 2048, 1, 6, 200.0, buf, None,
                                           None) == -1
    ctx = lib.mm_init()
    try:
        for bad in [dict(n_fft=1000), dict(n_bands=0), dict(n_bands=63), dict(sr=0)]:
            a = dict(sr=22050, n_fft=2048, n_bands=6)
            a.update(bad)
            assert lib.mm_spectral_contrast_device(ctx, buf, 4096, a["sr"], a["n_fft"], 512, 0, 1, a["n_bands"],
                                                   200.0, buf, None, None) == -1, bad
        assert lib.mm_spectral_descriptor_device(ctx, buf, 4096, 22050, 2048, 512, 0, 1, 9, 0.0, buf, None,
                                                 None) == -1   # no such descriptor kind
    finally:
        lib.mm_destroy(ctx)

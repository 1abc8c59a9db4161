"""ctypes binding of libmetalmom.so (the C ABI in include/metalmom.h).

Mirrors the reference loader ``python/metalmom/_native.py`` (cffi cdef of
``metalmom.h`` + ``dlopen`` of the built library). cffi is not a dependency
here, so the same declarations are made with ctypes. The library is looked up
in-tree next to this file (``libmetalmom.so``, built by ``metalmom_amd.build``)
or at ``$METALMOM_AMD_LIB``; a missing library raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

__all__ = ["lib", "MMBuffer", "MMSTFTParams", "MMSpecParams", "LIB_PATH", "load"]

LIB_PATH = os.environ.get("METALMOM_AMD_LIB",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmetalmom.so"))

MM_OK = 0
MM_ERR_INVALID_INPUT = -1
MM_ERR_METAL_UNAVAILABLE = -2
MM_ERR_ALLOC_FAILED = -3
MM_ERR_INTERNAL = -4

KIND_STFT_MAG, KIND_STFT_POWER, KIND_STFT_COMPLEX, KIND_MEL, KIND_MFCC, KIND_CHROMA = range(6)
PREC_F32, PREC_BF16, PREC_F16 = range(3)
PCM_U8, PCM_S16, PCM_S24, PCM_S32, PCM_F32, PCM_F64 = range(1, 7)   # MM_PCM_*
INPUT_F32, INPUT_F16 = range(2)
PAD_CONSTANT, PAD_REFLECT = range(2)


class MMBuffer(C.Structure):
    """``MMBuffer`` (reference ``metalmom.h:14-20``)."""
    _fields_ = [("data", C.POINTER(C.c_float)), ("shape", C.c_int64 * 8), ("ndim", C.c_int32),
                ("dtype", C.c_int32), ("count", C.c_int64)]


class MMSTFTParams(C.Structure):
    """``MMSTFTParams`` (reference ``metalmom.h:23-28``)."""
    _fields_ = [("n_fft", C.c_int32), ("hop_length", C.c_int32), ("win_length", C.c_int32),
                ("center", C.c_int32)]


class MMSpecParams(C.Structure):
    """``MMSpecParams`` (additive; full request description)."""
    _fields_ = [("kind", C.c_int32), ("sample_rate", C.c_int32), ("n_fft", C.c_int32),
                ("hop_length", C.c_int32), ("win_length", C.c_int32), ("center", C.c_int32),
                ("pad_mode", C.c_int32), ("power", C.c_float), ("n_mels", C.c_int32),
                ("f_min", C.c_float), ("f_max", C.c_float), ("n_mfcc", C.c_int32),
                ("top_db", C.c_float), ("precision", C.c_int32), ("input_dtype", C.c_int32)]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I64 = C.c_int64
_I32 = C.c_int32
_BUF = C.POINTER(MMBuffer)

# name -> (restype, argtypes); every symbol declared in include/metalmom.h
SIGNATURES = {
    "mm_init": (_P, []),
    "mm_destroy": (None, [_P]),
    "mm_stft": (_I32, [_P, _F, _I64, _I32, C.POINTER(MMSTFTParams), _BUF]),
    "mm_amplitude_to_db": (_I32, [_P, _F, _I64, C.c_float, C.c_float, C.c_float, _BUF]),
    "mm_power_to_db": (_I32, [_P, _F, _I64, C.c_float, C.c_float, C.c_float, _BUF]),
    "mm_mel_spectrogram": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, C.c_float, _I32,
                                  C.c_float, C.c_float, _BUF]),
    "mm_mfcc": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _I32, C.c_float, C.c_float,
                       _I32, _BUF]),
    "mm_buffer_free": (None, [_BUF]),
    "mm_stft_complex": (_I32, [_P, _F, _I64, _I32, C.POINTER(MMSTFTParams), _BUF]),
    "mm_istft": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _I32, _I64, _BUF]),
    "mm_spectral_centroid": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _BUF]),
    "mm_spectral_bandwidth": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, C.c_float, _BUF]),
    "mm_spectral_rolloff": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, C.c_float, _BUF]),
    "mm_spectral_contrast": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _I32, C.c_float, _BUF]),
    "mm_spectral_flatness": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _BUF]),
    "mm_chroma_stft": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, _I32, _I32, _BUF]),
    "mm_onset_strength": (_I32, [_P, _F, _I64, _I32, _I32, _I32, _I32, C.c_float, C.c_float, _I32, _I32, _BUF]),
    "mm_istft_length": (_I64, [_I32, _I32, _I32, _I32, _I64]),
    "mm_resample": (_I32, [_P, _F, _I64, _I32, _I32, _BUF]),
    "mm_get_duration": (_I32, [_P, C.c_char_p, C.POINTER(C.c_float)]),
    "mm_get_sample_rate": (_I32, [_P, C.c_char_p, C.POINTER(_I32)]),
    "mm_load": (_I32, [_P, C.c_char_p, _I32, _I32, C.c_float, C.c_float, _BUF, C.POINTER(_I32)]),
    "mm_resample_length": (_I64, [_I64, _I32, _I32]),
    "mm_resample_device": (_I32, [_P, _P, _I64, _I32, _I32, _P, _P]),
    "mm_spectral_descriptor_device": (_I32, [_P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _I32, C.c_float, _P,
                                             C.POINTER(_I64), _P]),
    "mm_spectral_contrast_device": (_I32, [_P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _I32, C.c_float, _P,
                                           C.POINTER(_I64), _P]),
    "mm_onset_strength_device": (_I32, [_P, _P, _I64, _I32, _I32, _I32, _I32, C.c_float, C.c_float, _I32, _I32,
                                        _P, C.POINTER(_I64), C.POINTER(_I64), _P]),
    "mm_istft_device": (_I32, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I64, _P, _P]),
    "mm_spectral": (_I32, [_P, C.POINTER(MMSpecParams), _F, _I64, _BUF]),
    "mm_spectral_device_noalloc": (_I32, [C.POINTER(MMSpecParams), _P, _I64, _P, C.POINTER(_I64), C.POINTER(_I64),
                                          _P]),
    "mm_spectral_batch": (_I32, [_P, C.POINTER(MMSpecParams), C.POINTER(_F), C.POINTER(_I64), _I32,
                                 _BUF]),
    "mm_batch_create": (_I32, [_P, C.POINTER(MMSpecParams), C.POINTER(_I64), _I32, C.POINTER(_P)]),
    "mm_batch_info": (_I32, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I32)]),
    "mm_batch_clip": (_I32, [_P, _I32, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64),
                             C.POINTER(_I64)]),
    "mm_batch_run": (_I32, [_P, _P, _P, _P]),
    "mm_batch_run_graph": (_I32, [_P, _P, _P, _P]),
    "mm_batch_set_grid": (_I32, [_P, _I32]),
    "mm_batch_set_shape": (_I32, [_P, _I32]),
    "mm_batch_launch_shape": (_I32, [_P, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32),
                                     C.POINTER(_I32)]),
    "mm_batch_autotune": (_I32, [_P, _P, _P, _I32, C.POINTER(_I32), C.POINTER(C.c_float)]),
    "mm_batch_time": (_I32, [_P, _P, _P, _P, _I32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "mm_batch_destroy": (_I32, [_P]),
    "mm_device_info": (_I32, [C.c_char_p, _I32, C.c_char_p, _I32, C.POINTER(_I32)]),
    "mm_server_create": (_I32, [_P, _I64, _I32, C.POINTER(_P)]),
    "mm_server_submit": (_I64, [_P, C.POINTER(MMSpecParams), _F, _I64]),
    "mm_server_submit_pcm": (_I64, [_P, C.POINTER(MMSpecParams), _P, _I64, _I32, _I32, _I32, _I32]),
    "mm_server_poll": (_I32, [_P, _I64, _BUF, C.POINTER(C.c_double)]),
    "mm_server_wait": (_I32, [_P, _I64, _BUF, C.POINTER(C.c_double)]),
    "mm_server_stats": (_I32, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64)]),
    "mm_server_device_bytes": (_I32, [_P, C.POINTER(_I64)]),
    "mm_server_destroy": (_I32, [_P]),
}


def _share_torch_hip_runtime():
    """One HIP runtime per process.

    libmetalmom.so links ``libamdhip64.so.7``. PyTorch-ROCm ships its own copy
    whose SONAME is also ``libamdhip64.so.7`` but which its libraries request
    by the file name ``libamdhip64.so``. If this library were loaded first,
    a later ``import torch`` would map a second HIP runtime and lose the GPU.
    Importing torch first makes the dynamic linker resolve this library's
    dependency to torch's already-loaded runtime. Set
    ``METALMOM_AMD_STANDALONE_RUNTIME=1`` to skip this in torch-free processes.
    """
    if os.environ.get("METALMOM_AMD_STANDALONE_RUNTIME") == "1":
        return
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads the library and declares every entry point; raises if absent."""
    _share_torch_hip_runtime()
    if not os.path.exists(path):
        raise RuntimeError(
            f"libmetalmom.so not found at {path}. Build it with `python -m metalmom_amd.build`.")
    handle = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


class _LazyLib:
    """Loads libmetalmom.so on first attribute access."""

    _lib = None

    def __getattr__(self, name):
        if _LazyLib._lib is None:
            _LazyLib._lib = load()
        return getattr(_LazyLib._lib, name)


lib = _LazyLib()


def float_ptr(arr) -> "C.POINTER(C.c_float)":
    """Pointer to the first element of a contiguous float32 numpy array."""
    return arr.ctypes.data_as(_F)

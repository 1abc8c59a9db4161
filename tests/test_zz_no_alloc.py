"""The allocation-free entry point mm_spectral_device_noalloc against the
allocating mm_spectral of the same configuration: the same kernels, so the
same bits, on two calls in a row (the second with a shorter clip, so the
shared device work list holds the first call's longer list when it starts),
and its size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, p, x, out):
    import torch
    rows, frames = C.c_int64(), C.c_int64()
    st = torch.cuda.current_stream().cuda_stream
    s = lib.mm_spectral_device_noalloc(C.byref(p), C.c_void_p(x.data_ptr()), x.numel(), C.c_void_p(out.data_ptr()),
                                       C.byref(rows), C.byref(frames), C.c_void_p(st))
    return s, rows.value, frames.value


@pytest.mark.parametrize("kind,prec", [("mfcc", "f32"), ("mfcc", "bf16"), ("mel", "f32"), ("stft", "f32")])
def test_noalloc_equals_mm_spectral(kind, prec):
    import torch
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    from metalmom_amd.core import make_params
    rng = np.random.default_rng(17)
    long_y = rng.standard_normal(22050 * 40).astype(np.float32) * 0.1
    short_y = rng.standard_normal(22050 * 3 + 123).astype(np.float32) * 0.1
    p = make_params(kind, 22050, 2048, 512, None, True, "constant", 2.0, 128, 0.0, None, 13, 80.0, prec, "f32")
    refs = [mm.spectral(y, kind=kind, n_mfcc=13, precision=prec) for y in (long_y, short_y)]   # tables resident
    xs = [torch.from_numpy(y).cuda() for y in (long_y, short_y)]
    outs = [torch.full((r.size + 64,), float("nan"), device="cuda") for r in refs]
    # the HIP runtime places the module's device array in memory on its first use
    assert _call(lib, p, xs[0], outs[0])[0] == 0
    for x, out, ref in zip(xs, outs, refs):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        s, rows, frames = _call(lib, p, x, out)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0   # the call took no device memory
        assert s == 0 and (rows, frames) == ref.shape
        got = out[:ref.size].cpu().numpy().reshape(ref.shape)
        assert np.array_equal(got, ref)
        assert torch.isnan(out[ref.size:]).all()   # nothing written past the result


def test_noalloc_limits():
    import torch
    from metalmom_amd._native import lib
    from metalmom_amd.core import make_params
    p = make_params("mfcc", 22050, 2048, 512, None, True, "constant", 2.0, 128, 0.0, None, 13, 80.0, "f32", "f32")
    x = torch.zeros(16, device="cuda")
    out = torch.zeros(16, device="cuda")
    # about 40,000 tiles of 16 frames: past the 1 MiB work list (rejected before any access)
    rows, frames = C.c_int64(), C.c_int64()
    s = lib.mm_spectral_device_noalloc(C.byref(p), C.c_void_p(x.data_ptr()), 512 * 16 * 40000,
                                       C.c_void_p(out.data_ptr()), C.byref(rows), C.byref(frames), None)
    assert s == -1
    q = make_params("mfcc", 22050, 8192, 512, None, True, "constant", 2.0, 128, 0.0, None, 13, 80.0, "f32", "f32")
    s = lib.mm_spectral_device_noalloc(C.byref(q), C.c_void_p(x.data_ptr()), 16, C.c_void_p(out.data_ptr()),
                                       None, None, None)
    assert s == -1   # n_fft 8192 runs on the generic path, which needs its own scratch

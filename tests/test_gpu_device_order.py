"""Device-memory entry points on one context from different streams.

mm_onset_strength_device, mm_spectral_descriptor_device and
mm_spectral_contrast_device keep the clip's work list and an intermediate
spectrogram in the context's buffers and return before the GPU has run. A
call on another stream (or a host-buffer call) on the same context must not
rewrite those buffers while the earlier call's kernels still read them: each
result must equal the same call made on a fresh context and synchronised.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _onset(lib, ctx, x, out, stream):
    rows, cols = C.c_int64(), C.c_int64()
    s = lib.mm_onset_strength_device(ctx, C.c_void_p(x.data_ptr()), x.numel(), 22050, 2048, 512, 128, 0.0, 0.0,
                                     1, 1, C.c_void_p(out.data_ptr()), C.byref(rows), C.byref(cols),
                                     C.c_void_p(stream.cuda_stream))
    assert s == 0
    return cols.value


def _centroid(lib, ctx, x, out, stream):
    s = lib.mm_spectral_descriptor_device(ctx, C.c_void_p(x.data_ptr()), x.numel(), 22050, 2048, 512, 2048, 1, 0,
                                          0.0, C.c_void_p(out.data_ptr()), None, C.c_void_p(stream.cuda_stream))
    assert s == 0


def test_device_calls_on_two_streams_one_context():
    import torch
    from metalmom_amd._native import MMBuffer, MMSTFTParams, lib
    rng = np.random.default_rng(23)
    long_y = torch.from_numpy(rng.standard_normal(22050 * 120).astype(np.float32) * 0.1).cuda()
    short_y = torch.from_numpy(rng.standard_normal(22050 * 7 + 31).astype(np.float32) * 0.1).cuda()
    torch.cuda.synchronize()
    # references: one fresh context per call, synchronised
    refs = {}
    for name, y in (("long", long_y), ("short", short_y)):
        ctx = lib.mm_init()
        st = torch.cuda.Stream()
        o = torch.zeros(1 + y.numel() // 512 + 8, device="cuda")
        n = _onset(lib, ctx, y, o, st)
        c = torch.zeros(1 + y.numel() // 512, device="cuda")
        _centroid(lib, ctx, y, c, st)
        torch.cuda.synchronize()
        refs[name] = (o[:n].cpu().numpy().copy(), c.cpu().numpy().copy())
        lib.mm_destroy(ctx)
    # one context, alternating streams, no synchronisation in between
    ctx = lib.mm_init()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        oa = torch.zeros(1 + long_y.numel() // 512 + 8, device="cuda")
        ob = torch.zeros(1 + short_y.numel() // 512 + 8, device="cuda")
        ca = torch.zeros(1 + long_y.numel() // 512, device="cuda")
        cb = torch.zeros(1 + short_y.numel() // 512, device="cuda")
        torch.cuda.synchronize()
        na = _onset(lib, ctx, long_y, oa, sa)
        nb = _onset(lib, ctx, short_y, ob, sb)
        _centroid(lib, ctx, long_y, ca, sa)
        _centroid(lib, ctx, short_y, cb, sb)
        # a host-buffer call on the same context (its lane's own stream) in between
        hx = np.ones(4096, np.float32)
        hb = MMBuffer()
        assert lib.mm_stft(ctx, hx.ctypes.data_as(C.POINTER(C.c_float)), hx.size, 22050,
                           C.byref(MMSTFTParams(2048, 512, 2048, 1)), C.byref(hb)) == 0
        assert (hb.shape[0], hb.shape[1]) == (1025, 9)
        lib.mm_buffer_free(C.byref(hb))
        torch.cuda.synchronize()
        assert np.array_equal(oa[:na].cpu().numpy(), refs["long"][0]), rep
        assert np.array_equal(ob[:nb].cpu().numpy(), refs["short"][0]), rep
        assert np.array_equal(ca.cpu().numpy(), refs["long"][1]), rep
        assert np.array_equal(cb.cpu().numpy(), refs["short"][1]), rep
    lib.mm_destroy(ctx)

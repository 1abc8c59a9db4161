"""Device-resident batches: many uneven clips in one persistent launch per step.

A :class:`SpectralBatch` wraps an ``mm_batch`` plan (include/metalmom.h). The
signal of all clips is one packed device buffer (clip ``i`` at
``input_offset(i)``); the output is one device buffer where clip ``i`` owns the
row-major ``[rows, frames]`` block at ``output_offset(i)``. Device memory can
be any pointer; :meth:`pack` and :meth:`empty_output` use torch tensors as
plumbing for it.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from ._native import MMBuffer, lib
from .core import make_params


class SpectralBatch:
    def __init__(self, lengths: Sequence[int], kind="mfcc", sr=22050, n_fft=2048, hop_length=512,
                 win_length=None, center=True, pad_mode="constant", power=2.0, n_mels=128, fmin=0.0,
                 fmax=None, n_mfcc=13, top_db=80.0, precision="f32", input_dtype="f32"):
        self.params = make_params(kind, sr, n_fft, hop_length, win_length, center, pad_mode, power,
                                  n_mels, fmin, fmax, n_mfcc, top_db, precision, input_dtype)
        self.input_dtype = input_dtype
        lens = np.ascontiguousarray(lengths, dtype=np.int64)
        self.n_clips = len(lens)
        h = C.c_void_p()
        s = lib.mm_batch_create(None, C.byref(self.params),
                                lens.ctypes.data_as(C.POINTER(C.c_int64)), self.n_clips, C.byref(h))
        if s != 0:
            raise RuntimeError(f"mm_batch_create failed with status {s}")
        self._h = h
        ti, to, tf, nt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        lib.mm_batch_info(h, C.byref(ti), C.byref(to), C.byref(tf), C.byref(nt))
        self.total_input = ti.value
        self.total_output = to.value
        self.total_frames = tf.value
        self.n_tiles = nt.value
        self.clips = [self.clip(i) for i in range(self.n_clips)]

    def clip(self, i):
        """(input_offset, output_offset_floats, rows, frames) of clip ``i``."""
        a, b, r, f = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        s = lib.mm_batch_clip(self._h, i, C.byref(a), C.byref(b), C.byref(r), C.byref(f))
        if s != 0:
            raise RuntimeError(f"mm_batch_clip failed with status {s}")
        return a.value, b.value, r.value, f.value

    # -- execution ---------------------------------------------------------
    @staticmethod
    def _ptr(x):
        return C.c_void_p(x.data_ptr() if hasattr(x, "data_ptr") else int(x))

    def run(self, d_signal, d_out, stream=None, graph=False):
        """Enqueues one step (async). ``stream``: a HIP stream handle (int) or a
        torch stream; ``None`` uses the plan's own stream."""
        st = C.c_void_p(getattr(stream, "cuda_stream", stream) or 0) if stream is not None else None
        fn = lib.mm_batch_run_graph if graph else lib.mm_batch_run
        s = fn(self._h, self._ptr(d_signal), self._ptr(d_out), st)
        if s != 0:
            raise RuntimeError(f"mm_batch_run failed with status {s}")

    def set_grid(self, wg_per_cu: int):
        if lib.mm_batch_set_grid(self._h, int(wg_per_cu)) != 0:
            raise RuntimeError("mm_batch_set_grid failed")

    def set_shape(self, shape: int):
        """Workgroup shape of the fused kernel (1: 16 waves x 1 frame, 2: 8 x 1,
        3: 8 x 2 frames, two workgroups per CU; 4: 16 x 1 pipelined, one
        barrier per tile, 16-bit MFCC at n_fft 2048 only)."""
        s = lib.mm_batch_set_shape(self._h, int(shape))
        if s != 0:
            raise RuntimeError(f"mm_batch_set_shape failed with status {s}")

    def launch_shape(self) -> dict:
        """What the next step launches: shape, waves per workgroup, frames per
        tile, persistent workgroups and tiles per workgroup."""
        v = [C.c_int32() for _ in range(5)]
        s = lib.mm_batch_launch_shape(self._h, *[C.byref(x) for x in v])
        if s != 0:
            raise RuntimeError(f"mm_batch_launch_shape failed with status {s}")
        keys = ("shape", "waves_per_wg", "tile_frames", "grid", "tiles_per_wg")
        return dict(zip(keys, (x.value for x in v)))

    def autotune(self, d_signal, d_out, iters=3):
        """Project autotuner: times every (workgroup shape, persistent grid)
        pair on this batch and keeps the fastest; returns (workgroups per CU,
        step ms)."""
        best, ms = C.c_int32(), C.c_float()
        s = lib.mm_batch_autotune(self._h, self._ptr(d_signal), self._ptr(d_out), int(iters),
                                  C.byref(best), C.byref(ms))
        if s != 0:
            raise RuntimeError(f"mm_batch_autotune failed with status {s}")
        return best.value, ms.value

    def time(self, d_signal, d_out, iters=10, stream=None):
        """(fused kernel ms, whole step ms) averaged over ``iters`` runs,
        measured with HIP events on the stream the kernels run on."""
        k, t = C.c_float(), C.c_float()
        st = C.c_void_p(getattr(stream, "cuda_stream", stream) or 0) if stream is not None else None
        s = lib.mm_batch_time(self._h, self._ptr(d_signal), self._ptr(d_out), st, int(iters),
                              C.byref(k), C.byref(t))
        if s != 0:
            raise RuntimeError(f"mm_batch_time failed with status {s}")
        return k.value, t.value

    # -- torch plumbing ----------------------------------------------------
    def pack(self, clips, device="cuda"):
        """Packs host clips into one device tensor in plan order."""
        import torch
        dt = torch.float16 if self.input_dtype in ("f16", "float16") else torch.float32
        host = np.empty(self.total_input, dtype=np.float16 if dt == torch.float16 else np.float32)
        for i, c in enumerate(clips):
            off = self.clips[i][0]
            host[off:off + len(c)] = c
        return torch.from_numpy(host).to(device)

    def empty_output(self, device="cuda"):
        import torch
        return torch.empty(max(self.total_output, 1), dtype=torch.float32, device=device)

    def unpack(self, out) -> list:
        """Splits a (host or device) output buffer into per-clip arrays."""
        host = out.detach().cpu().numpy() if hasattr(out, "detach") else np.asarray(out)
        res = []
        cplx = self.params.kind == 2
        for (_, off, rows, frames) in self.clips:
            n = rows * frames * (2 if cplx else 1)
            a = host[off:off + n]
            res.append(a.view(np.complex64).reshape(rows, frames) if cplx else a.reshape(rows, frames))
        return res

    def close(self):
        if getattr(self, "_h", None):
            lib.mm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spectral_batch(clips, kind="mfcc", **kw):
    """Host-buffer batch (``mm_spectral_batch``): list of float32 clips in,
    list of arrays out, all clips in one persistent launch."""
    from .core import make_params as _mk
    arrs = [np.ascontiguousarray(c, dtype=np.float32) for c in clips]
    n = len(arrs)
    p = _mk(kind, kw.get("sr", 22050), kw.get("n_fft", 2048), kw.get("hop_length", 512),
            kw.get("win_length"), kw.get("center", True), kw.get("pad_mode", "constant"),
            kw.get("power", 2.0), kw.get("n_mels", 128), kw.get("fmin", 0.0), kw.get("fmax"),
            kw.get("n_mfcc", 13), kw.get("top_db", 80.0), kw.get("precision", "f32"), "f32")
    ptrs = (C.POINTER(C.c_float) * n)(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    lens = (C.c_int64 * n)(*[len(a) for a in arrs])
    outs = (MMBuffer * n)()
    ctx = lib.mm_init()
    try:
        s = lib.mm_spectral_batch(ctx, C.byref(p), ptrs, lens, n, outs)
        if s != 0:
            raise RuntimeError(f"mm_spectral_batch failed with status {s}")
        from ._buffer import buffer_to_numpy
        return [buffer_to_numpy(outs[i]) for i in range(n)]
    finally:
        lib.mm_destroy(ctx)

"""Request server (continuous batching on the GPU) and an open-loop load generator.

:class:`SpectralServer` wraps ``mm_server_*`` (include/metalmom.h): clips of
any request type are submitted from any thread, admitted into device steps by
HBM budget, batched per configuration into persistent launches, and returned
with their submission-to-completion latency.

:func:`open_loop` drives it with Poisson arrivals at a fixed rate: log-normal
clip lengths (median 5 s, sigma 1, clipped to 0.5-300 s), sample rate drawn
from {22050, 44100}, request type uniform over {STFT magnitude, mel 128,
MFCC 13}, n_fft 2048, hop 512. Clip content is noise + linear chirp, sliced
from a pre-generated buffer per sample rate so generation stays off the
arrival clock.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from ._buffer import buffer_to_numpy
from ._native import MMBuffer, lib
from .core import make_params


class SpectralServer:
    def __init__(self, hbm_budget_bytes: int = 8 << 30, max_batch: int = 256):
        h = C.c_void_p()
        s = lib.mm_server_create(None, int(hbm_budget_bytes), int(max_batch), C.byref(h))
        if s != 0:
            raise RuntimeError(f"mm_server_create failed with status {s}")
        self._h = h
        self._params = {}

    def params(self, kind, sr=22050, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13,
               precision="f32", **kw):
        key = (kind, sr, n_fft, hop_length, n_mels, n_mfcc, precision, tuple(sorted(kw.items())))
        p = self._params.get(key)
        if p is None:
            p = make_params(kind, sr, n_fft, hop_length, kw.get("win_length"), kw.get("center", True),
                            kw.get("pad_mode", "constant"), kw.get("power", 2.0), n_mels,
                            kw.get("fmin", 0.0), kw.get("fmax"), n_mfcc, kw.get("top_db", 80.0),
                            precision, "f32")
            self._params[key] = p
        return p

    def submit(self, y, params) -> int:
        y = np.ascontiguousarray(y, dtype=np.float32)
        rid = lib.mm_server_submit(self._h, C.byref(params), y.ctypes.data_as(C.POINTER(C.c_float)),
                                   len(y))
        if rid <= 0:
            raise RuntimeError(f"mm_server_submit failed with status {rid}")
        return rid

    def submit_pcm(self, pcm, channels, pcm_format, source_sr, params, mono=True) -> int:
        """Queues raw interleaved PCM (``pcm_format`` one of ``_native.PCM_*``,
        ``channels`` samples per frame, at ``source_sr``); the GPU converts it
        to one channel and resamples it to the analysis rate of ``params``."""
        buf = np.ascontiguousarray(pcm)
        width = {1: 1, 2: 2, 3: 3, 4: 4, 5: 4, 6: 8}[int(pcm_format)]
        frames = buf.nbytes // (width * channels)
        if buf.nbytes != frames * width * channels:
            raise ValueError("PCM buffer is not a whole number of frames")
        rid = lib.mm_server_submit_pcm(self._h, C.byref(params), C.c_void_p(buf.ctypes.data), frames,
                                       int(channels), int(pcm_format), int(source_sr), 1 if mono else 0)
        if rid <= 0:
            raise RuntimeError(f"mm_server_submit_pcm failed with status {rid}")
        return rid

    def wait(self, rid: int):
        """Blocks until request ``rid`` is done; returns (array, latency_s)."""
        out, lat = MMBuffer(), C.c_double()
        s = lib.mm_server_wait(self._h, int(rid), C.byref(out), C.byref(lat))
        if s != 1:
            raise RuntimeError(f"request {rid} failed with status {s}")
        return buffer_to_numpy(out), lat.value

    def poll(self, rid: int):
        """(array, latency_s) when done, None while pending."""
        out, lat = MMBuffer(), C.c_double()
        s = lib.mm_server_poll(self._h, int(rid), C.byref(out), C.byref(lat))
        if s == 0:
            return None
        if s != 1:
            raise RuntimeError(f"request {rid} failed with status {s}")
        return buffer_to_numpy(out), lat.value

    def stats(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        lib.mm_server_stats(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"submitted": a.value, "completed": b.value, "steps": c.value, "frames": d.value}

    def device_bytes(self) -> int:
        """Device memory the server's streams hold for their steps."""
        b = C.c_int64()
        s = lib.mm_server_device_bytes(self._h, C.byref(b))
        if s != 0:
            raise RuntimeError(f"mm_server_device_bytes failed with status {s}")
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            lib.mm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


KINDS = ("stft", "mel", "mfcc")
SRS = (22050, 44100)


def make_sources(seconds: float = 330.0, seed: int = 4):
    """One noise + linear-chirp buffer per sample rate; clips are slices."""
    from .workloads import chirp_noise_clip
    rng = np.random.default_rng(seed)
    return {sr: chirp_noise_clip(int(seconds * sr), sr, rng) for sr in SRS}


def make_schedule(rate: float, duration: float, seed: int = 4, kinds=KINDS, n_requests: int = 0):
    """Poisson arrivals: (t_arrival, kind, sr, n_samples, source offset).

    ``rate`` = inf (or <= 0) gives a backlog: ``n_requests`` requests all
    arriving at t = 0 (the saturation point)."""
    rng = np.random.default_rng(seed)
    backlog = not (rate > 0 and np.isfinite(rate))
    n = max(1, int(n_requests) if backlog else int(rate * duration))
    t = np.zeros(n) if backlog else np.cumsum(rng.exponential(1.0 / rate, n))
    secs = np.clip(np.exp(np.log(5.0) + rng.standard_normal(n)), 0.5, 300.0)
    sr = np.array(SRS)[rng.integers(0, 2, n)]
    kind = rng.integers(0, len(kinds), n)
    lens = np.round(secs * sr).astype(np.int64)
    offs = np.array([rng.integers(0, int(330 * s) - L + 1) for s, L in zip(sr, lens)])
    return [(float(t[i]), kinds[kind[i]], int(sr[i]), int(lens[i]), int(offs[i])) for i in range(n)]


def pcm_sources(sources):
    """int16 stereo versions of the float sources (right channel at half the
    left's level), as a WAV file at the source's rate would hold them."""
    out = {}
    for sr, x in sources.items():
        left = np.clip(np.round(x * 16384.0), -32768, 32767)
        out[sr] = np.ascontiguousarray(np.stack([left, np.round(left * 0.5)], axis=1).astype("<i2"))
    return out


def open_loop(server: SpectralServer, rate: float, duration: float = 2.0, seed: int = 4,
              sources=None, keep_outputs: bool = False, pcm=None, analysis_sr: int = 22050,
              kinds=KINDS, precision: str = "f32", n_requests: int = 0):
    """Submits a Poisson schedule in real time and collects every result.

    Latency is measured from each request's scheduled arrival (t0 + its
    Poisson time), not from when the submitting thread got to it: a late
    submitter's lag, the copy of the clip into the server and the time to
    the result on the host all count (no coordinated omission). The server's
    own submit-to-completion time is reported beside it. ``rate`` = inf
    submits a backlog of ``n_requests`` at t = 0 (the saturation point).

    ``kinds`` restricts the request types (e.g. ``("mfcc",)``) and
    ``precision`` sets the mel / DCT contraction arithmetic (f32, bf16, f16).
    With ``pcm`` (from ``pcm_sources``) every clip is submitted as int16
    stereo PCM at its source rate and analysed at ``analysis_sr``: the GPU
    mixes it to mono and resamples it inside the request's step.

    Returns dict(latencies (arrival to result), p50, p99, server_p50,
    server_p99, frames, wall, frames_per_s, achieved_rate (measured
    submissions per second), n_requests, outputs (if kept), schedule).
    """
    sources = sources or make_sources(seed=seed)
    sched = make_schedule(rate, duration, seed, kinds=kinds, n_requests=n_requests)
    if pcm is not None:
        from ._native import PCM_S16
        params = {(k, sr): server.params(k, sr=analysis_sr, precision=precision) for k in kinds for sr in SRS}
    else:
        params = {(k, sr): server.params(k, sr=sr, precision=precision) for k in kinds for sr in SRS}
    ids = [None] * len(sched)
    ready = threading.Semaphore(0)
    lat = [0.0] * len(sched)
    t_sub = [0.0] * len(sched)
    frames = [0] * len(sched)
    outs = [None] * len(sched) if keep_outputs else None

    def collect():
        for i in range(len(sched)):
            ready.acquire()
            arr, l = server.wait(ids[i])
            lat[i] = l
            frames[i] = arr.shape[1] if arr.ndim == 2 else 0
            if keep_outputs:
                outs[i] = arr

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, (ta, kind, sr, n, off) in enumerate(sched):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if pcm is not None:
            ids[i] = server.submit_pcm(pcm[sr][off:off + n], 2, PCM_S16, sr, params[(kind, sr)])
        else:
            ids[i] = server.submit(sources[sr][off:off + n], params[(kind, sr)])
        # the server stamps its submit time as submit() returns (same monotonic
        # clock as perf_counter); the lag since the scheduled arrival is added
        t_sub[i] = time.perf_counter()
        ready.release()
    th.join()
    wall = time.perf_counter() - t0
    ls = np.array(lat)                                                          # submit -> result
    la = ls + np.array([t_sub[i] - (t0 + sched[i][0]) for i in range(len(sched))])   # arrival -> result
    span = t_sub[-1] - t0
    return {"latencies": la, "p50": float(np.percentile(la, 50)), "p99": float(np.percentile(la, 99)),
            "server_p50": float(np.percentile(ls, 50)), "server_p99": float(np.percentile(ls, 99)),
            "frames": int(sum(frames)), "wall": wall, "frames_per_s": sum(frames) / wall,
            "achieved_rate": len(sched) / span if span > 0 else float("inf"),
            "n_requests": len(sched), "outputs": outs, "schedule": sched}

"""Benchmark: MFCC frames/s of the batched BF16 path on one MI355X.

Workload (one step): the 256-clip uneven batch (lengths U[1 s, 60 s] at
22.05 kHz, noise + chirp, default_rng(1); 336,013 frames), n_fft 2048, hop
512, Hann, centred, 128 Slaney mels, power_to_db(top_db 80), 13 MFCC, with the
mel and DCT contractions on BF16 MFMA (FP32 accumulation, FFT in FP32). All
clips run in one persistent launch replayed from a hipGraph; the signal is in
HBM before the timed region starts.

Also reported on the same batch: the FP32 batched path and the plain path (one
clip per launch, FP32, no graph, clips in sequence), the fused kernel's
roofline position, and the float32 NumPy restatement of the reference's CPU
path timed on this host (bounded sample). The ``istft`` and
``onset_strength`` entries measure the inverse STFT and the onset strength of
one clip of the same total length on device-resident data.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MFCC frames/s batched+BF16 vs plain FP32 on 1 GPU; p50/p99 latency @ fixed req/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_tile_mfcc_bf16_uneven_batch_traffic.json")
ISTFT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_istft_tile_2048_long_clip_traffic.json")
RESAMPLE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_resample_dec2_44k_to_22k_traffic.json")
RESAMPLE_PHASE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_resample_phase_48k_to_22k_traffic.json")


def measured_traffic(bytes_launch, path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel from its committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a copy of known size), if it was taken on this same
    workload (same algorithmic bytes); else None."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if int(t.get("algorithmic_bytes_per_launch", -1)) != int(bytes_launch):
        return None
    return float(t["hbm_bytes_per_launch"])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(clips, max_frames=400000, workers=0):
    """Float32 NumPy restatement of the reference CPU path on a bounded sample
    of the same batch (whole clips until max_frames; the default takes the
    whole 336,013-frame batch), BLAS limited to one thread. workers = 0: one
    core, clips in sequence; workers > 0: that many threads over the clips
    (numpy's FFT, BLAS and elementwise kernels run without the GIL)."""
    from oracle.cpu_baseline import MfccF32
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:
        limiter = None
    f = MfccF32(22050, 2048, 512, 128, 13, 80.0)
    sample, frames = [], 0
    for c in clips:
        if frames >= max_frames:
            break
        sample.append(c)
        frames += 1 + len(c) // 512
    f(sample[0][:22050])  # warm-up
    t0 = time.perf_counter()
    if workers > 0:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(f, sample))
    else:
        for c in sample:
            f(c)
    dt = time.perf_counter() - t0
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    return {"value": frames / dt, "unit": "frames/s", "cores": max(workers, 1), "kind": "port",
            "sample": f"{len(sample)} whole clips of the uneven batch, {frames} frames, "
                      f"float32 numpy rfft + dense FB@P + dB + DCT, {dt:.1f} s"
                      + (f", {workers} threads" if workers else ", 1 thread"),
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


def istft_leg(torch, n_samples, iters, with_cpu):
    """Inverse STFT (mm_istft_device) of one long clip on device-resident data:
    n_fft 2048, hop 512, centred, the spectrogram of n_samples samples
    (synthetic complex64 values, [1025, T] row-major), n_samples samples out.
    Kernel time from HIP events on the launch stream; algorithmic bytes are
    the spectrogram read once plus the samples written."""
    import ctypes as C
    from metalmom_amd._native import lib
    n_fft, hop = 2048, 512
    F = n_fft // 2 + 1
    T = 1 + n_samples // hop
    ld = (T + 15) // 16 * 16   # rows padded to whole 128-byte lines of 16 frames
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    X = torch.randn(F * ld * 2, device="cuda", generator=g)
    n = int(lib.mm_istft_length(F, T, hop, 1, n_samples))
    Y = torch.empty(n, device="cuda")
    st = torch.cuda.Stream()
    ctx = lib.mm_init()

    def run():
        s = lib.mm_istft_device(ctx, C.c_void_p(X.data_ptr()), ld, F, T, 22050, hop, n_fft, 1, n_samples,
                                C.c_void_p(Y.data_ptr()), C.c_void_p(st.cuda_stream))
        if s != 0:
            raise RuntimeError(f"mm_istft_device failed with status {s}")

    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(iters):
        run()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / iters
    lib.mm_destroy(ctx)
    del X, Y
    bytes_launch = F * T * 8 + n * 4
    gbs = bytes_launch / (ms * 1e-3) / 1e9
    traffic = measured_traffic(bytes_launch, ISTFT_TRAFFIC_FILE)
    res = {"workload": f"one clip, spectrogram [1025, {T}] complex64 (rows padded to {ld} frames) -> {n} "
                       f"samples, n_fft 2048, hop 512, centred, Hann synthesis window",
           "frames": T, "samples": n, "kernel_ms": ms, "frames_per_s": T / (ms * 1e-3),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": bytes_launch}}
    if with_cpu:
        from oracle.cpu_baseline import IstftF32
        try:
            from threadpoolctl import threadpool_limits
            limiter = threadpool_limits(limits=1)
        except Exception:
            limiter = None
        Tc = 40000
        rng = np.random.default_rng(5)
        S = (rng.standard_normal((F, Tc), dtype=np.float32)
             + 1j * rng.standard_normal((F, Tc), dtype=np.float32)).astype(np.complex64)
        f = IstftF32(n_fft, hop)
        f(S[:, :64])
        t0 = time.perf_counter()
        f(S, (Tc - 1) * hop)
        dt = time.perf_counter() - t0
        if limiter is not None and hasattr(limiter, "unregister"):
            limiter.unregister()
        res["cpu_baseline"] = {"value": Tc / dt, "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": f"spectrogram [1025, {Tc}], float32 numpy irfft + overlap-add, "
                                         f"{dt:.1f} s, 1 thread", "host_cpus": os.cpu_count()}
    return res


def onset_leg(torch, n_samples, iters, with_cpu=False):
    """Onset strength (mm_onset_strength_device) of one long clip on
    device-resident data: n_fft 2048, hop 512, 128 mels, centred, aggregated;
    the signal is the uneven batch's clips joined end to end. Time from HIP
    events on the launch stream; algorithmic bytes are the samples read once
    plus the envelope written."""
    import ctypes as C
    from metalmom_amd._native import lib
    from metalmom_amd.workloads import uneven_batch
    y = np.concatenate(uneven_batch())[:n_samples]
    X = torch.from_numpy(y).cuda()
    T = 1 + len(y) // 512
    Y = torch.empty(T, device="cuda")
    rows, cols = C.c_int64(0), C.c_int64(0)
    st = torch.cuda.Stream()
    ctx = lib.mm_init()

    def run():
        s = lib.mm_onset_strength_device(ctx, C.c_void_p(X.data_ptr()), len(y), 22050, 2048, 512, 128, 0.0, 0.0,
                                         1, 1, C.c_void_p(Y.data_ptr()), C.byref(rows), C.byref(cols),
                                         C.c_void_p(st.cuda_stream))
        if s != 0:
            raise RuntimeError(f"mm_onset_strength_device failed with status {s}")

    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(iters):
        run()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / iters
    lib.mm_destroy(ctx)
    del X, Y
    bytes_launch = len(y) * 4 + cols.value * 4
    gbs = bytes_launch / (ms * 1e-3) / 1e9
    cpu = None
    if with_cpu:
        from oracle.cpu_baseline import OnsetF32
        try:
            from threadpoolctl import threadpool_limits
            limiter = threadpool_limits(limits=1)
        except Exception:
            limiter = None
        f = OnsetF32()
        f(y[:22050])
        sample = y[:20_000_000]
        t0 = time.perf_counter()
        f(sample)
        dt = time.perf_counter() - t0
        if limiter is not None and hasattr(limiter, "unregister"):
            limiter.unregister()
        fr = 1 + len(sample) // 512
        cpu = {"value": fr / dt, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"first {len(sample)} samples of the clip ({fr} frames), float32 numpy rfft + dense "
                         f"FB@P + dB + flux, {dt:.1f} s, 1 thread", "host_cpus": os.cpu_count()}
    res = {"workload": f"one clip of {len(y)} samples (the uneven batch joined), n_fft 2048, hop 512, 128 mels, "
                        f"centred, aggregated: fused mel kernel + dB with the clip maximum + flux kernel",
            "frames": T, "step_ms": ms, "frames_per_s": T / (ms * 1e-3),
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": bytes_launch}}
    if cpu is not None:
        res["cpu_baseline"] = cpu
    return res


def cpu_time_1thread(fn, warm, sample, frames, what):
    """One-thread CPU baseline entry: fn timed on `sample` after a warm-up
    call, frames/s over `frames`."""
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=1)
    except Exception:
        limiter = None
    fn(warm)
    t0 = time.perf_counter()
    fn(sample)
    dt = time.perf_counter() - t0
    if limiter is not None and hasattr(limiter, "unregister"):
        limiter.unregister()
    return {"value": frames / dt, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{what}, {dt:.1f} s, 1 thread", "host_cpus": os.cpu_count()}


def chroma_leg(torch, lens, d_in, steps, clips=None):
    """Batched STFT chroma (12 bins, no frame normalisation) of the uneven
    batch already in HBM, FP32 and BF16 contraction: one persistent launch
    per step replayed from a hipGraph; fused-kernel time from HIP events."""
    from metalmom_amd.batch import SpectralBatch
    res = {}
    for prec in ("f32", "bf16"):
        B = SpectralBatch(lens, kind="chroma", sr=22050, n_fft=2048, hop_length=512, n_mels=12, precision=prec)
        out = B.empty_output()
        B.autotune(d_in, out, iters=2)
        B.run(d_in, out, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            B.run(d_in, out, graph=True)
        torch.cuda.synchronize()
        fps = B.total_frames * steps / (time.perf_counter() - t0)
        k_ms, _ = B.time(d_in, out, iters=10)
        bytes_launch = 4 * B.total_input + 4 * 12 * B.total_frames
        gbs = bytes_launch / (k_ms * 1e-3) / 1e9
        res[prec] = {"frames_per_s": fps, "kernel_ms": k_ms,
                     "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": bytes_launch}}
        B.close()
    res["workload"] = ("uneven batch (256 clips, 336,013 frames), n_fft 2048, hop 512, 12 chroma bins, "
                       "dense chroma filterbank on the fused banded contraction")
    if clips is not None:
        from oracle.cpu_baseline import ChromaF32
        sample = np.concatenate(clips)[:20_000_000]
        fr = 1 + len(sample) // 512
        res["cpu_baseline"] = cpu_time_1thread(ChromaF32(), sample[:22050], sample, fr,
                                               f"first {len(sample)} samples of the batch joined ({fr} frames), "
                                               f"float32 numpy rfft + FB @ |X|^2")
    return res


def descriptor_leg(torch, n_samples, iters, with_cpu=False):
    """Spectral centroid / bandwidth / rolloff / flatness
    (mm_spectral_descriptor_device) of the uneven batch joined into one clip
    on device-resident data: n_fft 2048, hop 512, centred. Time per
    descriptor from HIP events on the launch stream (fused STFT kernel writing
    the magnitude spectrogram, then the per-frame reduction)."""
    import ctypes as C
    from metalmom_amd._native import lib
    from metalmom_amd.workloads import uneven_batch
    y = np.concatenate(uneven_batch())[:n_samples]
    X = torch.from_numpy(y).cuda()
    T = 1 + len(y) // 512
    Y = torch.empty(T, device="cuda")
    st = torch.cuda.Stream()
    ctx = lib.mm_init()
    res = {"workload": f"one clip of {len(y)} samples (the uneven batch joined), n_fft 2048, hop 512, centred",
           "frames": T}
    for kind, name, param in [(0, "centroid", 0.0), (1, "bandwidth", 2.0), (2, "rolloff", 0.85), (3, "flatness", 0.0)]:
        def run():
            s = lib.mm_spectral_descriptor_device(ctx, C.c_void_p(X.data_ptr()), len(y), 22050, 2048, 512, 2048, 1,
                                                  kind, param, C.c_void_p(Y.data_ptr()), None,
                                                  C.c_void_p(st.cuda_stream))
            if s != 0:
                raise RuntimeError(f"mm_spectral_descriptor_device failed with status {s}")
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(iters):
            run()
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / iters
        # algorithmic bytes: the signal in, one float per frame out (the fused
        # kernel keeps the spectrum on chip)
        gbs = (4 * len(y) + 4 * T) / (ms * 1e-3) / 1e9
        res[name] = {"step_ms": ms, "frames_per_s": T / (ms * 1e-3),
                     "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}
    # spectral contrast: 7 octave bands (magnitude STFT, then one thread per (frame, band))
    Yc = torch.empty(7 * T, device="cuda")

    def runc():
        s = lib.mm_spectral_contrast_device(ctx, C.c_void_p(X.data_ptr()), len(y), 22050, 2048, 512, 2048, 1, 6,
                                            200.0, C.c_void_p(Yc.data_ptr()), None, C.c_void_p(st.cuda_stream))
        if s != 0:
            raise RuntimeError(f"mm_spectral_contrast_device failed with status {s}")
    runc()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(iters):
        runc()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / iters
    gbs = (4 * len(y) + 4 * 7 * T) / (ms * 1e-3) / 1e9
    res["contrast"] = {"step_ms": ms, "frames_per_s": T / (ms * 1e-3),
                       "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}
    lib.mm_destroy(ctx)
    del X, Y
    if with_cpu:
        from oracle.cpu_baseline import CentroidF32
        sample = y[:20_000_000]
        fr = 1 + len(sample) // 512
        res["cpu_baseline_centroid"] = cpu_time_1thread(CentroidF32(), sample[:22050], sample, fr,
                                                        f"first {len(sample)} samples of the clip ({fr} frames), "
                                                        f"float32 numpy rfft + |X| + centroid")
        from oracle.cpu_baseline import ContrastF32
        small = y[:40_000_000]
        frc = 1 + len(small) // 512
        res["cpu_baseline_contrast"] = cpu_time_1thread(ContrastF32(), small[:22050], small, frc,
                                                        f"first {len(small)} samples of the clip ({frc} frames), "
                                                        f"float32 numpy rfft + |X| + per-band sort")
    return res


FP32_VECTOR_PEAK_TFLOPS = 157.3


def _fma_roofline(taps, ms):
    """FP32 FMA roofline of a launch of `taps` multiply-adds in `ms`."""
    tf = 2.0 * taps / (ms * 1e-3) / 1e12
    return {"achieved": tf, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP32_VECTOR_PEAK_TFLOPS,
            "frac_of_unpacked_peak": tf / (FP32_VECTOR_PEAK_TFLOPS / 2), "flops_per_launch": 2.0 * taps}


def resample_leg(torch, n_samples, iters, with_cpu=False):
    """Sample-rate conversion (mm_resample_device) of the uneven batch joined
    into one clip, taken as 44.1 kHz audio and converted to 22.05 kHz (the
    server's two rates): 256 taps per output. Kernel time from HIP events on
    the launch stream; algorithmic HBM bytes = 4 per input + 4 per output
    sample, algorithmic FLOPs = 2 per tap (one FMA). The kernels are bound by
    LDS reads and FMA issue, not by HBM: "fma" prices the taps against the
    FP32 vector peak (157.3 TFLOP/s counts packed FP32, which this build does
    not emit; one FMA per lane per cycle is half of it)."""
    import ctypes as C
    from metalmom_amd._native import lib
    from metalmom_amd.workloads import uneven_batch
    y = np.concatenate(uneven_batch())[:n_samples]
    src, dst = 44100, 22050
    m = lib.mm_resample_length(len(y), src, dst)
    X = torch.from_numpy(y).cuda()
    Y = torch.empty(m, device="cuda")
    st = torch.cuda.Stream()

    def run():
        s = lib.mm_resample_device(None, C.c_void_p(X.data_ptr()), len(y), src, dst, C.c_void_p(Y.data_ptr()),
                                   C.c_void_p(st.cuda_stream))
        if s != 0:
            raise RuntimeError(f"mm_resample_device failed with status {s}")
    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(iters):
        run()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / iters
    taps = 2 * int(np.ceil(64 / min(1.0, dst / src)))
    bytes_launch = 4 * (len(y) + m)
    gbs = bytes_launch / (ms * 1e-3) / 1e9
    res = {"workload": f"one clip of {len(y)} samples (the uneven batch joined) at 44.1 kHz -> 22.05 kHz, "
                       f"{m} output samples, {taps} taps each",
           "kernel_ms": ms, "out_samples_per_s": m / (ms * 1e-3), "in_samples_per_s": len(y) / (ms * 1e-3),
           "taps_per_s": m * taps / (ms * 1e-3),
           "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                   "traffic": measured_traffic(bytes_launch, RESAMPLE_TRAFFIC_FILE), "bytes_per_launch": bytes_launch},
           "fma": _fma_roofline(m * taps, ms)}
    # a pair with many phases: 48 kHz -> 22.05 kHz, 147 phases of 280 taps
    # (the phase-major kernel: two phases per wave)
    src2, dst2 = 48000, 22050
    m2 = lib.mm_resample_length(len(y), src2, dst2)
    Y2 = torch.empty(m2, device="cuda")

    def run2():
        s = lib.mm_resample_device(None, C.c_void_p(X.data_ptr()), len(y), src2, dst2, C.c_void_p(Y2.data_ptr()),
                                   C.c_void_p(st.cuda_stream))
        if s != 0:
            raise RuntimeError(f"mm_resample_device failed with status {s}")
    run2()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(iters):
        run2()
    e1.record(st)
    e1.synchronize()
    ms2 = e0.elapsed_time(e1) / iters
    taps2 = 2 * int(np.ceil(64 / (dst2 / src2)))
    res["48k_to_22k"] = {"kernel_ms": ms2, "out_samples_per_s": m2 / (ms2 * 1e-3),
                         "taps_per_s": m2 * taps2 / (ms2 * 1e-3), "taps": taps2, "phases": 147,
                         "fma": _fma_roofline(m2 * taps2, ms2)}
    bytes2 = 4 * (len(y) + m2)
    gbs2 = bytes2 / (ms2 * 1e-3) / 1e9
    res["48k_to_22k"]["hbm"] = {"achieved": gbs2, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs2 / HBM_PEAK_GBS,
                                "traffic": measured_traffic(bytes2, RESAMPLE_PHASE_TRAFFIC_FILE),
                                "bytes_per_launch": bytes2}
    del X, Y, Y2
    if with_cpu:
        from oracle.cpu_baseline import ResampleF32
        sample = y[:2_000_000]
        ms_ = len(sample) // 2
        r = cpu_time_1thread(ResampleF32(src, dst), sample[:20000], sample, ms_,
                             f"first {len(sample)} samples ({ms_} outputs), float32 numpy restatement")
        r["unit"] = "output samples/s"
        res["cpu_baseline"] = r
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-clip leg")
    ap.add_argument("--profile", action="store_true", help="short run for a profiler")
    ap.add_argument("--serve-rates", default="1000,5000,8000",
                    help="Poisson request rates (req/s) for the serving leg; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-pcm-rates", default="1000,5000",
                    help="rates for the raw-PCM serving leg (int16 stereo at 22.05 / 44.1 kHz, analysed at "
                         "22.05 kHz); empty to skip")
    ap.add_argument("--grid", type=int, default=0,
                    help="persistent workgroups per CU (skips the autotuner; 0 = autotune)")
    ap.add_argument("--no-istft", action="store_true", help="skip the inverse-STFT leg")
    ap.add_argument("--no-onset", action="store_true", help="skip the onset-strength leg")
    ap.add_argument("--no-chroma", action="store_true", help="skip the chroma leg")
    ap.add_argument("--no-descriptors", action="store_true", help="skip the spectral-descriptor leg")
    ap.add_argument("--no-resample", action="store_true", help="skip the sample-rate conversion leg")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="profiling aid: also run one device-to-device copy of this many bytes, "
                         "so HBM counters can be scaled by a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batch shards by clip with no exchange, "
              "so this build reports the single-GPU throughput (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from metalmom_amd.batch import SpectralBatch
    from metalmom_amd.workloads import uneven_batch

    torch.cuda.set_device(0)
    log("generating the 256-clip uneven batch ...")
    clips = uneven_batch()
    lens = [len(c) for c in clips]
    n_mfcc = 13

    def make(precision):
        return SpectralBatch(lens, kind="mfcc", sr=22050, n_fft=2048, hop_length=512, n_mels=128,
                             n_mfcc=n_mfcc, top_db=80.0, precision=precision)

    B = make("bf16")
    frames = B.total_frames
    d_in = B.pack(clips)
    out = B.empty_output()
    stream = torch.cuda.Stream()
    log(f"batch: {len(clips)} clips, {B.total_input} samples, {frames} frames, {B.n_tiles} tiles")

    if args.grid > 0:
        best_grid = args.grid
        B.set_grid(best_grid)
        log(f"persistent grid: {best_grid} workgroups/CU (fixed)")
    else:
        best_grid, tune_ms = B.autotune(d_in, out, iters=2)
        log(f"autotuned persistent grid: {best_grid} workgroups/CU ({tune_ms:.3f} ms/step)")

    steps = 3 if args.profile else args.steps
    warm = 1 if args.profile else args.warmup
    with torch.cuda.stream(stream):
        for _ in range(warm):
            B.run(d_in, out, stream=stream, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            B.run(d_in, out, stream=stream, graph=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    value = frames * steps / dt
    ms_step = dt * 1e3 / steps
    log(f"bf16 batched: {value / 1e6:.1f} M frames/s ({ms_step:.3f} ms/step)")

    if args.copy_bytes > 0:
        src = torch.empty(args.copy_bytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_bytes} bytes read, {args.copy_bytes} bytes written")
        del src, dst

    # fused kernel time from HIP events on its own stream (the plan's stream)
    k_ms, s_ms = B.time(d_in, out, iters=5 if args.profile else 20)
    bytes_launch = 4 * B.total_input + 4 * n_mfcc * frames
    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    log(f"fused kernel {k_ms:.3f} ms, step {s_ms:.3f} ms, {achieved:.0f} GB/s algorithmic")

    traffic = measured_traffic(bytes_launch)
    if traffic is not None:
        log(f"HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated): {traffic / 1e6:.1f} MB "
            f"= {traffic / bytes_launch:.2f}x algorithmic")

    extra = {}
    if not args.profile:
        # FP32 batched path on the same batch
        B32 = make("f32")
        out32 = B32.empty_output()
        B32.autotune(d_in, out32, iters=2)
        B32.run(d_in, out32, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(3, steps // 2)):
            B32.run(d_in, out32, graph=True)
        torch.cuda.synchronize()
        f32_fps = frames * max(3, steps // 2) / (time.perf_counter() - t0)
        extra["batched_fp32_frames_per_s"] = f32_fps
        # BF16 vs FP32 GPU agreement on the MFCCs (the FP64-oracle errors are in tests/)
        a, b = out.cpu().numpy(), out32.cpu().numpy()
        extra["bf16_vs_fp32_mfcc_max_rel"] = float(np.abs(a - b).max() / np.abs(b).max())
        log(f"fp32 batched: {f32_fps / 1e6:.1f} M frames/s")

        if not args.no_plain:
            # plain path: one clip per launch, FP32, no graph, in sequence on
            # one stream (each plan would otherwise use its own pooled stream
            # and clips could overlap across hardware queues)
            singles = [SpectralBatch([n], kind="mfcc", n_fft=2048, hop_length=512, n_mels=128,
                                     n_mfcc=n_mfcc, precision="f32") for n in lens]
            ins = [d_in[B.clips[i][0]:B.clips[i][0] + lens[i]] for i in range(len(lens))]
            outs = [s.empty_output() for s in singles]
            plain_stream = torch.cuda.Stream()
            for s, i_, o_ in zip(singles, ins, outs):
                s.run(i_, o_, stream=plain_stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s, i_, o_ in zip(singles, ins, outs):
                s.run(i_, o_, stream=plain_stream)
            torch.cuda.synchronize()
            plain_fps = frames / (time.perf_counter() - t0)
            extra["plain_fp32_frames_per_s"] = plain_fps
            extra["speedup_vs_plain_fp32"] = value / plain_fps
            log(f"plain fp32 single-clip: {plain_fps / 1e6:.2f} M frames/s; batched bf16 is "
                f"{value / plain_fps:.1f}x")
            for s in singles:
                s.close()

    serving = []
    if args.serve_rates and not args.profile:
        # continuous batching server: open-loop Poisson arrivals, mixed request
        # types (STFT / mel 128 / MFCC 13), sr in {22050, 44100}, log-normal
        # lengths (median 5 s); latency = submission to result on the host
        from metalmom_amd.serve import SpectralServer, make_sources, open_loop
        sources = make_sources()
        for rate in [float(r) for r in args.serve_rates.split(",") if r]:
            srv = SpectralServer(hbm_budget_bytes=32 << 30)
            # warm-up at the measured rate: grows the lane buffers and the
            # process heap that results are copied into to their steady size
            open_loop(srv, rate=rate, duration=0.5, seed=99, sources=sources)
            r = open_loop(srv, rate=rate, duration=args.serve_seconds, seed=4, sources=sources)
            st = srv.stats()
            srv.close()
            serving.append({"rate_req_s": rate, "achieved_rate": r["achieved_rate"],
                            "requests": r["n_requests"], "p50_ms": r["p50"] * 1e3, "p99_ms": r["p99"] * 1e3,
                            "frames_per_s": r["frames_per_s"], "device_steps": st["steps"]})
            log(f"serving @{rate:.0f} req/s: p50 {r['p50'] * 1e3:.2f} ms, p99 {r['p99'] * 1e3:.2f} ms, "
                f"{r['frames_per_s'] / 1e6:.2f} M frames/s, {st['steps']} steps for {r['n_requests']} requests")
        # the same schedule as raw int16 stereo PCM at the source rates, mixed
        # down and resampled to 22.05 kHz on the GPU inside each step
        from metalmom_amd.serve import pcm_sources
        psrc = pcm_sources(sources) if args.serve_pcm_rates else None
        for rate in [float(r) for r in args.serve_pcm_rates.split(",") if r]:
            srv = SpectralServer(hbm_budget_bytes=32 << 30)
            open_loop(srv, rate=rate, duration=0.5, seed=99, sources=sources, pcm=psrc)
            r = open_loop(srv, rate=rate, duration=args.serve_seconds, seed=4, sources=sources, pcm=psrc)
            st = srv.stats()
            srv.close()
            serving.append({"rate_req_s": rate, "input": "int16 stereo PCM, resampled to 22.05 kHz on the GPU",
                            "achieved_rate": r["achieved_rate"], "requests": r["n_requests"],
                            "p50_ms": r["p50"] * 1e3, "p99_ms": r["p99"] * 1e3,
                            "frames_per_s": r["frames_per_s"], "device_steps": st["steps"]})
            log(f"serving PCM @{rate:.0f} req/s: p50 {r['p50'] * 1e3:.2f} ms, p99 {r['p99'] * 1e3:.2f} ms, "
                f"{r['frames_per_s'] / 1e6:.2f} M frames/s, {st['steps']} steps for {r['n_requests']} requests")

    if not args.no_istft:
        ist = istft_leg(torch, B.total_input, 3 if args.profile else 20, not args.no_cpu and not args.profile)
        extra["istft"] = ist
        log(f"istft: {ist['frames_per_s'] / 1e6:.1f} M frames/s, kernel {ist['kernel_ms']:.3f} ms, "
            f"{ist['roofline']['achieved']:.0f} GB/s algorithmic ({100 * ist['roofline']['frac']:.1f} % of HBM)"
            + (f"; cpu {ist['cpu_baseline']['value']:.0f} frames/s" if "cpu_baseline" in ist else ""))

    if not args.no_onset:
        ons = onset_leg(torch, B.total_input, 3 if args.profile else 20, not args.no_cpu and not args.profile)
        extra["onset_strength"] = ons
        log(f"onset strength: {ons['frames_per_s'] / 1e6:.1f} M frames/s, {ons['step_ms']:.3f} ms per clip, "
            f"{ons['roofline']['achieved']:.0f} GB/s algorithmic"
            + (f"; cpu {ons['cpu_baseline']['value']:.0f} frames/s" if "cpu_baseline" in ons else ""))

    if not args.no_chroma and not args.profile:
        chr_ = chroma_leg(torch, lens, d_in, max(3, steps // 2), None if args.no_cpu else clips)
        extra["chroma"] = chr_
        log(f"chroma batched: fp32 {chr_['f32']['frames_per_s'] / 1e6:.1f} M frames/s "
            f"(kernel {chr_['f32']['kernel_ms']:.3f} ms), bf16 {chr_['bf16']['frames_per_s'] / 1e6:.1f} M frames/s "
            f"(kernel {chr_['bf16']['kernel_ms']:.3f} ms)")

    if not args.no_descriptors and not args.profile:
        dsc = descriptor_leg(torch, B.total_input, max(3, steps // 2), not args.no_cpu)
        extra["spectral_descriptors"] = dsc
        log("spectral descriptors: " + ", ".join(f"{k} {dsc[k]['frames_per_s'] / 1e6:.1f} M frames/s"
                                                 for k in ("centroid", "bandwidth", "rolloff", "flatness",
                                                           "contrast")))

    if not args.no_resample:
        rs = resample_leg(torch, B.total_input, 3 if args.profile else max(3, steps // 2),
                          not args.no_cpu and not args.profile)
        extra["resample"] = rs
        log(f"resample 44.1 -> 22.05 kHz: {rs['out_samples_per_s'] / 1e9:.2f} G output samples/s, "
            f"kernel {rs['kernel_ms']:.3f} ms, {rs['taps_per_s'] / 1e12:.2f} T taps/s "
            f"({100 * rs['fma']['frac']:.1f} % of FP32 FMA peak); 48 -> 22.05 kHz (147 phases) "
            f"{rs['48k_to_22k']['out_samples_per_s'] / 1e9:.2f} G output samples/s, "
            f"{rs['48k_to_22k']['kernel_ms']:.3f} ms ({100 * rs['48k_to_22k']['fma']['frac']:.1f} %)"
            + (f"; cpu {rs['cpu_baseline']['value'] / 1e6:.2f} M samples/s" if "cpu_baseline" in rs else ""))

    cpu = None
    if not args.no_cpu and not args.profile:
        log("cpu baseline ...")
        cpu = cpu_baseline(clips)
        log(f"cpu baseline: {cpu['value']:.0f} frames/s ({cpu['sample']})")
        # all host cores this job may use (at most 16 on a one-GPU share)
        nw = max(1, min(16, os.cpu_count() or 1))
        cpu_all = cpu_baseline(clips, workers=nw)
        extra["cpu_baseline_all_cores"] = cpu_all
        log(f"cpu baseline, {nw} threads: {cpu_all['value']:.0f} frames/s")

    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "uneven batch: 256 clips U[1,60] s @22.05 kHz, noise+chirp, "
                               "n_fft 2048 hop 512, 128 mels, 13 MFCC, top_db 80, BF16 mel/DCT MFMA "
                               "(FP32 FFT), one persistent hipGraph launch per step",
                   "clips": len(clips), "frames": frames, "samples": B.total_input,
                   "persistent_wg_per_cu": best_grid},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms": k_ms, "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
    }
    rec.update(extra)
    if serving:
        rec["serving"] = serving
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()

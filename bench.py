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
FP32_VECTOR_PEAK_TFLOPS = 157.3   # published FP32 vector peak (counts packed FP32, which this build does not emit)
METRIC = "MFCC frames/s batched+BF16 vs plain FP32 on 1 GPU; p50/p99 latency @ fixed req/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_tile_mfcc_bf16_uneven_batch_traffic.json")
ISTFT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_istft_tile_2048_long_clip_traffic.json")
RESAMPLE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_resample_dec2_44k_to_22k_traffic.json")
RESAMPLE_PHASE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_resample_phase_48k_to_22k_traffic.json")
LARGE_FFT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_tile_mfcc_f16_large_fft_traffic.json")


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


def cpu_baseline(clips, max_frames=400000, workers=0, iters=5):
    """Float32 NumPy restatement of the reference CPU path on a bounded sample
    of the same batch (whole clips until max_frames; the default takes the
    whole 336,013-frame batch), BLAS limited to one thread, timed over
    ``iters`` passes after a warm-up (mean, min and median pass). workers = 0:
    one core, clips in sequence; workers > 0: that many threads over the clips
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
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        if workers > 0:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=workers) as ex:
                list(ex.map(f, sample))
        else:
            for c in sample:
                f(c)
        times.append(time.perf_counter() - t0)
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    dt = float(np.mean(times))
    return {"value": frames / dt, "unit": "frames/s", "cores": max(workers, 1), "kind": "port",
            "sample": f"{len(sample)} whole clips of the uneven batch, {frames} frames per pass, "
                      f"float32 numpy rfft + dense FB@P + dB + DCT, {iters} passes "
                      f"(mean {dt:.2f} s, min {min(times):.2f} s, median {float(np.median(times)):.2f} s)"
                      + (f", {workers} threads" if workers else ", 1 thread"),
            "pass_s": {"mean": dt, "min": float(min(times)), "p50": float(np.median(times)), "n": iters},
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


def timing_stats(ms):
    ms = np.asarray(ms, dtype=np.float64)
    return {"mean": float(ms.mean()), "min": float(ms.min()), "p50": float(np.median(ms)), "n": int(ms.size)}


def fft_flops(n_fft):
    """Algorithmic FLOPs of one real FFT frame: 2.5 N log2 N."""
    return 2.5 * n_fft * np.log2(n_fft)


def mfma_flops_per_tile(sr, n_fft, n_mels, n_mfcc, prec):
    """Executed MFMA FLOPs of one tile of the fused kernel (16 output columns
    whatever the tile's frame count): each block of 8 mels contracts its bin
    band in chunks of 32 bins (16-bit) or 16 bins (FP32) as 16x16 outputs
    (rows 8-15 repeat rows 0-7), first bin rounded down to a multiple of 4;
    the DCT takes 3 16x16x16 MFMAs (hi/lo split) or 4 16x16x4 FP32 MFMAs per
    (block, 16 coefficients). The 16-bit kernels at n_fft 2048 also run each
    frame's FFT on the matrix cores: per stage two 32x32x32 real products per
    parity (a radix-2 step halves each 32-point DFT), each as 4 output tiles x
    3 (hi/lo) 16x16x32 FP16 MFMAs: 48 MFMAs per frame."""
    from metalmom_amd._mel_fb import mel_filterbank
    fb = mel_filterbank(sr, n_fft, n_mels)
    ch = 16 if prec == "f32" else 32
    chunks = 0
    for b in range((n_mels + 7) // 8):
        nz = np.nonzero(fb[8 * b:8 * b + 8].any(axis=0))[0]
        if nz.size:
            x0 = int(nz[0]) & ~3
            chunks += (int(nz[-1]) + 1 - x0 + ch - 1) // ch
    nmh, ndt = (n_mels + 7) // 8, (n_mfcc + 15) // 16
    dct = nmh * ndt * (4 * 2 * 16 * 16 * 4 if prec == "f32" else 3 * 2 * 16 * 16 * 16)
    fft = 16 * 48 * (2 * 16 * 16 * 32) if (prec != "f32" and n_fft == 2048) else 0
    return chunks * 2 * 16 * 16 * ch + dct + fft


def useful_mfma_flops_per_frame(sr, n_fft, n_mels, n_mfcc):
    """FLOPs of the contractions' nonzero work per frame: 2 x filterbank
    nonzeros + 2 x n_mels x n_mfcc (DCT)."""
    from metalmom_amd._mel_fb import mel_filterbank
    return 2 * int(np.count_nonzero(mel_filterbank(sr, n_fft, n_mels))) + 2 * n_mels * n_mfcc


MFMA_PEAK = {"bf16": 2500.0, "f16": 2500.0, "f32": 157.3}   # dense TFLOP/s, published


def kernel_roofline(frames, bytes_launch, n_fft, n_tiles, mfma_tile_flops, useful_mfma_frame, prec, k_ms,
                    traffic=None):
    """Position of one fused launch against the three ceilings: HBM (algorithmic
    bytes), FP32 vector (the FFT's 2.5 N log2 N per frame: the binding unit,
    as the kernel's time is VALU issue of the FFT) and MFMA (executed and
    useful contraction FLOPs)."""
    s = k_ms * 1e-3
    gbs = bytes_launch / s / 1e9
    fft_tf = frames * fft_flops(n_fft) / s / 1e12
    mf_tf = n_tiles * mfma_tile_flops / s / 1e12
    return {"bound": "valu", "achieved": fft_tf, "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": fft_tf / FP32_VECTOR_PEAK_TFLOPS, "frac_of_unpacked_fp32": fft_tf / (FP32_VECTOR_PEAK_TFLOPS / 2),
            "traffic": traffic, "kernel_ms": k_ms, "fft_flops_per_frame": fft_flops(n_fft),
            "hbm": {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                    "bytes_per_launch": bytes_launch, "traffic": traffic},
            "mfma": {"achieved": mf_tf, "peak": MFMA_PEAK[prec], "unit": "TFLOP/s", "frac": mf_tf / MFMA_PEAK[prec],
                     "executed_flops_per_launch": n_tiles * mfma_tile_flops,
                     "useful_flops_per_launch": frames * useful_mfma_frame, "precision": prec}}


def fp64_reference(y, sr=22050, n_fft=2048, hop=512, n_mels=128, n_mfcc=13, top_db=80.0):
    """Float64 STFT magnitude [F, T], mel power [n_mels, T] and MFCC of one
    clip with the reference's semantics (centre zero padding, periodic Hann,
    Slaney mel weights, whole-clip
    top_db, orthonormal DCT-II), for the bench's error report."""
    import scipy.fft
    from metalmom_amd._mel_fb import mel_filterbank
    n = np.arange(n_fft)
    w = 0.5 - 0.5 * np.cos(2 * np.pi * n / n_fft)
    x = np.pad(np.asarray(y, dtype=np.float64), (n_fft // 2, n_fft // 2))
    T = 1 + (len(x) - n_fft) // hop
    fr = np.lib.stride_tricks.as_strided(x, shape=(T, n_fft), strides=(x.strides[0] * hop, x.strides[0]))
    X = np.fft.rfft(fr * w, axis=1).T
    mel = mel_filterbank(sr, n_fft, n_mels, dtype=np.float64) @ (X.real ** 2 + X.imag ** 2)
    db = 10.0 * np.log10(np.maximum(mel, 1e-10))
    db = np.maximum(db, db.max() - top_db)
    return np.abs(X), mel, scipy.fft.dct(db, type=2, norm="ortho", axis=0)[:n_mfcc]


def _db(m, top_db=80.0):
    d = 10.0 * np.log10(np.maximum(np.asarray(m, dtype=np.float64), 1e-10))
    return np.maximum(d, d.max() - top_db)


def error_report(torch, clips, d_in, B, idx):
    """FP32 and BF16 outputs of sampled clips of the batch against the float64
    reference: STFT magnitude max-abs / max, mel max-abs dB, MFCC max-abs /
    max |MFCC|. Each clip runs through a one-clip plan on the batch's own
    device samples (bitwise equal to its block of the batch result)."""
    from metalmom_amd.batch import SpectralBatch
    common = dict(sr=22050, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13, top_db=80.0)
    res = {"f32": {"stft_rel": 0.0, "mel_db_abs": 0.0, "mfcc_rel": 0.0}, "bf16": {"mel_db_abs": 0.0, "mfcc_rel": 0.0}}

    def run(kind, prec, i):
        ioff = B.clips[i][0]
        n = len(clips[i])
        S = SpectralBatch([n], kind=kind, precision=prec, **common)
        o = S.empty_output()
        S.run(d_in[ioff:ioff + n], o)
        torch.cuda.synchronize()
        _, _, rows, frames = S.clips[0]
        a = o[:rows * frames].cpu().numpy().reshape(rows, frames).astype(np.float64)
        S.close()
        return a

    for i in idx:
        mag, mel, mf = fp64_reference(clips[i])
        res["f32"]["stft_rel"] = max(res["f32"]["stft_rel"], float(np.abs(run("stft", "f32", i) - mag).max() / mag.max()))
        for prec in ("f32", "bf16"):
            r = res[prec]
            r["mel_db_abs"] = max(r["mel_db_abs"], float(np.abs(_db(run("mel", prec, i)) - _db(mel)).max()))
            r["mfcc_rel"] = max(r["mfcc_rel"], float(np.abs(run("mfcc", prec, i) - mf).max() / np.abs(mf).max()))
    res["f32"]["within_targets"] = bool(res["f32"]["stft_rel"] <= 1e-5 and res["f32"]["mel_db_abs"] <= 1e-3
                                        and res["f32"]["mfcc_rel"] <= 1e-4)
    res["bf16"]["within_0.05_db"] = bool(res["bf16"]["mel_db_abs"] <= 0.05)
    res["sample"] = (f"clips {list(map(int, idx))} of the uneven batch (shortest, longest and four between), "
                     f"vs float64 numpy (rfft, Slaney mel, dB with top_db 80, DCT-II ortho)")
    return res


def large_fft_leg(torch, steps):
    """The large-FFT batch: 1024 clips x 10 s at 44.1 kHz, FP16 input, n_fft
    4096, hop 1024, 256 mels, 20 MFCC, FP16 mel / DCT contractions on MFMA
    (FP32 FFT): one persistent launch per step from a hipGraph, grid and shape
    autotuned."""
    from metalmom_amd.batch import SpectralBatch
    from metalmom_amd.workloads import large_fft_batch
    t0 = time.perf_counter()
    clips = large_fft_batch()
    gen_s = time.perf_counter() - t0
    lens = [len(c) for c in clips]
    common = dict(sr=44100, n_fft=4096, hop_length=1024, n_mels=256, n_mfcc=20, precision="f16", input_dtype="f16")
    B = SpectralBatch(lens, kind="mfcc", **common)
    d_in = B.pack(clips)
    err_idx = [0, 341, 682, 1023]
    err_clips = {i: np.asarray(clips[i], dtype=np.float64) for i in err_idx}
    del clips
    out = B.empty_output()
    st = torch.cuda.Stream()
    wpc, _ = B.autotune(d_in, out, iters=2)
    for _ in range(2):
        B.run(d_in, out, stream=st, graph=True)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(st)
        B.run(d_in, out, stream=st, graph=True)
        b.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [a.elapsed_time(b) for a, b in ev]
    k_ms, _ = B.time(d_in, out, iters=10)
    frames = B.total_frames
    bytes_launch = 2 * B.total_input + 4 * 20 * frames
    rl = kernel_roofline(frames, bytes_launch, 4096, B.n_tiles, mfma_flops_per_tile(44100, 4096, 256, 20, "f16"),
                         useful_mfma_flops_per_frame(44100, 4096, 256, 20), "f16", k_ms,
                         measured_traffic(bytes_launch, LARGE_FFT_TRAFFIC_FILE))
    sh = B.launch_shape()
    # FP64 error of sampled clips (outside the timed region): the batch's own
    # MFCC blocks, and the mel of one-clip plans on the same device samples
    B.run(d_in, out, stream=st, graph=True)
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    err = {"mel_db_abs": 0.0, "mfcc_rel": 0.0}
    for i in err_idx:
        _, mel, mf = fp64_reference(err_clips[i], sr=44100, n_fft=4096, hop=1024, n_mels=256, n_mfcc=20)
        _, ooff, rows, nfr = B.clips[i]
        got = res[ooff:ooff + rows * nfr].reshape(rows, nfr).astype(np.float64)
        err["mfcc_rel"] = max(err["mfcc_rel"], float(np.abs(got - mf).max() / np.abs(mf).max()))
        ioff, n = B.clips[i][0], lens[i]
        S = SpectralBatch([n], kind="mel", **common)
        o = S.empty_output()
        S.run(d_in[ioff:ioff + n], o)
        torch.cuda.synchronize()
        _, _, mrows, mfr = S.clips[0]
        gm = o[:mrows * mfr].cpu().numpy().reshape(mrows, mfr)
        S.close()
        err["mel_db_abs"] = max(err["mel_db_abs"], float(np.abs(_db(gm) - _db(mel)).max()))
    err["within_0.05_db"] = bool(err["mel_db_abs"] <= 0.05)
    err["sample"] = (f"clips {err_idx} of the batch vs float64 numpy (rfft, Slaney mel, dB with top_db 80, "
                     f"DCT-II ortho); MFCC from the batch result, mel from one-clip plans")
    B.close()
    return {"workload": "1024 clips x 10 s @44.1 kHz, FP16 input, n_fft 4096 hop 1024, 256 mels, 20 MFCC, "
                        "FP16 MFMA mel/DCT (FP32 FFT), one persistent hipGraph launch per step",
            "frames": frames, "frames_per_s": frames * steps / wall, "step_ms": timing_stats(step_ms),
            "kernel_ms": k_ms, "persistent_wg_per_cu": wpc, "launch": sh, "roofline": rl,
            "error_vs_fp64": err, "input_generation_s": gen_s}


def single_clip_leg(iters=20):
    """One clip per call through the drop-in Python API (a context per call,
    host buffers in and out, as the reference's wrappers do): the plain 30 s
    clip at 22.05 kHz (MFCC 13, zero and reflect padding), and the reference's
    own benchmark condition, a 30 s 440 Hz sine at 44.1 kHz (STFT magnitude,
    128-mel spectrogram, MFCC 20), whose published Apple M4 Max timings are
    7.9 / 7.1 / 7.3 ms. Wall time per call, after 3 warm-up calls."""
    import metalmom_amd as mm
    from metalmom_amd import feature
    from metalmom_amd.workloads import plain_clip
    y0 = plain_clip()
    t = np.arange(30 * 44100) / 44100
    y1 = np.sin(2 * np.pi * 440 * t).astype(np.float32)

    def timed(fn):
        for _ in range(3):
            r = fn()
        ms = []
        for _ in range(iters):
            t0 = time.perf_counter()
            r = fn()
            ms.append((time.perf_counter() - t0) * 1e3)
        return timing_stats(ms), r.shape[-1]

    out = {}
    for name, fn in [
        ("mfcc_30s_22k_constant", lambda: feature.mfcc(y=y0, sr=22050, n_mfcc=13)),
        ("mfcc_30s_22k_reflect", lambda: mm.spectral(y0, kind="mfcc", sr=22050, n_mfcc=13, pad_mode="reflect")),
        ("stft_30s_44k_sine", lambda: mm.stft(y1, n_fft=2048)),
        ("mel_30s_44k_sine", lambda: feature.melspectrogram(y=y1, sr=44100)),
        ("mfcc20_30s_44k_sine", lambda: feature.mfcc(y=y1, sr=44100, n_mfcc=20)),
    ]:
        st, frames = timed(fn)
        out[name] = {"ms": st, "frames": int(frames), "frames_per_s": frames / (st["p50"] * 1e-3)}
    out["reference_published_ms_m4max"] = {"stft_30s_44k_sine": 7.9, "mel_30s_44k_sine": 7.1,
                                           "mfcc20_30s_44k_sine": 7.3}
    return out


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


def serve_leg(rates, seconds, sources, pcm=None, kinds=None, precision="f32", backlog=0):
    """Open-loop Poisson serving at each rate (after a 0.5 s warm-up at that
    rate) and, with backlog > 0, a saturation point: that many requests
    queued at t = 0. Latency runs from each request's scheduled arrival to
    its result on the host."""
    from metalmom_amd.serve import KINDS, SpectralServer, open_loop
    kinds = kinds or KINDS
    out = []
    for rate in list(rates) + ([float("inf")] if backlog else []):
        srv = SpectralServer(hbm_budget_bytes=32 << 30)
        kw = dict(sources=sources, pcm=pcm, kinds=kinds, precision=precision)
        if np.isfinite(rate):
            open_loop(srv, rate=rate, duration=0.5, seed=99, **kw)
            r = open_loop(srv, rate=rate, duration=seconds, seed=4, **kw)
        else:
            open_loop(srv, rate=float("inf"), n_requests=max(1, backlog // 4), seed=99, **kw)
            r = open_loop(srv, rate=float("inf"), n_requests=backlog, seed=4, **kw)
        st = srv.stats()
        srv.close()
        e = {"rate_req_s": rate if np.isfinite(rate) else "saturation",
             "achieved_rate": r["achieved_rate"], "requests": r["n_requests"],
             "p50_ms": r["p50"] * 1e3, "p99_ms": r["p99"] * 1e3,
             "server_p50_ms": r["server_p50"] * 1e3, "server_p99_ms": r["server_p99"] * 1e3,
             "frames_per_s": r["frames_per_s"], "device_steps": st["steps"],
             "types": "+".join(kinds), "precision": precision}
        if pcm is not None:
            e["input"] = "int16 stereo PCM, resampled to 22.05 kHz on the GPU"
        out.append(e)
        log(f"serving {e['types']} {precision}{' PCM' if pcm is not None else ''} @{e['rate_req_s']}: "
            f"submitted {r['achieved_rate']:.0f} req/s, p50 {e['p50_ms']:.2f} ms, p99 {e['p99_ms']:.2f} ms "
            f"(server {e['server_p50_ms']:.2f} / {e['server_p99_ms']:.2f}), {r['frames_per_s'] / 1e6:.2f} M frames/s, "
            f"{st['steps']} steps for {r['n_requests']} requests")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the GPU ramps up over ~40 steps after any idle period (even
    # 50 ms), so 50 warm-up steps reach the sustained rate before the timed ones
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain single-clip leg")
    ap.add_argument("--profile", action="store_true", help="short run for a profiler")
    ap.add_argument("--serve-rates", default="200,500,1000,2000,5000",
                    help="Poisson request rates (req/s) for the serving legs; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-backlog", type=int, default=4000,
                    help="requests queued at once for the saturation point (0: none)")
    ap.add_argument("--serve-bf16-rates", default="1000,5000",
                    help="rates of the BF16 MFCC-only serving leg; empty to skip")
    ap.add_argument("--serve-pcm-rates", default="1000,5000",
                    help="rates for the raw-PCM serving leg (int16 stereo at 22.05 / 44.1 kHz, analysed at "
                         "22.05 kHz); empty to skip")
    ap.add_argument("--grid", type=int, default=0,
                    help="persistent workgroups per CU (skips the autotuner; 0 = autotune)")
    ap.add_argument("--shape", type=int, default=0,
                    help="fused-kernel workgroup shape with --grid (1: 16x1, 2: 8x1, 3: 8x2, 4: 16x1 pipelined; 0: default)")
    ap.add_argument("--no-large-fft", action="store_true", help="skip the large-FFT FP16 batch leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the FP64 error report")
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
        if args.shape > 0:
            B.set_shape(args.shape)
        best_grid = args.grid
        B.set_grid(best_grid)
        log(f"persistent grid: {best_grid} workgroups/CU (fixed), launch {B.launch_shape()}")
    else:
        best_grid, tune_ms = B.autotune(d_in, out, iters=2)
        log(f"autotuned: {B.launch_shape()} ({tune_ms:.3f} ms/step)")
    launch = B.launch_shape()

    # a profiler run keeps the warm-up (its trace is read at steady state)
    # and times few steps
    steps = 10 if args.profile else args.steps
    warm = args.warmup
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    with torch.cuda.stream(stream):
        for _ in range(warm):
            B.run(d_in, out, stream=stream, graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            B.run(d_in, out, stream=stream, graph=True)
            b.record(stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    value = frames * steps / dt
    ms_step = dt * 1e3 / steps
    step_stats = timing_stats([a.elapsed_time(b) for a, b in ev])
    log(f"bf16 batched: {value / 1e6:.1f} M frames/s ({ms_step:.3f} ms/step; events: mean {step_stats['mean']:.3f}, "
        f"min {step_stats['min']:.3f}, p50 {step_stats['p50']:.3f} ms)")

    if args.copy_bytes > 0:
        src = torch.empty(args.copy_bytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_bytes} bytes read, {args.copy_bytes} bytes written")
        del src, dst

    # fused kernel time from HIP events on its own stream (the plan's stream),
    # after the same warm-up
    B.time(d_in, out, iters=warm)
    k_ms, s_ms = B.time(d_in, out, iters=20 if args.profile else 50)
    bytes_launch = 4 * B.total_input + 4 * n_mfcc * frames
    traffic = measured_traffic(bytes_launch)
    roofline = kernel_roofline(frames, bytes_launch, 2048, B.n_tiles, mfma_flops_per_tile(22050, 2048, 128, n_mfcc, "bf16"),
                               useful_mfma_flops_per_frame(22050, 2048, 128, n_mfcc), "bf16", k_ms, traffic)
    log(f"fused kernel {k_ms:.3f} ms, step {s_ms:.3f} ms: FFT {roofline['achieved']:.1f} TFLOP/s "
        f"({100 * roofline['frac']:.1f} % of {FP32_VECTOR_PEAK_TFLOPS} FP32 vector), HBM {roofline['hbm']['achieved']:.0f} GB/s "
        f"({100 * roofline['hbm']['frac']:.1f} %), MFMA {roofline['mfma']['achieved']:.1f} TFLOP/s executed "
        f"({100 * roofline['mfma']['frac']:.2f} % of {MFMA_PEAK['bf16']:.0f})")
    if traffic is not None:
        log(f"HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated): {traffic / 1e6:.1f} MB "
            f"= {traffic / bytes_launch:.2f}x algorithmic")

    extra = {"step_ms_events": step_stats}
    if not args.profile:
        # FP32 batched path on the same batch
        B32 = make("f32")
        out32 = B32.empty_output()
        B32.autotune(d_in, out32, iters=2)
        for _ in range(warm):
            B32.run(d_in, out32, graph=True)
        torch.cuda.synchronize()
        n32 = max(5, steps // 2)
        t0 = time.perf_counter()
        for _ in range(n32):
            B32.run(d_in, out32, graph=True)
        torch.cuda.synchronize()
        f32_fps = frames * n32 / (time.perf_counter() - t0)
        k32, _ = B32.time(d_in, out32, iters=50)
        extra["batched_fp32_frames_per_s"] = f32_fps
        extra["batched_fp32_kernel_ms"] = k32
        extra["batched_fp32_launch"] = B32.launch_shape()
        log(f"fp32 batched: {f32_fps / 1e6:.1f} M frames/s (kernel {k32:.3f} ms)")
        B32.close()

        if not args.no_errors:
            order = np.argsort(lens)
            idx = sorted({int(order[0]), int(order[-1])} | {int(order[len(order) * q // 5]) for q in (1, 2, 3, 4)})
            err = error_report(torch, clips, d_in, B, idx)
            extra["error_vs_fp64"] = err
            log(f"error vs fp64 on {len(idx)} clips: FP32 STFT {err['f32']['stft_rel']:.2e} rel, mel "
                f"{err['f32']['mel_db_abs']:.2e} dB, MFCC {err['f32']['mfcc_rel']:.2e} rel; BF16 mel "
                f"{err['bf16']['mel_db_abs']:.2e} dB, MFCC {err['bf16']['mfcc_rel']:.2e} rel")

        if not args.no_plain:
            # plain path: one clip per launch, FP32, no graph, in sequence on
            # one stream (each plan would otherwise use its own pooled stream
            # and clips could overlap across hardware queues); 5 timed passes
            singles = [SpectralBatch([n], kind="mfcc", n_fft=2048, hop_length=512, n_mels=128,
                                     n_mfcc=n_mfcc, precision="f32") for n in lens]
            ins = [d_in[B.clips[i][0]:B.clips[i][0] + lens[i]] for i in range(len(lens))]
            outs = [s.empty_output() for s in singles]
            plain_stream = torch.cuda.Stream()
            for s, i_, o_ in zip(singles, ins, outs):
                s.run(i_, o_, stream=plain_stream)
            torch.cuda.synchronize()
            passes = []
            for _ in range(5):
                t0 = time.perf_counter()
                for s, i_, o_ in zip(singles, ins, outs):
                    s.run(i_, o_, stream=plain_stream)
                torch.cuda.synchronize()
                passes.append((time.perf_counter() - t0) * 1e3)
            pst = timing_stats(passes)
            plain_fps = frames / (pst["mean"] * 1e-3)
            extra["plain_fp32_frames_per_s"] = plain_fps
            extra["plain_fp32_pass_ms"] = pst
            extra["speedup_vs_plain_fp32"] = value / plain_fps
            log(f"plain fp32 single-clip: {plain_fps / 1e6:.2f} M frames/s (pass mean {pst['mean']:.2f} ms, min "
                f"{pst['min']:.2f}, p50 {pst['p50']:.2f}); batched bf16 is {value / plain_fps:.2f}x")
            for s in singles:
                s.close()

        if not args.profile:
            # end to end from host buffers (mm_spectral_batch: clips copied
            # into pinned memory, one H2D copy, the persistent launch, one D2H
            # copy, per-clip results copied out), reported beside the
            # device-resident figure
            from metalmom_amd.batch import spectral_batch
            e2e = []
            for it in range(4):
                t0 = time.perf_counter()
                res = spectral_batch(clips, kind="mfcc", n_fft=2048, hop_length=512, n_mels=128, n_mfcc=n_mfcc,
                                     precision="bf16")
                if it:
                    e2e.append((time.perf_counter() - t0) * 1e3)
            assert sum(r.shape[1] for r in res) == frames
            est = timing_stats(e2e)
            extra["end_to_end_host_batch_bf16"] = {
                "ms": est, "frames_per_s": frames / (est["p50"] * 1e-3),
                "input_bytes": int(4 * B.total_input),
                "path": "host float32 clips -> pinned staging -> one H2D copy -> persistent launch -> one D2H copy "
                        "-> per-clip host arrays (mm_spectral_batch, plan created per call)"}
            log(f"end to end from host buffers (BF16): {frames / (est['p50'] * 1e-3) / 1e6:.2f} M frames/s "
                f"(p50 {est['p50']:.1f} ms per batch)")

    if not args.no_large_fft and not args.profile:
        lf = large_fft_leg(torch, max(5, steps // 2))
        extra["large_fft_f16"] = lf
        log(f"large-FFT FP16 batch: {lf['frames_per_s'] / 1e6:.1f} M frames/s, kernel {lf['kernel_ms']:.3f} ms, "
            f"{lf['roofline']['hbm']['achieved']:.0f} GB/s, FFT {lf['roofline']['achieved']:.1f} TFLOP/s "
            f"(inputs generated in {lf['input_generation_s']:.1f} s)")
    if not args.profile:
        sc = single_clip_leg()
        extra["single_clip_api"] = sc
        log("single clip through the Python API (p50 ms): " +
            ", ".join(f"{k} {v['ms']['p50']:.2f}" for k, v in sc.items() if isinstance(v, dict) and "ms" in v))

    serving = []
    if args.serve_rates and not args.profile:
        from metalmom_amd.serve import make_sources, pcm_sources
        sources = make_sources()
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        serving += serve_leg(rates, args.serve_seconds, sources, backlog=args.serve_backlog)
        if args.serve_bf16_rates:
            serving += serve_leg([float(r) for r in args.serve_bf16_rates.split(",") if r], args.serve_seconds,
                                 sources, kinds=("mfcc",), precision="bf16", backlog=args.serve_backlog)
        if args.serve_pcm_rates:
            serving += serve_leg([float(r) for r in args.serve_pcm_rates.split(",") if r], args.serve_seconds,
                                 sources, pcm=pcm_sources(sources))

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
                   "persistent_wg_per_cu": best_grid, "launch": launch},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    rec.update(extra)
    if serving:
        rec["serving"] = serving
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()

"""GPU parity at the benchmark's own shapes, through the persistent batch path.

The batched workloads run every clip in one persistent launch: workgroup g
processes tiles [g*C, (g+1)*C), C = ceil(tiles / grid), so at these sizes
each workgroup runs many tiles in sequence (the next tile's samples are
prefetched during the current tile's contraction). These tests check that
path at full size:

* the 256-clip uneven batch (lengths U[1 s, 60 s] at 22.05 kHz, 336,013
  frames, n_fft 2048, hop 512, 128 mels, 13 MFCC) in FP32 and with BF16
  mel/DCT contractions, with the autotuned workgroup shape and grid, with one
  workgroup per CU (~80 tiles each) and with every other shape: all bitwise
  equal, and every clip bitwise equal to the same clip run alone;
* a sample of its clips (the shortest, the longest and clips whose tiles are
  split between two workgroups) against the float64 oracle at the FP32
  targets (STFT magnitude max-abs error / max <= 1e-5, mel <= 1e-3 dB, MFCC
  max-abs error / max |MFCC| <= 1e-4) and the 16-bit tier (mel <= 0.05 dB,
  MFCC <= 2e-3);
* the large-FFT batch (1024 clips x 10 s at 44.1 kHz, FP16 input, n_fft 4096,
  hop 1024, 256 mels, 20 MFCC, FP16 contractions) the same way;
* the reference's stress cases (Tests/test_memory_stress.py:13-93: 500
  repeated STFTs without growth, a 5-minute signal with the exact frame
  count; ThreadSafetyTests.swift:264-283: 1000 context create/destroy pairs).
"""
import ctypes as C
import tracemalloc

import numpy as np
import pytest

from oracle import spectral as o

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def db_err(mel, ref_mel):
    return float(np.abs(o.power_to_db(np.asarray(mel, np.float64)) - o.power_to_db(ref_mel)).max())


def sync():
    import torch
    torch.cuda.synchronize()


def tile_ranges(B, tile_frames):
    """[first tile, end tile) of every clip in the plan's tile list."""
    r, t = [], 0
    for (_, _, _, frames) in B.clips:
        n = (frames + tile_frames - 1) // tile_frames
        r.append((t, t + n))
        t += n
    return r


def sample_clips(B, lens, shape, k=6):
    """Shortest clip, longest clip and k clips whose tiles are split between
    two persistent workgroups (spread over the batch)."""
    C_ = shape["tiles_per_wg"]
    split = [i for i, (a, b) in enumerate(tile_ranges(B, shape["tile_frames"]))
             if (a // C_) != ((b - 1) // C_)]
    assert len(split) >= k, (len(split), shape)
    pick = [split[j * (len(split) - 1) // (k - 1)] for j in range(k)]
    out = [int(np.argmin(lens)), int(np.argmax(lens))] + pick
    return sorted(set(out))


def run_plan(B, d_in, graph=True):
    out = B.empty_output()
    B.run(d_in, out, graph=graph)
    sync()
    return out


def clip_out(B, out, i):
    _, off, rows, frames = B.clips[i]
    return out[off:off + rows * frames].cpu().numpy().reshape(rows, frames)


def assert_single_clip_bitwise(B, d_in, res, lens, make, stride=1):
    """Every stride-th clip run alone (its own one-clip plan on the same
    device samples) equals its block of the batch result bit for bit."""
    for i in range(0, len(lens), stride):
        ioff, ooff, rows, frames = B.clips[i]
        S = make([int(lens[i])])
        o1 = S.empty_output()
        S.run(d_in[ioff:ioff + int(lens[i])], o1)
        sync()
        alone = o1[:rows * frames].cpu().numpy()
        assert np.array_equal(alone, res[ooff:ooff + rows * frames]), i
        S.close()


# -- top_db clamping inside persistent batches, and the pipelined shape ------------

def quiet_section_clips(n=64, seed=7, sr=22050):
    """n clips of up to 8 s of noise (every 9th shorter than 3000 samples);
    every 3rd has its middle third 120 dB down, so the whole-clip top_db
    floor clamps those frames."""
    rng = np.random.default_rng(seed)
    clips = []
    for i in range(n):
        m = int(rng.integers(200, 8 * sr)) if i % 9 else int(rng.integers(1, 3000))
        x = (0.1 * rng.standard_normal(m)).astype(np.float32)
        if i % 3 == 0:
            x[m // 3:2 * m // 3] *= np.float32(1e-6)
        clips.append(x)
    return clips


@pytest.mark.parametrize("prec,tol", [("f32", 1e-4), ("bf16", 2e-3), ("f16", 2e-3)])
def test_batched_top_db_clamp(prec, tol):
    """The second (clamping) pass of a persistent batch recomputes every tile
    whose frames fall below its clip's top_db floor, also when a workgroup's
    other tiles need nothing (regression: a workgroup whose first tile was
    unclamped used to leave without recomputing the rest). Every clip with a
    quiet section matches the float64 oracle and its own single-clip run; for
    the 16-bit contractions the pipelined 16-wave shape (staging rows and DCT
    partials double-buffered, one barrier per tile, each tile's block sum done
    after the next tile's barrier or after the loop) gives the bits of the
    plain 16-wave shape, with long persistent loops (one workgroup per CU) and
    short ones, graph or direct launches, and clips shorter than one tile."""
    from metalmom_amd.batch import SpectralBatch
    SR = 22050
    clips = quiet_section_clips(sr=SR)
    lens = np.array([len(c) for c in clips], dtype=np.int64)

    def make(ls):
        return SpectralBatch(ls, kind="mfcc", sr=SR, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13,
                             top_db=80.0, precision=prec)

    B = make(lens)
    d_in = B.pack(clips)
    ref = None
    for shape in ((1, 4) if prec != "f32" else (1,)):
        B.set_shape(shape)
        for g in (0, 1):
            B.set_grid(g)
            sh = B.launch_shape()
            assert (sh["shape"], sh["waves_per_wg"], sh["tile_frames"]) == (shape, 16, 16), sh
            if g == 1:
                assert sh["tiles_per_wg"] >= 3, sh
            for graph in (False, True):
                r = run_plan(B, d_in, graph=graph).cpu().numpy()
                if ref is None:
                    ref = r
                    assert np.isfinite(ref).all()
                else:
                    assert np.array_equal(r, ref), (shape, g, graph)
    quiet = [i for i in range(len(clips)) if i % 3 == 0]
    for i in quiet:
        _, off, rows, frames = B.clips[i]
        got = ref[off:off + rows * frames].reshape(rows, frames)
        want = o.mfcc(clips[i], SR, 13, 2048, 512)
        assert got.shape == want.shape
        assert rel_err(got, want) <= tol, i
    B.set_shape(1)
    assert_single_clip_bitwise(B, d_in, ref, lens, make, stride=3)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_plan_steps_restore_clip_maxima(prec):
    """A plan's steps launch no fill of the top_db clip maxima: they
    alternate between two sets, and each step's first pass sets the other one
    to -inf for the next step. Steps alternate between two inputs whose clip
    maxima differ by 60 dB (the loud parts of the second are 1000x quieter,
    so its floor sits 60 dB lower and clamps other frames), through graph
    replays of both sets and direct launches, in an order that gives each
    input both sets; every step must equal a fresh plan's first step on the
    same input bit for bit (a maximum left over from an earlier step would
    raise the floor)."""
    from metalmom_amd.batch import SpectralBatch
    SR = 22050
    clips = quiet_section_clips(n=48, seed=11, sr=SR)
    quiet = []
    for i, c in enumerate(clips):
        q = c.copy()
        if i % 3 == 0:   # loud thirds 60 dB down, the quiet middle unchanged
            m = len(q)
            q[:m // 3] *= np.float32(1e-3)
            q[2 * m // 3:] *= np.float32(1e-3)
        quiet.append(q)
    lens = np.array([len(c) for c in clips], dtype=np.int64)

    def make():
        return SpectralBatch(lens, kind="mfcc", sr=SR, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13,
                             top_db=80.0, precision=prec)

    refs = []
    for data in (clips, quiet):
        F = make()
        refs.append(run_plan(F, F.pack(data)).cpu().numpy())
        F.close()
    assert not np.array_equal(refs[0], refs[1])
    B = make()
    B.set_grid(1)   # long persistent loops
    d = [B.pack(clips), B.pack(quiet)]
    outs = [B.empty_output(), B.empty_output()]   # one graph per (input, output): replays
    for step, (k, graph) in enumerate([(0, True), (1, True), (0, True), (1, False), (1, True), (0, False),
                                       (0, True), (1, True)]):
        outs[k].fill_(float("nan"))
        B.run(d[k], outs[k], graph=graph)
        sync()
        assert np.array_equal(outs[k].cpu().numpy(), refs[k]), (step, k, graph)
    B.close()


# -- the 256-clip uneven batch ---------------------------------------------------

@pytest.fixture(scope="module")
def uneven():
    from metalmom_amd.workloads import uneven_batch
    clips = uneven_batch()
    return clips, np.array([len(c) for c in clips], dtype=np.int64)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-4), ("bf16", 2e-3)])
def test_uneven_batch_mfcc_persistent(uneven, prec, tol):
    from metalmom_amd.batch import SpectralBatch
    clips, lens = uneven

    def make(ls):
        return SpectralBatch(ls, kind="mfcc", sr=22050, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13,
                             top_db=80.0, precision=prec)

    B = make(lens)
    assert B.total_frames == 336013 and B.n_clips == 256
    d_in = B.pack(clips)
    B.autotune(d_in, B.empty_output(), iters=1)
    tuned = B.launch_shape()
    print("autotuned launch:", tuned)
    assert tuned["grid"] > 0 and tuned["tiles_per_wg"] >= 2
    res = run_plan(B, d_in).cpu().numpy()
    assert np.isfinite(res).all()

    # one workgroup per CU: the longest persistent loops
    B.set_grid(1)
    one = B.launch_shape()
    assert one["tiles_per_wg"] >= 40, one
    assert np.array_equal(run_plan(B, d_in).cpu().numpy(), res)
    # every workgroup shape gives the same bits (the pipelined 16-wave shape
    # exists for the 16-bit contractions only)
    shapes = [(1, 16, 16), (2, 8, 8), (3, 8, 16)] + ([(4, 16, 16)] if prec != "f32" else [])
    for shape, waves, frames in shapes:
        B.set_shape(shape)
        B.set_grid(0)
        sh = B.launch_shape()
        assert (sh["shape"], sh["waves_per_wg"], sh["tile_frames"]) == (shape, waves, frames), sh
        assert np.array_equal(run_plan(B, d_in, graph=False).cpu().numpy(), res), shape
    if prec == "f32":
        with pytest.raises(RuntimeError):
            B.set_shape(4)

    # every clip alone
    B.set_shape(tuned["shape"])
    assert_single_clip_bitwise(B, d_in, res, lens, make)

    # a sample against the float64 oracle
    errs = {}
    for i in sample_clips(B, lens, tuned):
        _, off, rows, frames = B.clips[i]
        got = res[off:off + rows * frames].reshape(rows, frames)
        ref = o.mfcc(clips[i], 22050, 13, 2048, 512, n_mels=128)
        assert got.shape == ref.shape
        errs[i] = rel_err(got, ref)
    print(f"{prec} MFCC max rel error vs fp64 per sampled clip:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= tol
    B.close()


def test_uneven_batch_stft_and_mel(uneven):
    """STFT magnitude and mel (FP32 and BF16) of the whole batch; sampled
    clips against the oracle at the FP32 targets and the 16-bit tier."""
    from metalmom_amd.batch import SpectralBatch
    clips, lens = uneven
    common = dict(sr=22050, n_fft=2048, hop_length=512, n_mels=128)
    Bs = SpectralBatch(lens, kind="stft", **common)
    d_in = Bs.pack(clips)
    out_s = run_plan(Bs, d_in)
    shape = Bs.launch_shape()
    assert shape["tiles_per_wg"] >= 2
    idx = sample_clips(Bs, lens, shape)
    mel = {}
    for prec in ("f32", "bf16"):
        Bm = SpectralBatch(lens, kind="mel", precision=prec, **common)
        assert Bm.launch_shape()["tiles_per_wg"] >= 2
        out_m = run_plan(Bm, d_in)
        mel[prec] = {i: clip_out(Bm, out_m, i) for i in idx}
        del out_m
        Bm.close()
    e_stft, e_f32, e_bf16 = [], [], []
    for i in idx:
        X = o.stft_complex(clips[i], 2048, 512)
        ref_mag = np.abs(X)
        e_stft.append(rel_err(clip_out(Bs, out_s, i), ref_mag))
        ref_mel = o.mel_filterbank(22050, 2048, 128) @ (X.real ** 2 + X.imag ** 2)
        e_f32.append(db_err(mel["f32"][i], ref_mel))
        e_bf16.append(db_err(mel["bf16"][i], ref_mel))
    print(f"sampled clips {idx}: STFT rel {max(e_stft):.2e}, mel FP32 {max(e_f32):.2e} dB, "
          f"mel BF16 {max(e_bf16):.2e} dB")
    assert max(e_stft) <= 1e-5
    assert max(e_f32) <= 1e-3
    assert max(e_bf16) <= 0.05
    Bs.close()


# -- the large-FFT FP16 batch ------------------------------------------------------

def test_large_fft_f16_batch():
    from metalmom_amd.batch import SpectralBatch
    from metalmom_amd.workloads import large_fft_batch
    clips = large_fft_batch()
    lens = np.array([len(c) for c in clips], dtype=np.int64)
    sr = 44100
    common = dict(sr=sr, n_fft=4096, hop_length=1024, n_mels=256, precision="f16", input_dtype="f16")

    def make(ls):
        return SpectralBatch(ls, kind="mfcc", n_mfcc=20, **common)

    B = make(lens)
    assert B.total_frames == 441344 and B.n_clips == 1024
    d_in = B.pack(clips)
    B.autotune(d_in, B.empty_output(), iters=1)
    shape = B.launch_shape()
    print("autotuned launch:", shape)
    assert shape["tiles_per_wg"] >= 2
    res = run_plan(B, d_in).cpu().numpy()
    assert np.isfinite(res).all()
    B.set_grid(1)
    assert np.array_equal(run_plan(B, d_in).cpu().numpy(), res)
    # 54 tiles per clip: with 3 workgroups per CU the chunks of 72 tiles
    # split clips between workgroups (the tuned grid may not)
    B.set_grid(3)
    split_shape = B.launch_shape()
    assert np.array_equal(run_plan(B, d_in).cpu().numpy(), res)
    B.set_grid(0)
    assert_single_clip_bitwise(B, d_in, res, lens, make, stride=7)

    Bm = SpectralBatch(lens, kind="mel", **common)
    out_m = run_plan(Bm, d_in)
    e_db, e_mfcc = [], []
    fb = o.mel_filterbank(sr, 4096, 256)
    for i in sample_clips(B, lens, split_shape):
        y = clips[i].astype(np.float64)
        X = o.stft_complex(y, 4096, 1024)
        ref_mel = fb @ (X.real ** 2 + X.imag ** 2)
        e_db.append(db_err(clip_out(Bm, out_m, i), ref_mel))
        _, off, rows, frames = B.clips[i]
        assert (rows, frames) == (20, 431)
        ref = o.mfcc_from_logmel(o.power_to_db(ref_mel), 20)
        e_mfcc.append(rel_err(res[off:off + rows * frames].reshape(rows, frames), ref))
    print(f"FP16 large-FFT batch: mel {max(e_db):.2e} dB, MFCC rel {max(e_mfcc):.2e}")
    assert max(e_db) <= 0.05
    assert max(e_mfcc) <= 2e-3
    Bm.close()
    B.close()


# -- the reference's stress cases --------------------------------------------------

def test_repeated_stft_no_growth():
    """Tests/test_memory_stress.py:13-51: 500 STFTs of a 1 s tone; Python-side
    growth under 50 MB, and the device's free memory does not shrink."""
    import torch
    import metalmom_amd as mm
    sr = 22050
    t = np.linspace(0, 1.0, sr, endpoint=False, dtype=np.float32)
    y = np.sin(2 * np.pi * 440 * t).astype(np.float32)
    first = mm.stft(y, n_fft=2048, hop_length=512)
    free0 = torch.cuda.mem_get_info()[0]
    tracemalloc.start()
    before = tracemalloc.take_snapshot()
    for _ in range(500):
        r = mm.stft(y, n_fft=2048, hop_length=512)
        assert r.shape == (1025, 44)
    after = tracemalloc.take_snapshot()
    tracemalloc.stop()
    grow = sum(s.size_diff for s in after.compare_to(before, "lineno") if s.size_diff > 0) / 2 ** 20
    assert grow < 50.0, grow
    assert free0 - torch.cuda.mem_get_info()[0] < 64 << 20
    assert np.array_equal(r, first)


def test_five_minute_signal():
    """Tests/test_memory_stress.py:54-93: 5 minutes at 22.05 kHz (808 tiles of
    16 frames, so the single-clip launch also runs several tiles per
    workgroup): exact frame count, finite, and the oracle's values."""
    import metalmom_amd as mm
    sr, n = 22050, 22050 * 300
    y = np.sin(2 * np.pi * 440 * (np.arange(n, dtype=np.float32) / sr)).astype(np.float32)
    R = mm.stft(y, n_fft=2048, hop_length=512)
    assert R.shape == (1025, 1 + (n + 2048 - 2048) // 512) and R.dtype == np.float32
    assert np.isfinite(R).all() and R.max() > 0
    assert rel_err(R, o.stft_magnitude(y, 2048, 512)) <= 1e-5
    M = mm.mfcc(y=y, sr=sr, n_mfcc=13)
    assert rel_err(M, o.mfcc(y, sr, 13, 2048, 512)) <= 1e-4


def test_rapid_context_create_destroy():
    """ThreadSafetyTests.swift:264-283: 1000 mm_init / mm_destroy pairs, a
    second mm_destroy of the same context is a no-op, and the library still
    computes the same result afterwards."""
    import metalmom_amd as mm
    from metalmom_amd._native import lib
    y = np.random.default_rng(7).standard_normal(30000).astype(np.float32)
    before = mm.mfcc(y=y, n_mfcc=13)
    for i in range(1000):
        ctx = lib.mm_init()
        assert ctx, i
        lib.mm_destroy(ctx)
    ctx = lib.mm_init()
    lib.mm_destroy(ctx)
    lib.mm_destroy(ctx)
    assert np.array_equal(mm.mfcc(y=y, n_mfcc=13), before)


def test_config0_reflect_full_length():
    """The plain 30 s clip with numpy-style reflect padding (the additive
    pad_mode), at full length, against the oracle."""
    import metalmom_amd as mm
    from metalmom_amd.workloads import plain_clip
    y = plain_clip()
    R = mm.spectral(y, kind="mfcc", n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13, pad_mode="reflect")
    O = o.mfcc(y, 22050, 13, 2048, 512, n_mels=128, pad_mode="reflect")
    assert R.shape == O.shape == (13, 1292)
    assert rel_err(R, O) <= 1e-4
    S = mm.spectral(y, kind="stft", n_fft=2048, hop_length=512, pad_mode="reflect")
    assert rel_err(S, o.stft_magnitude(y, 2048, 512, pad_mode="reflect")) <= 1e-5

"""The matrix-core FFT of the 16-bit mel / MFCC kernels at n_fft 2048
(mfft2048: a radix-2 step and two 16-point contractions per 32-point DFT,
FP16 hi/lo operands on v_mfma_f32_16x16x32_f16, per-frame power-of-two
input scale).

* Exactness of the input scale: frames whose largest sample lies outside
  [2^-4, 2^9) are normalised by a power of two before the transform and the
  factor is removed from the frame's exponent afterwards. Two inputs that
  differ by a power of two and are both normalised therefore run the very
  same transform, so their mel spectra agree bit for bit after the exact
  rescale (a property the FP64 oracle does not need to be consulted for).
* Accuracy at extreme amplitudes (1e-9 to 3e4) against the FP64 oracle, to
  the same bound as the other 16-bit tests (0.05 dB on mel, 2e-3 relative on
  MFCC).
* Every frame-loading path of the kernels (even and odd frame starts, zero
  and reflect padding, clips shorter than a frame, windows shorter than
  n_fft, FP16 samples in device memory) and the mel powers 1 and 1.5 with
  the input scale active, against the oracle at the same tier.
"""
import numpy as np
import pytest

from oracle import spectral as o

pytestmark = pytest.mark.gpu

SR = 22050


def clips(seed, n=4):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        m = int(rng.integers(SR, 3 * SR))
        t = np.arange(m) / SR
        x = 0.1 * rng.standard_normal(m)
        for f in rng.uniform(80, 8000, 3):
            x += 0.3 * np.sin(2 * np.pi * f * t)
        out.append(x.astype(np.float32))
    return out


@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_power_of_two_input_scale_is_exact(prec):
    from metalmom_amd.batch import spectral_batch
    base = clips(1)
    for lo, hi in ((-20, -30), (20, 26), (-12, 14)):
        a = spectral_batch([c * np.float32(2.0 ** lo) for c in base], kind="mel", precision=prec)
        b = spectral_batch([c * np.float32(2.0 ** hi) for c in base], kind="mel", precision=prec)
        for ma, mb in zip(a, b):
            # power spectra scale by 2^(2 k); the mel values stay normal floats here
            assert np.array_equal(ma * np.float32(2.0 ** (-2 * lo)), mb * np.float32(2.0 ** (-2 * hi))), (lo, hi)


@pytest.mark.parametrize("prec", ["bf16", "f16"])
@pytest.mark.parametrize("amp", [1e-9, 1e-3, 1.0, 3e4])
def test_extreme_amplitudes_against_oracle(prec, amp):
    from metalmom_amd.batch import spectral_batch
    cs = [c * np.float32(amp) for c in clips(2, 2)]
    mel = spectral_batch(cs, kind="mel", precision=prec)
    mf = spectral_batch(cs, kind="mfcc", n_mfcc=13, precision=prec)
    for c, m, f in zip(cs, mel, mf):
        Om = o.melspectrogram(c, SR, 2048, 512)
        # dB relative to each clip's own peak, floor 80 dB below it (amin does not bite)
        ref = 10 * np.log10(np.maximum(Om, Om.max() * 1e-8))
        got = 10 * np.log10(np.maximum(m.astype(np.float64), Om.max() * 1e-8))
        assert np.abs(got - ref).max() <= 0.05, amp
        if amp >= 1e-3:   # MFCC: power_to_db's amin (1e-10) keeps quieter clips on its floor
            Of = o.mfcc(c, SR, 13, 2048, 512)
            assert np.abs(f - Of).max() / np.abs(Of).max() <= 2e-3, amp


FRAMING = [
    # (label, clip length, spectral kwargs, oracle kwargs)
    ("odd hop", 3 * SR, dict(hop_length=333), dict(hop=333)),
    ("win 1500", 3 * SR, dict(win_length=1500), dict(win_length=1500)),
    ("uncentred", 3 * SR, dict(center=False), dict(center=False)),
    ("reflect", 3 * SR, dict(pad_mode="reflect"), dict(pad_mode="reflect")),
    ("shorter than a frame", 1500, dict(), dict()),
    ("odd hop, reflect, win 1800", 2 * SR + 7, dict(hop_length=441, pad_mode="reflect", win_length=1800),
     dict(hop=441, pad_mode="reflect", win_length=1800)),
]


@pytest.mark.parametrize("prec", ["bf16", "f16"])
@pytest.mark.parametrize("case", FRAMING, ids=[c[0] for c in FRAMING])
def test_framing_paths(prec, case):
    """Every way the matrix-core kernels load a frame: interior frames with
    an even start (pair loads, prefetched in batches), interior frames with
    an odd start (odd hops: single-sample loads), frames over the centre
    padding (zero or reflect index maps, clips shorter than a frame), and a
    window shorter than n_fft (zero-padded centrally in the window table);
    mel in dB and MFCC against the float64 oracle at the 16-bit tier."""
    import metalmom_amd as mm
    label, n, kw, okw = case
    y = clips(31, 1)[0]
    y = np.resize(y, n).astype(np.float32)
    hop = okw.get("hop", 512)
    mel = mm.spectral(y, kind="mel", n_fft=2048, n_mels=128, precision=prec, **kw)
    ref = o.melspectrogram(y, SR, 2048, hop, okw.get("win_length"), okw.get("center", True),
                           pad_mode=okw.get("pad_mode", "constant"))
    assert mel.shape == ref.shape, label
    assert np.abs(o.power_to_db(mel.astype(np.float64)) - o.power_to_db(ref)).max() <= 0.05, label
    mf = mm.spectral(y, kind="mfcc", n_fft=2048, n_mels=128, n_mfcc=13, precision=prec, **kw)
    want = o.mfcc(y, SR, 13, 2048, hop, okw.get("win_length"), center=okw.get("center", True),
                  pad_mode=okw.get("pad_mode", "constant"))
    assert mf.shape == want.shape, label
    assert np.abs(mf - want).max() / np.abs(want).max() <= 2e-3, label


@pytest.mark.parametrize("prec", ["bf16", "f16"])
@pytest.mark.parametrize("power,amp", [(1.0, 1e3), (1.5, 1e3), (1.5, 1e-3), (2.0, 1e-6)])
def test_power_with_input_scale(prec, power, amp):
    """Mel powers other than 2 on the matrix-core kernels, with frames that
    need the power-of-two input scale (largest sample outside [2^-4, 2^9)):
    power 1 takes the square root of (2|X|)^2 and folds the scale into the
    frame exponent, other powers remove the scale before the power."""
    import metalmom_amd as mm
    y = (amp * clips(37, 1)[0]).astype(np.float32)
    mel = mm.spectral(y, kind="mel", n_fft=2048, n_mels=128, power=power, precision=prec)
    ref = o.melspectrogram(y.astype(np.float64), SR, 2048, 512, power=power)
    # dB of a power-p spectrum: 10 log10 over the same amin floor as the oracle's comparison
    d = np.abs(o.power_to_db(mel.astype(np.float64), amin=1e-30) - o.power_to_db(ref, amin=1e-30))
    assert d.max() <= 0.05 * max(power, 1.0), (power, amp, float(d.max()))


@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_f16_input_batch(prec):
    """FP16 samples in device memory through the matrix-core kernel (the
    kernel's 16-bit sample loads, prefetch and padding paths), persistent
    batch with several tiles per workgroup, against the oracle on the same
    FP16 values."""
    import torch
    from metalmom_amd.batch import SpectralBatch
    cs = [c.astype(np.float16) for c in clips(41, 6)]
    B = SpectralBatch([len(c) for c in cs], kind="mfcc", sr=SR, n_fft=2048, hop_length=512, n_mels=128, n_mfcc=13,
                      precision=prec, input_dtype="f16")
    B.set_grid(1)
    d_in = B.pack(cs)
    out = B.empty_output()
    B.run(d_in, out, graph=True)
    torch.cuda.synchronize()
    for c, r in zip(cs, B.unpack(out)):
        want = o.mfcc(c.astype(np.float64), SR, 13, 2048, 512)
        assert r.shape == want.shape
        assert np.abs(r - want).max() / np.abs(want).max() <= 2e-3

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

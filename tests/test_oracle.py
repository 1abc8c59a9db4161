"""Pins the float64 oracle before anything is checked against it.

Each oracle function is checked against the reference's librosa golden vectors
(Tests/golden, copied to tests/golden) and against what the mathematics fixes
(known answers, brute-force DFT, Parseval, scipy's DCT, invariants), chosen so
that a dropped term, wrong sign, wrong index or transposed operand fails.
"""
import numpy as np
import pytest
import scipy.fft

from conftest import load_golden
from oracle import spectral as o


# -- window and framing ------------------------------------------------------

def test_hann_known_answer():
    # WindowTests.swift:60-66
    np.testing.assert_allclose(o.hann_periodic(4), [0.0, 0.5, 1.0, 0.5], atol=1e-15)
    assert o.hann_periodic(0).size == 0
    np.testing.assert_array_equal(o.hann_periodic(1), [1.0])


def test_prepare_window_pad_and_truncate():
    w = o.prepare_window(4, 8)          # centred: 2 zeros, hann(4), 2 zeros
    np.testing.assert_allclose(w, [0, 0, 0, 0.5, 1, 0.5, 0, 0], atol=1e-15)
    w = o.prepare_window(5, 8)          # (8-5)//2 = 1 zero in front
    np.testing.assert_allclose(w[:1], [0]); np.testing.assert_allclose(w[1:6], o.hann_periodic(5))
    w = o.prepare_window(8, 4)          # centre 4 of hann(8): indices 2..5
    np.testing.assert_allclose(w, o.hann_periodic(8)[2:6])
    assert not o.prepare_window(0, 8).any() and not o.prepare_window(-3, 8).any()


def test_frame_counts_known_answers():
    assert o.frame_count(22050, 2048, 512, True) == 44        # STFTTests.swift:8-24
    assert o.frame_count(661500, 2048, 512, True) == 1292     # 30 s at 22.05 kHz
    assert o.frame_count(0, 2048, 512, True) == 1             # EdgeCaseTests.swift:69-81
    assert o.frame_count(0, 2048, 512, False) == 0
    assert o.frame_count(2047, 2048, 512, False) == 0
    assert o.frame_count(2048, 2048, 512, False) == 1
    assert o.frame_count(441000, 4096, 1024, True) == 431


def test_reflect_pad_matches_numpy():
    y = np.arange(5.0)
    np.testing.assert_array_equal(o.pad_signal(y, 4, True, "reflect"), np.pad(y, 2, mode="reflect"))


# -- STFT --------------------------------------------------------------------

def test_stft_golden(signal_440):
    G = load_golden("stft_440hz_default_magnitude")
    M = o.stft_magnitude(signal_440, 2048, 512)
    assert M.shape == G.shape == (1025, 44)
    np.testing.assert_allclose(M, G, rtol=1e-4, atol=1e-4)   # test_parity_stft.py tier
    assert np.abs(M - G).max() / G.max() < 1e-6


@pytest.mark.parametrize("n_fft", [2, 4, 8, 32, 256])
def test_stft_bruteforce_dft(n_fft):
    rng = np.random.default_rng(n_fft)
    y = rng.standard_normal(3 * n_fft + 5)
    hop = max(1, n_fft // 4)
    X = o.stft_complex(y, n_fft, hop, n_fft, True)
    xp = np.pad(y, n_fft // 2)
    w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(n_fft) / n_fft)
    k = np.arange(n_fft // 2 + 1)[:, None]
    n = np.arange(n_fft)[None, :]
    D = np.exp(-2j * np.pi * k * n / n_fft)
    for f in range(X.shape[1]):
        ref = D @ (xp[f * hop:f * hop + n_fft] * w)
        np.testing.assert_allclose(X[:, f], ref, atol=1e-10 * max(1, np.abs(ref).max()))


def test_stft_parseval_and_symmetries():
    rng = np.random.default_rng(7)
    y = rng.standard_normal(5000)
    n = 512
    X = o.stft_complex(y, n, 128, n, True)
    xp = np.pad(y, n // 2)
    w = o.hann_periodic(n)
    for f in (0, 5, X.shape[1] - 1):
        fr = xp[f * 128:f * 128 + n] * w
        full = np.concatenate([X[:, f], np.conj(X[1:-1, f][::-1])])
        assert np.isclose(np.sum(np.abs(full) ** 2) / n, np.sum(fr ** 2))
    assert np.allclose(X[0].imag, 0) and np.allclose(X[-1].imag, 0)


def test_stft_tone_peak_bin():
    sr, n = 22050, 2048
    y = np.sin(2 * np.pi * 440 * np.arange(sr) / sr)
    M = o.stft_magnitude(y, n, 512)
    assert np.argmax(M[:, 20]) == round(440 * n / sr)          # STFTTests.swift:28-66


# -- mel ---------------------------------------------------------------------

@pytest.mark.parametrize("name,args", [
    ("mel_filterbank_128_2048", (22050, 2048, 128)),
    ("mel_filterbank_40_1024", (22050, 1024, 40)),
    ("mel_filterbank_64_2048_300_8000", (22050, 2048, 64, 300, 8000)),
])
def test_mel_filterbank_golden(name, args):
    G = load_golden(name)
    fb = o.mel_filterbank(*args)
    assert fb.shape == G.shape
    np.testing.assert_allclose(fb, G, rtol=1e-5, atol=1e-6)   # test_parity_mel_filterbank.py tier
    assert np.abs(fb - G).max() < 1e-8


def test_mel_scale_closed_form():
    assert np.isclose(o.hz_to_mel(1000.0), 15.0)
    assert np.isclose(o.hz_to_mel(200.0), 3.0)
    assert np.isclose(o.mel_to_hz(o.hz_to_mel(3456.0)), 3456.0)


@pytest.mark.parametrize("name,kw", [
    ("mel_spectrogram_440hz_default", {}),
    ("mel_spectrogram_440hz_40_1024", dict(n_fft=1024, n_mels=40)),
    ("mel_spectrogram_440hz_64_fmin300_fmax8000", dict(n_mels=64, fmin=300, fmax=8000)),
])
def test_melspectrogram_golden(signal_440, name, kw):
    G = load_golden(name)
    M = o.melspectrogram(signal_440, 22050, **kw)
    np.testing.assert_allclose(M, G, rtol=1e-4, atol=1e-4)
    assert np.abs(M - G).max() / G.max() < 1e-6


# -- dB ----------------------------------------------------------------------

def test_db_known_answers():
    # ScalingTests.swift:39-56
    np.testing.assert_allclose(o.power_to_db(np.array([0.5]), top_db=None), [-3.0102999566], atol=1e-9)
    np.testing.assert_allclose(o.power_to_db(np.array([100.0]), top_db=None), [20.0], atol=1e-12)
    d = o.power_to_db(np.array([1.0, 1e-12]), top_db=80.0)
    np.testing.assert_allclose(d, [0.0, -80.0])


def test_db_goldens(signal_440):
    M = o.stft_magnitude(signal_440, 2048, 512)
    P = M ** 2
    np.testing.assert_allclose(o.power_to_db(P, ref=P.max()), load_golden("stft_440hz_power_db"), atol=1e-4)
    np.testing.assert_allclose(o.amplitude_to_db(M, ref=M.max()), load_golden("stft_440hz_amplitude_db"),
                               atol=1e-4)
    np.testing.assert_allclose(o.amplitude_to_db(M, ref=1.0), load_golden("stft_440hz_amplitude_db_ref1"),
                               atol=1e-4)


# -- DCT / MFCC ---------------------------------------------------------------

def test_dct_basis_matches_scipy_and_is_orthonormal():
    B = o.dct_ortho_basis(128, 128)
    x = np.random.default_rng(3).standard_normal(128)
    np.testing.assert_allclose(B @ x, scipy.fft.dct(x, type=2, norm="ortho"), atol=1e-12)
    np.testing.assert_allclose(B @ B.T, np.eye(128), atol=1e-12)
    c = o.mfcc_from_logmel(np.full((40, 3), 2.0), 13)          # constant column: c0 only
    assert np.allclose(c[1:], 0) and np.allclose(c[0], 2.0 * np.sqrt(40))


@pytest.mark.parametrize("name,kw,tol", [
    ("mfcc_440hz_default", dict(n_mfcc=20), 1e-4),
    ("mfcc_440hz_13_1024", dict(n_mfcc=13, n_fft=1024, n_mels=40), 1e-4),
    ("mfcc_440hz_20_64_fmin300_fmax8000", dict(n_mfcc=20, n_mels=64, fmin=300, fmax=8000), 5e-4),
])
def test_mfcc_golden(signal_440, name, kw, tol):
    G = load_golden(name)
    C = o.mfcc(signal_440, 22050, **kw)
    np.testing.assert_allclose(C, G, rtol=tol, atol=tol)       # test_parity_mfcc.py tiers


def test_mfcc_caps_n_mfcc_at_n_mels():
    y = np.random.default_rng(0).standard_normal(4000)
    assert o.mfcc(y, n_mfcc=50, n_mels=20).shape[0] == 20       # EdgeCaseTests.swift:185-213


def test_planted_mistakes_are_caught(signal_440):
    """A wrong sign, dropped Slaney factor or transposed DCT fails the goldens."""
    G = load_golden("mfcc_440hz_default")
    M = o.melspectrogram(signal_440, 22050)
    good = o.mfcc_from_logmel(o.power_to_db(M), 20)
    assert np.abs(good - G).max() < 1e-3
    bad_sign = o.dct_ortho_basis(20, 128) @ (-o.power_to_db(M))
    bad_T = o.dct_ortho_basis(128, 128)[:, :20].T @ o.power_to_db(M)
    fb = o.mel_filterbank(22050, 2048, 128)
    lo = np.argmax(fb > 0, axis=1)
    assert np.abs(bad_sign - G).max() > 1.0 and np.abs(bad_T - G).max() > 1.0
    assert np.abs(fb - load_golden("mel_filterbank_128_2048")).max() < 1e-8
    assert lo[0] >= 1  # first triangle starts above DC


# -- iSTFT (STFT.swift:686-861) ----------------------------------------------

def _istft_bruteforce(S, hop, win_length, center, length):
    """Direct inverse DFT of each column (O(N^2), DC and Nyquist imaginary parts
    left out as in vDSP's packed format), then overlap-add, window-sum division
    and the trim / length rule, written out with loops. Small sizes only."""
    F, T = S.shape
    N = 2 * (F - 1)
    w = o.prepare_window(win_length, N)
    total = N + (T - 1) * hop
    out = np.zeros(total)
    ws = np.zeros(total)
    n = np.arange(N)
    for f in range(T):
        X = S[:, f]
        x = X[0].real + X[F - 1].real * np.cos(np.pi * n)
        for k in range(1, F - 1):
            x = x + 2.0 * (X[k].real * np.cos(2 * np.pi * k * n / N) - X[k].imag * np.sin(2 * np.pi * k * n / N))
        x = x / N
        for i in range(N):
            out[f * hop + i] += x[i] * w[i]
            ws[f * hop + i] += w[i] * w[i]
    for i in range(total):
        if ws[i] > 1e-10:
            out[i] /= ws[i]
    start = N // 2 if center else 0
    if length is None:
        end = total - N // 2 if center else total
        return out[start:end] if end > start else np.zeros(0)
    res = np.zeros(length)
    m = max(min(length, total - start), 0)
    res[:m] = out[start:start + m]
    return res


@pytest.mark.parametrize("F,T,hop,win,center,length", [
    (5, 6, 3, 8, True, None), (5, 6, 2, 6, False, None), (9, 4, 5, 12, True, 11),
    (9, 7, 4, 16, False, 40), (17, 3, 8, 32, True, 30), (2, 3, 1, 2, False, None)])
def test_istft_bruteforce(F, T, hop, win, center, length):
    rng = np.random.default_rng(F * 100 + T)
    S = rng.standard_normal((F, T)) + 1j * rng.standard_normal((F, T))   # DC/Nyquist imag set on purpose
    got = o.istft(S, hop, win, center, length)
    ref = _istft_bruteforce(S, hop, win, center, length)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0)


@pytest.mark.parametrize("n_fft,hop,win,center", [
    (2048, 512, 2048, True), (1024, 256, 1024, False), (256, 64, 200, True), (512, 100, 512, True),
    (4096, 1024, 4096, False), (64, 16, 64, True), (2048, 128, 2048, True)])
def test_istft_inverts_the_stft(n_fft, hop, win, center):
    """istft(stft_complex(y)) = y wherever the summed squared window is well
    away from zero (everywhere with centre padding; the interior without)."""
    y = np.random.default_rng(n_fft + hop).standard_normal(12000)
    S = o.stft_complex(y, n_fft, hop, win, center)
    r = o.istft(S, hop, win, center, length=len(y))
    assert r.shape == y.shape
    sl = slice(0, len(y)) if center and win == n_fft else slice(n_fft, len(y) - n_fft)
    assert np.abs(r[sl] - y[sl]).max() < 1e-9


def test_istft_length_known_answers():
    # ISTFTTests.swift:8-37: 44 hops of a 2048-point STFT (45 frames) come back whole
    assert o.istft_length(1025, 45, 512, True) == 44 * 512
    assert o.istft_length(1025, 45, 512, True, 22050) == 22050
    assert o.istft_length(1025, 45, 512, False) == 2048 + 44 * 512
    assert o.istft_length(1025, 1, 512, True) == 0            # nothing left after the trim
    assert o.istft_length(4, 3, 2, True) == 6 + 2 * 2         # n_fft 6: zeros, untrimmed
    assert o.istft_length(2, 3, 0, True) == 0                 # n_fft 2 at hop 0: buffer of 2
    z = o.istft(np.ones((4, 3), complex), 2, 6, True, None)
    assert z.shape == (10,) and not z.any()
    S = o.stft_complex(np.random.default_rng(5).standard_normal(22528), 2048, 512)
    assert S.shape == (1025, 45)
    assert o.istft(S, 512, 2048, True).shape == (22528,)
    short = o.istft(S, 512, 2048, True, 1000)                  # ISTFTTests.swift:66-89
    assert short.shape == (1000,)
    longer = o.istft(S, 512, 2048, True, 30000)
    assert longer.shape == (30000,) and not longer[24576 - 1024:].any()


def test_istft_planted_mistakes_are_caught():
    """Reading the Nyquist bin from the DC imaginary slot (vDSP's packed
    layout misread), skipping the window-sum division or a missing 1/N fails
    the brute-force comparison."""
    rng = np.random.default_rng(9)
    S = rng.standard_normal((9, 5)) + 1j * rng.standard_normal((9, 5))
    ref = _istft_bruteforce(S, 4, 16, True, None)
    assert np.abs(o.istft(S, 4, 16, True) - ref).max() < 1e-12
    N, w = 16, o.prepare_window(16, 16)
    Sb = S.copy()
    Sb[-1] = S[0].imag
    frames = [np.fft.irfft(S.T, n=N, axis=1) * w, np.fft.irfft(Sb.T, n=N, axis=1) * w,
              np.fft.irfft(S.T, n=N, axis=1) * w * N]
    outs = []
    for k, fr in enumerate(frames):
        out, ws = np.zeros(N + 4 * 4), np.zeros(N + 4 * 4)
        for f in range(5):
            out[4 * f:4 * f + N] += fr[f]
            ws[4 * f:4 * f + N] += w * w
        if k != 2:
            nz = ws > 1e-10
            out[nz] /= ws[nz]
        outs.append(out[N // 2:len(out) - N // 2])
    assert np.abs(outs[0] - ref).max() < 1e-12
    assert np.abs(outs[1] - ref).max() > 1e-3 and np.abs(outs[2] - ref).max() > 1e-3


# -- onset strength (OnsetDetection.swift:23-128) ------------------------------

def test_onset_strength_golden(signal_440):
    G = load_golden("onset_strength_440hz_default")          # librosa.onset.onset_strength
    env = o.onset_strength(signal_440, 22050)
    assert env.shape == (1, 44)
    assert np.abs(env[0] - G).max() <= 1e-6 * max(np.abs(G).max(), 1.0)


def test_onset_strength_known_answers():
    rng = np.random.default_rng(2)
    y = rng.standard_normal(30000)
    env = o.onset_strength(y)
    T = o.frame_count(len(y), 2048, 512, True)
    assert env.shape == (1, T) and (env >= 0).all() and not env[0, :3].any()   # 1 + 2048 // 1024 zeros
    per = o.onset_strength(y, aggregate=False)
    assert per.shape == (128, T) and np.allclose(per.mean(axis=0), env[0])
    assert not o.onset_strength(np.zeros(22050)).any()                  # silence (test_onset_strength.py:23-26)
    un = o.onset_strength(y, center=False)
    assert un.shape == (1, o.frame_count(len(y), 2048, 512, False)) and not un[0, 0]
    assert o.onset_strength(np.ones(100), center=False).shape == (1, 1)   # one frame or none: a zero column
    assert o.onset_strength(y, hop=300).shape == (1, o.frame_count(len(y), 2048, 300, True))


def test_onset_strength_planted_mistakes_are_caught(signal_440):
    """No half-wave rectification, or a pad without the centre term, fails
    the golden."""
    G = load_golden("onset_strength_440hz_default")
    db = o.power_to_db(o.melspectrogram(signal_440, 22050))
    raw = (db[:, 1:] - db[:, :-1]).mean(axis=0)
    no_rect = np.concatenate([np.zeros(3), raw])[:44]
    rect = np.maximum(0, db[:, 1:] - db[:, :-1]).mean(axis=0)
    short_pad = np.concatenate([np.zeros(1), rect])[:44]
    assert np.abs(np.concatenate([np.zeros(3), rect])[:44] - G).max() < 1e-6
    assert np.abs(no_rect - G).max() > 1e-3 and np.abs(short_pad - G).max() > 1e-3


# -- chroma (Chroma.swift:30-275) ---------------------------------------------

@pytest.mark.parametrize("name,kw", [
    ("chroma_stft_440hz_no_norm", dict(tuning=0.02)),
    ("chroma_stft_440hz_default", dict(tuning=0.02, norm=np.inf)),
    ("chroma_stft_440hz_1024", dict(n_fft=1024, tuning=-0.07)),
])
def test_chroma_goldens(signal_440, name, kw):
    """librosa.feature.chroma_stft estimated the tuning of the 440 Hz golden
    signal on its 0.01-bin grid (0.02 at n_fft 2048, -0.07 at 1024); with
    those tunings the restated filterbank reproduces the goldens to float32
    rounding (at tuning 0, which the reference uses, they differ by 1-7 %)."""
    G = load_golden(name)
    C = o.chroma_stft(signal_440, 22050, **kw)
    assert C.shape == G.shape == (12, 44)
    assert np.abs(C - G).max() <= 2e-6 * np.abs(G).max()
    kw0 = dict(kw, tuning=0.0)
    assert np.abs(o.chroma_stft(signal_440, 22050, **kw0) - G).max() > 1e-2 * np.abs(G).max()


def test_chroma_filterbank_properties():
    fb = o.chroma_filterbank(22050, 2048)
    assert fb.shape == (12, 1025) and (fb > 0).all()
    # before octave weighting every column has unit L2 norm
    k = np.arange(1, 2048)
    frq = np.empty(2048)
    frq[1:] = 12 * np.log2(k * 22050 / 2048 / 27.5)
    frq[0] = frq[1] - 18
    ow = np.exp(-0.5 * ((frq / 12 - 5.0) / 2.0) ** 2)[:1025]
    assert np.allclose(np.sqrt(((fb / ow) ** 2).sum(axis=0)), 1.0)
    # a 261.63 Hz (C4) partial lands in row 0 (C) after the roll; 440 Hz in row 9 (A)
    for f, row in [(261.63, 0), (440.0, 9)]:
        assert np.argmax(fb[:, int(round(f * 2048 / 22050))]) == row


def test_normalize_frames_norms():
    rng = np.random.default_rng(4)
    C = np.abs(rng.standard_normal((12, 20)))
    C[:, 3] = 0
    for p, f in [(1.0, lambda c: np.abs(c).sum(0)), (2.0, lambda c: np.sqrt((c * c).sum(0))),
                 (np.inf, lambda c: np.abs(c).max(0))]:
        N = o.normalize_frames(C, p)
        n = f(N)
        assert np.allclose(np.delete(n, 3), 1.0) and not N[:, 3].any()


# -- spectral descriptors (SpectralDescriptors.swift) ---------------------------

@pytest.mark.parametrize("name,tol", [("spectral_centroid", 1e-7), ("spectral_bandwidth", 1e-7),
                                      ("spectral_rolloff", 0.0), ("spectral_flatness", 2e-6)])
def test_descriptor_goldens(signal_440, name, tol):
    G = load_golden(f"{name}_440hz_default")
    fn = getattr(o, name)
    v = fn(signal_440) if name == "spectral_flatness" else fn(signal_440, 22050)
    assert v.shape == G.shape == (1, 44)
    assert np.abs(v - G).max() <= tol * np.abs(G).max()


def test_descriptor_known_answers():
    sr, n = 22050, 22050
    k = 100                                   # a tone exactly on bin 100 of a 2048-point frame
    f0 = k * sr / 2048
    t = np.arange(n) / sr
    tone = np.sin(2 * np.pi * f0 * t)
    c = o.spectral_centroid(tone, sr)[0, 4:-4]
    assert np.allclose(c, f0, rtol=1e-3)      # Hann leakage is symmetric about the tone
    assert (o.spectral_bandwidth(tone, sr)[0, 4:-4] < 30.0).all()   # within the window's main lobe
    assert (o.spectral_flatness(tone)[0, 4:-4] < 1e-3).all()
    noise = np.random.default_rng(0).standard_normal(n)
    assert (o.spectral_flatness(noise)[0, 4:-4] > 0.3).all()
    z = np.zeros(n)
    assert not o.spectral_centroid(z).any() and not o.spectral_rolloff(z).any()
    assert not o.spectral_flatness(z).any() and not o.spectral_bandwidth(z).any()
    # rolloff against a direct search on one frame
    S = o.stft_magnitude(noise, 2048, 512)
    fr = 10
    cum, tot = 0.0, S[:, fr].sum()
    for kk in range(S.shape[0]):
        cum += S[kk, fr]
        if cum >= 0.85 * tot:
            break
    assert o.spectral_rolloff(noise, sr)[0, fr] == kk * sr / 2048


# -- spectral contrast (SpectralDescriptors.swift:137-204) ---------------------

def test_contrast_bands_by_hand():
    """sr 22050, n_fft 2048: bin width 10.7666 Hz; 200, 400, ..., 6400 Hz ->
    18.58, 37.15, 74.3, 148.6, 297.2, 594.4 -> 19, 37, 74, 149, 297, 594;
    Nyquist -> 1024. Band sizes 19, 18, 37, 75, 148, 297, 430 -> 2 % of them
    rounded half away from zero, at least 1: 1, 1, 1, 2, 3, 6, 9."""
    assert o.contrast_bands(22050, 2048, 1025) == ([0, 19, 37, 74, 149, 297, 594, 1024], [1, 1, 1, 2, 3, 6, 9])
    e, q = o.contrast_bands(22050, 4096, 2049)
    assert e == [0, 37, 74, 149, 297, 594, 1189, 2048] and q == [1, 1, 2, 3, 6, 12, 17]
    e, q = o.contrast_bands(8000, 512, 257, n_bands=7)   # upper octaves past Nyquist: empty bands
    assert e[-3:] == [257, 257, 256] and q[-2:] == [0, 0]


def test_contrast_known_answers():
    rng = np.random.default_rng(2)
    S = np.ones((1025, 3))
    S[149:297, 0] = rng.permutation(np.arange(1.0, 149.0))   # band 4: 148 bins, q = 3
    c = o.contrast_from_magnitude(S, 22050, 2048)
    assert abs(c[4, 0] - 10 * np.log10(147.0 / 2.0)) < 1e-12    # mean(146..148) / mean(1..3)
    assert np.all(c[:, 1:] == 0) and np.all(np.delete(c[:, 0], 4) == 0)   # flat bands: 0
    S2 = S.copy()
    S2[149:297, 0] = S2[149:297, 0][::-1]                          # order within a band is irrelevant
    assert np.array_equal(o.contrast_from_magnitude(S2, 22050, 2048), c)
    z = o.contrast_from_magnitude(np.zeros((1025, 2)), 22050, 2048)
    assert np.all(z == 0)                                          # 10 log10(1e-10) - itself


def test_contrast_planted_mistakes_are_caught():
    """Swapping peak and valley flips the sign; rounding the 75-bin band's
    1.5 down keeps one bin per end instead of two."""
    S = np.ones((1025, 1))
    S[74:149, 0] = np.arange(1.0, 76.0)
    good = o.contrast_from_magnitude(S, 22050, 2048)[3, 0]
    assert abs(good - 10 * np.log10(74.5 / 1.5)) < 1e-12
    assert abs(-good - good) > 1.0                                  # swapped
    assert abs(10 * np.log10(75.0 / 1.0) - good) > 1.0             # q = floor(1.5) = 1


def test_contrast_golden_at_reference_tolerance(signal_440):
    """librosa's spectral_contrast (the golden) differs from the reference's
    algorithm; the reference checks itself against it at rtol 0.5, atol 15 dB
    (test_parity_spectral.py:158-172), and so does the oracle."""
    G = load_golden("spectral_contrast_440hz_default")
    c = o.spectral_contrast(signal_440, 22050)
    assert c.shape == G.shape
    np.testing.assert_allclose(c, G, rtol=0.5, atol=15.0)
    assert np.all(np.isfinite(c))


def test_contrast_cpu_baseline_port_matches_oracle():
    """The float32 contrast timed as the CPU baseline is the oracle's
    computation (FP32 rounding only, on a noise-rich signal)."""
    from oracle.cpu_baseline import ContrastF32
    y = np.random.default_rng(3).standard_normal(30000).astype(np.float32)
    a = ContrastF32()(y)
    b = o.spectral_contrast(y.astype(np.float64), 22050)
    assert a.shape == b.shape and np.abs(a - b).max() < 0.01

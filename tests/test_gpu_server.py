"""Serving: continuous batching returns, for every request, exactly what the
same clip gives alone; admission respects the HBM budget; oversized requests
fail cleanly without disturbing the others."""
import numpy as np
import pytest

from oracle import spectral as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sources():
    from metalmom_amd.serve import make_sources
    return make_sources(seconds=40.0, seed=9)


def _alone(kind, sr, y):
    import metalmom_amd
    return metalmom_amd.spectral(y, kind=kind, sr=sr, n_fft=2048, hop_length=512, n_mels=128,
                                 n_mfcc=13)


def test_mixed_requests_equal_alone(sources):
    from metalmom_amd.serve import SpectralServer
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(40):
        sr = (22050, 44100)[i % 2]
        kind = ("stft", "mel", "mfcc")[i % 3]
        n = int(rng.integers(2000, 8 * sr))
        off = int(rng.integers(0, len(sources[sr]) - n))
        reqs.append((kind, sr, sources[sr][off:off + n]))
    srv = SpectralServer(hbm_budget_bytes=2 << 30)
    ids = [srv.submit(y, srv.params(kind, sr=sr)) for kind, sr, y in reqs]    # burst: all in flight
    outs = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["steps"] >= 1
    for (kind, sr, y), (arr, lat) in zip(reqs, outs):
        ref = _alone(kind, sr, y)
        assert arr.shape == ref.shape
        assert np.array_equal(arr, ref), (kind, sr, len(y))
        assert lat > 0


def test_chroma_and_mel_of_equal_shape_stay_apart(sources):
    """Chroma is the banded contraction with another filterbank: requests of
    both kinds with the same n_mels, n_fft and hop in one step must be
    launched as separate plans."""
    from metalmom_amd.serve import SpectralServer
    import metalmom_amd
    y = sources[22050][:5 * 22050]
    srv = SpectralServer(hbm_budget_bytes=2 << 30)
    ids = [srv.submit(y, srv.params(k, sr=22050, n_mels=12)) for k in ("mel", "chroma", "mel", "chroma")]
    outs = [srv.wait(i)[0] for i in ids]
    srv.close()
    mel = metalmom_amd.spectral(y, kind="mel", sr=22050, n_fft=2048, hop_length=512, n_mels=12)
    chroma = metalmom_amd.chroma_stft(y=y, sr=22050, n_fft=2048, hop_length=512)
    for k, arr in zip(("mel", "chroma", "mel", "chroma"), outs):
        assert np.array_equal(arr, mel if k == "mel" else chroma), k


def test_results_match_oracle(sources):
    from metalmom_amd.serve import SpectralServer
    y = sources[22050][:3 * 22050]
    srv = SpectralServer()
    a, _ = srv.wait(srv.submit(y, srv.params("mfcc", sr=22050)))
    b, _ = srv.wait(srv.submit(y, srv.params("mel", sr=22050)))
    srv.close()
    O = o.mfcc(y, 22050, 13, 2048, 512)
    assert np.abs(a - O).max() / np.abs(O).max() <= 1e-4
    M = o.melspectrogram(y, 22050, 2048, 512)
    assert np.abs(b - M).max() / M.max() <= 1e-5


def test_small_budget_serialises_and_rejects_oversized(sources):
    from metalmom_amd.serve import SpectralServer
    y = sources[22050][:5 * 22050]
    big = sources[22050][:30 * 22050]
    srv = SpectralServer(hbm_budget_bytes=4 << 20, max_batch=4)   # fits ~1-2 small requests
    p = srv.params("mfcc", sr=22050)
    ids = [srv.submit(y, p) for _ in range(6)]
    bad = srv.submit(big, srv.params("stft", sr=22050))          # needs > 4 MiB: rejected
    outs = [srv.wait(i)[0] for i in ids]
    with pytest.raises(RuntimeError, match="status -3"):
        srv.wait(bad)
    srv.close()
    ref = _alone("mfcc", 22050, y)
    for a in outs:
        assert np.array_equal(a, ref)


def test_device_memory_stays_within_budget_multiple(sources):
    """The server's streams keep their step buffers only up to a third of
    the HBM budget each once a step is done, so the device memory they hold
    stays within 2.5 x the budget while steps run and within the budget once
    they are done, whatever the mix of small and near-budget requests; the
    results still equal each clip run alone."""
    from metalmom_amd.serve import SpectralServer
    budget = 16 << 20
    srv = SpectralServer(hbm_budget_bytes=budget, max_batch=8)
    assert srv.device_bytes() == 0
    rng = np.random.default_rng(21)
    y22 = sources[22050]
    reqs = []
    for i in range(36):
        kind = ("stft", "mel", "mfcc")[i % 3]
        # every 4th request is long (30-38 s): as an STFT magnitude request
        # ~12 bytes per sample, 8-10 MB, over half the budget
        n = int(rng.integers(30 * 22050, 38 * 22050)) if i % 4 == 0 else int(rng.integers(2000, 3 * 22050))
        off = int(rng.integers(0, len(y22) - n))
        reqs.append((kind, y22[off:off + n]))
    ids = [srv.submit(y, srv.params(kind, sr=22050)) for kind, y in reqs]
    peak = 0
    outs = []
    for i in ids:
        peak = max(peak, srv.device_bytes())
        outs.append(srv.wait(i)[0])
        peak = max(peak, srv.device_bytes())
    st = srv.stats()
    idle = srv.device_bytes()
    srv.close()
    assert st["completed"] == len(reqs)
    assert 0 < peak <= 2.5 * budget, peak
    assert idle <= budget, idle
    for (kind, y), a in zip(reqs, outs):
        assert np.array_equal(a, _alone(kind, 22050, y)), (kind, len(y))


def test_open_loop_latency_report(sources):
    from metalmom_amd.serve import SpectralServer, open_loop
    srv = SpectralServer()
    r = open_loop(srv, rate=200.0, duration=0.5, seed=3)
    srv.close()
    assert r["n_requests"] >= 50
    assert 0 < r["p50"] <= r["p99"] < 5.0
    assert r["frames"] > 0


def test_open_loop_pcm_requests():
    """The serving schedule as int16 stereo PCM at 22.05 / 44.1 kHz, analysed
    at 22.05 kHz: every request completes, and one kept result equals the
    same clip mixed down, resampled and analysed alone."""
    import metalmom_amd as mm
    from metalmom_amd.serve import SpectralServer, make_sources, open_loop, pcm_sources
    src = make_sources()   # the schedule's offsets reach 330 s into each source
    psrc = pcm_sources(src)
    srv = SpectralServer()
    r = open_loop(srv, rate=100.0, duration=0.3, seed=3, sources=src, pcm=psrc, keep_outputs=True)
    srv.close()
    assert r["n_requests"] >= 10 and 0 < r["p50"] <= r["p99"] < 5.0
    t, kind, sr, n, off = r["schedule"][0]
    x = psrc[sr][off:off + n]
    clip = ((x[:, 0].astype(np.float64) / 32768.0 + x[:, 1].astype(np.float64) / 32768.0) / 2).astype(np.float32)
    clip = mm.resample(clip, sr, 22050)
    alone = mm.spectral(clip, kind=kind, sr=22050, n_fft=2048, hop_length=512)
    assert np.array_equal(r["outputs"][0], alone)


def test_full_rings_fall_back_to_heap_paths():
    """More input than the pinned input ring holds (32 MiB at this budget)
    sends later clips down the heap path; results left unpolled fill the
    output ring, so later steps stage results in the lane buffers. Every
    result still equals the clip run alone."""
    import metalmom_amd as mm
    from metalmom_amd.serve import SpectralServer
    srv = SpectralServer(hbm_budget_bytes=64 << 20, max_batch=4)
    try:
        rng = np.random.default_rng(11)
        clips = [(0.3 * rng.standard_normal(2_000_000 + 1000 * i)).astype(np.float32) for i in range(12)]  # 8 MB each
        p = srv.params("stft", sr=22050)
        ids = [srv.submit(c, p) for c in clips]
        # let every request finish before collecting any result
        import time
        t0 = time.time()
        while srv.stats()["completed"] < len(ids) and time.time() - t0 < 60:
            time.sleep(0.01)
        assert srv.stats()["completed"] == len(ids)
        for rid, c in zip(ids, clips):
            got, _ = srv.wait(rid)
            alone = mm.spectral(c, kind="stft", sr=22050, n_fft=2048, hop_length=512)
            assert np.array_equal(got, alone)
    finally:
        srv.close()


def test_two_waiters_on_one_request(sources):
    """Two threads waiting on the same request id: one collects the result,
    the other returns the invalid-id status instead of touching the
    collected request or blocking."""
    import threading
    from metalmom_amd.serve import SpectralServer
    y = sources[22050][:4 * 22050]
    srv = SpectralServer(hbm_budget_bytes=1 << 30)
    for rep in range(5):
        rid = srv.submit(y, srv.params("mfcc", sr=22050))
        got, errs = [], []

        def waiter():
            try:
                got.append(srv.wait(rid)[0])
            except RuntimeError as e:
                errs.append(str(e))

        ths = [threading.Thread(target=waiter) for _ in range(2)]
        for t in ths:
            t.start()
        for t in ths:
            t.join(timeout=60)
        assert not any(t.is_alive() for t in ths), rep
        assert len(got) == 1 and len(errs) == 1 and "status -1" in errs[0], (rep, errs)
    srv.close()
    assert np.array_equal(got[0], _alone("mfcc", 22050, y))

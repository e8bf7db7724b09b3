"""GPU surveillance parity: FFT against numpy's transform, Welch PSD against
the FP64 oracle, and CFAR detection lists identical to the oracle's on the
same spectrum array.  Every call goes through libargus_hip.so."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FS = 10_000_000


@pytest.fixture()
def surv(gpu):
    from argus import surveillance
    return surveillance


@pytest.mark.parametrize("n", [4096, 32768])
def test_fft_matches_numpy(surv, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)).astype(np.complex64)
    s = surv.Surveillance(n, FS)
    got = s.Exec(x).reshape(3, n)
    want = np.fft.fft(x.reshape(3, n).astype(np.complex128), axis=1)
    err = np.abs(got - want) / np.sqrt(n)
    assert float(np.max(err)) < 2e-5, float(np.max(err))
    w = O.Hann(n)
    got_w = s.Exec(x, window=True).reshape(3, n)
    want_w = np.fft.fft(x.reshape(3, n).astype(np.complex128) * w, axis=1)
    assert float(np.max(np.abs(got_w - want_w) / np.sqrt(n))) < 2e-5
    # impulse -> flat spectrum; tone -> single bin
    imp = np.zeros(n, np.complex64)
    imp[0] = 1
    np.testing.assert_allclose(s.Exec(imp), np.ones(n), atol=1e-5)
    tone = np.exp(2j * np.pi * 37 * np.arange(n) / n).astype(np.complex64)
    X = s.Exec(tone)
    assert int(np.argmax(np.abs(X))) == 37 and abs(abs(X[37]) - n) < 1e-3 * n
    s.Close()


def test_welch_matches_fp64_oracle(surv):
    from argus.synth import surveillance_scene
    n, hop, segs, frames = 32768, 16384, 8, 3
    total = (frames - 1) * segs * hop + (segs - 1) * hop + n
    iq, _ = surveillance_scene(total, FS, n_carriers=12, seed=3)
    s = surv.Surveillance(n, FS)
    got = s.WelchPSD(iq, segs, hop, segs * hop, frames)
    w = O.Hann(n)
    for f in range(frames):
        st = f * segs * hop
        want = O.WelchPSD(iq[st:st + (segs - 1) * hop + n], n, 0.5, w)
        err = np.abs(got[f] - want)
        print(f"frame {f}: max |dB err| {err.max():.2e}", flush=True)
        assert float(err.max()) < 0.01
    s.Close()


def _same_lists(gpu_sigs, orc_sigs):
    a = [(g.FirstBin, g.LastBin, g.PeakBin) for g in gpu_sigs]
    b = [(g.FirstBin, g.LastBin, g.PeakBin) for g in orc_sigs]
    return a == b


@pytest.mark.parametrize("mode", ["CA", "GOSCA", "CASO", "OS"])
def test_detection_list_identical_on_same_spectrum(surv, mode):
    from argus.synth import surveillance_scene
    n, hop, segs, frames = 32768, 16384, 8, 4
    total = (frames - 1) * segs * hop + (segs - 1) * hop + n
    iq, _ = surveillance_scene(total, FS, n_carriers=40, seed=11)
    s = surv.Surveillance(n, FS)
    cfar = surv.cfar_params(mode, guard=4, train=16, scale_db=12.3748, wrap=True, hysteresis_db=3.0, rank=24)
    det = surv.detector_params(cfar_scale_db=12.3748, center_hz=100e6)
    lists, psd = s.Detect(iq, segs, hop, segs * hop, frames, cfar, det)
    near = 0
    for f in range(frames):
        spec = psd[f].astype(np.float64)
        thr = O.CFARThresholds(mode, spec, 4, 16, 12.3748, wrap=True, rank=24)
        near += int(np.sum(np.abs(spec - thr) < 1e-9) + np.sum(np.abs(spec - (thr - 3.0)) < 1e-9))
        want = O.DetectSignals(spec, thr, sample_rate=FS, center_hz=100e6, cfar_scale_db=12.3748)
        print(f"frame {f}: gpu {len(lists[f])} signals, oracle {len(want)}", flush=True)
        assert _same_lists(lists[f], want), (f, [(g.FirstBin, g.LastBin, g.PeakBin) for g in lists[f]][:8],
                                             [(g.FirstBin, g.LastBin, g.PeakBin) for g in want][:8])
        for g, w in zip(lists[f], want):
            assert abs(g.CenterHz - w.CenterHz) < 1e-3 and abs(g.NoiseDb - w.NoiseDb) < 1e-9
    assert near == 0  # no threshold near-ties in this scene (identity is exact, not luck)
    s.Close()


@pytest.mark.parametrize("mode", ["CA", "GOSCA", "CASO", "OS"])
def test_cfar_ambiguity_flags_ties(surv, mode):
    """A flat spectrum with hysteresis equal to the CFAR scale puts every
    bin on the threshold less the hysteresis (up to rounding): each bin's
    decision is a tie the reference's running sums may break either way, so
    every bin is ambiguous -- except for OS-CFAR, which ranks dB values
    without arithmetic.  With any other hysteresis, no bin is ambiguous.  On
    the 40-carrier scene no frame has an ambiguous bin, and the runs equal
    the oracle's (running-sum) runs."""
    from argus.synth import surveillance_scene
    n = 4096
    s = surv.Surveillance(n, FS)
    flat = np.full((2, n), -87.25, np.float32)
    for hyst, want in ((6.0, 0 if mode == "OS" else n), (3.0, 0)):
        cfar = surv.cfar_params(mode, guard=4, train=16, scale_db=6.0, wrap=True, hysteresis_db=hyst, rank=24)
        _, _, amb = s.CFARRuns(flat, cfar, ambiguous=True)
        assert list(amb) == [want, want], (mode, hyst, list(amb))
    s.Close()
    n, hop, segs, frames = 32768, 16384, 8, 3
    total = (frames - 1) * segs * hop + (segs - 1) * hop + n
    iq, _ = surveillance_scene(total, FS, n_carriers=40, seed=11)
    s = surv.Surveillance(n, FS)
    psd = s.WelchPSD(iq, segs, hop, segs * hop, frames)
    cfar = surv.cfar_params(mode, guard=4, train=16, scale_db=12.3748, wrap=True, hysteresis_db=3.0, rank=24)
    counts, runs, amb = s.CFARRuns(psd, cfar, ambiguous=True)
    assert list(amb) == [0] * frames
    for f in range(frames):
        spec = psd[f].astype(np.float64)
        thr = O.CFARThresholds(mode, spec, 4, 16, 12.3748, wrap=True, rank=24)
        want, on, start = [], False, 0  # threshold crossing with hysteresis (detector.go:436-480)
        for b in range(n):
            nxt = True if spec[b] >= thr[b] else (False if spec[b] < thr[b] - 3.0 else on)
            if nxt and not on:
                start = b
            if on and not nxt:
                want.append((start, b - 1))
            on = nxt
        if on:
            want.append((start, n - 1))
        assert [(r.first_bin, r.last_bin) for r in runs[f]] == want, f
    s.Close()


def _runs_with_hysteresis(spec, thr, hyst):
    """Threshold crossing with hysteresis (detector.go:436-480) on the
    oracle's running-sum thresholds: [(first, last)] in order."""
    want, on, start = [], False, 0
    for b in range(len(spec)):
        nxt = True if spec[b] >= thr[b] else (False if spec[b] < thr[b] - hyst else on)
        if nxt and not on:
            start = b
        if on and not nxt:
            want.append((start, b - 1))
        on = nxt
    if on:
        want.append((start, len(spec) - 1))
    return want


@pytest.mark.parametrize("n,guard,train,wrap", [(4096, 4, 16, True), (4096, 0, 3, False), (32768, 40, 260, True),
                                                 (1024, 2, 8, False)])
def test_cfar_tiles_halos_and_edges_against_oracle(surv, n, guard, train, wrap):
    """The status kernel works on tiles of 1024 bins with a halo of guard +
    train cells on each side: runs that cross tile boundaries, halos larger
    than a wavefront's row, clamped and wrapped edges and runs open at both
    ends of a frame give the oracle's runs, peaks and noise estimates; the
    per-frame run count is the true count when max_runs truncates the list."""
    rng = np.random.default_rng(n + guard + train)
    frames = 5
    psd = (-90.0 + 3.0 * rng.standard_normal((frames, n))).astype(np.float32)
    for f in range(frames):
        for _ in range(n // 256):
            c, w = int(rng.integers(0, n)), int(rng.integers(1, 80))
            psd[f, c:c + w] += np.float32(rng.uniform(6.0, 40.0))
        # overlapping carriers stack up; capped, the frame's largest cell stays
        # within 60 dB of the noise and the near-tie bound far below every
        # bin's margin (at least 8 times, computed on the CPU)
        np.minimum(psd[f], np.float32(-48.0), out=psd[f])
        for b in range(1024, n, 1024):  # plateaus across every tile boundary
            psd[f, b - 3:b + 4] = np.float32(-40.0 - f)
    psd[1, :5] = -30.0
    psd[1, -4:] = -30.0
    s = surv.Surveillance(n, FS)
    for mode in ("CA", "GOSCA", "CASO", "OS"):
        if mode == "OS" and train > 128:
            continue
        cfar = surv.cfar_params(mode, guard=guard, train=train, scale_db=9.0, wrap=wrap, hysteresis_db=3.0,
                                rank=train)
        counts, runs, amb = s.CFARRuns(psd, cfar, ambiguous=True)
        for f in range(frames):
            spec = psd[f].astype(np.float64)
            thr = O.CFARThresholds(mode, spec, guard, train, 9.0, wrap=wrap, rank=train)
            want = _runs_with_hysteresis(spec, thr, 3.0)
            assert amb[f] == 0, (mode, f, int(amb[f]))
            got = [(r.first_bin, r.last_bin) for r in runs[f]]
            assert int(counts[f]) == len(want) and got == want, (mode, f, got[:6], want[:6])
            for r, (a, b) in zip(runs[f], want):
                seg = np.arange(a, b + 1)
                above = seg[spec[seg] >= thr[seg]]
                pk = int(above[np.argmax(spec[above])])
                assert r.peak_bin == pk and r.peak_db == spec[pk], (mode, f, r.peak_bin, pk)
                assert abs(r.noise_db - (thr[pk] - 9.0)) < 1e-9, (mode, f, r.noise_db, thr[pk] - 9.0)
        # truncated lists keep the true count and the first max_runs runs
        c2, runs2 = s.CFARRuns(psd, cfar, max_runs=3)
        assert list(c2) == list(counts)
        for f in range(frames):
            assert [(r.first_bin, r.last_bin) for r in runs2[f]] == [(r.first_bin, r.last_bin) for r in runs[f][:3]]
    s.Close()


def test_cfar_runs_known_plateau(surv):
    n = 4096
    s = surv.Surveillance(n, FS)
    spec = np.full(n, -100.0, np.float32)
    spec[1000:1011] = -20
    spec[3000:3002] = -30
    counts, runs = s.CFARRuns(spec, surv.cfar_params("CA", guard=2, train=16, scale_db=6.0, wrap=True))
    got = [(r.first_bin, r.last_bin, r.peak_bin) for r in runs[0]]
    assert got == [(1000, 1010, 1000), (3000, 3001, 3000)], got
    s.Close()


@pytest.mark.parametrize("n", [4096, 32768])
def test_single_fft_spectrum_db_matches_fp64_oracle(surv, n):
    """fft.go:24-73 + pipeline_runtime.go:315-341: Hann, FFT, 20 log10(|X|/N +
    1e-12), fftshift; against the oracle's FP64 SpectrumWithPlan."""
    rng = np.random.default_rng(12)
    t = np.arange(2 * n)
    iq = (0.5 * np.exp(2j * np.pi * 0.123 * t) + 0.01 * (rng.standard_normal(2 * n)
                                                          + 1j * rng.standard_normal(2 * n))).astype(np.complex64)
    s = surv.Surveillance(n, FS)
    got = s.SpectrumDB(iq, window=True)
    w = O.Hann(n)
    for f in range(2):
        want = O.Spectrum(iq[f * n:(f + 1) * n], w)
        # FP32 FFT: error bounded relative to the peak magnitude (about log2 N
        # roundings of 6e-8), so dB error grows as a bin falls below the peak
        lg, lw = 10 ** (got[f] / 20.0), 10 ** (want / 20.0)
        assert np.max(np.abs(lg - lw)) <= 2e-6 * lw.max(), float(np.max(np.abs(lg - lw)) / lw.max())
        near = want > want.max() - 40.0
        assert np.max(np.abs(got[f][near] - want[near])) < 1e-3
        assert int(np.argmax(got[f])) == int(np.argmax(want))
    s.Close()


def test_spectra_sanitise_non_finite_input(surv):
    """pipeline_runtime.go:391-397: NaN / Inf bins read -200 dB."""
    n = 4096
    s = surv.Surveillance(n, FS)
    iq = np.zeros(n, np.complex64)
    iq[7] = np.nan
    db = s.SpectrumDB(iq, window=False)
    assert np.all(db == -200.0)
    hop = n // 2
    iq2 = np.zeros(3 * hop + n, np.complex64)
    iq2[5] = np.inf
    psd = s.WelchPSD(iq2, 2, hop, 2 * hop, 1)
    assert np.all(np.isfinite(psd)) and np.all(psd[0] == -200.0)
    s.Close()


def _compare_finish(s, psd, cfar, det, cap):
    dev_lists, dev_counts = s.DetectSpectra(psd, cfar, det, cap=cap)
    counts, runs = s.CFARRuns(psd, cfar)
    for f in range(psd.shape[0]):
        want, m = s.FinishFrameHost(psd[f], runs[f], det, cap=cap)
        print(f"frame {f}: {len(runs[f])} runs -> device {dev_counts[f]}, host {m} signals", flush=True)
        assert dev_counts[f] == m, (f, dev_counts[f], m)
        assert len(dev_lists[f]) == len(want)
        for g, w in zip(dev_lists[f], want):
            assert (g.FirstBin, g.LastBin, g.PeakBin) == (w.FirstBin, w.LastBin, w.PeakBin), (f, g, w)
            assert (g.PeakDb, g.SNRDb, g.NoiseDb, g.BWHz) == (w.PeakDb, w.SNRDb, w.NoiseDb, w.BWHz), (f, g, w)
            assert abs(g.CenterHz - w.CenterHz) <= 1e-6, (f, g.CenterHz, w.CenterHz)


def test_device_detection_list_equals_host_on_a_scene(surv):
    """gpud_surv_finish_device against the host detection list
    (gpud_surv_finish_frame) on the same Welch spectra and CFAR runs."""
    from argus.synth import surveillance_scene
    n, hop, segs, frames = 32768, 16384, 8, 3
    total = (frames - 1) * segs * hop + (segs - 1) * hop + n
    iq, _ = surveillance_scene(total, FS, n_carriers=40, seed=21)
    s = surv.Surveillance(n, FS)
    psd = s.WelchPSD(iq, segs, hop, segs * hop, frames)
    for mode in ("CA", "OS"):
        cfar = surv.cfar_params(mode, guard=4, train=16, scale_db=12.3748, wrap=True, hysteresis_db=3.0, rank=24)
        det = surv.detector_params(cfar_scale_db=12.3748, center_hz=100e6)
        _compare_finish(s, psd, cfar, det, cap=1024)
    s.Close()


def test_device_detection_list_edges_merges_and_cap(surv):
    """Crafted spectra: runs within 50 bins of either end (the global-median
    path: the mean of the two middle values of the 4096 bins), neighbours
    closer than the merge gap, expansions that cross a neighbour, and more
    signals than the cap."""
    n = 4096
    rng = np.random.default_rng(5)
    s = surv.Surveillance(n, FS)
    frames = []
    for f in range(2):
        spec = (-100.0 + rng.standard_normal(n) * 0.5).astype(np.float32)
        spec[10:20] = -40 + f             # near the left end
        spec[n - 30:n - 22] = -45         # near the right end
        spec[1000:1010] = -50             # two plateaus closer than the merge gap
        spec[1012:1020] = -52
        spec[2000:2040] = -60             # a wide one with a dip and a shoulder
        spec[2015:2018] = -70
        spec[2040:2060] = -75
        for k in range(40):               # many narrow carriers
            c = 2500 + 35 * k
            spec[c:c + 3] = -55 - (k % 5)
        frames.append(spec)
    psd = np.stack(frames)
    cfar = surv.cfar_params("CA", guard=2, train=12, scale_db=10.0, wrap=False, hysteresis_db=3.0)
    for gap_hz, cap in ((5000.0, 1024), (0.0, 8), (60_000.0, 1024)):
        det = surv.detector_params(cfar_scale_db=10.0, center_hz=0.0, merge_gap_hz=gap_hz)
        _compare_finish(s, psd, cfar, det, cap=cap)
    s.Close()


@pytest.mark.parametrize("n,decim", [(32768, 2), (4096, 2), (4096, 4)])
def test_level_spectrum_matches_fp64_oracle(surv, n, decim):
    """The multi-resolution derived level (gpud_surv_level_spectrum_device):
    low-pass + decimate + n/decim-point spectrum per frame, against the
    oracle's FP64 composition (ApplyFIR, Decimate, Spectrum), on two frames
    with a tone inside the level, one beyond its band and noise."""
    rng = np.random.default_rng(n + decim)
    frames = 2
    t = np.arange(frames * n)
    fs = FS
    iq = (0.5 * np.exp(2j * np.pi * 0.0371 * t) + 0.2 * np.exp(2j * np.pi * 0.46 * t)
          + 0.01 * (rng.standard_normal(frames * n) + 1j * rng.standard_normal(frames * n))).astype(np.complex64)
    s = surv.Surveillance(n, fs)
    got = s.LevelSpectrumDB(iq, decim=decim, frame_step=n, n_frames=frames)
    n_d = n // decim
    assert got.shape == (frames, n_d)
    for f in range(frames):
        want = O.SurveillanceLevelSpectrum(iq[f * n:(f + 1) * n], fs, decim)
        lg, lw = 10 ** (got[f] / 20.0), 10 ** (want / 20.0)
        err = float(np.max(np.abs(lg - lw)) / lw.max())
        print(f"frame {f}: max |mag err| / peak {err:.2e}", flush=True)
        assert err <= 2e-6
        near = want > want.max() - 40.0
        assert np.max(np.abs(got[f][near] - want[near])) < 1e-3
        assert int(np.argmax(got[f])) == int(np.argmax(want))
    with pytest.raises(Exception):
        s.LevelSpectrumDB(iq, decim=3)
    s.Close()


def test_level_detection_on_the_device(surv):
    """Derived-level detection (detectDerivedCandidates,
    pipeline_runtime.go:742-789): the level spectra of the base engine go to
    a detection-only engine at the level's rate and size, whose CFAR runs and
    device detection list equal the host list and the oracle's."""
    n, decim, frames = 32768, 2, 3
    rng = np.random.default_rng(8)
    t = np.arange(frames * n)
    iq = 0.01 * (rng.standard_normal(frames * n) + 1j * rng.standard_normal(frames * n))
    for k in range(12):  # carriers inside the level's +-2.5 MHz, one beyond it
        f0 = -2.2e6 + k * 0.4e6 + rng.uniform(-20e3, 20e3)
        iq = iq + 10 ** (-rng.uniform(0, 1.5)) * np.exp(2j * np.pi * (f0 / FS * t + rng.uniform()))
    iq = (iq + 0.5 * np.exp(2j * np.pi * 4.0e6 / FS * t)).astype(np.complex64)
    base = surv.Surveillance(n, FS)
    lvl = base.LevelSpectrumDB(iq, decim=decim, frame_step=n, n_frames=frames)
    det_engine = surv.Surveillance(n // decim, FS // decim)
    with pytest.raises(Exception):
        det_engine.SpectrumDB(iq[:n // decim])  # no transform at this size
    cfar = surv.cfar_params("CA", guard=4, train=16, scale_db=12.3748, wrap=True, hysteresis_db=3.0)
    det = surv.detector_params(cfar_scale_db=12.3748, center_hz=100e6)
    _compare_finish(det_engine, lvl, cfar, det, cap=1024)
    lists, _ = det_engine.DetectSpectra(lvl, cfar, det)
    for f in range(frames):
        spec = lvl[f].astype(np.float64)
        thr = O.CFARThresholds("CA", spec, 4, 16, 12.3748, wrap=True)
        want = O.DetectSignals(spec, thr, sample_rate=FS // decim, center_hz=100e6, cfar_scale_db=12.3748)
        assert _same_lists(lists[f], want), f
        assert len(want) > 5
    base.Close()
    det_engine.Close()

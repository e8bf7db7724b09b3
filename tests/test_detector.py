"""The detector's settings and smoothing (detector.go:101-278, 806-817).

CPU: the oracle pinned by the reference's TestGuardTrainHzScaling shapes
(detector_test.go:142-176: 2.048 Msps, 500 Hz guard, 5 kHz train at FFT 2048
and 65536), the config defaults (config.go:470-481), closed forms of the
adaptive alpha (dt = 1/15 s gives EmaAlpha, 2/15 s gives 1 - (1 - a)^2, the
clamps) and of the EMA (first frame copied, the step response
1 - (1 - a)^k, a constant is a fixed point); then the library's host
functions equal to the oracle.  GPU (marked): the smoothing kernel bit-exact
against the oracle EMA, and the device detection chain on its output."""
import numpy as np
import pytest

import oracle as O


def test_settings_hz_to_cells_known_answers():
    s = O.DetectorSettings(2_048_000, 2048, 500.0, 5000.0)     # binWidth 1000 Hz
    assert (s["guard"], s["train"], s["rank"]) == (1, 5, 8)   # rank round(7.5) = 8, half away from zero
    s = O.DetectorSettings(2_048_000, 65536, 500.0, 5000.0)    # binWidth 31.25 Hz
    assert (s["guard"], s["train"], s["rank"]) == (16, 160, 240)
    s = O.DetectorSettings(10_000_000, 32768, 0.0, 0.0)        # train at least 1, guard at least 0
    assert (s["guard"], s["train"], s["rank"]) == (0, 1, 2)    # round(1.5) = 2
    s = O.DetectorSettings(2_048_000, 0, 500.0, 5000.0, rank=36)  # fft out of range -> 65536; rank kept
    assert s["fft"] == 65536 and s["rank"] == 36
    s = O.DetectorSettings(2_048_000, 2048, 500.0, 3000.0, rank=99)  # rank > 2 train -> default round(4.5) = 5
    assert s["rank"] == 5
    d = O.DetectorSettings(1_000_000, 1024, 1.0, 1.0)
    assert (d["ema_alpha"], d["hysteresis_db"], d["scale_db"], d["edge_margin_db"], d["max_signal_bw_hz"],
            d["merge_gap_hz"]) == (0.2, 3.0, 6.0, 3.0, 150000.0, 5000.0)


def test_adaptive_alpha_closed_forms():
    a = 0.2
    assert O.AdaptiveAlpha(a, 1 / 15) == pytest.approx(a, rel=1e-15)
    assert O.AdaptiveAlpha(a, 2 / 15) == pytest.approx(1 - (1 - a) ** 2, rel=1e-15)
    assert O.AdaptiveAlpha(a, 0.0) == O.AdaptiveAlpha(a, 1 / 15) == O.AdaptiveAlpha(a, 5.0)
    assert O.AdaptiveAlpha(1.0, 1 / 15) == 0.99
    assert O.AdaptiveAlpha(0.001, 1 / 150) == 0.01


def test_smooth_spectrum_closed_forms():
    sm = O.SmoothSpectrum()
    a = 0.3
    first = sm(np.zeros(5), a)
    assert np.array_equal(first, np.zeros(5))
    for k in range(1, 20):
        got = sm(np.ones(5), a)
        assert np.allclose(got, 1 - (1 - a) ** k, rtol=0, atol=1e-15)
    c = O.SmoothSpectrum()
    c(np.full(3, -80.0), a)
    assert np.array_equal(c(np.full(3, -80.0), a), np.full(3, -80.0))
    assert np.array_equal(c(np.full(4, 1.0), a), np.ones(4))  # a size change restarts with a copy


def test_library_host_functions_match_oracle():
    from argus import surveillance as S
    rng = np.random.default_rng(4)
    for _ in range(300):
        sr = int(rng.choice([2_048_000, 10_000_000, 20_000_000, 250_000, int(rng.integers(8000, 60_000_000))]))
        fft = int(rng.choice([1024, 2048, 4096, 32768, 65536, 0, -5, 1 << 23]))
        g, t = float(rng.uniform(0, 20000)), float(rng.uniform(0, 50000))
        rank = int(rng.choice([0, 1, 5, 36, 1000]))
        cfg = S.DetectorConfig(CFARGuardHz=g, CFARTrainHz=t, CFARRank=rank, EmaAlpha=float(rng.choice([0.0, 0.2, 2.0])),
                               HysteresisDb=0.0, CFARScaleDb=float(rng.choice([0.0, 12.3748])))
        cf, det, alpha, n = S.DetectorSettings(cfg, sr, fft)
        w = O.DetectorSettings(sr, fft, g, t, rank=rank, ema_alpha=cfg.EmaAlpha, scale_db=cfg.CFARScaleDb)
        assert (n, cf.guard, cf.train, cf.rank) == (w["fft"], w["guard"], w["train"], w["rank"])
        assert (alpha, cf.scale_db, cf.hysteresis_db, det.edge_margin_db, det.max_signal_bw_hz, det.merge_gap_hz) \
            == (w["ema_alpha"], w["scale_db"], w["hysteresis_db"], w["edge_margin_db"], w["max_signal_bw_hz"],
                w["merge_gap_hz"])
        dt = float(rng.choice([0.0, -1.0, 1 / 15, 0.01, 0.5, 1.0, 1.5, float(rng.uniform(0, 1))]))
        assert S.AdaptiveAlpha(alpha, dt) == pytest.approx(O.AdaptiveAlpha(alpha, dt), rel=1e-15)


@pytest.mark.gpu
def test_device_smoothing_and_detection(gpu):
    from argus import surveillance as S
    from argus.synth import surveillance_scene
    n, sr = 4096, 10_000_000
    iq, _ = surveillance_scene(n * 24, sr, n_carriers=20, seed=7)
    eng = S.Surveillance(n, sr)
    psd = eng.WelchPSD(iq, n_segs=2, seg_hop=n // 2, frame_step=3 * n // 2, n_frames=12)
    cfg = S.DetectorConfig(CFARMode="CA", CFARGuardHz=4 * sr / n, CFARTrainHz=16 * sr / n, CFARScaleDb=12.3748,
                           EmaAlpha=0.35)
    d = S.Detector(cfg, sr, n)
    times = np.cumsum(np.r_[0.0, np.full(11, 1 / 15)]) + np.r_[0, 0, 0.02, 0, 0, 0, 0.3, 0, 0, 0, 0, 0]
    ref = O.SmoothSpectrum()
    last = None
    for lo, hi in ((0, 1), (1, 5), (5, 12)):  # the state carries across calls
        print(f"smoothing frames {lo}..{hi}", flush=True)
        lists, sm = d.Process(psd[lo:hi], times[lo:hi])
        for f in range(lo, hi):
            dt = 0.0 if last is None else times[f] - last
            last = times[f]
            want = ref(psd[f].astype(np.float64), O.AdaptiveAlpha(0.35, dt))
            assert np.array_equal(sm[f - lo], want.astype(np.float32)), f  # FP64 EMA, bit for bit
            thr = O.CFARThresholds("CA", sm[f - lo].astype(np.float64), d.cfar.guard, d.cfar.train, 12.3748)
            w = O.DetectSignals(sm[f - lo].astype(np.float64), thr, sample_rate=sr, threshold_db=-20.0,
                                hysteresis_db=3.0, cfar_scale_db=12.3748)
            assert [(s.FirstBin, s.LastBin, s.PeakBin) for s in lists[f - lo]] \
                == [(s.FirstBin, s.LastBin, s.PeakBin) for s in w], f
        assert np.array_equal(d.EMA(), ref.ema)
    d.Close()
    eng.Close()

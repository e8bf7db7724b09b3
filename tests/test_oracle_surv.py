"""Pins the surveillance oracle: FFT, Hann, Welch, CFAR and the detection list.

The FFT is third-party in the reference (gonum v0.15.0 dsp/fourier), so it is
pinned by the DFT definition (brute force), numpy's FFT, Parseval and the tone
position the reference's synth tests assert.  CFAR known answers are
cfar_test.go:21-94; detection-list behaviour follows detector_test.go.
"""
import numpy as np
import pytest

import oracle as O


def test_fft_matches_definition_and_numpy():
    rng = np.random.default_rng(1)
    for n in [1, 2, 8, 60, 256, 4096]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        got = O.FFT(x)
        if n <= 256:
            k = np.arange(n)
            brute = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
            np.testing.assert_allclose(got, brute, rtol=0, atol=1e-9 * max(1, n))
        np.testing.assert_allclose(got, np.fft.fft(x), rtol=0, atol=1e-10 * n)
        # Parseval
        assert abs(np.sum(np.abs(got) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-9 * n


def test_hann_symmetric_closed_form():
    w = O.Hann(32768)
    assert w[0] == 0 and w[-1] < 1e-15 and abs(w[16383] - w[16384]) < 1e-15
    i = np.arange(32768)
    np.testing.assert_allclose(w, 0.5 * (1 - np.cos(2 * np.pi * i / 32767)), atol=1e-15)
    assert list(O.Hann(1)) == [1.0] and len(O.Hann(0)) == 0


def test_tone_lands_at_shifted_bin():
    # synth_test.go:16-21,49-80: a tone at +f sits at bin N/2 + f/df
    n, fs, f = 4096, 4_096_000, 100_000
    t = np.arange(n)
    x = np.exp(2j * np.pi * f / fs * t).astype(np.complex64)
    s = O.Spectrum(x, O.Hann(n))
    assert int(np.argmax(s)) == n // 2 + f // (fs // n)
    p = O.WelchPSD(np.tile(x, 3), n, 0.5, O.Hann(n))
    assert int(np.argmax(p)) == n // 2 + f // (fs // n)


def test_welch_against_numpy_and_variance_reduction():
    rng = np.random.default_rng(2)
    n, seg = 8 * 16384 + 16384, 32768
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    w = O.Hann(seg)
    got = O.WelchPSD(x, seg, 0.5, w)
    acc = np.zeros(seg)
    segs = 0
    for st in range(0, n - seg + 1, seg // 2):
        X = np.fft.fft(x[st:st + seg].astype(np.complex128) * w)
        acc += (np.abs(X) / seg) ** 2
        segs += 1
    assert segs == 8
    want = 10 * np.log10(np.fft.fftshift(acc / segs) + 1e-24)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)
    # welch_test.go:30-68: averaging lowers the per-bin spread vs one periodogram
    single = O.Spectrum(x[:seg], w)
    assert np.std(got) < 0.6 * np.std(single)


def test_welch_short_input_falls_back_to_spectrum():
    x = np.ones(100, np.complex64)
    np.testing.assert_allclose(O.WelchPSD(x, 256, 0.5), O.Spectrum(x), atol=0)


def test_cfar_known_answers():
    # cfar_test.go:65-94
    spec = [-100, -100, -100, -80, -100, -90, -100, -100, -100]
    ca = O.CFARThresholds("CA", spec, 0, 2, 0.0 + 1e-300, wrap=False)[5]
    go = O.CFARThresholds("GOSCA", spec, 0, 2, 1e-300, wrap=False)[5]
    so = O.CFARThresholds("CASO", spec, 0, 2, 1e-300, wrap=False)[5]
    assert abs(ca - 10 * np.log10((1e-8 + 3e-10) / 4)) < 1e-9      # -85.892
    assert abs(go - 10 * np.log10(5.05e-9)) < 1e-9                 # -82.967
    assert abs(so - (-100.0)) < 1e-9
    assert go > ca > so
    # scale <= 0 defaults to 6 dB (cfar.go:13-15)
    assert abs(O.CFARThresholds("CA", spec, 0, 2, 0.0, wrap=False)[5] - ca - 6) < 1e-9
    # cfar_test.go:21-51: a plateau is detected, noise is not, wrap edges finite
    s = np.full(1024, -100.0)
    s[500:511] = -20
    for m in ["CA", "OS", "GOSCA", "CASO"]:
        th = O.CFARThresholds(m, s, 2, 16, 6, wrap=True, rank=24)
        assert s[505] >= th[505] and s[200] < th[200]
    s2 = np.full(256, -100.0)
    s2[0:6] = -20
    th = O.CFARThresholds("CA", s2, 2, 8, 6, wrap=True)
    assert -200 < th[0] <= 0 and -200 < th[255] <= 0
    assert O.CFARThresholds("CA", s2, 2, 0, 6) is None


def test_cfar_ca_brute_force_window_sums():
    rng = np.random.default_rng(3)
    s = rng.uniform(-110, -60, 4096)
    g, t = 4, 16
    th = O.CFARThresholds("CA", s, g, t, 12.3748, wrap=True)
    lin = 10 ** (s / 10)
    idx = np.arange(len(s))
    left = sum(lin[(idx - g - k) % len(s)] for k in range(1, t + 1))
    right = sum(lin[(idx + g + k) % len(s)] for k in range(1, t + 1))
    want = 10 * np.log10((left + right) / (2 * t)) + 12.3748
    np.testing.assert_allclose(th, want, rtol=0, atol=1e-9)


def test_pfa_scale_closed_form():
    # CA-CFAR, exponential noise, N = 2t = 32 cells: T = N (Pfa^(-1/N) - 1)
    N, pfa = 32, 1e-6
    assert abs(10 * np.log10(N * (pfa ** (-1 / N) - 1)) - 12.3748) < 1e-4


def test_detector_creates_signal_fixed_threshold():
    # detector_test.go:11-40 (CFAR OFF, ThresholdDb -10, fs 1000, 10 bins)
    s = [-30, -30, -30, -5, -5, -30, -30, -30, -30, -30]
    sig = O.DetectSignals(s, None, sample_rate=1000, threshold_db=-10)
    assert len(sig) == 1 and sig[0].BWHz > 0


def test_signal_bandwidth_expansion():
    # detector_test.go:57-105
    n, fs = 2048, 2_048_000
    s = np.full(n, -100.0)
    s[1000:1013] = -20
    for j in range(1, 6):
        s[1000 - j] = -20 - 3 * j
        s[1012 + j] = -20 - 3 * j
    sig = O.DetectSignals(s, None, sample_rate=fs, center_hz=434e6, threshold_db=-20)
    assert sig and sig[0].BWHz >= 18_000


def test_merge_gap():
    # detector_test.go:107-140
    n, fs = 2048, 2_048_000
    s = np.full(n, -100.0)
    s[500:506] = -20
    s[509:515] = -20
    sig = O.DetectSignals(s, None, sample_rate=fs, center_hz=434e6, threshold_db=-20, merge_gap_hz=5000)
    assert len(sig) == 1 and sig[0].BWHz >= 14_000


def test_cfar_detection_on_noise_floor():
    # a narrow carrier over a noisy floor passes CA-CFAR (t=16, g=4, Pfa 1e-6)
    rng = np.random.default_rng(0)
    n, fs = 8192, 8_192_000
    s = -100 + 10 * np.log10(rng.exponential(1.0, n))
    s[3000:3004] = -60
    th = O.CFARThresholds("CA", s, 4, 16, 12.3748, wrap=True)
    sig = O.DetectSignals(s, th, sample_rate=fs, center_hz=100e6, cfar_scale_db=12.3748)
    hits = [g for g in sig if g.FirstBin <= 3003 and g.LastBin >= 3000]
    assert len(hits) == 1
    assert abs(hits[0].CenterHz - (100e6 + (3001.5 - 4096) * 1000.0)) < 2e3


def test_detection_list_hysteresis_keeps_run_open():
    s = np.full(1000, -100.0)
    s[100:110] = -50
    s[110:115] = -52.5   # below on (-51) but above off (-54): stays in the run
    s[115:120] = -50
    th = np.full(1000, -51.0)
    sig = O.DetectSignals(s, th, sample_rate=1_000_000, hysteresis_db=3.0, max_signal_bw_hz=0.0)
    assert len(sig) == 1 and sig[0].FirstBin == 100 and sig[0].LastBin == 119


def test_level_spectrum_composition_and_known_answers():
    """The derived surveillance level (pipeline_runtime.go:399-426, 624-645)
    as the oracle composes it: equal to an independent numpy restatement
    (causal FP64 convolution from zero history, every decim-th sample, Hann,
    FFT, dB, fftshift); a tone inside the level lands at bin n_d/2 + f/df
    (the +f -> upper-half orientation of synth_test.go:16-21); a tone beyond
    the level's band is suppressed by the decimation low-pass."""
    fs, decim, n = 2_000_000, 2, 8192
    n_d = n // decim
    df = fs / decim / n_d
    t = np.arange(n)
    f_in = 40 * df
    x = (np.exp(2j * np.pi * f_in * t / fs) + 1e-3 * np.exp(2j * np.pi * 0.45 * t)).astype(np.complex64)
    got = O.SurveillanceLevelSpectrum(x, fs, decim)
    taps = O.LowpassFIR(float(fs // decim) * 0.5 * 0.8, fs, 101)
    y = np.convolve(x.astype(np.complex128), taps)[:n].astype(np.complex64)[::decim]
    w = 0.5 * (1 - np.cos(2 * np.pi * np.arange(n_d) / (n_d - 1)))
    X = np.fft.fft(y.astype(np.complex128) * w)
    want = 20 * np.log10(np.abs(np.fft.fftshift(X)) / n_d + 1e-12)
    assert len(got) == n_d
    np.testing.assert_allclose(got, want, atol=1e-6)
    assert int(np.argmax(got)) == n_d // 2 + 40
    # 0.45 fs aliases to 0.45 fs - fs/2 = -0.05 fs in the level: far below the tone
    alias = n_d // 2 + int(round((0.45 * fs - fs / 2) / df))
    assert got[alias] < got.max() - 80

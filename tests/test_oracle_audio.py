"""The oracle's post-demodulation audio stage (internal/dsp/resample.go,
internal/recorder/streamer.go:1408-1470), pinned by:
- the reference's own resample_test.go cases (gcd, ratio, identity, output
  length, chunked streaming, tone preservation, stereo);
- an independent restatement of the prototype (numpy's Kaiser window,
  sin(2 pi fc x)/x) and of the output as a direct FP64 convolution;
- the closed form of one call (resample_block below), including the phase
  step the reference loses when a call's input ends mid-consumption;
- scipy's lfilter for the de-emphasis recursion."""
import numpy as np
import pytest
from scipy.signal import lfilter

import oracle as O


@pytest.mark.parametrize("a,b,want", [(48000, 51200, 3200), (48000, 44100, 300), (48000, 48000, 48000),
                                      (48000, 96000, 48000), (48000, 200000, 8000)])
def test_gcd(a, b, want):
    """resample_test.go:8-24"""
    assert O.gcd(a, b) == want


@pytest.mark.parametrize("fin,fout,L,M", [(51200, 48000, 15, 16), (44100, 48000, 160, 147), (48000, 48000, 1, 1),
                                          (96000, 48000, 1, 2), (250000, 48000, 24, 125), (12500, 48000, 96, 25)])
def test_ratio(fin, fout, L, M):
    """resample_test.go:26-44 (+ the service's demod rates)"""
    assert O.Resampler(fin, fout, 32).Ratio() == (L, M)


def test_identity_copies():
    """resample_test.go:46-62"""
    x = np.sin(2 * np.pi * 440 * np.arange(1000) / 48000).astype(np.float32)
    np.testing.assert_array_equal(O.Resampler(48000, 48000, 32).Process(x), x)


@pytest.mark.parametrize("fin,fout,n", [(51200, 48000, 5120), (51200, 48000, 10240), (44100, 48000, 4410),
                                        (96000, 48000, 9600), (200000, 48000, 20000), (16000, 48000, 333),
                                        (250000, 48000, 5000)])
def test_output_length(fin, fout, n):
    """resample_test.go:64-88: within 3 of n * out / in."""
    r = O.Resampler(fin, fout, 32)
    y = r.Process(np.sin(2 * np.pi * 1000 * np.arange(n) / fin).astype(np.float32))
    assert abs(len(y) - n * fout / fin) <= 3


def test_prototype_is_kaiser_windowed_sinc():
    """polyBank[p][t] = proto[p + t L], proto = sin(2 pi fc x) / x * kaiser(beta 6) * L
    (resample.go:59-84), restated with numpy's Kaiser window."""
    for fin, fout in [(250000, 48000), (12500, 48000), (51200, 48000)]:
        r = O.Resampler(fin, fout, 32)
        L, M = r.Ratio()
        n = 32 * L + 1
        x = np.arange(n) - (n - 1) / 2
        fc = 0.90 / max(L, M)
        with np.errstate(invalid="ignore", divide="ignore"):
            s = np.where(np.abs(x) < 1e-12, 2 * np.pi * fc, np.sin(2 * np.pi * fc * x) / x)
        proto = s * np.kaiser(n, 6.0) * L
        bank = np.zeros((L, r.taps))
        for p in range(L):
            idx = p + L * np.arange(r.taps)
            ok = idx < n
            bank[p, ok] = proto[idx[ok]]
        np.testing.assert_allclose(r.Bank(), bank, rtol=1e-9, atol=1e-12 * np.abs(proto).max())


def resample_block(bank, L, M, t0, hist, x):
    """Closed form of one Process call: (outputs, new t0, new history).  Output
    k consumes D_k = floor((t0 + k M) / L) samples of this block and uses arm
    t0 + k M - L D_k over the last `taps` samples; with k_n the first k whose
    D_k reaches n, the call emits k_n + 1 outputs if D_{k_n} == n, else k_n
    and loses one phase step (resample.go:124-134 decrements before finding
    the input exhausted)."""
    n, T = len(x), bank.shape[1]
    if n == 0:
        return np.zeros(0, np.float32), t0, hist
    kn = max(0, -(-(n * L - t0) // M))
    D = (t0 + kn * M) // L
    K = kn + 1 if D == n else kn
    t_new = t0 + K * M - L * (n + (0 if D == n else 1))
    xp = np.concatenate([hist, x.astype(np.float64)])
    out = np.empty(K, np.float32)
    for k in range(K):
        Dk = (t0 + k * M) // L
        out[k] = bank[t0 + k * M - L * Dk] @ xp[len(hist) + Dk - 1 - np.arange(T)]
    return out, t_new, xp[-T:]


def test_stream_closed_form_and_aligned_chunks():
    """Any chunking of the input gives, call by call, what the closed form
    above predicts (counts, values, carried phase); chunks of whole multiples
    of M give exactly the one-block output (resample_test.go:90-140)."""
    rng = np.random.default_rng(3)
    for fin, fout in [(250000, 48000), (51200, 48000), (16000, 48000), (25000, 48000), (2500000, 48000)]:
        x = rng.standard_normal(fin // 20).astype(np.float32)
        r = O.Resampler(fin, fout, 32)
        L, M = r.Ratio()
        bank, T = r.Bank(), r.taps
        one = r.Process(x)
        for chunk in (M * 20, 777, 1, 4096, 5000):
            rc = O.Resampler(fin, fout, 32)
            t0, hist, parts = 0, np.zeros(T), []
            for i in range(0, len(x), chunk):
                got = rc.Process(x[i:i + chunk])
                want, t0, hist = resample_block(bank, L, M, t0, hist, x[i:i + chunk])
                assert len(got) == len(want), (fin, chunk, i)
                np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * max(1.0, np.abs(want).max(initial=0)))
                parts.append(got)
            if chunk % M == 0:
                np.testing.assert_array_equal(np.concatenate(parts), one)


def test_tone_preservation():
    """resample_test.go:176-202: a 1 kHz tone keeps its frequency (zero crossings)."""
    fin, fout = 51200, 48000
    x = np.sin(2 * np.pi * 1000 * np.arange(fin) / fin).astype(np.float32)
    y = O.Resampler(fin, fout, 32).Process(x)
    z = y[100:]
    crossings = np.count_nonzero((z[:-1] <= 0) & (z[1:] > 0) | (z[:-1] >= 0) & (z[1:] < 0))
    assert abs(crossings / 2 * fout / (len(y) - 101) - 1000) < 10


def test_dc_gain_is_the_arm_sum():
    """A constant input settles to the sum of the arm in use: the reference's
    sinc omits the 1/pi, so every arm sums to about pi."""
    r = O.Resampler(250000, 48000, 32)
    L, M = r.Ratio()
    y = r.Process(np.ones(5000, np.float32))
    sums = r.Bank().sum(axis=1)
    j = np.arange(200, len(y))
    np.testing.assert_allclose(y[200:], sums[(j * M) % L], rtol=1e-6)
    assert np.all(np.abs(sums - np.pi) < 0.01)


def test_stereo_is_two_monos():
    """resample_test.go:204-233: left and right resampled independently, interleaved."""
    fin = 51200
    t = np.arange(fin // 2) / fin
    st = np.empty(2 * len(t), np.float32)
    st[0::2], st[1::2] = np.sin(2 * np.pi * 440 * t), np.sin(2 * np.pi * 880 * t)
    out = O.StereoResampler(fin, 48000).Process(st)
    np.testing.assert_array_equal(out[0::2], O.Resampler(fin, 48000).Process(st[0::2]))
    np.testing.assert_array_equal(out[1::2], O.Resampler(fin, 48000).Process(st[1::2]))
    assert abs(len(out) // 2 - len(t) * 48000 / fin) <= 3


def test_deemphasis_is_one_pole_lowpass():
    """y = a y + (1 - a) x (streamer.go:1443-1467) equals lfilter([1-a], [1, -a])
    with the carried state as initial condition; stereo channels run apart."""
    rng = np.random.default_rng(4)
    a = np.exp(-1.0 / (48000 * 50e-6))
    x = rng.standard_normal(960).astype(np.float32)
    got, y = O.Deemphasis(x, a, 0.3)
    want, zf = lfilter([1 - a], [1, -a], x.astype(np.float64), zi=[a * 0.3])
    np.testing.assert_allclose(got, want.astype(np.float32), rtol=0, atol=1e-6)
    assert y == pytest.approx(want[-1], rel=1e-12)
    st = rng.standard_normal(2 * 480).astype(np.float32)
    g1, _ = O.Deemphasis(st, a, 0.0, stride=2, offset=0)
    g2, _ = O.Deemphasis(g1, a, 0.0, stride=2, offset=1)
    np.testing.assert_allclose(g2[0::2], lfilter([1 - a], [1, -a], st[0::2].astype(np.float64)), atol=1e-6)
    np.testing.assert_allclose(g2[1::2], lfilter([1 - a], [1, -a], st[1::2].astype(np.float64)), atol=1e-6)

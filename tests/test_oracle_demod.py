"""Pins of the oracle demodulators (internal/demod fm.go, am.go, ssb.go,
cw.go) by closed forms: a known input whose demodulated output the
mathematics fixes, and chunked == monolithic for the carried states."""
import numpy as np
import pytest

import oracle as O


def _tone(n, w, phi0=0.0, amp=1.0):
    i = np.arange(n)
    return (amp * np.exp(1j * (w * i + phi0))).astype(np.complex64)


@pytest.mark.parametrize("w", [0.3, -1.1, 2.9, -3.05])
def test_fm_discriminator_of_a_tone_is_its_frequency(w):
    # fm.go:179-196: atan2(Im(conj(p) c), Re(conj(p) c)) = w for c = p e^{jw}
    y = O.fmDiscrim(_tone(4000, w, 0.4, 0.8))
    assert len(y) == 3999
    assert np.max(np.abs(y - w)) < 2e-5  # float32 IQ rounding only


def test_fm_discriminator_of_a_linear_chirp():
    # phase k^2 * c -> frequency (2k + 1) c, wrapped into (-pi, pi]
    c = 1e-4
    k = np.arange(5000)
    iq = np.exp(1j * c * k.astype(np.float64) ** 2).astype(np.complex64)
    want = np.angle(np.exp(1j * c * (2 * k[1:] - 1)))
    assert np.max(np.abs(O.fmDiscrim(iq) - want)) < 1e-4


def test_bridged_fm_equals_monolithic():
    # streamer.go:1287-1302: prepending the previous block's last sample makes
    # the block stream equal to one discriminator over the whole stream
    rng = np.random.default_rng(5)
    iq = (rng.standard_normal(10000) + 1j * rng.standard_normal(10000)).astype(np.complex64)
    mono = O.fmDiscrim(iq)
    br = O.BridgedFMDiscrim()
    parts = [br.Demod(iq[a:b]) for a, b in [(0, 1), (1, 17), (17, 4096), (4096, 4097), (4097, 10000)]]
    np.testing.assert_array_equal(np.concatenate(parts), mono)


def test_am_of_a_modulated_tone():
    # am.go:11-27: |A (1 + m cos(v i))| - mean = A m cos(v i) - mean of that
    n, A, m, v = 6000, 0.7, 0.5, 2 * np.pi / 600.0  # 10 whole periods
    i = np.arange(n)
    env = A * (1 + m * np.cos(v * i))
    iq = (env * np.exp(1j * 0.37 * i)).astype(np.complex64)
    want = env - env.mean()
    assert np.max(np.abs(O.AMDemod(iq) - want)) < 1e-6
    assert abs(float(np.mean(O.AMDemod(iq)))) < 1e-6


def test_am_of_constant_envelope_is_zero():
    y = O.AMDemod(_tone(1000, 0.9, amp=0.25))
    assert np.max(np.abs(y)) < 1e-7


@pytest.mark.parametrize("usb", [True, False])
def test_ssb_product_detector_closed_form(usb):
    # ssb.go:19-39: Re(x e^{j phi_i}), phi_i = (i + 1) * 2 pi (+-700) / fs
    fs, n, w = 48_000, 3000, 2 * np.pi * 1100 / 48_000
    x = _tone(n, w, 0.2, 0.9)
    got, ph = O.SSBDemod(x, fs, usb=usb)
    bfo = 2 * np.pi * (700.0 if usb else -700.0) / fs
    i = np.arange(n)
    want = 0.9 * np.cos(w * i + 0.2 + (i + 1) * bfo)
    assert np.max(np.abs(got - want)) < 1e-5
    assert abs(np.angle(np.exp(1j * (ph - n * bfo)))) < 1e-9


def test_cw_closed_form():
    # cw.go:11-26: Re(x) cos(phi_i), +700 Hz
    fs, n = 8000, 2000
    x = _tone(n, 0.05, 0.0, 1.0)
    got, _ = O.CWDemod(x, fs)
    i = np.arange(n)
    want = np.cos(0.05 * i) * np.cos((i + 1) * 2 * np.pi * 700 / fs)
    assert np.max(np.abs(got - want)) < 1e-5


def test_ssb_and_cw_carried_phase_equal_monolithic():
    rng = np.random.default_rng(6)
    x = (rng.standard_normal(9000) + 1j * rng.standard_normal(9000)).astype(np.complex64)
    for fn in (lambda v, p: O.SSBDemod(v, 16000, True, p), lambda v, p: O.SSBDemod(v, 16000, False, p),
               lambda v, p: O.CWDemod(v, 16000, p)):
        whole, _ = fn(x, 0.0)
        ph, parts = 0.0, []
        for a, b in [(0, 1), (1, 1000), (1000, 1001), (1001, 9000)]:
            y, ph = fn(x[a:b], ph)
            parts.append(y)
        assert np.max(np.abs(np.concatenate(parts) - whole)) < 1e-5

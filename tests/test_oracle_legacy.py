"""Pins of the oracle's legacy block chain (LegacyShiftFilterDecimate,
LegacyDemod) by library routines and closed forms, so the GPU parity tests
of the legacy path compare against something fixed independently:

* the FIR and the kept indices against numpy's convolution, decimated by
  slicing (exports.cu:114-161: causal FIR, keep 0, D, 2D, ...);
* a tone at the job's offset is shifted to DC: after the filter's T-1
  samples the output is the constant e^{j phase_start} * sum(h);
* the phase end is phase_start + inc * n (batch_runner.go:87-90);
* the discriminator of a tone is its constant phase step; the envelope of a
  constant-amplitude carrier is that amplitude; the USB / LSB product
  detector moves a tone at f to f + 700 / f - 700 Hz (spectral peak)."""
import numpy as np

import oracle as O


def test_chain_against_numpy_convolution():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(5003) + 1j * rng.standard_normal(5003)).astype(np.complex64)
    taps = rng.standard_normal(37).astype(np.float32)
    sr, off, ph0 = 1_000_000, 12_345.0, 0.75
    for D in (1, 3, 8, 40):
        y, ph_end = O.LegacyShiftFilterDecimate(x, sr, off, ph0, taps, D)
        inc = -2 * np.pi * off / sr
        sh = x.astype(np.complex128) * np.exp(1j * (ph0 + inc * np.arange(len(x))))
        want = np.convolve(sh, taps.astype(np.float64))[:len(x)][::D][:len(x) // D]
        assert len(y) == len(x) // D
        assert np.max(np.abs(y - want)) < 1e-9
        assert ph_end == ph0 + inc * len(x)


def test_tone_at_offset_lands_at_dc():
    sr, off, ph0, n = 2_048_000, 300_000.0, -1.1, 4096
    x = np.exp(2j * np.pi * off * np.arange(n) / sr)
    h = O.LowpassFIR(50_000, sr, 101)
    y, _ = O.LegacyShiftFilterDecimate(x, sr, off, ph0, h, 4)
    steady = y[(len(h) + 3) // 4:]
    assert np.max(np.abs(steady - np.exp(1j * ph0) * h.sum())) < 1e-9


def test_demodulators_closed_forms():
    sr = 48_000
    n = 4800
    f = 1234.0
    tone = np.exp(2j * np.pi * f * np.arange(n) / sr)
    fm, _ = O.LegacyDemod(tone, 0, sr)
    assert len(fm) == n - 1 and np.max(np.abs(fm - 2 * np.pi * f / sr)) < 1e-9
    am, _ = O.LegacyDemod(0.37 * tone, 2, sr)
    assert len(am) == n and np.max(np.abs(am - 0.37)) < 1e-12
    for mode, shift in ((3, 700.0), (4, -700.0), (5, 700.0)):
        out, bfo = O.LegacyDemod(tone, mode, sr, 0.0)
        spec = np.abs(np.fft.rfft(out * np.hanning(n)))
        peak = np.argmax(spec) * sr / n
        assert abs(peak - abs(f + shift)) <= sr / n
        assert bfo == 2 * np.pi * shift / sr * n

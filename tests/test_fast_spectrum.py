"""The fast form's per-channel spectrum (gpud_fast_spectrum, host code):
equal to numpy's FFT of the modulated taps to float rounding, for the
lengths and offsets the service uses; the phase increment is applied as a
fixed-point turn, so k * inc wraps exactly.  Host code only."""
import numpy as np
import pytest


@pytest.mark.parametrize("T,inc", [(127, -0.3), (511, 2.9), (64, 0.0), (3, -3.1)])
def test_fast_spectrum_matches_numpy(T, inc):
    from argus import serve
    rng = np.random.default_rng(T)
    h = rng.standard_normal(T).astype(np.float32)
    got = serve.FastSpectrum(h, inc)
    turns = inc / (2 * np.pi)
    k = np.arange(T)
    g = h.astype(np.float64) * np.exp(-2j * np.pi * ((turns * k) % 1.0))
    want = np.fft.fft(g, 4096) / 4096
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert err < 1e-6, err


def test_fast_spectrum_rejects_long_filters():
    from argus import serve
    from argus._lib import GpudError
    with pytest.raises(GpudError):
        serve.FastSpectrum(np.ones(513, np.float32), 0.1)

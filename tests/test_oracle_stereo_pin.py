"""Pins of the oracle's stereo decoder (streamer.go:1496-1643) beyond the
reference's own tests (which check only a 1.3x level difference):

1. Orientation, derived from the reference's conventions.  The PLL mixes
   the band-passed pilot with cos(phi) and -sin(phi) and drives
   atan2(Q, I) to zero, so for a pilot sin(psi) it locks at
   phi = psi - pi/2.  The L-R carrier is regenerated as 2 sin(2 phi') from
   the phase phi' after the sample's update, one step w = 2 pi 19 kHz / fs
   ahead: 2 sin(2 psi - pi + 2w) = -2 sin(2 psi + 2w).  On a broadcast
   multiplex 1/2 (L+R) + 1/2 (L-R) sin(2 psi) + 0.09 sin(psi) that gives
   lr = -g 1/2 (L-R) with g = cos(2w), delayed by the band-pass and the
   L-R low-pass (50 + 50 samples) against 1/2 (L+R) behind one low-pass
   (50 samples).  Output 0 = 1/2 (lpr + lr), output 1 = 1/2 (lpr - lr): an
   L-only tone of frequency f and amplitude 1 comes out with amplitude
   |1 - g e^{-jx}| / 4 on output 0 and |1 + g e^{-jx}| / 4 on output 1,
   x = 2 pi f 50 / fs -- the tone lands on output 1.  A flipped carrier,
   swapped matrix, a phase taken before the update or a different delay
   all change these two numbers.
2. Loop dynamics.  The linearised loop -- phase error u, the 200 Hz one-pole
   on the I/Q products acting on it, the type-II update f += beta e,
   phi += f + alpha e with (alpha, beta) from the closed form of
   pllCoefficients(50, 0.707, fs) -- predicts the frequency estimate after
   a 20 Hz pilot step; the oracle must follow it, settle on the new
   frequency, and stay locked."""
import numpy as np
import pytest

import oracle as O

FS = 250_000


def pll_coefficients(loop_bw, damping, fs):
    """streamer.go:1496-1506 restated from its formula."""
    bl = loop_bw / fs
    theta = bl / (damping + 0.25 / damping)
    d = 1 + 2 * damping * theta + theta * theta
    return 4 * damping * theta / d, 4 * theta * theta / d


def mpx(left, right, t):
    return (0.5 * (left + right) + 0.5 * (left - right) * np.sin(2 * np.pi * 38000 * t)
            + 0.09 * np.sin(2 * np.pi * 19000 * t))


@pytest.mark.parametrize("f", [100.0, 400.0, 1000.0])
@pytest.mark.parametrize("side", ["L", "R"])
def test_orientation_and_separation_closed_form(f, side):
    n = FS
    t = np.arange(n) / FS
    tone = np.sin(2 * np.pi * f * t)
    zero = np.zeros(n)
    x = mpx(tone, zero, t) if side == "L" else mpx(zero, tone, t)
    d = O.StereoDecoder(FS)
    out, locked = d.Decode(x.astype(np.float32))
    assert locked  # this block's lock test (enabling takes 4 locked blocks)
    rms = [float(np.sqrt(np.mean(out[c::2][FS // 2:] ** 2))) for c in (0, 1)]
    g = np.cos(2 * 2 * np.pi * 19000 / FS)
    xph = 2 * np.pi * f * 50 / FS
    a0 = abs(1 - g * np.exp(-1j * xph)) / 4 / np.sqrt(2)
    a1 = abs(1 + g * np.exp(-1j * xph)) / 4 / np.sqrt(2)
    want = [a0, a1] if side == "L" else [a1, a0]  # R only: lr changes sign
    assert rms == pytest.approx(want, rel=0.01), (rms, want)
    loud = 1 if side == "L" else 0
    assert rms[loud] > rms[1 - loud]


def test_pll_coefficients_closed_form():
    a, b = O.pllCoefficients(50, 0.707, FS)
    wa, wb = pll_coefficients(50, 0.707, FS)
    assert a == pytest.approx(wa, rel=1e-14) and b == pytest.approx(wb, rel=1e-14)
    assert O.pllCoefficients(0, 0.707, FS) == (0.0, 0.0)


def test_pll_frequency_step_follows_linear_loop():
    w0 = 2 * np.pi * 19000 / FS
    dw = 2 * np.pi * 20 / FS
    n0, n1 = FS // 2, FS // 10
    ph = np.concatenate([w0 * np.arange(n0), w0 * (n0 - 1) + (w0 + dw) * np.arange(1, n1 + 1)])
    x = (0.09 * np.sin(ph)).astype(np.float32)
    d = O.StereoDecoder(FS)
    d.Decode(x[:n0])
    got = np.zeros(n1)
    for i in range(n1):
        d.Decode(x[n0 + i:n0 + i + 1])
        got[i] = d.st.freq
    alpha, beta = pll_coefficients(50, 0.707, FS)
    a = 1 - np.exp(-2 * np.pi * 200 / FS)
    u, v, fr = 0.0, 0.0, w0
    want = np.zeros(n1)
    for i in range(n1):
        v += a * (u - v)
        fr += beta * v
        u += (w0 + dw) - (fr + alpha * v)
        want[i] = fr
    dev = np.max(np.abs(got - want)) / dw
    assert dev < 0.03, dev
    assert abs(got[-1] - (w0 + dw)) < 0.005 * dw
    half = int(np.argmax(want - w0 >= dw / 2))
    assert 2500 < half < 4500  # the 50 Hz loop: half the step in about 14 ms
    assert d.Enabled

"""Seeded synthetic IQ scenes for tests and benchmarks.

Signal models follow the reference's internal/synth (synth.go:68-356): the
in-band SNR definition P_s = 10^(SNR/10) * 2 sigma^2 * B / fs (synth.go:84-90),
multitone FM within Carson bandwidth, and the WFM multiplex (pilot 19 kHz,
L-R DSB-SC at 38 kHz, RDS BPSK at 57 kHz).  The random streams are numpy's,
not Go's math/rand, so scenes match the reference's models, not its samples.
"""
from __future__ import annotations

import numpy as np


def awgn(n: int, sigma: float, rng) -> np.ndarray:
    """Complex AWGN with per-component standard deviation sigma."""
    return (sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)


def carrier_power(snr_db: float, sigma: float, bw_hz: float, fs: float) -> float:
    """synth.go:84-90 in-band SNR definition."""
    return 10 ** (snr_db / 10) * 2 * sigma * sigma * bw_hz / fs


def multitone_fm(t: np.ndarray, f0: float, bw_hz: float, rng) -> np.ndarray:
    """Unit-power FM carrier whose Carson bandwidth is about bw_hz."""
    fm = max(bw_hz / 20.0, 50.0)
    dev = max(bw_hz / 2.0 - fm, bw_hz / 10.0)
    tones = rng.uniform(0.3 * fm, fm, 3)
    amps = rng.dirichlet(np.ones(3))
    ph = 2 * np.pi * f0 * t
    for a, f in zip(amps, tones):
        ph = ph + a * dev / f * np.sin(2 * np.pi * f * t + rng.uniform(0, 2 * np.pi))
    return np.exp(1j * ph)


def surveillance_scene(n: int, fs: float, n_carriers: int = 40, seed: int = 0, sigma: float = 0.01,
                       span_hz: float = 4.9e6, bw_range=(3e3, 200e3), snr_range=(6.0, 40.0), chunk: int = 1 << 20):
    """40 carriers, centres uniform in +-span_hz, bandwidths and SNR uniform,
    over AWGN.  Returns (iq complex64, carrier table)."""
    rng = np.random.default_rng(seed)
    cars = []
    for _ in range(n_carriers):
        f0 = float(rng.uniform(-span_hz, span_hz))
        bw = float(rng.uniform(*bw_range))
        snr = float(rng.uniform(*snr_range))
        cars.append((f0, bw, snr, int(rng.integers(0, 2**31))))
    out = np.empty(n, np.complex64)
    nrng = np.random.default_rng(seed + 1)
    for s0 in range(0, n, chunk):
        m = min(chunk, n - s0)
        t = (np.arange(m) + s0) / fs
        x = awgn(m, sigma, nrng).astype(np.complex128)
        for f0, bw, snr, cs in cars:
            amp = np.sqrt(carrier_power(snr, sigma, bw, fs))
            x += amp * multitone_fm(t, f0, bw, np.random.default_rng(cs))
        out[s0:s0 + m] = x.astype(np.complex64)
    return out, cars

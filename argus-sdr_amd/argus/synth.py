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


RDS_BAUD = 1187.5


_RDS_OFFSETS = (0x0FC, 0x198, 0x168, 0x1B4)  # A, B, C, D offset words (IEC 62106)


def _rds_block(data: int, offset: int) -> int:
    """26-bit block: 16 information bits, then the 10-bit checkword (the
    remainder of data * x^10 modulo x^10+x^8+x^7+x^5+x^4+x^3+1) xor the
    offset word."""
    reg = data << 10
    for bit in range(25, 9, -1):
        if reg >> bit & 1:
            reg ^= 0x5B9 << (bit - 10)
    return (data << 10) | ((reg & 0x3FF) ^ offset)


def rds_group_bits(pi: int, ps: str, rt: str, n_groups: int) -> np.ndarray:
    """Data bits (MSB first) of n_groups RDS groups alternating 0A (two PS
    characters per group) and 2A (four RT characters per group)."""
    ps8 = ps.encode("latin-1").ljust(8)[:8]
    rt64 = rt.encode("latin-1").ljust(64)[:64]
    out = np.zeros(104 * n_groups, np.uint8)
    for g in range(n_groups):
        if g % 2 == 0:
            a = (g // 2) % 4
            words = (pi, a, 0xE0CD, ps8[2 * a] << 8 | ps8[2 * a + 1])
        else:
            a = (g // 2) % 16
            words = (pi, 2 << 12 | a, rt64[4 * a] << 8 | rt64[4 * a + 1], rt64[4 * a + 2] << 8 | rt64[4 * a + 3])
        for k, (w, off) in enumerate(zip(words, _RDS_OFFSETS)):
            blk = _rds_block(w, off)
            base = 104 * g + 26 * k
            out[base:base + 26] = [(blk >> (25 - i)) & 1 for i in range(26)]
    return out


def rds_symbols(bits: np.ndarray) -> np.ndarray:
    """Differentially encode data bits (s_i = s_{i-1} xor d_i) and map to NRZ
    levels +-1; the decoder recovers d_i = h_i xor h_{i-1}
    (rds.go:205-224)."""
    s = np.bitwise_xor.accumulate(bits.astype(np.uint8)) if len(bits) else np.zeros(0, np.uint8)
    return np.where(s > 0, 1.0, -1.0)


class WFMStation:
    """Broadcast stereo WFM station after genWFMDynamic (synth.go:269-356):
    multitone L/R program, 19 kHz pilot (0.09), L-R DSB-SC at 38 kHz, 57 kHz
    RDS NRZ BPSK (0.04) from a given symbol stream, +-75 kHz deviation.
    Generated block by block with continuous phase."""

    def __init__(self, fs: float, f0: float, symbols: np.ndarray, seed: int = 0, amp: float = 1.0):
        rng = np.random.default_rng(seed)
        self.fs, self.f0, self.amp = fs, f0, amp
        self.sym = symbols
        k = 12
        self.loud = 0.25 + 0.75 * rng.random()
        bw = 4000.0 + 11000.0 * rng.random()
        self.lf = bw * np.arange(1, k + 1) / k
        self.la, self.lp = rng.random(k), rng.random(k) * 2 * np.pi
        self.ra, self.rp = rng.random(k), rng.random(k) * 2 * np.pi
        peak = np.sum(self.la) if np.sum(self.la) > np.sum(self.ra) else np.sum(self.ra)
        self.g = self.loud / max(peak, 1e-9)
        self.phase = 0.0
        self.n = 0

    def block(self, m: int) -> np.ndarray:
        t = (np.arange(m) + self.n) / self.fs
        l = (self.la[None, :] * np.cos(2 * np.pi * self.lf[None, :] * t[:, None] + self.lp[None, :])).sum(1) * self.g
        r = (self.ra[None, :] * np.cos(2 * np.pi * self.lf[None, :] * t[:, None] + self.rp[None, :])).sum(1) * self.g
        mono, diff = 0.5 * (l + r), 0.5 * (l - r)
        bi = np.minimum((t * RDS_BAUD).astype(np.int64), len(self.sym) - 1)
        comp = (mono + diff * np.sin(2 * np.pi * 38000 * t) + 0.09 * np.sin(2 * np.pi * 19000 * t)
                + 0.04 * self.sym[bi] * np.sin(2 * np.pi * 57000 * t))
        inst = self.f0 + 75000.0 * comp
        ph = self.phase + np.cumsum(2 * np.pi * inst / self.fs)
        self.phase = float(np.mod(ph[-1], 2 * np.pi)) if m else self.phase
        self.n += m
        return (self.amp * np.exp(1j * ph)).astype(np.complex64)


class StereoScene:
    """Wideband span of stereo WFM stations with RDS (the 20-station
    multiplex shape), generated block by block.  Each station's composite is
    synthesised at a low rate (WFMStation at `low_rate`, whose 200 kHz
    occupied band it represents), its phase is interpolated to the span rate
    and the carrier offset added; AWGN sets each station's in-band SNR
    (synth.go:84-90 definition over 200 kHz).  Returns complex64 blocks."""

    def __init__(self, fs: float, offsets, snr_db, stations_rds, seed: int = 0, sigma: float = 0.01,
                 low_rate: float = 500_000.0):
        self.fs, self.low = fs, low_rate
        self.r = int(round(fs / low_rate))
        assert self.r * low_rate == fs
        self.offsets = list(offsets)
        self.sigma = sigma
        self.amps = [np.sqrt(carrier_power(s, sigma, 200e3, fs)) for s in snr_db]
        self.st = [WFMStation(low_rate, 0.0, rds_symbols(bits), seed=seed + 1 + k)
                   for k, bits in enumerate(stations_rds)]
        self.last_ph = [0.0] * len(self.st)
        self.n = 0
        self.rng = np.random.default_rng(seed)

    def block(self, m: int) -> np.ndarray:
        assert m % self.r == 0
        ml = m // self.r
        t = (np.arange(m) + self.n) / self.fs
        x = awgn(m, self.sigma, self.rng).astype(np.complex64)
        xi = np.arange(m, dtype=np.float64) / self.r + (1.0 - 1.0 / self.r)
        lows = [st.block(ml) for st in self.st]

        def station(k):
            ph = np.unwrap(np.concatenate([[self.last_ph[k]], np.angle(lows[k])]))
            # low-rate sample j sits at span index (j + 1) * r - 1 of this block
            p = np.interp(xi, np.arange(ml + 1), ph)
            p += 2 * np.pi * self.offsets[k] * t
            return ph[-1], (self.amps[k] * np.exp(1j * p)).astype(np.complex64)

        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=min(8, len(self.st) or 1)) as ex:
            parts = list(ex.map(station, range(len(self.st))))
        for k, (last, v) in enumerate(parts):
            self.last_ph[k] = last
            x += v
        self.n += m
        return x

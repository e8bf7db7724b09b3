"""Pins the WFM stereo / RDS oracle: the RDS checkword algebra (IEC 62106),
encode -> decode round trips, the reference's stereo-lock and separation
behaviour (stereo_lock_test.go, streamer_test.go), the RDS baseband
scratch == allocating identity (fm_rds_test.go) and the full station chain."""
import numpy as np
import pytest

import oracle as O

OFFS = {"A": 0x0FC, "B": 0x198, "C": 0x168, "Cp": 0x350, "D": 0x1B4}


def test_syndrome_of_encoded_blocks_is_the_offset():
    rng = np.random.default_rng(0)
    for data in list(rng.integers(0, 1 << 16, 50)) + [0, 0xFFFF, 0x1234]:
        for off in OFFS.values():
            blk = O.RDSBlock(int(data), off)
            assert O.crcSyndrome(blk) == off
            assert blk >> 10 == data
            # any single-bit error leaves a syndrome that is not this offset
            for b in range(26):
                assert O.crcSyndrome(blk ^ (1 << b)) != off


def test_syndrome_is_division_by_the_rds_generator():
    # g(x) = x^10 + x^8 + x^7 + x^5 + x^4 + x^3 + 1 (0x5B9): a brute-force
    # GF(2) remainder must match the reference's reduction (rds.go:590-599).
    def rem(v):
        for i in range(25, 9, -1):
            if (v >> i) & 1:
                v ^= 0x5B9 << (i - 10)
        return v & 0x3FF
    rng = np.random.default_rng(1)
    for raw in rng.integers(0, 1 << 26, 200):
        assert O.crcSyndrome(int(raw)) == rem(int(raw))


def _rds_baseband_from_bits(bits, fs_in=250_000, phase0=0.7, noise=0.02, seed=0):
    """NRZ symbols at fs_in through the chain's 7.5 kHz low-pass and /13
    decimation (fm.go:253-269), so the pulses have the shape the decoder's
    timing loop expects."""
    from argus.synth import rds_symbols, RDS_BAUD
    sym = rds_symbols(bits)
    n = int(len(sym) * fs_in / RDS_BAUD)
    t = np.arange(n) / fs_in
    bi = np.minimum((t * RDS_BAUD).astype(int), len(sym) - 1)
    rng = np.random.default_rng(seed)
    x = sym[bi] * np.exp(1j * phase0) + noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    y = O.ApplyFIR(x.astype(np.complex64), O.LowpassFIR(7500, fs_in, 101))
    d = fs_in // 19000
    return O.Decimate(y, d), fs_in // d


@pytest.mark.parametrize("chunk", [None, 4000, 20000])
def test_rds_encode_decode_round_trip(chunk):
    # The decoder finds groups inside one call's bits only (rds.go:365-430):
    # groups straddling call boundaries are lost, so chunked calls keep fewer.
    bits = O.RDSEncodeGroups(0xC0DE, "ARGUSFM1", "MI355X native extraction path - RDS test text 0123456789", 64)
    x, fs = _rds_baseband_from_bits(bits)
    d = O.RDSDecoder()
    if chunk is None:
        d.Decode(x, fs)
    else:
        for i in range(0, len(x), chunk):
            d.Decode(x[i:i + chunk], fs)
    assert d.PI == 0xC0DE
    assert d.PS == "ARGUSFM1"
    assert d.RT.startswith("MI355X native extraction path - RDS test text")
    groups = d.Groups()
    assert len(groups) >= (56 if chunk is None else 32)
    # every decoded group is one that was sent, in order
    sent = [tuple(int(bits[g * 104 + k * 26:(g * 104 + (k + 1) * 26)].dot(1 << np.arange(25, -1, -1))) >> 10
                  for k in range(4)) for g in range(64)]
    pos = 0
    for g in groups:
        while pos < len(sent) and sent[pos] != g:
            pos += 1
        assert pos < len(sent), g
        pos += 1


def test_rds_baseband_scratch_equals_allocating():
    # fm_rds_test.go:10-35: byte-identical for fs in {200k, 250k, 240k}
    rng = np.random.default_rng(2)
    for fs in [200_000, 250_000, 240_000]:
        for n in [0, 1, 2, 17, 999, 4096]:
            iq = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
            a, ra = O.RDSBasebandComplex(iq, fs)
            b, rb = O.RDSBasebandComplex(iq, fs)
            assert ra == rb and np.array_equal(a, b)
            if n >= 2:
                assert ra == fs // (fs // 19000)
                assert len(a) == len(range(0, n - 1, fs // 19000))


def test_rds_stream_matches_per_call_on_one_block():
    """Streaming form == reference per-call form when fed one whole block.

    The streaming mixer takes its phase from the exact sample count where the
    per-call form adds the increment in FP64, so the two differ by rounding
    only (about 1e-12 rad over 5000 samples)."""
    rng = np.random.default_rng(3)
    fs = 250_000
    iq = (rng.standard_normal(5000) + 1j * rng.standard_normal(5000)).astype(np.complex64)
    want, rate = O.RDSBasebandComplex(iq, fs)
    s = O.RDSBasebandStream(fs)
    got = s.Process(O.fmDiscrim(iq))
    assert s.rate == rate and got.shape == want.shape
    scale = float(np.max(np.abs(want)))
    assert float(np.max(np.abs(got - want))) <= 1e-6 * scale


def test_rds_stream_mixer_phase_does_not_drift():
    """A 57 kHz discriminator tone lands at DC with phase (omega - theta),
    also after 10 million samples: the carried phase is exact."""
    fs, w, theta, a = 250_000, 2 * np.pi * 57000 / 250_000, 0.7, 0.5
    s = O.RDSBasebandStream(fs)
    n0 = 10_000_000
    # skip n0 samples of silence: the mixer phase advances, the FIR forgets
    z = np.zeros(1_000_000, np.float32)
    for _ in range(n0 // len(z)):
        s.Process(z)
    i = np.arange(n0, n0 + 20000)
    disc = (a * np.cos(w * i + theta)).astype(np.float32)
    y = s.Process(disc)[20:]  # past the FIR transient (101 taps / 13)
    g = float(np.sum(O.LowpassFIR(7500, fs, 101)))
    want = 0.5 * a * g * np.exp(1j * (w - theta))
    # the 114 kHz image is suppressed by the Hann-windowed FIR, not removed
    assert np.max(np.abs(y - want)) <= 1e-2 * abs(want)


def _station_mono(fs_iq, seconds, seed=0, rds_bits=None, block=5000):
    from argus.synth import WFMStation, rds_symbols
    bits = rds_bits if rds_bits is not None else np.zeros(2000, np.uint8)
    st = WFMStation(fs_iq, 0.0, rds_symbols(bits), seed=seed)
    fm = O.BridgedFMDiscrim()
    n = int(seconds * fs_iq)
    out = []
    for i in range(0, n, block):
        out.append(fm.Demod(st.block(min(block, n - i))))
    return np.concatenate(out)


def _gen_mpx(n, sr, audio_amp, pilot_amp, t0):
    # stereo_lock_test.go:13-25 genMPX
    t = t0 + np.arange(n) / sr
    x = (audio_amp * np.sin(2 * np.pi * 1000 * t) + pilot_amp * np.sin(2 * np.pi * 19000 * t)
         + 0.3 * audio_amp * np.sin(2 * np.pi * 2000 * t) * np.sin(2 * np.pi * 38000 * t))
    return x.astype(np.float32), t0 + n / sr


def test_stereo_lock_holds_through_loud_audio():
    # stereo_lock_test.go:31-62
    sr, blk = 256_000, 4096
    amps = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.9, 0.15, 0.95, 0.1, 0.9, 0.95, 0.12, 0.9, 0.95, 0.9, 0.1, 0.95]
    dec = O.StereoDecoder(sr)
    t0, loud, loud_locked = 0.0, 0, 0
    for b, amp in enumerate(amps):
        mpx, t0 = _gen_mpx(blk, sr, amp, 0.08, t0)
        _, locked = dec.Decode(mpx)
        if b >= 6 and amp >= 0.8:
            loud += 1
            loud_locked += locked
    assert loud > 0 and loud_locked == loud


def test_stereo_no_false_lock_on_noise():
    # stereo_lock_test.go:66-84 (numpy's random stream, not Go's)
    sr, blk = 256_000, 4096
    for seed in (1, 2, 3):
        rng = np.random.default_rng(seed)
        dec = O.StereoDecoder(sr)
        false = sum(dec.Decode((rng.standard_normal(blk) * 0.3).astype(np.float32))[1] for _ in range(15))
        assert false == 0, (seed, false)


def test_stereo_locks_on_a_synthesised_station():
    fs = 250_000
    mono = _station_mono(fs, 0.6, seed=4)
    dec = O.StereoDecoder(fs)
    locks = [dec.Decode(mono[i:i + 5000])[1] for i in range(0, len(mono), 5000)]
    assert dec.Enabled and sum(locks[5:]) == len(locks[5:])


def test_stereo_separates_left_and_right():
    # streamer_test.go:8-67: L-only program -> |L| dominates |R| by >= 1.3x
    fs = 250_000
    n = int(0.4 * fs)
    t = np.arange(n) / fs
    left = 0.5 * np.sin(2 * np.pi * 1000 * t)
    mpx = 0.5 * left + 0.5 * left * np.sin(2 * np.pi * 38000 * t) + 0.09 * np.sin(2 * np.pi * 19000 * t)
    dec = O.StereoDecoder(fs)
    outs = [dec.Decode(mpx[i:i + 5000].astype(np.float32))[0] for i in range(0, n, 5000)]
    lr = np.concatenate(outs).reshape(-1, 2)[n // 2:]
    L, R = np.std(lr[:, 0]), np.std(lr[:, 1])
    assert max(L, R) >= 1.3 * min(L, R) and max(L, R) >= 0.05


def test_station_chain_recovers_rds():
    fs = 250_000
    bits = O.RDSEncodeGroups(0x2F1A, "STATION7", "RT text of station seven", 40)
    mono = _station_mono(fs, 3.6, seed=7, rds_bits=bits)
    s = O.RDSBasebandStream(fs)
    d = O.RDSDecoder()
    win = []
    for i in range(0, len(mono), 5000):
        win.append(s.Process(mono[i:i + 5000]))
        if sum(len(w) for w in win) >= s.rate:  # decode 1 s windows
            d.Decode(np.concatenate(win), s.rate)
            win = []
    assert d.PI == 0x2F1A and d.PS == "STATION7"


def test_stereo_pilot_present_on_station_absent_on_mono():
    """classifier/pll.go:82-131: a stereo station's discriminator shows the
    19 kHz pilot (max/median of the 201 Goertzel bins far above 4); a mono FM
    carrier (1 kHz tone, 75 kHz deviation) and noise do not."""
    from argus.synth import WFMStation, rds_symbols
    fs = 250_000
    st = WFMStation(fs, 0.0, rds_symbols(np.zeros(800, np.uint8)), seed=3)
    ok, ratio = O.StereoPilotPresent(st.block(fs // 2), fs)
    assert ok and ratio > 20, ratio
    t = np.arange(fs // 2) / fs
    mono = np.exp(1j * (75000 / 1000) * np.sin(2 * np.pi * 1000 * t)).astype(np.complex64)
    ok, ratio = O.StereoPilotPresent(mono, fs)
    assert not ok, ratio
    rng = np.random.default_rng(4)
    noise = (rng.standard_normal(fs // 2) + 1j * rng.standard_normal(fs // 2)).astype(np.complex64)
    assert not O.StereoPilotPresent(noise, fs)[0]


def test_stereo_pilot_uses_the_last_half_second():
    """Only the last 0.5 s of IQ is examined (pll.go:90-96): a pilot in the
    first second followed by 0.5 s of mono is not detected."""
    from argus.synth import WFMStation, rds_symbols
    fs = 250_000
    st = WFMStation(fs, 0.0, rds_symbols(np.zeros(2000, np.uint8)), seed=5)
    t = np.arange(fs // 2) / fs
    mono = np.exp(1j * 75 * np.sin(2 * np.pi * 1000 * t)).astype(np.complex64)
    assert O.StereoPilotPresent(st.block(fs), fs)[0]
    assert not O.StereoPilotPresent(np.concatenate([st.block(fs), mono]), fs)[0]


def test_pilot_goertzel_bins_are_dft_magnitudes():
    """The 201 Goertzel magnitudes (pll.go:178-193) equal |sum d[i] e^{-j 2pi f i / fs}|
    of the FP64 discriminator, computed here as a direct DFT; the reported
    offset is the bin of a pilot placed 300 Hz above 19 kHz."""
    fs = 200_000
    n = fs // 2
    t = np.arange(n) / fs
    rng = np.random.default_rng(12)
    mpx = 0.3 * np.sin(2 * np.pi * 1200 * t) + 0.08 * np.sin(2 * np.pi * 19300 * t) + 0.01 * rng.standard_normal(n)
    iq = np.exp(1j * 2 * np.pi * 75000 * np.cumsum(mpx) / fs).astype(np.complex64)
    ok, ratio, off, mags = O.StereoPilotEstimate(iq, fs)
    x = iq.astype(np.complex128)
    d = np.angle(np.conj(x[:-1]) * x[1:])
    freqs = 19000.0 + 10.0 * np.arange(-100, 101)
    ph = np.exp(-2j * np.pi * np.outer(freqs, np.arange(n - 1)) / fs)
    want = np.abs(ph @ d)
    np.testing.assert_allclose(mags, want, rtol=1e-7, atol=1e-9 * want.max())
    assert ok and off == 300.0
    assert ratio == pytest.approx(want.max() / np.sort(want)[100], rel=1e-7)


def test_batch_stereo_decoder_locks_and_separates():
    """wfm_batch_stereo.go:76-157: a station's MPX locks in one call (pilot
    ratio > 0.003) and an L-only program comes out on one channel (the
    separation criterion of streamer_test.go:8-67); noise does not lock."""
    fs = 250_000
    n = int(0.4 * fs)
    t = np.arange(n) / fs
    left = 0.5 * np.sin(2 * np.pi * 1000 * t)
    mpx = (0.5 * left + 0.5 * left * np.sin(2 * np.pi * 38000 * t) + 0.09 * np.sin(2 * np.pi * 19000 * t)).astype(np.float32)
    out, locked = O.StereoDecodeBatch(mpx, fs)
    assert locked
    lr = out.reshape(-1, 2)[n // 2:]
    L, R = np.std(lr[:, 0]), np.std(lr[:, 1])
    assert max(L, R) >= 1.3 * min(L, R) and max(L, R) >= 0.05
    rng = np.random.default_rng(8)
    _, locked = O.StereoDecodeBatch((rng.standard_normal(n) * 0.3).astype(np.float32), fs)
    assert not locked


def test_batch_stereo_equals_stateful_decoder_on_first_block():
    """Both decoders start from the same zero state: on one block their L/R
    outputs agree exactly (only the lock rule differs)."""
    fs = 250_000
    mono = _station_mono(fs, 0.1, seed=9)
    a, _ = O.StereoDecodeBatch(mono, fs)
    b, _ = O.StereoDecoder(fs).Decode(mono)
    np.testing.assert_array_equal(a, b)

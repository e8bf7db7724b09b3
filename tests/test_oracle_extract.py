"""Pins the CPU oracle of the streaming extractor.

Known answers and invariants are the reference's own tests
(internal/demod/gpudemod/*_test.go); the FP64 brute force and the FNV-1a
published vectors pin what those leave open.
"""
import numpy as np
import pytest

import oracle as O
from helpers import rel_rms


def test_polyphase_known_answer():
    # polyphase_test.go:5-22
    got = O.BuildPolyphaseTapsPhaseMajor(np.arange(1, 8, dtype=np.float32), 3)
    np.testing.assert_array_equal(got, [1, 4, 7, 2, 5, 0, 3, 6, 0])


def test_exact_integer_decimation():
    # cpu_oracle_test.go:60-67
    assert O.ExactIntegerDecimation(4_000_000, 200_000) == 20
    with pytest.raises(ValueError):
        O.ExactIntegerDecimation(4_000_000, 192_000)
    # the snapped rates of the 20 Msps workloads
    assert O.ExactIntegerDecimation(20_000_000, 250_000) == 80
    with pytest.raises(ValueError):
        O.ExactIntegerDecimation(20_000_000, 256_000)


def _mk(numtaps=65, decim=20):
    taps = O.MakeLowpassTaps(numtaps)
    return O.CPUOracleState(SignalID=1, ConfigHash=123, NCOPhase=0, Decim=decim, PhaseCount=0, NumTaps=numtaps,
                            BaseTaps=taps, PolyphaseTaps=O.BuildPolyphaseTapsPhaseMajor(taps, decim))


def test_monolithic_vs_chunked():
    # cpu_oracle_test.go:39-58
    iq = O.MakeDeterministicIQ(200000)
    mono = O.CPUOracleExtract(iq, _mk(), 0.017)
    st = _mk()
    parts, pos = [], 0
    for n in [4096, 5000, 8192, 27307]:
        parts.append(O.CPUOracleExtract(iq[pos:pos + n], st, 0.017))
        pos += n
    parts.append(O.CPUOracleExtract(iq[pos:], st, 0.017))
    chunked = np.concatenate(parts)
    assert len(mono) == len(chunked) == 200000 // 20
    assert np.max(np.abs(mono - chunked)) <= 1e-5


def test_direct_vs_polyphase():
    # cpu_oracle_test.go:69-89
    iq = O.MakeDeterministicIQ(50000)
    d = O.CPUOracleExtract(iq, _mk(), 0.017)
    p = O.CPUOracleExtractPolyphase(iq, _mk(), 0.017)
    assert np.max(np.abs(d - p)) <= 1e-5


def test_host_core_matches_polyphase_oracle():
    # streaming_host_core_test.go:5-40
    iq = O.MakeDeterministicIQ(12000)
    st = _mk()
    want = O.CPUOracleExtractPolyphase(iq, st, 0.017)
    st2 = _mk()
    out, phase, pc, hist = O.runStreamingPolyphaseHostCore(iq, 4_000_000, -0.017 * 4_000_000 / (2 * np.pi),
                                                           st2.NCOPhase, st2.PhaseCount, st2.NumTaps, st2.Decim,
                                                           st2.ShiftedHistory, st2.PolyphaseTaps)
    assert np.max(np.abs(want - out)) <= 1e-5
    assert phase != 0 and 0 <= pc < 20 and len(hist) == 64


def _brute_force(iq, sr, off, taps, decim, phase0=0.0):
    """FP64 shift + causal FIR + emit at D-1, 2D-1, ... (independent of the oracle)."""
    n = np.arange(len(iq))
    inc = -2 * np.pi * off / sr
    s = iq.astype(np.complex128) * np.exp(1j * (phase0 + inc * n))
    y = np.convolve(s, np.asarray(taps, np.float64))[:len(iq)]
    return y[decim - 1::decim]


@pytest.mark.parametrize("decim,ntaps", [(8, 127), (20, 65), (80, 301), (400, 63), (1600, 511), (1, 5)])
def test_host_core_vs_fp64_brute_force(decim, ntaps):
    rng = np.random.default_rng(decim * 1000 + ntaps)
    iq = (rng.standard_normal(20000) + 1j * rng.standard_normal(20000)).astype(np.complex64)
    sr, off = 20_000_000, 1_234_567.0
    base = O.LowpassFIR(max(100000 / 2, 200), sr, ntaps).astype(np.float32)
    poly = O.BuildPolyphaseTapsPhaseMajor(base, decim)
    outs, ph, pc, hist = [], 0.0, 0, np.zeros(0, np.complex64)
    pos = 0
    for n in [0, 1, 7, ntaps - 1, ntaps, ntaps + 1, 3333, 20000]:
        o, ph, pc, hist = O.runStreamingPolyphaseHostCore(iq[pos:pos + n], sr, off, ph, pc, ntaps, decim, hist, poly)
        outs.append(o)
        pos = min(pos + n, len(iq))
    got = np.concatenate(outs)
    want = _brute_force(iq[:pos], sr, off, base, decim)
    assert len(got) == len(want)
    assert rel_rms(got, want) < 2e-6
    # a wrong NCO sign or emission offset fails by orders of magnitude
    assert rel_rms(got, _brute_force(iq[:pos], sr, -off, base, decim)) > 1e-2 or decim == 1
    if decim > 1:
        wrong_phase = np.convolve(iq[:pos].astype(np.complex128) * np.exp(-2j * np.pi * off / sr * np.arange(pos)),
                                  base.astype(np.float64))[:pos][0::decim][:len(want)]
        assert rel_rms(got, wrong_phase) > 1e-3


def test_config_hash_is_fnv1a64():
    def fnv(b: bytes) -> int:
        h = 0xcbf29ce484222325
        for c in b:
            h ^= c
            h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
        return h
    # published FNV-1a 64 vectors
    assert fnv(b"") == 0xcbf29ce484222325
    assert fnv(b"a") == 0xaf63dc4c8601ec8c
    assert fnv(b"foobar") == 0x85944171f73967e8
    # streaming_types.go:70: "sig=%d|out=%d|taps=%d|sr=%d"; offset/bw ignored
    want = fnv(b"sig=7|out=200000|taps=65|sr=4000000")
    assert O.StreamingConfigHash(7, 12500.0, 20000.0, 200000, 65, 4000000) == want
    assert O.StreamingConfigHash(7, -1.0, 1.0, 200000, 65, 4000000) == want
    assert O.StreamingConfigHash(-3, 0, 0, 1, 2, 3) == fnv(b"sig=-3|out=1|taps=2|sr=3")


def test_lowpass_fir_closed_form():
    # dsp/fir.go:6-28
    for taps, cutoff, sr in [(101, 10000, 4_000_000), (64, 7500, 250_000), (127, 100000, 2_048_000)]:
        h = O.LowpassFIR(cutoff, sr, taps)
        t = taps + (taps % 2 == 0)
        assert len(h) == t
        np.testing.assert_allclose(h, h[::-1], rtol=0, atol=1e-15)
        fc = cutoff / sr
        m = (t - 1) / 2
        x = np.arange(t) - m
        with np.errstate(invalid="ignore", divide="ignore"):
            sinc = np.where(x == 0, 2 * fc, np.sin(2 * np.pi * fc * x) / (np.pi * x))
        w = 0.5 * (1 - np.cos(2 * np.pi * np.arange(t) / (t - 1)))
        np.testing.assert_allclose(h, sinc * w, rtol=1e-12, atol=1e-16)
        if fc * t > 5:  # enough taps for the windowed sinc to reach unit DC gain
            assert abs(h.sum() - 1.0) < 0.05
    assert np.all(O.LowpassFIR(0, 1000, 11) == 0)
    hm = O.LowpassFIRHamming(100000, 2_048_000, 127)
    assert len(hm) == 127 and abs(hm[63] - 2 * 100000 / 2_048_000) < 1e-15


def test_runner_state_init():
    # stream_state_test.go:5-31: fs 4 MHz, out 200 k -> D 20, 65 taps, history cap >= 64
    r = O.CPUOracleRunner(4_000_000)
    job = type("J", (), dict(SignalID=7, OffsetHz=12500, Bandwidth=20000, OutRate=200000, NumTaps=65,
                             ConfigHash=555))
    res = r.StreamingExtract(O.MakeDeterministicIQ(1000), [job])
    st = r.States[7]
    assert st.Decim == 20 and st.NumTaps == 65 and len(st.ShiftedHistory) == 64
    assert res[0].NOut == 50 and res[0].PhaseCount == 0 and res[0].HistoryLen == 64


def test_runner_lifecycle_and_threads_agree():
    # streaming_gpu_exec_test.go:74-121 step list; the threaded runner must
    # give the single-threaded answer exactly.
    class J:
        def __init__(self, **kw):
            self.__dict__.update(kw)
    A = J(SignalID=11, OffsetHz=12500, Bandwidth=20000, OutRate=200000, NumTaps=65, ConfigHash=1001)
    B = J(SignalID=22, OffsetHz=-18750, Bandwidth=16000, OutRate=100000, NumTaps=33, ConfigHash=2002)
    A2 = J(**{**A.__dict__, "ConfigHash": 1002})
    steps = [(O.MakeDeterministicIQ(512), [A, B]), (None, [A2, B]), (O.makeToneNoiseIQ(96, 0.041), [A2]),
             (O.MakeDeterministicIQ(160), [A2, B]), (O.makeToneNoiseIQ(65, 0.017), [A2, B])]
    r1, r4 = O.CPUOracleRunner(4_000_000, 1), O.CPUOracleRunner(4_000_000, 4)
    for iq, jobs in steps:
        a = r1.StreamingExtract(iq, jobs)
        b = r4.StreamingExtract(iq, jobs)
        for x, y in zip(a, b):
            assert x.NOut == y.NOut and np.array_equal(x.IQ, y.IQ)
        assert set(r1.States) == {j.SignalID for j in jobs}
    # config reset with zero new samples zeroes phase, count and history
    r = O.CPUOracleRunner(4_000_000)
    r.StreamingExtract(O.MakeDeterministicIQ(512), [A])
    r.StreamingExtract(None, [A2])
    st = r.States[11]
    assert st.NCOPhase == 0 and st.PhaseCount == 0 and len(st.ShiftedHistory) == 0

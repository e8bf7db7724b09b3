"""GPU parity of the streaming extractor against the CPU oracle.

Sequences follow the reference's own GPU validation tests
(streaming_gpu_exec_test.go:33-121, streaming_gpu_native_prepare_test.go:
47-206): chunk patterns, config reset with zero new samples, disappearance,
reappearance, capacity growth and the history boundary.  Tolerances are the
reference's native ones: samples 1e-4 absolute, NCO phase 1e-8 rad; counts
and state keys exact.  Every call goes through libargus_hip.so.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import phase_diff, rel_rms

pytestmark = pytest.mark.gpu

SAMPLE_TOL = 1e-4
PHASE_TOL = 1e-8


class J:
    def __init__(self, **kw):
        self.__dict__.update(kw)


JOB = dict(SignalID=1, OffsetHz=12500, Bandwidth=20000, OutRate=200000, NumTaps=65, ConfigHash=777)


def steps_from_chunks(iq, chunks, jobs):
    steps, pos = [], 0
    for n in chunks:
        end = min(pos + max(n, 0), len(iq))
        steps.append((iq[pos:end], jobs))
        pos = end
    if pos < len(iq):
        steps.append((iq[pos:], jobs))
    return steps


def run_sequence(gpu, mode, steps, sr=4_000_000):
    r = gpu.NewBatchRunner(32768, sr, exec_mode=mode)
    orc = O.CPUOracleRunner(sr)
    try:
        for idx, (iq, jobs) in enumerate(steps):
            print(f"step {idx} mode={mode} n={0 if iq is None else len(iq)} jobs={[j.SignalID for j in jobs]}",
                  flush=True)
            got = r.StreamingExtractGPU(iq, jobs)
            want = orc.StreamingExtract(iq, jobs)
            assert len(got) == len(want)
            ostates = orc.States
            for g, w, job in zip(got, want, jobs):
                assert g.SignalID == w.SignalID and g.Rate == w.Rate
                assert g.NOut == w.NOut, (idx, g.NOut, w.NOut)
                assert g.PhaseCount == w.PhaseCount
                assert g.HistoryLen == w.HistoryLen
                if w.NOut:
                    assert np.max(np.abs(g.IQ - w.IQ)) <= SAMPLE_TOL, (idx, np.max(np.abs(g.IQ - w.IQ)))
                gs, os_ = r.SignalState(job.SignalID), ostates[job.SignalID]
                assert gs.ConfigHash == os_.ConfigHash and gs.Decim == os_.Decim and gs.NumTaps == os_.NumTaps
                assert gs.PhaseCount == os_.PhaseCount
                assert phase_diff(gs.NCOPhase, os_.NCOPhase) <= PHASE_TOL
                assert len(gs.ShiftedHistory) == len(os_.ShiftedHistory)
                if len(os_.ShiftedHistory):
                    assert np.max(np.abs(gs.ShiftedHistory - os_.ShiftedHistory)) <= SAMPLE_TOL
            assert sorted(r.StateIDs()) == sorted(ostates)
    finally:
        r.Close()


@pytest.mark.parametrize("mode", [0, 1], ids=["batched", "per_signal"])
def test_chunk_patterns_deterministic(gpu, mode):
    steps = steps_from_chunks(O.MakeDeterministicIQ(8192), [0, 1, 2, 17, 63, 64, 65, 129, 511, 2048], [J(**JOB)])
    run_sequence(gpu, mode, steps)


@pytest.mark.parametrize("mode", [0, 1], ids=["batched", "per_signal"])
def test_chunk_patterns_tone_noise(gpu, mode):
    steps = steps_from_chunks(O.makeToneNoiseIQ(12288, 0.023), [7, 20, 3, 63, 64, 65, 777, 2048, 4096], [J(**JOB)])
    run_sequence(gpu, mode, steps)


@pytest.mark.parametrize("mode", [0, 1], ids=["batched", "per_signal"])
def test_lifecycle_reset_and_capacity(gpu, mode):
    A = J(SignalID=11, OffsetHz=12500, Bandwidth=20000, OutRate=200000, NumTaps=65, ConfigHash=3001)
    B = J(SignalID=22, OffsetHz=-18750, Bandwidth=16000, OutRate=100000, NumTaps=33, ConfigHash=4002)
    A2 = J(**{**A.__dict__, "ConfigHash": 3002})
    steps = [
        (O.MakeDeterministicIQ(256), [A, B]),          # prime both
        (O.makeToneNoiseIQ(4096, 0.037), [A, B]),      # grow capacity
        (None, [A2, B]),                               # config reset, zero new samples
        (O.MakeDeterministicIQ(64), [A2]),             # B disappears
        (O.makeToneNoiseIQ(96, 0.017), [A2, B]),       # B reappears fresh
        (O.MakeDeterministicIQ(65), [A2, B]),          # history boundary
    ]
    run_sequence(gpu, mode, steps)


def test_uneven_mix_long_stream(gpu):
    """The batched uneven extraction shape: 20 Msps, 64 concurrent signals of
    the four modes with decimations 8..1600 and taps 63..511, 65 536-sample
    chunks (ragged), one fused launch per chunk."""
    sr = 20_000_000
    rng = np.random.default_rng(0)
    modes = [(200_000, [8, 16, 40, 80]), (12_500, [400, 800]), (10_000, [500, 1000]), (3_000, [1250, 1600])]
    jobs = []
    for i in range(64):
        bw, ds = modes[i % 4]
        d = int(rng.choice(ds))
        t = int(rng.integers(31, 256)) * 2 + 1
        jobs.append(J(SignalID=100 + i, OffsetHz=float(rng.uniform(-9e6, 9e6)), Bandwidth=bw, OutRate=sr // d,
                      NumTaps=t, ConfigHash=1000 + i))
    r = gpu.NewBatchRunner(65536, sr)
    orc = O.CPUOracleRunner(sr, threads=8)
    n = 65536
    acc_g = {j.SignalID: [] for j in jobs}
    acc_w = {j.SignalID: [] for j in jobs}
    for step in range(6):
        iq = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        k = n - 1000 * step  # ragged chunk lengths
        print(f"mix step {step} n={k}", flush=True)
        got = r.StreamingExtractGPU(iq[:k], jobs)
        want = orc.StreamingExtract(iq[:k], jobs)
        for g, w in zip(got, want):
            assert g.NOut == w.NOut and g.PhaseCount == w.PhaseCount and g.HistoryLen == w.HistoryLen
            acc_g[g.SignalID].append(g.IQ)
            acc_w[w.SignalID].append(w.IQ)
    for sid in acc_g:
        e = rel_rms(np.concatenate(acc_g[sid]), np.concatenate(acc_w[sid]))
        assert e <= 1e-4, (sid, e)
    r.Close()


def test_tile_geometries(gpu):
    """Channels that take each tile shape: register-blocked dense (D=2, 80),
    many-plane dense with one output per thread (D=400 < T=511), the narrow
    fallback for a dense tile that does not fit the LDS budget (D=3000 <
    T=3301), no decimation (D=1) and narrow windows (D >= T)."""
    sr = 12_000_000
    shapes = [(2, 63), (80, 301), (400, 511), (3000, 3301), (1, 101), (6000, 1001), (8, 9)]
    jobs = [J(SignalID=200 + i, OffsetHz=-1.1e6 + 0.4e6 * i, Bandwidth=sr / d / 2, OutRate=sr // d, NumTaps=t,
              ConfigHash=5000 + i) for i, (d, t) in enumerate(shapes)]
    r = gpu.NewBatchRunner(70000, sr)
    orc = O.CPUOracleRunner(sr, threads=8)
    rng = np.random.default_rng(3)
    acc_g = {j.SignalID: [] for j in jobs}
    acc_w = {j.SignalID: [] for j in jobs}
    for k, n in enumerate([5000, 65536, 12345, 1]):
        iq = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        print(f"geometry step {k} n={n}", flush=True)
        got = r.StreamingExtractGPU(iq, jobs)
        want = orc.StreamingExtract(iq, jobs)
        for g, w in zip(got, want):
            assert g.NOut == w.NOut and g.PhaseCount == w.PhaseCount and g.HistoryLen == w.HistoryLen
            acc_g[g.SignalID].append(g.IQ)
            acc_w[w.SignalID].append(w.IQ)
    for (d, t), sid in zip(shapes, acc_g):
        e = rel_rms(np.concatenate(acc_g[sid]), np.concatenate(acc_w[sid]))
        assert e <= 1e-4, (d, t, e)
    r.Close()


def test_stateful_entry_point_direct(gpu):
    """gpud_launch_streaming_polyphase_stateful_hip against the host core,
    including its argument-error contract (exports.cu:552-606)."""
    from argus._lib import lib
    L = lib()
    sr, off, T, Dd = 4_000_000, 12500.0, 65, 20
    base = O.LowpassFIR(10000, sr, T).astype(np.float32)
    poly = O.BuildPolyphaseTapsPhaseMajor(base, Dd)
    iq = O.makeToneNoiseIQ(5000, 0.023)
    nb = 5000 * 8
    ptrs = {}
    for name, size in [("in", nb), ("tmp", nb), ("taps", poly.nbytes), ("hist", 8 * 64), ("scr", 8 * 64),
                       ("out", nb)]:
        p = C.c_void_p()
        assert L.gpud_malloc(C.byref(p), size) == 0
        ptrs[name] = p.value
    try:
        assert L.gpud_memcpy_h2d(ptrs["taps"], poly.ctypes.data, poly.nbytes) == 0
        hl, pc, ph, nout = C.c_int(0), C.c_int(0), C.c_double(0.0), C.c_int(0)
        inc = -2 * np.pi * off / sr
        oh, oph, opc = np.zeros(0, np.complex64), 0.0, 0
        pos = 0
        for n in [0, 1, 64, 65, 1000, 3870]:
            chunk = np.ascontiguousarray(iq[pos:pos + n])
            pos += n
            if n:
                assert L.gpud_memcpy_h2d(ptrs["in"], chunk.ctypes.data, chunk.nbytes) == 0
            rc = L.gpud_launch_streaming_polyphase_stateful_hip(
                ptrs["in"], n, ptrs["tmp"], ptrs["taps"], len(poly), Dd, T, ptrs["hist"], ptrs["scr"], 64,
                C.byref(hl), C.byref(pc), C.byref(ph), inc, ptrs["out"], n // Dd + 2, C.byref(nout))
            assert rc == 0
            want, oph, opc, oh = O.runStreamingPolyphaseHostCore(chunk, sr, off, oph, opc, T, Dd, oh, poly)
            got = np.zeros(nout.value, np.complex64)
            if nout.value:
                assert L.gpud_memcpy_d2h(got.ctypes.data, ptrs["out"], got.nbytes) == 0
            assert nout.value == len(want) and pc.value == opc and hl.value == len(oh)
            assert phase_diff(ph.value, oph) <= PHASE_TOL
            if len(want):
                assert np.max(np.abs(got - want)) <= SAMPLE_TOL
            hist = np.zeros(hl.value, np.complex64)
            if hl.value:
                assert L.gpud_memcpy_d2h(hist.ctypes.data, ptrs["hist"], hist.nbytes) == 0
                assert np.max(np.abs(hist - oh)) <= SAMPLE_TOL
        # error contract (no launch happens on these paths)
        args = lambda **kw: [kw.get("in_", ptrs["in"]), kw.get("n", 10), kw.get("tmp", ptrs["tmp"]),
                             kw.get("taps", ptrs["taps"]), kw.get("plen", len(poly)), Dd, T, ptrs["hist"],
                             ptrs["scr"], 64, C.byref(hl), C.byref(pc), C.byref(ph), inc,
                             kw.get("out", ptrs["out"]), kw.get("cap", 100), C.byref(nout)]
        f = L.gpud_launch_streaming_polyphase_stateful_hip
        assert f(*args(taps=None)) == -10
        assert f(*args(n=-1)) == -11
        assert f(*args(plen=len(poly) - 1)) == -12
        assert f(*args(in_=None)) == -13
        assert f(*args(tmp=None)) == -13
        assert f(*args(n=5000, cap=1)) == -14
        assert f(*args(n=5000, out=None, cap=1000)) == -15
    finally:
        for p in ptrs.values():
            L.gpud_free(p)


def test_primitives_against_closed_forms(gpu):
    """The successor primitive launchers (exports.cu:37-326)."""
    from argus._lib import lib
    L = lib()
    rng = np.random.default_rng(5)
    n = 10007
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    bufs = []

    def dev(nbytes):
        p = C.c_void_p()
        assert L.gpud_malloc(C.byref(p), nbytes) == 0
        bufs.append(p.value)
        return p.value

    def down(p, count, dtype):
        a = np.zeros(count, dtype)
        assert L.gpud_memcpy_d2h(a.ctypes.data, p, a.nbytes) == 0
        return a
    try:
        d_in, d_out, d_f = dev(x.nbytes), dev(x.nbytes), dev(4 * n)
        assert L.gpud_memcpy_h2d(d_in, x.ctypes.data, x.nbytes) == 0
        inc, ph0 = -0.0123, 1e5  # unreduced start phase, like the legacy callers pass
        assert L.gpud_launch_freq_shift_hip(d_in, d_out, n, inc, ph0) == 0
        want = x * np.exp(1j * (ph0 + inc * np.arange(n)))
        assert rel_rms(down(d_out, n, np.complex64), want) < 2e-6
        assert L.gpud_launch_fm_discrim_hip(d_in, d_f, n) == 0
        want = np.angle(np.conj(x[:-1].astype(np.complex128)) * x[1:])
        assert np.max(np.abs(down(d_f, n - 1, np.float32) - want)) < 1e-5
        assert L.gpud_launch_decimate_hip(d_in, d_out, n // 7, 7) == 0
        assert np.array_equal(down(d_out, n // 7, np.complex64), x[::7][:n // 7])
        taps = (rng.standard_normal(101) * 0.1).astype(np.float32)
        assert L.gpud_upload_fir_taps_hip(taps.ctypes.data, 101) == 0
        assert L.gpud_launch_fir_hip(d_in, d_out, n, 101) == 0
        want = np.convolve(x.astype(np.complex128), taps.astype(np.float64))[:n]
        assert rel_rms(down(d_out, n, np.complex64), want) < 2e-6
        d_t = dev(taps.nbytes)
        assert L.gpud_memcpy_h2d(d_t, taps.ctypes.data, taps.nbytes) == 0
        assert L.gpud_launch_fir_v2_stream_hip(d_in, d_out, d_t, n, 101, None) == 0
        assert L.gpud_stream_sync(None) == 0
        assert rel_rms(down(d_out, n, np.complex64), want) < 2e-6
        assert L.gpud_upload_fir_taps_hip(taps.ctypes.data, 257) == -1
        assert L.gpud_launch_am_envelope_hip(d_in, d_f, n) == 0
        assert np.max(np.abs(down(d_f, n, np.float32) - np.abs(x))) < 1e-5
        bfo = 2 * np.pi * 700 / 48000
        assert L.gpud_launch_ssb_product_hip(d_in, d_f, n, bfo, 0.25) == 0
        want = np.real(x * np.exp(1j * (0.25 + bfo * np.arange(n))))
        assert np.max(np.abs(down(d_f, n, np.float32) - want)) < 2e-5
        # n <= 0 launches are no-ops
        assert L.gpud_launch_freq_shift_hip(d_in, d_out, 0, inc, 0.0) == 0
        assert L.gpud_launch_fm_discrim_hip(d_in, d_f, 1) == 0
    finally:
        for p in bufs:
            L.gpud_free(p)

"""GPU parity of the legacy block path and the entry points the streaming
runner does not use, through libargus_hip.so:

* BatchRunner.ShiftFilterDecimateBatch[WithPhase] (batch_runner.go:61-98,
  batch_runner_windows.go:87-182) on jobs from the legacy job builder;
* the demod Engine: ShiftFilterDecimate with carried phase and sticky taps,
  DemodFused and Demod for every mode with the carried BFO phase
  (gpudemod_windows.go:320-654), SetFIR;
* gpud_launch_streaming_polyphase_prepare_hip (exports.cu:356-460) over a
  chunk sequence with history in/out arrays;
* the _stream primitive variants (exports.cu:62-75, 183-196, 259-271).

Against oracle.LegacyShiftFilterDecimate / LegacyDemod (FP64 restatement of
the reference's kernels) and the streaming host core; baseband and audio
within 1e-4 relative RMS, counts, rates and carried phases exact."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import rel_rms
from serve_oracle import synth_scene

pytestmark = pytest.mark.gpu

SR = 2_048_000
TOL = 1e-4


def legacy_taps(bw, sr=SR):
    return O.LowpassFIR(max(bw / 2, 200), sr, 101).astype(np.float32)


@pytest.fixture()
def scene():
    kinds = [("wfm", 300_000.0, 0.8), ("fm", -250_000.0, 0.6), ("am", 100_000.0, 0.5), ("ssb", -600_000.0, 0.4)]
    return synth_scene(3 * 65536, SR, kinds, seed=11)


def test_shift_filter_decimate_batch(gpu, scene):
    from argus import jobs as J
    sigs = [J.DetectedSignal(1, 145.3e6, 180e3, "WFM"), J.DetectedSignal(2, 144.75e6, 12e3, "NFM"),
            J.DetectedSignal(3, 145.1e6, 8e3), J.DetectedSignal(4, 144.4e6, 3e3, "WFM")]
    jobs = J.buildExtractJobs(SR, 145.0e6, sigs)
    assert [j.OutRate for j in jobs] == [512000, 204800, 204800, 512000]
    jobs.append(gpu.ExtractJob(-600_000.0, 3000.0, 32000, 12345.678))  # D = 64, unreduced phase start
    jobs.append(gpu.ExtractJob(100_000.0, 10000.0, 8000, -3.0))         # D = 256 > T
    r = gpu.NewBatchRunner(65536, SR)
    for k, n in enumerate([65536, 4099, 60001]):
        x = scene[k * 65536:k * 65536 + n]
        print(f"batch block {k} n={n}", flush=True)
        res = r.ShiftFilterDecimateBatchWithPhase(x, jobs)
        outs, rates = r.ShiftFilterDecimateBatch(x, jobs)
        for j, g, o, rate in zip(jobs, res, outs, rates):
            D = SR // j.OutRate
            want, ph_end = O.LegacyShiftFilterDecimate(x, SR, j.OffsetHz, j.PhaseStart, legacy_taps(j.BW), D)
            assert g.Rate == rate == SR // D and len(g.IQ) == n // D
            assert np.array_equal(g.IQ, o)
            assert rel_rms(g.IQ, want) <= TOL, (j, rel_rms(g.IQ, want))
            assert g.PhaseEnd == pytest.approx(ph_end, rel=1e-15, abs=1e-12)
    # n = 0: empty outputs, zero rates (batch_runner_windows.go:93-96)
    outs, rates = r.ShiftFilterDecimateBatch(np.zeros(0, np.complex64), jobs)
    assert all(len(o) == 0 for o in outs) and rates == [0] * len(jobs)
    with pytest.raises(ValueError, match="integer decimation"):
        r.ShiftFilterDecimateBatch(scene[:4096], [gpu.ExtractJob(0.0, 1e4, 300000)])
    with pytest.raises(ValueError, match="not enough output"):
        r.ShiftFilterDecimateBatch(scene[:100], [gpu.ExtractJob(0.0, 1e4, 8000)])
    r.Close()


def test_engine_shift_filter_decimate_carries_phase(gpu, scene):
    e = gpu.New(65536, SR)
    off, bw = 300_000.0, 180_000.0
    taps = legacy_taps(bw)
    ph = 0.0
    for k, n in enumerate([65536, 1000, 33333]):
        x = scene[k * 65536:k * 65536 + n]
        got, rate = e.ShiftFilterDecimate(x, off, 5e3 if k else bw, 192000)  # later bandwidths ignored: taps stick
        D = round(SR / 192000)
        want, ph = O.LegacyShiftFilterDecimate(x, SR, off, ph, taps, D)
        assert rate == SR // D and len(got) == n // D
        assert rel_rms(got, want) <= TOL, (k, rel_rms(got, want))
        assert e.State()[0] == pytest.approx(ph, rel=1e-14) and e.State()[2] == 101
    # SetFIR: a custom set, cut at 256 taps
    custom = (np.hanning(300) / 150).astype(np.float32)
    e.SetFIR(custom)
    assert e.State()[2] == 256
    x = scene[:20000]
    got, _ = e.ShiftFilterDecimate(x, -250_000.0, 1e4, 48000)
    want, _ = O.LegacyShiftFilterDecimate(x, SR, -250_000.0, ph, custom[:256], 43)
    assert rel_rms(got, want) <= TOL
    with pytest.raises(ValueError, match="invalid output"):
        e.ShiftFilterDecimate(x, 0.0, 1e4, 0)
    e.Close()


@pytest.mark.parametrize("fused", [True, False])
def test_engine_demod_every_mode(gpu, scene, fused):
    cases = [(gpu.DemodNFM, -250_000.0, 12_500.0), (gpu.DemodWFM, 300_000.0, 180_000.0),
             (gpu.DemodAM, 100_000.0, 10_000.0), (gpu.DemodUSB, -600_000.0, 3_000.0),
             (gpu.DemodLSB, -600_000.0, 3_000.0), (gpu.DemodCW, 100_000.0, 500.0)]
    for mode, off, bw in cases:
        e = gpu.New(65536, SR)
        D = round(SR / (192000 if mode == gpu.DemodWFM else 48000))
        taps = legacy_taps(bw)
        ph, bfo = 0.0, 0.0
        for k, n in enumerate([65536, 7777]):
            x = scene[k * 65536:k * 65536 + n]
            print(f"engine fused={fused} mode={mode} block {k}", flush=True)
            got, rate = (e.DemodFused if fused else e.Demod)(x, off, bw, mode)
            dec, ph = O.LegacyShiftFilterDecimate(x, SR, off, ph, taps, D)
            want, bfo = O.LegacyDemod(dec, mode, SR, bfo)
            assert rate == SR // D and len(got) == len(want), (mode, len(got), len(want))
            assert rel_rms(got, want) <= TOL, (mode, k, rel_rms(got, want))
            st = e.State()
            assert st[0] == pytest.approx(ph, rel=1e-14) and st[1] == pytest.approx(bfo, rel=1e-14, abs=1e-14)
        with pytest.raises(ValueError, match="unsupported"):
            e.Demod(scene[:1000], off, bw, 9)
        with pytest.raises(ValueError, match="not enough"):
            e.DemodFused(scene[:D + 1], off, bw, mode)  # one output: fused needs two
        e.Close()


def _dev(L, bufs, nbytes):
    p = C.c_void_p()
    assert L.gpud_malloc(C.byref(p), max(nbytes, 1)) == 0
    bufs.append(p.value)
    return p.value


def _down(L, p, count, dtype):
    a = np.zeros(count, dtype)
    if count:
        assert L.gpud_memcpy_d2h(a.ctypes.data, p, a.nbytes) == 0
    return a


def test_streaming_polyphase_prepare_sequence(gpu):
    """The legacy one-shot entry point carried by its caller across chunks
    (history in/out arrays, phase count and phase in and out) against the
    streaming host core (offset = -inc fs / 2 pi, streaming_host_core_test.go:5-40)."""
    from argus._lib import lib
    L = lib()
    sr, D, T, inc = 4_000_000, 20, 65, 0.017
    base = O.LowpassFIR(12500, sr, T).astype(np.float32)
    poly = O.BuildPolyphaseTapsPhaseMajor(base, D)
    x = O.makeToneNoiseIQ(20000, inc)
    bufs = []
    try:
        d_poly = _dev(L, bufs, poly.nbytes)
        assert L.gpud_memcpy_h2d(d_poly, poly.ctypes.data, poly.nbytes) == 0
        d_hist = [_dev(L, bufs, 8 * T), _dev(L, bufs, 8 * T)]
        d_in, d_out = _dev(L, bufs, 8 * 20000), _dev(L, bufs, 8 * 2000)
        hl, pc, ph = 0, 0, 0.0
        ohist, ophase, opc = np.zeros(0, np.complex64), 0.0, 0
        pos, sel = 0, 0
        for n in [0, 1, 2, 17, 63, 64, 65, 129, 511, 2048, 7, 4096]:
            chunk = np.ascontiguousarray(x[pos:pos + n])
            pos += n
            if n:
                assert L.gpud_memcpy_h2d(d_in, chunk.ctypes.data, chunk.nbytes) == 0
            n_out, pc_out, ph_out = C.c_int(), C.c_int(), C.c_double()
            rc = L.gpud_launch_streaming_polyphase_prepare_hip(d_in, n, d_hist[sel], hl, d_poly, len(poly), D, T, pc,
                                                               ph, inc, d_out, C.byref(n_out), C.byref(pc_out),
                                                               C.byref(ph_out), d_hist[sel ^ 1])
            assert rc == 0
            assert L.gpud_device_sync() == 0
            want, ophase, opc, ohist = O.runStreamingPolyphaseHostCore(chunk, sr, -inc * sr / (2 * np.pi), ophase,
                                                                       opc, T, D, ohist, poly)
            got = _down(L, d_out, n_out.value, np.complex64)
            assert n_out.value == len(want) and pc_out.value == opc, (n, n_out.value, len(want))
            if len(want):
                assert np.max(np.abs(got - want)) <= TOL
            assert abs(np.angle(np.exp(1j * (ph_out.value - ophase)))) <= 1e-8
            hl = min(hl + n, T - 1)
            assert hl == len(ohist)
            assert np.max(np.abs(_down(L, d_hist[sel ^ 1], hl, np.complex64) - ohist), initial=0) <= TOL
            sel ^= 1
            pc, ph = pc_out.value, ph_out.value
        # argument errors keep the reference's numbering (exports.cu:374-408)
        assert L.gpud_launch_streaming_polyphase_prepare_hip(d_in, 5, None, 0, None, len(poly), D, T, 0, 0.0, inc,
                                                             d_out, None, None, None, None) == -1
        assert L.gpud_launch_streaming_polyphase_prepare_hip(d_in, 5, None, 0, d_poly, D * 3, D, T, 0, 0.0, inc,
                                                             d_out, None, None, None, None) == -2
        assert L.gpud_launch_streaming_polyphase_prepare_hip(None, 5, None, 0, d_poly, len(poly), D, T, 0, 0.0, inc,
                                                             d_out, None, None, None, None) == -3
        assert L.gpud_launch_streaming_polyphase_prepare_hip(d_in, 40, None, 0, d_poly, len(poly), D, T, 0, 0.0,
                                                             inc, None, None, None, None, None) == -4
        assert L.gpud_launch_streaming_polyphase_prepare_hip(d_in, 5, None, 3, d_poly, len(poly), D, T, 0, 0.0, inc,
                                                             d_out, None, None, None, None) == -5
    finally:
        for p in bufs:
            L.gpud_free(p)


def test_stream_primitive_variants(gpu):
    """freq_shift / fir / decimate on a caller's stream, chained without a
    device-wide sync, against closed forms."""
    from argus._lib import lib
    L = lib()
    rng = np.random.default_rng(9)
    n = 30011
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    taps = (rng.standard_normal(77) * 0.1).astype(np.float32)
    st = C.c_void_p()
    assert L.gpud_stream_create(C.byref(st)) == 0
    bufs = []
    try:
        d_in, d_sh, d_fi, d_de = (_dev(L, bufs, x.nbytes) for _ in range(4))
        assert L.gpud_memcpy_h2d(d_in, x.ctypes.data, x.nbytes) == 0
        assert L.gpud_upload_fir_taps_hip(taps.ctypes.data, len(taps)) == 0
        inc, ph0 = 0.00731, -2.5
        assert L.gpud_launch_freq_shift_stream_hip(d_in, d_sh, n, inc, ph0, st) == 0
        assert L.gpud_launch_fir_stream_hip(d_sh, d_fi, n, len(taps), st) == 0
        assert L.gpud_launch_decimate_stream_hip(d_fi, d_de, n // 9, 9, st) == 0
        assert L.gpud_stream_sync(st) == 0
        sh = x.astype(np.complex128) * np.exp(1j * (ph0 + inc * np.arange(n)))
        fi = np.convolve(sh, taps.astype(np.float64))[:n]
        assert rel_rms(_down(L, d_sh, n, np.complex64), sh) < 2e-6
        assert rel_rms(_down(L, d_fi, n, np.complex64), fi) < 2e-6
        assert rel_rms(_down(L, d_de, n // 9, np.complex64), fi[::9][:n // 9]) < 2e-6
        # no-op launches and a null taps pointer
        assert L.gpud_launch_decimate_stream_hip(d_fi, d_de, 0, 9, st) == 0
        assert L.gpud_launch_fir_v2_stream_hip(d_sh, d_fi, None, n, 5, st) == -1
    finally:
        assert L.gpud_stream_destroy(st) == 0
        for p in bufs:
            L.gpud_free(p)

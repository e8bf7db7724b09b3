"""Host mirror of the reference's gpudemod streaming surface over the C ABI.

Names, argument meaning and error behaviour follow
internal/demod/gpudemod (batch_runner.go, streaming_types.go,
stream_state.go, streaming_gpu_stub.go, polyphase.go, cpu_oracle.go):

    r = NewBatchRunner(max_samples, sample_rate)
    results = r.StreamingExtractGPU(iq_new, jobs)    # list[StreamingExtractResult]
    r.ResetSignalState(signal_id); r.ResetAllSignalStates(); r.Close()
    outs, rates = r.ShiftFilterDecimateBatch(iq, extract_jobs)   # legacy block path
    e = New(max_samples, sample_rate); e.Demod(iq, offset, bw, DemodNFM)

All work runs in libargus_hip.so; this module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (ErrUnavailable, GpudError, check, gpud_streaming_job, gpud_streaming_result, lib, ptr, register)

P_, I_, L_ = C.c_void_p, C.c_int, C.c_int64
register("gpud_out_rings_create", I_, [C.POINTER(P_)])
register("gpud_out_rings_destroy", I_, [P_])
register("gpud_out_rings_next", I_, [P_, L_, L_, C.POINTER(P_)])
register("gpud_out_rings_prune", I_, [P_, P_, I_])
register("gpud_out_rings_pool_get", I_, [P_, L_, C.POINTER(P_), C.POINTER(L_)])
register("gpud_out_rings_pool_put", I_, [P_, P_, L_])
register("gpud_out_rings_stats", I_, [P_, C.POINTER(I_), C.POINTER(I_), C.POINTER(L_), C.POINTER(L_)])
register("gpud_out_rings_slot", I_, [P_, L_, C.POINTER(I_), P_])
register("gpud_runner_streaming_extract_ring", I_, [P_, P_, I_, P_, I_, P_, P_])
register("gpud_runner_out_rings", P_, [P_])
register("gpud_runner_set_fast", I_, [P_, I_])

OUT_RING_DEPTH = 4  # streaming_out_ring.go:11 streamOutRingDepth


class OutRings:
    """Per-signal output rings and their shared free pool (streamOutRing /
    outBufPool, streaming_out_ring.go:22-112).  Next(id, n) returns the host
    address of the buffer that takes signal id's next n outputs; Prune(active)
    returns the rings of other signals to the pool.  A runner's own rings
    (BatchRunner.OutRings()) are borrowed and never destroyed here."""

    def __init__(self, handle=None):
        self._own = handle is None
        if handle is None:
            h = C.c_void_p()
            check("gpud_out_rings_create", lib().gpud_out_rings_create(C.byref(h)))
            handle = h.value
        self._h = handle

    def Close(self) -> None:
        if self._h and self._own:
            lib().gpud_out_rings_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Next(self, signal_id: int, n: int) -> int:
        b = C.c_void_p()
        check("gpud_out_rings_next", lib().gpud_out_rings_next(self._h, int(signal_id), int(n), C.byref(b)))
        return b.value or 0

    def Prune(self, active) -> None:
        a = np.ascontiguousarray(list(active), np.int64)
        check("gpud_out_rings_prune", lib().gpud_out_rings_prune(self._h, a.ctypes.data if len(a) else None, len(a)))

    def PoolGet(self, n: int) -> tuple[int, int]:
        b, cap = C.c_void_p(), C.c_int64()
        check("gpud_out_rings_pool_get", lib().gpud_out_rings_pool_get(self._h, int(n), C.byref(b), C.byref(cap)))
        return b.value, cap.value

    def PoolPut(self, buf: int, cap: int) -> None:
        check("gpud_out_rings_pool_put", lib().gpud_out_rings_pool_put(self._h, buf, int(cap)))

    def Stats(self) -> dict:
        nr, nf, a, r = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        check("gpud_out_rings_stats", lib().gpud_out_rings_stats(self._h, C.byref(nr), C.byref(nf), C.byref(a),
                                                                  C.byref(r)))
        return {"rings": nr.value, "free": nf.value, "allocated": a.value, "released": r.value}

    def Slot(self, signal_id: int):
        """(next slot index, capacities of the ring's buffers) or None."""
        idx = C.c_int()
        caps = np.zeros(OUT_RING_DEPTH, np.int64)
        rc = lib().gpud_out_rings_slot(self._h, int(signal_id), C.byref(idx), caps.ctypes.data)
        return None if rc != 0 else (idx.value, [int(c) for c in caps])


def ring_view(addr: int, n: int) -> np.ndarray:
    """A complex64 view of n samples at a host address (no copy)."""
    if not addr or n <= 0:
        return np.zeros(0, np.complex64)
    return np.ctypeslib.as_array((C.c_float * (2 * n)).from_address(addr)).view(np.complex64)


EXEC_BATCHED = 0
EXEC_PER_SIGNAL = 1
WINDOW_HANN = 0
WINDOW_HAMMING = 1


# gpudemod.go:81-90 DemodType
DemodNFM, DemodWFM, DemodAM, DemodUSB, DemodLSB, DemodCW = range(6)


class gpud_extract_job(C.Structure):
    _fields_ = [("offset_hz", C.c_double), ("bw", C.c_double), ("out_rate", C.c_int32), ("pad", C.c_int32),
                ("phase_start", C.c_double)]


class gpud_extract_result(C.Structure):
    _fields_ = [("rate", C.c_int32), ("n_out", C.c_int32), ("out_offset", C.c_int64), ("phase_end", C.c_double)]


_P, _I, _D = C.c_void_p, C.c_int, C.c_double
register("gpud_runner_shift_filter_decimate_batch", _I,
         [_P, _P, _I, C.POINTER(gpud_extract_job), _I, _P, C.c_int64, C.POINTER(gpud_extract_result)])
register("gpud_engine_create", _I, [_I, _I, C.POINTER(_P)])
register("gpud_engine_destroy", _I, [_P])
register("gpud_engine_set_fir", _I, [_P, _P, _I])
register("gpud_engine_state", _I, [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I)])
register("gpud_engine_shift_filter_decimate", _I, [_P, _P, _I, _D, _D, _I, _P, _I, C.POINTER(_I), C.POINTER(_I)])
register("gpud_engine_demod", _I, [_P, _P, _I, _D, _D, _I, _I, _P, _I, C.POINTER(_I), C.POINTER(_I)])

_LEGACY_ERRORS = {-2: "sample count exceeds engine capacity", -3: "invalid output sample rate",
                  -4: "not enough output samples after decimation", -5: "unsupported demod type"}


@dataclass
class ExtractJob:
    """batch.go:5-10"""
    OffsetHz: float
    BW: float
    OutRate: int
    PhaseStart: float = 0.0


@dataclass
class ExtractResult:
    """batch.go:14-18"""
    IQ: np.ndarray
    Rate: int
    PhaseEnd: float


@dataclass
class StreamingExtractJob:
    """streaming_types.go:8-15"""
    SignalID: int
    OffsetHz: float
    Bandwidth: float
    OutRate: int
    NumTaps: int
    ConfigHash: int


@dataclass
class StreamingExtractResult:
    """streaming_types.go:17-24 (+ the NCO phase the runner carries)."""
    SignalID: int
    IQ: np.ndarray
    Rate: int
    NOut: int
    PhaseCount: int
    HistoryLen: int
    NCOPhase: float = 0.0


@dataclass
class ExtractStreamState:
    """streaming_types.go:26-44 (the fields a test inspects)."""
    SignalID: int
    ConfigHash: int
    NCOPhase: float
    Decim: int
    PhaseCount: int
    NumTaps: int
    ShiftedHistory: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex64))


def Available() -> bool:
    """gpudemod.go:112-118"""
    try:
        n = C.c_int(0)
        return lib().gpud_device_count(C.byref(n)) == 0 and n.value > 0
    except ErrUnavailable:
        return False


def StreamingConfigHash(signal_id, offset_hz, bandwidth, out_rate, num_taps, sample_rate) -> int:
    """streaming_types.go:58-72"""
    return int(lib().gpud_streaming_config_hash(signal_id, offset_hz, bandwidth, out_rate, num_taps, sample_rate))


def ExactIntegerDecimation(sample_rate: int, out_rate: int) -> int:
    """cpu_oracle.go:155-163; raises ValueError like the Go error return."""
    d = lib().gpud_exact_integer_decimation(sample_rate, out_rate)
    if d < 0:
        raise ValueError(f"streaming polyphase extractor requires integer decimation: "
                         f"sampleRate={sample_rate} outRate={out_rate}")
    return d


def LowpassFIR(cutoff_hz: float, sample_rate: int, taps: int) -> np.ndarray:
    """dsp/fir.go:6-28"""
    out = np.zeros(taps + 1, np.float64)
    n = lib().gpud_lowpass_fir(cutoff_hz, sample_rate, taps, ptr(out))
    return out[:n]


def PolyphasePhaseLen(base_len: int, decim: int) -> int:
    """polyphase.go:24-28"""
    if decim <= 0 or base_len <= 0:
        return 0
    return (base_len + decim - 1) // decim


def BuildPolyphaseTapsPhaseMajor(base, decim: int) -> np.ndarray:
    """polyphase.go:6-22"""
    base = np.ascontiguousarray(base, np.float32)
    if decim <= 0 or len(base) == 0:
        return np.zeros(0, np.float32)
    out = np.zeros(decim * PolyphasePhaseLen(len(base), decim), np.float32)
    n = lib().gpud_build_polyphase_taps(ptr(base), len(base), decim, ptr(out))
    return out[:n]


def _jobs_array(jobs):
    arr = (gpud_streaming_job * max(1, len(jobs)))()
    for i, j in enumerate(jobs):
        arr[i].signal_id = int(j.SignalID)
        arr[i].offset_hz = float(j.OffsetHz)
        arr[i].bandwidth = float(j.Bandwidth)
        arr[i].out_rate = int(j.OutRate)
        arr[i].num_taps = int(j.NumTaps)
        arr[i].config_hash = int(j.ConfigHash) & 0xFFFFFFFFFFFFFFFF
    return arr


class BatchRunner:
    """batch_runner.go:12-50 (streaming surface)."""

    def __init__(self, max_samples: int, sample_rate: int, exec_mode: int = EXEC_BATCHED,
                 window: int = WINDOW_HANN):
        h = C.c_void_p()
        rc = lib().gpud_runner_create(max_samples, sample_rate, C.byref(h))
        if rc != 0:
            raise ErrUnavailable(f"gpud_runner_create failed ({rc})")
        self._h = h
        self.sample_rate = sample_rate
        self.max_samples = max_samples
        self.SetExecMode(exec_mode)
        check("gpud_runner_set_window", lib().gpud_runner_set_window(self._h, window))

    # -- lifecycle --------------------------------------------------------
    def Close(self) -> None:
        if self._h:
            lib().gpud_runner_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def SetExecMode(self, mode: int) -> None:
        check("gpud_runner_set_exec_mode", lib().gpud_runner_set_exec_mode(self._h, mode))

    def stream(self) -> int:
        return lib().gpud_runner_stream(self._h)

    def SetFast(self, enable: bool) -> None:
        """Fast-convolution form of the batched launch (on by default)."""
        check("gpud_runner_set_fast", lib().gpud_runner_set_fast(self._h, 1 if enable else 0))

    # -- streaming extraction ----------------------------------------------
    def StreamingExtractGPU(self, iq_new, jobs) -> list[StreamingExtractResult]:
        """streaming_gpu_stub.go:21-26"""
        if not self._h:
            raise ErrUnavailable("runner closed")
        iq = np.ascontiguousarray(iq_new if iq_new is not None else np.zeros(0, np.complex64), np.complex64)
        n = len(iq)
        jarr = _jobs_array(jobs)
        res = (gpud_streaming_result * max(1, len(jobs)))()
        cap = sum(n // max(1, 1) for _ in jobs) + 2 * len(jobs) + 1  # upper bound (D >= 1)
        out = np.zeros(max(cap, 1), np.complex64)
        rc = lib().gpud_runner_streaming_extract(self._h, ptr(iq) if n else None, n, jarr, len(jobs), ptr(out),
                                                 len(out), res)
        if rc == -1:
            raise ValueError("streaming extract rejected the jobs (bad arguments or non-integer decimation)")
        check("gpud_runner_streaming_extract", rc)
        results = []
        for i in range(len(jobs)):
            r = res[i]
            iqo = out[r.out_offset:r.out_offset + r.n_out].copy()
            results.append(StreamingExtractResult(r.signal_id, iqo, r.rate, r.n_out, r.phase_count, r.history_len,
                                                  r.nco_phase))
        return results

    StreamingExtractGPUExec = StreamingExtractGPU

    def StreamingExtractGPURing(self, iq_new, jobs) -> list[StreamingExtractResult]:
        """StreamingExtractGPU with the reference's output ownership
        (streaming_gpu_native_prepare.go:78-104): each result's IQ is a view
        of its signal's ring buffer, valid until that signal's ring comes
        round again (OUT_RING_DEPTH calls) or the signal is pruned."""
        if not self._h:
            raise ErrUnavailable("runner closed")
        iq = np.ascontiguousarray(iq_new if iq_new is not None else np.zeros(0, np.complex64), np.complex64)
        n = len(iq)
        jarr = _jobs_array(jobs)
        res = (gpud_streaming_result * max(1, len(jobs)))()
        outs = (C.c_void_p * max(1, len(jobs)))()
        rc = lib().gpud_runner_streaming_extract_ring(self._h, ptr(iq) if n else None, n, jarr, len(jobs), outs, res)
        if rc == -1:
            raise ValueError("streaming extract rejected the jobs (bad arguments or non-integer decimation)")
        check("gpud_runner_streaming_extract_ring", rc)
        results = []
        for i in range(len(jobs)):
            r = res[i]
            results.append(StreamingExtractResult(r.signal_id, ring_view(outs[i] or 0, r.n_out), r.rate, r.n_out,
                                                  r.phase_count, r.history_len, r.nco_phase))
        return results

    def OutRings(self) -> OutRings | None:
        h = lib().gpud_runner_out_rings(self._h)
        return OutRings(h) if h else None

    def StreamingExtractDevice(self, iq_dev_ptr: int, n_new: int, jobs):
        """Device-resident input; returns (device output pointer, raw results).
        Completion is ordered on self.stream()."""
        jarr = _jobs_array(jobs)
        res = (gpud_streaming_result * max(1, len(jobs)))()
        out = C.c_void_p()
        check("gpud_runner_streaming_extract_device",
              lib().gpud_runner_streaming_extract_device(self._h, iq_dev_ptr, n_new, jarr, len(jobs),
                                                         C.byref(out), res))
        return out.value, res

    # -- state ------------------------------------------------------------
    def SignalState(self, signal_id: int) -> ExtractStreamState | None:
        h = C.c_uint64()
        ph = C.c_double()
        dec, pc, nt = C.c_int(), C.c_int(), C.c_int()
        hist = np.zeros(4096, np.complex64)
        n = lib().gpud_runner_get_signal_state(self._h, signal_id, C.byref(h), C.byref(ph), C.byref(dec),
                                               C.byref(pc), C.byref(nt), ptr(hist), len(hist))
        if n < 0:
            return None
        return ExtractStreamState(signal_id, h.value, ph.value, dec.value, pc.value, nt.value, hist[:n].copy())

    def StateIDs(self) -> list[int]:
        n = lib().gpud_runner_num_states(self._h)
        ids = np.zeros(max(n, 1), np.int64)
        lib().gpud_runner_state_ids(self._h, ptr(ids), len(ids))
        return [int(x) for x in ids[:n]]

    def ResetSignalState(self, signal_id: int) -> None:
        """stream_state.go:9-16"""
        check("gpud_runner_reset_signal_state", lib().gpud_runner_reset_signal_state(self._h, signal_id))

    def ResetAllSignalStates(self) -> None:
        """stream_state.go:18-24"""
        check("gpud_runner_reset_all_signal_states", lib().gpud_runner_reset_all_signal_states(self._h))

    # -- legacy block path ------------------------------------------------
    def _sfd_batch(self, iq, jobs):
        iq = np.ascontiguousarray(iq, np.complex64)
        n = len(iq)
        arr = (gpud_extract_job * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            arr[i] = gpud_extract_job(float(j.OffsetHz), float(j.BW), int(j.OutRate), 0, float(j.PhaseStart))
        res = (gpud_extract_result * max(1, len(jobs)))()
        out = np.zeros(max(1, n * max(1, len(jobs))), np.complex64)
        rc = lib().gpud_runner_shift_filter_decimate_batch(self._h, ptr(iq) if n else None, n, arr, len(jobs),
                                                           ptr(out), len(out), res)
        if rc == -3:
            raise ValueError("streaming polyphase extractor requires integer decimation")
        if rc in _LEGACY_ERRORS:
            raise ValueError(_LEGACY_ERRORS[rc])
        check("gpud_runner_shift_filter_decimate_batch", rc)
        return [(out[r.out_offset:r.out_offset + r.n_out].copy(), r.rate, r.phase_end) for r in res[:len(jobs)]]

    def ShiftFilterDecimateBatch(self, iq, jobs):
        """batch_runner.go:61-67: (outputs, rates)."""
        got = self._sfd_batch(iq, jobs)
        return [g[0] for g in got], [g[1] for g in got]

    def ShiftFilterDecimateBatchWithPhase(self, iq, jobs) -> list[ExtractResult]:
        """batch_runner.go:69-98"""
        return [ExtractResult(*g) for g in self._sfd_batch(iq, jobs)]


class Engine:
    """gpudemod_windows.go Engine: carried shift and BFO phases, a FIR tap
    set that sticks once designed or set, ShiftFilterDecimate, Demod and
    DemodFused."""

    def __init__(self, max_samples: int, sample_rate: int):
        h = C.c_void_p()
        rc = lib().gpud_engine_create(max_samples, sample_rate, C.byref(h))
        if rc != 0:
            raise ErrUnavailable(f"gpud_engine_create failed ({rc})")
        self._h = h
        self.max_samples = max_samples
        self.sample_rate = sample_rate

    def Close(self) -> None:
        if self._h:
            lib().gpud_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def SetFIR(self, taps) -> None:
        taps = np.ascontiguousarray(taps if taps is not None else np.zeros(0), np.float32)
        check("gpud_engine_set_fir", lib().gpud_engine_set_fir(self._h, ptr(taps) if len(taps) else None, len(taps)))

    def State(self):
        """(shift phase, BFO phase, tap count)"""
        ph, bfo, nt = C.c_double(), C.c_double(), C.c_int()
        check("gpud_engine_state", lib().gpud_engine_state(self._h, C.byref(ph), C.byref(bfo), C.byref(nt)))
        return ph.value, bfo.value, nt.value

    def ShiftFilterDecimate(self, iq, offset_hz: float, bw: float, out_rate: int):
        """gpudemod_windows.go:320-393: (IQ, rate)."""
        iq = np.ascontiguousarray(iq, np.complex64)
        out = np.zeros(max(1, len(iq)), np.complex64)
        n_out, rate = C.c_int(), C.c_int()
        rc = lib().gpud_engine_shift_filter_decimate(self._h, ptr(iq) if len(iq) else None, len(iq), offset_hz, bw,
                                                     out_rate, ptr(out), len(out), C.byref(n_out), C.byref(rate))
        if rc in _LEGACY_ERRORS:
            raise ValueError(_LEGACY_ERRORS[rc])
        check("gpud_engine_shift_filter_decimate", rc)
        return out[:n_out.value].copy(), rate.value

    def _demod(self, iq, offset_hz, bw, mode, fused):
        iq = np.ascontiguousarray(iq, np.complex64)
        out = np.zeros(max(1, len(iq)), np.float32)
        n_out, rate = C.c_int(), C.c_int()
        rc = lib().gpud_engine_demod(self._h, ptr(iq) if len(iq) else None, len(iq), offset_hz, bw, mode,
                                     1 if fused else 0, ptr(out), len(out), C.byref(n_out), C.byref(rate))
        if rc in _LEGACY_ERRORS:
            raise ValueError(_LEGACY_ERRORS[rc])
        check("gpud_engine_demod", rc)
        return out[:n_out.value].copy(), rate.value

    def Demod(self, iq, offset_hz: float, bw: float, mode: int):
        """gpudemod_windows.go:523-654: (audio, rate)."""
        return self._demod(iq, offset_hz, bw, mode, False)

    def DemodFused(self, iq, offset_hz: float, bw: float, mode: int):
        """gpudemod_windows.go:395-521: (audio, rate)."""
        return self._demod(iq, offset_hz, bw, mode, True)


def New(max_samples: int, sample_rate: int) -> Engine:
    """gpudemod_windows.go:112-162"""
    return Engine(max_samples, sample_rate)


def NewBatchRunner(max_samples: int, sample_rate: int, **kw) -> BatchRunner:
    """batch_runner.go:23-34"""
    return BatchRunner(max_samples, sample_rate, **kw)


__all__ = [
    "Available", "BatchRunner", "NewBatchRunner", "StreamingExtractJob", "StreamingExtractResult",
    "ExtractStreamState", "StreamingConfigHash", "ExactIntegerDecimation", "BuildPolyphaseTapsPhaseMajor",
    "PolyphasePhaseLen", "LowpassFIR", "ErrUnavailable", "GpudError", "EXEC_BATCHED", "EXEC_PER_SIGNAL",
    "WINDOW_HANN", "WINDOW_HAMMING", "ExtractJob", "ExtractResult", "Engine", "New", "DemodNFM", "DemodWFM",
    "DemodAM", "DemodUSB", "DemodLSB", "DemodCW", "OutRings", "OUT_RING_DEPTH", "ring_view",
]

"""oracle -- ctypes wrapper of the CPU restatement of the reference path.

TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this; the product (libargus_hip.so and the argus
Python binding) never does.  Function names follow the reference Go code so
tests read like the reference's own tests.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB_PATH = _DIR / "liboracle.so"
_lib = None


class orc_job(C.Structure):
    _fields_ = [("signal_id", C.c_int64), ("offset_hz", C.c_double), ("bandwidth", C.c_double),
                ("out_rate", C.c_int32), ("num_taps", C.c_int32), ("config_hash", C.c_uint64)]


class orc_result(C.Structure):
    _fields_ = [("signal_id", C.c_int64), ("rate", C.c_int32), ("n_out", C.c_int32),
                ("phase_count", C.c_int32), ("history_len", C.c_int32), ("out_offset", C.c_int64),
                ("nco_phase", C.c_double)]


P, I, D = C.c_void_p, C.c_int, C.c_double
PI_, PD_ = C.POINTER(C.c_int), C.POINTER(C.c_double)
_SIGS = {
    "orc_lowpass_fir": (I, [D, I, I, P]),
    "orc_lowpass_fir_hamming": (I, [D, I, I, P]),
    "orc_apply_fir": (None, [P, I, P, I, P]),
    "orc_decimate": (I, [P, I, I, P]),
    "orc_freq_shift": (None, [P, I, I, D, P]),
    "orc_stateful_fir_real": (None, [P, I, P, P, I, P]),
    "orc_stateful_decim_fir_complex": (I, [P, I, I, P, P, PI_, P, I, P]),
    "orc_polyphase_phase_len": (I, [I, I]),
    "orc_build_polyphase": (I, [P, I, I, P]),
    "orc_exact_decim": (I, [I, I]),
    "orc_config_hash": (C.c_uint64, [C.c_int64, D, D, I, I, I]),
    "orc_host_core": (I, [P, I, I, D, PD_, PI_, I, I, P, PI_, P, I, P, I]),
    "orc_extract_direct": (I, [P, I, D, PD_, PI_, I, I, P, P, PI_, P, I]),
    "orc_extract_polyphase": (I, [P, I, D, PD_, PI_, I, I, P, P, P, PI_, P, I]),
    "orc_runner_new": (P, [I]),
    "orc_runner_free": (None, [P]),
    "orc_runner_streaming_extract": (I, [P, P, I, C.POINTER(orc_job), I, P, C.c_int64, C.POINTER(orc_result)]),
    "orc_runner_num_states": (I, [P]),
    "orc_runner_state_ids": (I, [P, P, I]),
    "orc_runner_get_state": (I, [P, C.c_int64, C.POINTER(C.c_uint64), PD_, PI_, PI_, PI_, P, I]),
    "orc_runner_reset_signal": (None, [P, C.c_int64]),
    "orc_runner_reset_all": (None, [P]),
    "orc_runner_set_threads": (I, [P, I]),
    "orc_fm_discrim": (I, [P, I, P]),
    "orc_fm_discrim_bridged": (I, [P, I, P, PI_, P]),
    "orc_am_demod": (I, [P, I, P]),
    "orc_ssb_demod": (I, [P, I, I, I, PD_, P]),
    "orc_cw_demod": (I, [P, I, I, PD_, P]),
    "orc_rds_baseband": (I, [P, I, I, P, I, PI_]),
}


def register(name, restype, argtypes):
    _SIGS[name] = (restype, argtypes)
    if _lib is not None:
        f = getattr(_lib, name)
        f.restype, f.argtypes = restype, argtypes


def lib():
    global _lib
    if _lib is None:
        h = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def c64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex64)


# ---------------------------------------------------------------- dsp
def LowpassFIR(cutoff_hz, sample_rate, taps):
    out = np.zeros(taps + 1)
    n = lib().orc_lowpass_fir(cutoff_hz, sample_rate, taps, _p(out))
    return out[:n]


def LowpassFIRHamming(cutoff_hz, sample_rate, taps):
    out = np.zeros(taps + 1)
    n = lib().orc_lowpass_fir_hamming(cutoff_hz, sample_rate, taps, _p(out))
    return out[:n]


def ApplyFIR(iq, taps):
    iq = c64(iq)
    taps = np.ascontiguousarray(taps, np.float64)
    out = np.zeros(len(iq), np.complex64)
    if len(iq) and len(taps):
        lib().orc_apply_fir(_p(iq), len(iq), _p(taps), len(taps), _p(out))
    return out


def Decimate(iq, factor):
    iq = c64(iq)
    out = np.zeros(len(iq) + 1, np.complex64)
    m = lib().orc_decimate(_p(iq), len(iq), factor, _p(out))
    return out[:m]


def FreqShift(iq, sample_rate, offset_hz):
    iq = c64(iq)
    out = np.zeros(len(iq), np.complex64)
    lib().orc_freq_shift(_p(iq), len(iq), sample_rate, offset_hz, _p(out))
    return out


class StatefulDecimatingFIRComplex:
    """dsp/decimating_fir.go:6-95"""

    def __init__(self, taps, factor):
        self.taps = np.ascontiguousarray(taps, np.float64)
        self.factor = max(1, factor)
        self.dr = np.zeros(len(self.taps))
        self.di = np.zeros(len(self.taps))
        self.phase = C.c_int(0)

    def Process(self, iq):
        iq = c64(iq)
        out = np.zeros(len(iq) + 1, np.complex64)
        m = lib().orc_stateful_decim_fir_complex(_p(self.taps), len(self.taps), self.factor, _p(self.dr),
                                                 _p(self.di), C.byref(self.phase), _p(iq), len(iq), _p(out))
        return out[:m]


class StatefulFIRReal:
    """dsp/fir_stateful.go:6-54"""

    def __init__(self, taps):
        self.taps = np.ascontiguousarray(taps, np.float64)
        self.delay = np.zeros(len(self.taps))

    def Process(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x), np.float32)
        if len(x):
            lib().orc_stateful_fir_real(_p(self.taps), len(self.taps), _p(self.delay), _p(x), len(x), _p(out))
        return out


# ---------------------------------------------------------------- extractor
def PolyphasePhaseLen(base_len, decim):
    return lib().orc_polyphase_phase_len(base_len, decim)


def BuildPolyphaseTapsPhaseMajor(base, decim):
    base = np.ascontiguousarray(base, np.float32)
    if decim <= 0 or len(base) == 0:
        return np.zeros(0, np.float32)
    out = np.zeros(decim * PolyphasePhaseLen(len(base), decim), np.float32)
    n = lib().orc_build_polyphase(_p(base), len(base), decim, _p(out))
    return out[:n]


def ExactIntegerDecimation(sample_rate, out_rate):
    d = lib().orc_exact_decim(sample_rate, out_rate)
    if d < 0:
        raise ValueError("integer decimation required")
    return d


def StreamingConfigHash(signal_id, offset_hz, bandwidth, out_rate, num_taps, sample_rate):
    return int(lib().orc_config_hash(signal_id, offset_hz, bandwidth, out_rate, num_taps, sample_rate))


@dataclass
class CPUOracleState:
    """cpu_oracle.go:8-18"""
    SignalID: int = 0
    ConfigHash: int = 0
    NCOPhase: float = 0.0
    Decim: int = 1
    PhaseCount: int = 0
    NumTaps: int = 0
    ShiftedHistory: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex64))
    BaseTaps: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    PolyphaseTaps: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))


def _hist_buf(state):
    cap = max(1, state.NumTaps - 1) + 1
    h = np.zeros(max(cap, len(state.ShiftedHistory) + 1), np.complex64)
    h[:len(state.ShiftedHistory)] = state.ShiftedHistory
    return h


def CPUOracleExtract(iq_new, state: CPUOracleState, phase_inc: float):
    """cpu_oracle.go:32-75 (mutates state)."""
    iq = c64(iq_new)
    if state.NumTaps <= 0 or state.Decim <= 0 or len(state.BaseTaps) < state.NumTaps:
        return np.zeros(0, np.complex64)
    h = _hist_buf(state)
    hl, pc, ph = C.c_int(len(state.ShiftedHistory)), C.c_int(state.PhaseCount), C.c_double(state.NCOPhase)
    out = np.zeros(len(iq) // state.Decim + 2, np.complex64)
    base = np.ascontiguousarray(state.BaseTaps, np.float32)
    n = lib().orc_extract_direct(_p(iq), len(iq), phase_inc, C.byref(ph), C.byref(pc), state.NumTaps, state.Decim,
                                 _p(base), _p(h), C.byref(hl), _p(out), len(out))
    state.NCOPhase, state.PhaseCount = ph.value, pc.value
    state.ShiftedHistory = h[:hl.value].copy()
    return out[:n]


def CPUOracleExtractPolyphase(iq_new, state: CPUOracleState, phase_inc: float):
    """cpu_oracle.go:79-132 (mutates state)."""
    iq = c64(iq_new)
    if state.NumTaps <= 0 or state.Decim <= 0 or len(state.BaseTaps) < state.NumTaps:
        return np.zeros(0, np.complex64)
    if len(state.PolyphaseTaps) == 0:
        state.PolyphaseTaps = BuildPolyphaseTapsPhaseMajor(state.BaseTaps, state.Decim)
    h = _hist_buf(state)
    hl, pc, ph = C.c_int(len(state.ShiftedHistory)), C.c_int(state.PhaseCount), C.c_double(state.NCOPhase)
    out = np.zeros(len(iq) // state.Decim + 2, np.complex64)
    base = np.ascontiguousarray(state.BaseTaps, np.float32)
    poly = np.ascontiguousarray(state.PolyphaseTaps, np.float32)
    n = lib().orc_extract_polyphase(_p(iq), len(iq), phase_inc, C.byref(ph), C.byref(pc), state.NumTaps,
                                    state.Decim, _p(base), _p(poly), _p(h), C.byref(hl), _p(out), len(out))
    state.NCOPhase, state.PhaseCount = ph.value, pc.value
    state.ShiftedHistory = h[:hl.value].copy()
    return out[:n]


def runStreamingPolyphaseHostCore(iq_new, sample_rate, offset_hz, nco_phase, phase_count, num_taps, decim,
                                  history, poly_taps):
    """streaming_host_core.go:5-64 -> (out, phase, phaseCount, history)."""
    iq = c64(iq_new)
    cap = max(1, num_taps - 1) + 1
    h = np.zeros(max(cap, len(history) + 1), np.complex64)
    h[:len(history)] = history
    hl, pc, ph = C.c_int(len(history)), C.c_int(phase_count), C.c_double(nco_phase)
    poly = np.ascontiguousarray(poly_taps, np.float32)
    out = np.zeros(len(iq) // max(1, decim) + 2, np.complex64)
    n = lib().orc_host_core(_p(iq), len(iq), sample_rate, offset_hz, C.byref(ph), C.byref(pc), num_taps, decim,
                            _p(h), C.byref(hl), _p(poly), len(poly), _p(out), len(out))
    return out[:n], ph.value, pc.value, h[:hl.value].copy()


@dataclass
class StreamingExtractResult:
    SignalID: int
    IQ: np.ndarray
    Rate: int
    NOut: int
    PhaseCount: int
    HistoryLen: int
    NCOPhase: float = 0.0


class CPUOracleRunner:
    """streaming_oracle_extract.go:9-111"""

    def __init__(self, sample_rate: int, threads: int = 1):
        self.SampleRate = sample_rate
        self._h = lib().orc_runner_new(sample_rate)
        lib().orc_runner_set_threads(self._h, threads)

    def __del__(self):
        try:
            if self._h:
                lib().orc_runner_free(self._h)
                self._h = None
        except Exception:
            pass

    def StreamingExtract(self, iq_new, jobs):
        iq = c64(iq_new if iq_new is not None else np.zeros(0, np.complex64))
        jarr = (orc_job * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            jarr[i].signal_id, jarr[i].offset_hz, jarr[i].bandwidth = j.SignalID, j.OffsetHz, j.Bandwidth
            jarr[i].out_rate, jarr[i].num_taps = j.OutRate, j.NumTaps
            jarr[i].config_hash = int(j.ConfigHash) & 0xFFFFFFFFFFFFFFFF
        res = (orc_result * max(1, len(jobs)))()
        out = np.zeros(len(iq) * max(1, len(jobs)) + 2 * len(jobs) + 1, np.complex64)
        rc = lib().orc_runner_streaming_extract(self._h, _p(iq), len(iq), jarr, len(jobs), _p(out), len(out), res)
        if rc != 0:
            raise ValueError(f"oracle StreamingExtract failed ({rc})")
        return [StreamingExtractResult(r.signal_id, out[r.out_offset:r.out_offset + r.n_out].copy(), r.rate,
                                       r.n_out, r.phase_count, r.history_len, r.nco_phase)
                for r in (res[i] for i in range(len(jobs)))]

    def StreamingExtractRaw(self, iq, jobs_arr, n_jobs, out, results):
        """Raw call for the CPU baseline timing (preallocated ctypes arrays)."""
        return lib().orc_runner_streaming_extract(self._h, _p(iq), len(iq), jobs_arr, n_jobs, _p(out), len(out),
                                                  results)

    @property
    def States(self) -> dict:
        n = lib().orc_runner_num_states(self._h)
        ids = np.zeros(max(n, 1), np.int64)
        lib().orc_runner_state_ids(self._h, _p(ids), len(ids))
        out = {}
        for sid in ids[:n]:
            h = C.c_uint64()
            ph = C.c_double()
            dec, pc, nt = C.c_int(), C.c_int(), C.c_int()
            hist = np.zeros(8192, np.complex64)
            hl = lib().orc_runner_get_state(self._h, int(sid), C.byref(h), C.byref(ph), C.byref(dec), C.byref(pc),
                                            C.byref(nt), _p(hist), len(hist))
            out[int(sid)] = CPUOracleState(int(sid), h.value, ph.value, dec.value, pc.value, nt.value,
                                           hist[:hl].copy())
        return out

    def ResetSignalState(self, sid):
        lib().orc_runner_reset_signal(self._h, sid)

    def ResetAllStates(self):
        lib().orc_runner_reset_all(self._h)


# ---------------------------------------------------------------- demods
def fmDiscrim(iq):
    iq = c64(iq)
    out = np.zeros(max(len(iq) - 1, 0), np.float32)
    n = lib().orc_fm_discrim(_p(iq), len(iq), _p(out))
    return out[:n]


class BridgedFMDiscrim:
    """streamer.go:1287-1302 block-bridged discriminator."""

    def __init__(self):
        self.last = np.zeros(2, np.float32)
        self.have = C.c_int(0)

    def Demod(self, iq):
        iq = c64(iq)
        out = np.zeros(len(iq) + 1, np.float32)
        n = lib().orc_fm_discrim_bridged(_p(iq), len(iq), _p(self.last), C.byref(self.have), _p(out))
        return out[:n]


def AMDemod(iq):
    iq = c64(iq)
    out = np.zeros(len(iq), np.float32)
    n = lib().orc_am_demod(_p(iq), len(iq), _p(out))
    return out[:n]


def SSBDemod(iq, sample_rate, usb=True, phase=0.0):
    """Returns (audio, carried phase)."""
    iq = c64(iq)
    out = np.zeros(len(iq), np.float32)
    ph = C.c_double(phase)
    n = lib().orc_ssb_demod(_p(iq), len(iq), sample_rate, 1 if usb else 0, C.byref(ph), _p(out))
    return out[:n], ph.value


def CWDemod(iq, sample_rate, phase=0.0):
    iq = c64(iq)
    out = np.zeros(len(iq), np.float32)
    ph = C.c_double(phase)
    n = lib().orc_cw_demod(_p(iq), len(iq), sample_rate, C.byref(ph), _p(out))
    return out[:n], ph.value


def RDSBasebandComplex(iq, sample_rate):
    """fm.go:143-163 / 234-271 -> (samples, rate)."""
    iq = c64(iq)
    out = np.zeros(len(iq) + 1, np.complex64)
    rate = C.c_int(0)
    n = lib().orc_rds_baseband(_p(iq), len(iq), sample_rate, _p(out), len(out), C.byref(rate))
    return out[:n], rate.value


# ---------------------------------------------------------------- reference test inputs
def MakeDeterministicIQ(n):
    """test_harness.go:16-24"""
    i = np.arange(n, dtype=np.float64)
    a, b = 0.017 * i, 0.031 * i
    return (np.cos(a) + 0.2 * np.cos(b) + 1j * (np.sin(a) + 0.15 * np.sin(b))).astype(np.complex64)


def MakeToneIQ(n, phase_inc):
    """test_harness.go:26-34 (sequential accumulation of the phase)."""
    ph = np.cumsum(np.concatenate([[0.0], np.full(max(n - 1, 0), phase_inc)]))[:n]
    return (np.cos(ph) + 1j * np.sin(ph)).astype(np.complex64)


def MakeLowpassTaps(n):
    """test_harness.go:36-42"""
    return np.full(n, np.float32(1.0) / np.float32(n), np.float32)


def makeToneNoiseIQ(n, phase_inc):
    """streaming_gpu_validation_helpers_test.go:16-27"""
    i = np.arange(n, dtype=np.float64)
    ph = np.cumsum(np.concatenate([[0.0], np.full(max(n - 1, 0), phase_inc)]))[:n]
    tone = np.cos(ph) + 1j * np.sin(ph)
    ni = 0.17 * np.cos(0.113 * i + 0.31) + 0.07 * np.sin(0.071 * i)
    nq = 0.13 * np.sin(0.097 * i + 0.11) - 0.05 * np.cos(0.043 * i)
    return (0.85 * tone + 0.15 * (ni + 1j * nq)).astype(np.complex64)


# ---------------------------------------------------------------- surveillance
CFAR_CA, CFAR_GOSCA, CFAR_CASO, CFAR_OS = 0, 1, 2, 3
_CFAR_MODES = {"CA": CFAR_CA, "GOSCA": CFAR_GOSCA, "CASO": CFAR_CASO, "OS": CFAR_OS}


class orc_detector_cfg(C.Structure):
    _fields_ = [("threshold_db", C.c_double), ("hysteresis_db", C.c_double), ("cfar_scale_db", C.c_double),
                ("edge_margin_db", C.c_double), ("max_signal_bw_hz", C.c_double), ("merge_gap_hz", C.c_double),
                ("bin_width", C.c_double), ("center_hz", C.c_double), ("nbins", C.c_int), ("pad", C.c_int)]


class orc_signal(C.Structure):
    _fields_ = [("first_bin", C.c_int), ("last_bin", C.c_int), ("peak_bin", C.c_int), ("pad", C.c_int),
                ("center_hz", C.c_double), ("bw_hz", C.c_double), ("peak_db", C.c_double), ("snr_db", C.c_double),
                ("noise_db", C.c_double)]


register("orc_hann", None, [I, P])
register("orc_fft", None, [P, I])
register("orc_spectrum", None, [P, I, P, I, P])
register("orc_spectrum_from_fft", None, [P, I, P])
register("orc_welch_psd", I, [P, I, I, D, P, I, P])
register("orc_cfar_thresholds", I, [I, I, I, I, D, I, P, I, P])
register("orc_median", D, [P, I])
register("orc_detect_signals", I, [C.POINTER(orc_detector_cfg), P, P, I, C.POINTER(orc_signal), I, PD_])


def Hann(n):
    w = np.zeros(max(n, 0))
    lib().orc_hann(n, _p(w))
    return w


def FFT(x):
    a = np.ascontiguousarray(x, np.complex128).copy()
    lib().orc_fft(_p(a), len(a))
    return a


def Spectrum(iq, window=None):
    iq = c64(iq)
    w = np.ascontiguousarray(window if window is not None else np.zeros(0), np.float64)
    out = np.zeros(len(iq))
    lib().orc_spectrum(_p(iq), len(iq), _p(w), len(w), _p(out))
    return out


def SpectrumFromFFT(X):
    X = c64(X)
    out = np.zeros(len(X))
    lib().orc_spectrum_from_fft(_p(X), len(X), _p(out))
    return out


def SurveillanceLevelSpectrum(iq, sample_rate, decim):
    """pipeline_runtime.go:624-645 (derived level) with decimationTaps /
    decimateSurveillanceIQ (399-426): ApplyFIR with LowpassFIR((sample_rate //
    decim) * 0.5 * 0.8, sample_rate, 101), Decimate, then the level's
    spectrum (spectrumForLevel, 379-389: Hann of the level's size, FP64 FFT,
    dB, fftshift).  iq is the level's n_fft * decim input samples."""
    taps = LowpassFIR(float(sample_rate // decim) * 0.5 * 0.8, sample_rate, 101)
    d = Decimate(ApplyFIR(iq, taps), decim)
    return Spectrum(d, Hann(len(d)))


def WelchPSD(iq, seg, overlap, window=None):
    iq = c64(iq)
    w = np.ascontiguousarray(window if window is not None else np.zeros(0), np.float64)
    out = np.zeros(max(seg, len(iq)))
    segs = lib().orc_welch_psd(_p(iq), len(iq), seg, overlap, _p(w), len(w), _p(out))
    return out[:seg] if segs > 0 else out[:len(iq)]


def CFARThresholds(mode, spectrum, guard, train, scale_db, wrap=True, rank=0):
    s = np.ascontiguousarray(spectrum, np.float64)
    out = np.zeros(len(s))
    m = _CFAR_MODES[mode] if isinstance(mode, str) else mode
    rc = lib().orc_cfar_thresholds(m, guard, train, rank, scale_db, 1 if wrap else 0, _p(s), len(s), _p(out))
    if rc != 0:
        return None
    return out


def median(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().orc_median(_p(v), len(v))


@dataclass
class Signal:
    FirstBin: int
    LastBin: int
    PeakBin: int
    CenterHz: float
    BWHz: float
    PeakDb: float
    SNRDb: float
    NoiseDb: float


def DetectSignals(spectrum, thresholds, *, sample_rate, center_hz=0.0, threshold_db=-20.0, hysteresis_db=3.0,
                  cfar_scale_db=6.0, edge_margin_db=3.0, max_signal_bw_hz=150000.0, merge_gap_hz=5000.0):
    """detector.go:408-481 threshold path on an already smoothed spectrum."""
    s = np.ascontiguousarray(spectrum, np.float64)
    n = len(s)
    cfg = orc_detector_cfg(threshold_db, hysteresis_db, cfar_scale_db, edge_margin_db, max_signal_bw_hz,
                           merge_gap_hz, sample_rate / n, center_hz, n, 0)
    thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.float64)
    cap = 4096
    out = (orc_signal * cap)()
    ng = C.c_double()
    m = lib().orc_detect_signals(C.byref(cfg), _p(s), _p(thr), n, out, cap, C.byref(ng))
    return [Signal(o.first_bin, o.last_bin, o.peak_bin, o.center_hz, o.bw_hz, o.peak_db, o.snr_db, o.noise_db)
            for o in out[:min(m, cap)]]


# ---------------------------------------------------------------- WFM stereo / RDS
class orc_fir_state(C.Structure):
    _fields_ = [("taps", C.c_double * 103), ("delay", C.c_double * 103), ("nt", C.c_int), ("pad", C.c_int)]


class orc_stereo_state(C.Structure):
    _fields_ = [("lpf", orc_fir_state), ("bp_hi", orc_fir_state), ("bp_lo", orc_fir_state),
                ("lr_lpf", orc_fir_state), ("pilot_hi", orc_fir_state), ("pilot_lo", orc_fir_state),
                ("phase", C.c_double), ("freq", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("err_avg", C.c_double), ("i_state", C.c_double), ("q_state", C.c_double),
                ("lp_alpha", C.c_double), ("rate", C.c_int), ("on_count", C.c_int), ("off_count", C.c_int),
                ("enabled", C.c_int)]


RDS_MAX_GROUPS = 4096


class orc_rds_decoder(C.Structure):
    _fields_ = [("ps", C.c_char * 8), ("rt", C.c_char * 64), ("pi", C.c_uint16), ("pad0", C.c_uint16),
                ("costas_phase", C.c_double), ("costas_freq", C.c_double), ("mu", C.c_double),
                ("prev", C.c_float * 2), ("prev_dec", C.c_float), ("has_prev", C.c_int), ("last_rate", C.c_int),
                ("last_hard", C.c_int), ("has_last_hard", C.c_int), ("total_decodes", C.c_int),
                ("groups_found", C.c_int), ("n_groups", C.c_int),
                ("groups", (C.c_uint16 * 4) * RDS_MAX_GROUPS)]


register("orc_pll_coefficients", None, [D, D, I, PD_, PD_])
register("orc_stereo_init", None, [C.POINTER(orc_stereo_state), I])
register("orc_stereo_decode", I, [C.POINTER(orc_stereo_state), P, I, P])
register("orc_stereo_decode_batch", I, [P, I, I, P])
register("orc_stereo_pilot_present", I, [P, I, I, PD_])
register("orc_stereo_pilot_estimate", I, [P, I, I, PD_, PD_, P])
register("orc_rds_syndrome", C.c_uint16, [C.c_uint32])
register("orc_rds_init", None, [C.POINTER(orc_rds_decoder)])
register("orc_rds_decode", I, [C.POINTER(orc_rds_decoder), P, I, I])
register("orc_rds_block", C.c_uint32, [C.c_uint16, C.c_uint16])
register("orc_rds_encode_groups", I, [C.c_uint16, C.c_char_p, C.c_char_p, I, P])


def pllCoefficients(loop_bw, damping, sample_rate):
    a, b = C.c_double(), C.c_double()
    lib().orc_pll_coefficients(loop_bw, damping, sample_rate, C.byref(a), C.byref(b))
    return a.value, b.value


class StereoDecoder:
    """streamer.go:1525-1643 stereoDecodeStateful + lock hysteresis."""

    def __init__(self, sample_rate):
        self.st = orc_stereo_state()
        lib().orc_stereo_init(C.byref(self.st), sample_rate)

    def Decode(self, mono):
        mono = np.ascontiguousarray(mono, np.float32)
        out = np.zeros(2 * len(mono), np.float32)
        locked = lib().orc_stereo_decode(C.byref(self.st), _p(mono), len(mono), _p(out))
        return out, bool(locked)

    @property
    def Enabled(self):
        return bool(self.st.enabled)


def StereoDecodeBatch(mono, sample_rate):
    """wfm_batch_stereo.go:76-157 wfmStereoDecodeBatch -> (L/R interleaved, locked)."""
    mono = np.ascontiguousarray(mono, np.float32)
    out = np.zeros(2 * len(mono), np.float32)
    locked = lib().orc_stereo_decode_batch(_p(mono), len(mono), sample_rate, _p(out))
    return out, bool(locked)


def StereoPilotPresent(iq, sample_rate):
    iq = c64(iq)
    r = C.c_double()
    ok = lib().orc_stereo_pilot_present(_p(iq), len(iq), sample_rate, C.byref(r))
    return bool(ok), r.value


def StereoPilotEstimate(iq, sample_rate):
    """(present, ratio, offset_hz, 201 Goertzel magnitudes) of estimateWFMPilot
    on the last sample_rate/2 samples (pll.go:82-131)."""
    iq = c64(iq)
    r, o = C.c_double(), C.c_double()
    mags = np.zeros(201, np.float64)
    ok = lib().orc_stereo_pilot_estimate(_p(iq), len(iq), sample_rate, C.byref(r), C.byref(o),
                                         mags.ctypes.data)
    return bool(ok), r.value, o.value, mags


def crcSyndrome(raw):
    return int(lib().orc_rds_syndrome(raw))


def RDSBlock(data, offset):
    return int(lib().orc_rds_block(data, offset))


def RDSEncodeGroups(pi, ps, rt, n_groups):
    bits = np.zeros(104 * n_groups, np.uint8)
    ps8 = ps.encode().ljust(8)[:8]
    rt64 = rt.encode().ljust(64)[:64]
    n = lib().orc_rds_encode_groups(pi, ps8, rt64, n_groups, _p(bits))
    return bits[:n]


class RDSDecoder:
    """rds.go:10-264 Decoder."""

    def __init__(self):
        self.d = orc_rds_decoder()
        lib().orc_rds_init(C.byref(self.d))

    def Decode(self, samples, sample_rate):
        s = c64(samples)
        return lib().orc_rds_decode(C.byref(self.d), _p(s), len(s), sample_rate)

    @property
    def PI(self):
        return int(self.d.pi)

    @property
    def PS(self):
        return bytes(self.d.ps).decode("latin-1").rstrip(" ")

    @property
    def RT(self):
        return bytes(self.d.rt).decode("latin-1").rstrip(" ")

    def Groups(self):
        n = min(self.d.n_groups, RDS_MAX_GROUPS)
        return [tuple(self.d.groups[i]) for i in range(n)]


class orc_rds_stream(C.Structure):
    _fields_ = [("taps", C.c_double * 103), ("hr", C.c_double * 103), ("hi", C.c_double * 103),
                ("mix_num", C.c_longlong), ("count", C.c_longlong), ("nt", C.c_int),
                ("decim", C.c_int), ("rate", C.c_int), ("pad", C.c_int)]


register("orc_rds_stream_init", None, [C.POINTER(orc_rds_stream), I])
register("orc_rds_stream_process", I, [C.POINTER(orc_rds_stream), P, I, P])


class RDSBasebandStream:
    """fm.go:234-271 carried across blocks (mixer phase, FIR history, decimation phase)."""

    def __init__(self, sample_rate):
        self.s = orc_rds_stream()
        lib().orc_rds_stream_init(C.byref(self.s), sample_rate)
        self.rate = sample_rate // self.s.decim

    def Process(self, disc):
        d = np.ascontiguousarray(disc, np.float32)
        out = np.zeros(len(d) // max(self.s.decim, 1) + 2, np.complex64)
        m = lib().orc_rds_stream_process(C.byref(self.s), _p(d), len(d), _p(out))
        return out[:m]


# ---- post-demodulation audio (internal/dsp/resample.go, streamer.go:1408-1470) ----
class orc_resampler(C.Structure):
    _fields_ = [("l", I), ("m", I), ("taps", I), ("out_time", I), ("bank", P), ("delay", P)]


register("orc_gcd", I, [I, I])
register("orc_windowed_sinc", None, [I, D, D, P])
register("orc_resampler_init", I, [C.POINTER(orc_resampler), I, I, I])
register("orc_resampler_free", None, [C.POINTER(orc_resampler)])
register("orc_resampler_process", I, [C.POINTER(orc_resampler), P, I, P])
register("orc_deemphasis", D, [P, I, I, D, D])


def gcd(a, b):
    return int(lib().orc_gcd(a, b))


def WindowedSinc(length, fc, gain):
    out = np.zeros(length, np.float64)
    lib().orc_windowed_sinc(length, fc, gain, _p(out))
    return out


class Resampler:
    """dsp.Resampler (resample.go:28-170): stateful rational polyphase L/M."""

    def __init__(self, in_rate, out_rate, taps_per_phase=32):
        self._r = orc_resampler()
        lib().orc_resampler_init(C.byref(self._r), in_rate, out_rate, taps_per_phase)

    def __del__(self):
        try:
            lib().orc_resampler_free(C.byref(self._r))
        except Exception:
            pass

    def Ratio(self):
        return self._r.l, self._r.m

    @property
    def taps(self):
        return self._r.taps

    def Bank(self):
        """polyBank[phase][tap] as an (l, taps) array."""
        n = self._r.l * self._r.taps
        return np.ctypeslib.as_array(C.cast(self._r.bank, C.POINTER(C.c_double)), (n,)).reshape(self._r.l, -1).copy()

    def Process(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x) * self._r.l // self._r.m + 8, np.float32)
        k = lib().orc_resampler_process(C.byref(self._r), _p(x), len(x), _p(out))
        return out[:k].copy()


class StereoResampler:
    """dsp.StereoResampler (resample.go:174-233): two resamplers, interleaved
    in and out, the shorter length kept."""

    def __init__(self, in_rate, out_rate, taps_per_phase=32):
        self.left = Resampler(in_rate, out_rate, taps_per_phase)
        self.right = Resampler(in_rate, out_rate, taps_per_phase)

    def Process(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n = len(x) // 2
        if n == 0:
            return np.zeros(0, np.float32)
        a = self.left.Process(x[0:2 * n:2])
        b = self.right.Process(x[1:2 * n:2])
        k = min(len(a), len(b))
        out = np.empty(2 * k, np.float32)
        out[0::2] = a[:k]
        out[1::2] = b[:k]
        return out


def Deemphasis(x, alpha, y, stride=1, offset=0):
    """streamer.go:1443-1467 one channel (in place on a float32 copy); returns
    (samples, new state)."""
    x = np.ascontiguousarray(x, np.float32).copy()
    if len(x) == 0:
        return x, y
    y = lib().orc_deemphasis(x[offset:].ctypes.data, (len(x) - offset + stride - 1) // stride, stride, alpha, y)
    return x, y


# ---------------------------------------------------------------------------
# Job building of the DSP loop (cmd/sdrd) and the legacy block path
# ---------------------------------------------------------------------------
WFM_STREAM_OUT_RATE = 512000   # helpers.go:279
WFM_STREAM_MIN_BW = 250000.0   # helpers.go:280


def snapOutRate(sample_rate, target):
    """helpers.go:292-310: the divisor d <= 256 of sample_rate nearest
    sample_rate / target (first on ties), as sample_rate / d."""
    if sample_rate <= 0 or target <= 0:
        return target
    ideal = sample_rate / target
    best, best_err = 0, float("inf")
    for d in range(1, min(256, sample_rate) + 1):
        if sample_rate % d:
            continue
        e = abs(d - ideal)
        if e < best_err:
            best, best_err = d, e
    return target if best < 1 else sample_rate // best


def nearestExactDecimationRate(sample_rate, target_rate, min_rate):
    """streaming_refactor.go:105-137"""
    if sample_rate <= 0 or target_rate <= 0:
        return target_rate
    if sample_rate % target_rate == 0:
        return target_rate
    td = sample_rate // target_rate
    best_rate, best_dist = 0, sample_rate
    for d in range(max(1, td - 2), td + 3):
        rate = sample_rate // d
        if rate < min_rate or rate <= 0 or sample_rate % rate:
            continue
        dist = abs(target_rate - rate)
        if dist < best_dist:
            best_dist, best_rate = dist, rate
    return best_rate if best_rate > 0 else target_rate


def isWFM(center_hz, mod_type=""):
    """streaming_refactor.go:29-31: FM broadcast band or a WFM class."""
    mhz = center_hz / 1e6
    return (87.5 <= mhz <= 108.0) or mod_type in ("WFM", "WFM_STEREO")


def buildStreamingJobs(sample_rate, center_hz, signals, fir_taps=0, bw_mult=0.0, snap_wfm=False):
    """streaming_refactor.go:16-60.  signals: (id, center_hz, bw_hz, mod_type).
    Returns (jobs, failed_index): jobs as (id, offset, bw, out_rate, taps,
    hash), or (None, i) where the reference returns the
    ExactIntegerDecimation error.  snap_wfm snaps the WFM out-rate with
    snapOutRate."""
    if bw_mult <= 0:
        bw_mult = 1.0
    if fir_taps <= 0:
        fir_taps = 101
    jobs = []
    for i, (sid, cen, bw_hz, mod) in enumerate(signals):
        bw = bw_hz * bw_mult
        if isWFM(cen, mod):
            out_rate = snapOutRate(sample_rate, WFM_STREAM_OUT_RATE) if snap_wfm else WFM_STREAM_OUT_RATE
            if bw < WFM_STREAM_MIN_BW:
                bw = WFM_STREAM_MIN_BW
        else:
            out_rate = nearestExactDecimationRate(sample_rate, 200000, 128000)
            if bw < 20000:
                bw = 20000.0
        if lib().orc_exact_decim(sample_rate, out_rate) < 0:
            return None, i
        off = cen - center_hz
        jobs.append((sid, off, bw, out_rate, fir_taps,
                     StreamingConfigHash(sid, off, bw, out_rate, fir_taps, sample_rate)))
    return jobs, -1


def buildExtractJobs(sample_rate, center_hz, signals):
    """helpers.go:175-195: ExtractJob (offset, bw, out_rate, phase_start 0)
    of extractSignalIQBatch."""
    out = []
    for _, cen, bw, mod in signals:
        rate = 200000
        if isWFM(cen, mod):
            rate = WFM_STREAM_OUT_RATE
            if bw < WFM_STREAM_MIN_BW:
                bw = WFM_STREAM_MIN_BW
        elif bw < 20000:
            bw = 20000.0
        out.append((cen - center_hz, bw, snapOutRate(sample_rate, rate), 0.0))
    return out


def LegacyShiftFilterDecimate(iq, sample_rate, offset_hz, phase_start, taps, decim, n_out=None):
    """The legacy GPU chain in FP64 (test reference for the block path):
    gpud_freq_shift_kernel (exports.cu:37-60: sample i rotated by
    phase_start + inc * i, inc = -2 pi offset / fs), gpud_fir_kernel
    (exports.cu:127-161: causal FIR from zero history) and
    gpud_decimate_kernel (exports.cu:114-124: keep indices 0, D, 2D, ...),
    n / D outputs (batch_runner_windows.go:159-162).  Returns (IQ, phase_end)."""
    x = np.asarray(iq, np.complex128)
    n = len(x)
    inc = -2.0 * np.pi * offset_hz / sample_rate
    m = n // decim if n_out is None else n_out
    idx = np.arange(m) * decim
    s = x * np.exp(1j * (phase_start + inc * np.arange(n)))
    h = np.asarray(taps, np.float64)
    y = np.zeros(m, np.complex128)
    for k in range(len(h)):
        j = idx - k
        ok = j >= 0
        y[ok] += h[k] * s[j[ok]]
    return y, phase_start + inc * n


def LegacyDemod(dec, mode, sample_rate, bfo_phase=0.0):
    """The engine's demodulators on the decimated chain output (FP64):
    gpud_fm_discrim_kernel (exports.cu:87-112, n-1 outputs),
    gpud_am_envelope_kernel (|x|, exports.cu:273-294), and
    gpud_ssb_product_kernel (Re(x e^{j(bfo_phase + inc i)}),
    exports.cu:296-326) with inc = 2 pi (+-700) / sample_rate, the engine's
    input rate (gpudemod_windows.go:304, 508).  mode: DemodType order
    NFM, WFM, AM, USB, LSB, CW.  Returns (audio, bfo_phase_end)."""
    d = np.asarray(dec, np.complex128)
    if mode in (0, 1):
        return np.angle(np.conj(d[:-1]) * d[1:]), bfo_phase
    if mode == 2:
        return np.abs(d), bfo_phase
    inc = 2.0 * np.pi * (-700.0 if mode == 4 else 700.0) / sample_rate
    out = np.real(d * np.exp(1j * (bfo_phase + inc * np.arange(len(d)))))
    return out, bfo_phase + inc * len(d)


# ---------------------------------------------------------------------------
# Detector settings and smoothing (detector.go:101-278, 806-817)
# ---------------------------------------------------------------------------
def go_round(x):
    """math.Round: half away from zero."""
    return int(np.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def DetectorSettings(sample_rate, fft_size, guard_hz, train_hz, rank=0, ema_alpha=0.0, hysteresis_db=0.0,
                     scale_db=0.0, edge_margin_db=0.0, max_signal_bw_hz=0.0, merge_gap_hz=0.0):
    """detector.New (detector.go:101-230): dict of the mapped settings."""
    n = fft_size if 0 < fft_size <= 4194304 else 65536
    bw = sample_rate / n
    guard = max(0, int(np.ceil(guard_hz / bw)))
    train = max(1, int(np.ceil(train_hz / bw)))
    if rank <= 0 or rank > 2 * train:
        rank = go_round(0.75 * float(2 * train))
        if rank <= 0:
            rank = 1
    return dict(fft=n, guard=guard, train=train, rank=rank,
                ema_alpha=ema_alpha if 0 < ema_alpha <= 1 else 0.2,
                hysteresis_db=hysteresis_db if hysteresis_db > 0 else 3.0,
                scale_db=scale_db if scale_db > 0 else 6.0,
                edge_margin_db=edge_margin_db if edge_margin_db > 0 else 3.0,
                max_signal_bw_hz=max_signal_bw_hz if max_signal_bw_hz > 0 else 150000.0,
                merge_gap_hz=merge_gap_hz if merge_gap_hz > 0 else 5000.0)


def AdaptiveAlpha(ema_alpha, dt):
    """detector.go:263-278 (dt <= 0 or > 1 s -> 1/15 s)."""
    if not dt > 0 or dt > 1.0:
        dt = 1.0 / 15.0
    a = 1.0 - (1.0 - ema_alpha) ** (dt / (1.0 / 15.0))
    return min(max(a, 0.01), 0.99)


class SmoothSpectrum:
    """detector.go:806-817: the EMA of a detector, FP64 per bin; the first
    frame (or a size change) is copied."""

    def __init__(self):
        self.ema = None

    def __call__(self, spectrum, alpha):
        v = np.asarray(spectrum, np.float64)
        if self.ema is None or len(self.ema) != len(v):
            self.ema = v.copy()
            return self.ema.copy()
        self.ema = alpha * v + (1 - alpha) * self.ema
        return self.ema.copy()


# ---------------------------------------------------------------------------
# Output rings (streaming_out_ring.go:22-112), pure Python
# ---------------------------------------------------------------------------
OUT_RING_DEPTH = 4          # streaming_out_ring.go:11
OUT_POOL_MAX = 4 * 32       # streaming_out_ring.go:82


class OutBuf:
    """A host buffer of the model: an identity and a capacity."""
    _next = 0

    def __init__(self, cap):
        OutBuf._next += 1
        self.uid = OutBuf._next
        self.cap = cap


class OutBufPool:
    """outBufPool (streaming_out_ring.go:51-89); `allocated` counts make()s."""

    def __init__(self):
        self.free = []
        self.allocated = 0

    def get(self, n):
        best = -1
        for i, b in enumerate(self.free):
            if b.cap >= n:
                best = i
                break
            if best < 0 or b.cap > self.free[best].cap:
                best = i
        if best >= 0:
            b = self.free[best]
            self.free[best] = self.free[-1]
            self.free.pop()
            if b.cap >= n:
                return b
        self.allocated += 1
        return OutBuf(n)

    def put(self, bufs):
        for b in bufs:
            if b is None or len(self.free) >= OUT_POOL_MAX:
                continue
            self.free.append(b)


class OutRings:
    """streamOutRing per signal ID with outRingFor / pruneOutRings
    (streaming_out_ring.go:22-45, 91-112)."""

    def __init__(self):
        self.rings = {}
        self.pool = OutBufPool()

    def next(self, sid, n):
        ring = self.rings.setdefault(sid, [[None] * OUT_RING_DEPTH, 0])
        bufs, idx = ring
        b = bufs[idx]
        if b is None or b.cap < n:
            if not (b is None and n == 0):  # cap(nil) < 0 is false: a nil slot serves n = 0
                b = self.pool.get(n)
        bufs[idx] = b
        ring[1] = (idx + 1) % OUT_RING_DEPTH
        return b

    def prune(self, active):
        # Go visits the map in random order; the product and this model use
        # ascending IDs so the bounded pool's contents are reproducible
        for sid in sorted(s for s in self.rings if s not in set(active)):
            self.pool.put(self.rings[sid][0])
            del self.rings[sid]


# ---------------------------------------------------------------------------
# The RDS caller flow (pipeline_runtime.go:1044-1176, helpers.go:117-160)
# ---------------------------------------------------------------------------
class RDSRuntime:
    """updateRDS and maintenance's prune over the oracle chain.  Update takes
    the signals of one tick in order; each due signal's goroutine body runs
    right after the bookkeeping of the tick (its result visible on return).
    `last[id]` keeps (baseband, rate, rds baseband, rds rate) of the last
    decode, `dec[id]` the decoder."""

    def __init__(self):
        self.states = {}
        self.last = {}

    def _extract(self, iq, fs, offset):
        """extractRDS (helpers.go:122-160) via ShiftFilterDecimateBatch, or
        the fallback chain of pipeline_runtime.go:1104-1124."""
        d = max(1, fs // 250000)
        out_rate = fs // d
        if out_rate > 0 and fs % out_rate == 0:
            taps = LowpassFIR(max(200000 / 2, 200), fs, 101)
            y, _ = LegacyShiftFilterDecimate(iq, fs, offset, 0.0, taps, d)
            return y.astype(np.complex64), fs // d
        shifted = FreqShift(iq, fs, offset)
        d1 = max(1, fs // 1000000)
        f1 = ApplyFIR(shifted, LowpassFIR(float(fs // d1) / 2.0 * 0.8, fs, 51))
        x1 = Decimate(f1, d1)
        rate1 = fs // d1
        d2 = max(1, rate1 // 250000)
        f2 = ApplyFIR(x1, LowpassFIR(float(rate1 // d2) / 2.0 * 0.8, rate1, 101))
        return Decimate(f2, d2), rate1 // d2

    def Update(self, now, signals, window, fs, center):
        """-> [(stereo, station, turn)] per (id, center_hz)."""
        due = []
        turns = []
        for sid, hz in signals:
            if sid == 0:
                turns.append(-1)
                continue
            st = self.states.get(sid)
            if st is None:
                st = self.states[sid] = {"last_seen": 0.0, "last_decode": -np.inf, "stereo": 0, "ps": None,
                                         "pi": 0, "rt": "", "dec": RDSDecoder()}
            st["last_seen"] = now
            for other in [k for k, v in self.states.items() if now - v["last_seen"] > 30.0]:
                del self.states[other]
            if now - st["last_decode"] >= 4.0:
                st["last_decode"] = now
                due.append((sid, hz))
                turns.append(1)
            else:
                turns.append(0)
        if due and fs > 0 and len(window) >= fs:
            for sid, hz in due:
                st = self.states[sid]
                base, rate = self._extract(window, fs, hz - center)
                st["stereo"] = int(StereoPilotPresent(base, rate)[0])
                rds, rrate = RDSBasebandComplex(base, rate)
                self.last[sid] = (base, rate, rds, rrate)
                if len(rds) == 0:
                    continue
                st["dec"].Decode(rds, rrate)
                if st["dec"].PS:
                    st["ps"], st["pi"], st["rt"] = st["dec"].PS, st["dec"].PI, st["dec"].RT
        out = []
        for (sid, hz), turn in zip(signals, turns):
            st = self.states.get(sid) if sid else None
            stereo = st["stereo"] if st else 0
            station = st["ps"].strip(" ") if st and st["ps"] else ""
            out.append((stereo, station, turn))
        return out

    def Maintain(self, active):
        for sid in [s for s in self.states if s not in set(active)]:
            del self.states[sid]

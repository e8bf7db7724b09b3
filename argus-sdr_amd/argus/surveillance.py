"""Surveillance binding: Welch PSD, FFT, CFAR runs and the detection list
(gpud_surv_* in include/argus_gpu.h).  Mirrors internal/fft (WelchPSD,
Spectrum), internal/fft/gpufft (Engine.Exec) and the threshold path of
internal/detector."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, register

CFAR_CA, CFAR_GOSCA, CFAR_CASO, CFAR_OS = 0, 1, 2, 3
_MODES = {"CA": CFAR_CA, "GOSCA": CFAR_GOSCA, "CASO": CFAR_CASO, "OS": CFAR_OS}


class gpud_cfar_params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("guard", C.c_int32), ("train", C.c_int32), ("wrap", C.c_int32),
                ("scale_db", C.c_double), ("hysteresis_db", C.c_double), ("rank", C.c_int32), ("pad", C.c_int32)]


class gpud_surv_run(C.Structure):
    _fields_ = [("first_bin", C.c_int32), ("last_bin", C.c_int32), ("peak_bin", C.c_int32), ("pad", C.c_int32),
                ("peak_db", C.c_double), ("noise_db", C.c_double)]


class gpud_surv_signal(C.Structure):
    _fields_ = [("first_bin", C.c_int32), ("last_bin", C.c_int32), ("peak_bin", C.c_int32), ("pad", C.c_int32),
                ("center_hz", C.c_double), ("bw_hz", C.c_double), ("peak_db", C.c_double), ("snr_db", C.c_double),
                ("noise_db", C.c_double)]


class gpud_detector_params(C.Structure):
    _fields_ = [("threshold_db", C.c_double), ("hysteresis_db", C.c_double), ("cfar_scale_db", C.c_double),
                ("edge_margin_db", C.c_double), ("max_signal_bw_hz", C.c_double), ("merge_gap_hz", C.c_double),
                ("center_hz", C.c_double)]


P, I, L64 = C.c_void_p, C.c_int, C.c_int64
register("gpud_surv_create", I, [I, I, C.POINTER(P)])
register("gpud_surv_destroy", I, [P])
register("gpud_surv_stream", P, [P])
register("gpud_surv_welch_device", I, [P, P, L64, I, I, I, P])
register("gpud_surv_fft_device", I, [P, P, L64, I, I, P])
register("gpud_surv_spectrum_device", I, [P, P, L64, I, I, P])
register("gpud_surv_cfar_runs_device", I, [P, P, I, C.POINTER(gpud_cfar_params), P, P, I])
register("gpud_surv_cfar_runs_device_ex", I, [P, P, I, C.POINTER(gpud_cfar_params), P, P, I, P])
register("gpud_surv_finish_frame", I, [P, I, I, P, I, C.POINTER(gpud_detector_params), P, I])
register("gpud_surv_finish_device", I, [P, P, I, P, P, I, C.POINTER(gpud_detector_params), P, I, P])
register("gpud_surv_level_spectrum_device", I, [P, P, L64, I, I, P])
register("gpud_surv_detect", I, [P, P, L64, I, I, I, C.POINTER(gpud_cfar_params), C.POINTER(gpud_detector_params),
                                 P, I, P, P])


class gpud_detector_config(C.Structure):
    _fields_ = [("threshold_db", C.c_double), ("ema_alpha", C.c_double), ("hysteresis_db", C.c_double),
                ("cfar_guard_hz", C.c_double), ("cfar_train_hz", C.c_double), ("cfar_scale_db", C.c_double),
                ("edge_margin_db", C.c_double), ("max_signal_bw_hz", C.c_double), ("merge_gap_hz", C.c_double),
                ("cfar_mode", C.c_int32), ("cfar_rank", C.c_int32), ("cfar_wrap", C.c_int32), ("pad", C.c_int32)]


register("gpud_detector_settings", I, [C.POINTER(gpud_detector_config), I, I, C.POINTER(gpud_cfar_params),
                                       C.POINTER(gpud_detector_params), C.POINTER(C.c_double), C.POINTER(I)])
register("gpud_detector_adaptive_alpha", C.c_double, [C.c_double, C.c_double])
register("gpud_surv_smooth_device", I, [P, P, I, P, I, P, P])


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


class DeviceBuffer:
    """Device allocation through the C ABI (no torch)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check("gpud_malloc", lib().gpud_malloc(C.byref(p), max(nbytes, 1)))
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check("gpud_memcpy_h2d", lib().gpud_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self, dtype, count) -> np.ndarray:
        out = np.zeros(count, dtype)
        check("gpud_memcpy_d2h", lib().gpud_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().gpud_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cfar_params(mode="CA", guard=4, train=16, scale_db=12.3748, wrap=True, hysteresis_db=3.0, rank=0):
    """rank (OS only): 1-based rank in the 2*train cells; <= 0 takes the
    detector's default round(0.75 * 2 * train) (detector.go:203-207)."""
    if rank <= 0:
        rank = max(1, int(math.floor(0.75 * 2 * train + 0.5)))
    return gpud_cfar_params(_MODES[mode] if isinstance(mode, str) else mode, guard, train, 1 if wrap else 0,
                            scale_db, hysteresis_db, rank, 0)


def detector_params(threshold_db=-20.0, hysteresis_db=3.0, cfar_scale_db=12.3748, edge_margin_db=3.0,
                    max_signal_bw_hz=150000.0, merge_gap_hz=5000.0, center_hz=0.0):
    return gpud_detector_params(threshold_db, hysteresis_db, cfar_scale_db, edge_margin_db, max_signal_bw_hz,
                                merge_gap_hz, center_hz)


class Surveillance:
    """gpufft.Engine successor plus the Welch / CFAR / detection-list pass."""

    def __init__(self, n_fft: int, sample_rate: int):
        h = C.c_void_p()
        rc = lib().gpud_surv_create(n_fft, sample_rate, C.byref(h))
        check("gpud_surv_create", rc)
        self._h = h
        self.n = n_fft
        self.sample_rate = sample_rate

    def Close(self):
        if self._h:
            lib().gpud_surv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def sync(self):
        check("gpud_stream_sync", lib().gpud_stream_sync(lib().gpud_surv_stream(self._h)))

    def Exec(self, iq, window=False) -> np.ndarray:
        """gpu.go:67-91: forward unnormalised FFT of one or more frames."""
        iq = np.ascontiguousarray(iq, np.complex64)
        frames = len(iq) // self.n
        din = DeviceBuffer(iq.nbytes).upload(iq)
        dout = DeviceBuffer(iq.nbytes)
        check("gpud_surv_fft_device", lib().gpud_surv_fft_device(self._h, din.ptr, self.n, frames, 1 if window else 0,
                                                                 dout.ptr))
        self.sync()
        return dout.download(np.complex64, frames * self.n)

    def SpectrumDB(self, iq, window=True) -> np.ndarray:
        """spectrumFromIQWithPlan's GPU branch: window, FFT, SpectrumFromFFT
        (dB, fftshifted) and NaN/Inf sanitising, for each n-sample frame."""
        iq = np.ascontiguousarray(iq, np.complex64)
        frames = len(iq) // self.n
        din = DeviceBuffer(iq.nbytes).upload(iq)
        dout = DeviceBuffer(4 * self.n * frames)
        check("gpud_surv_spectrum_device", lib().gpud_surv_spectrum_device(self._h, din.ptr, self.n, frames,
                                                                           1 if window else 0, dout.ptr))
        self.sync()
        return dout.download(np.float32, frames * self.n).reshape(frames, self.n)

    def LevelSpectrumDB(self, iq, decim: int = 2, frame_step: int = 0, n_frames: int = 1) -> np.ndarray:
        """The decimated surveillance level's dB spectra (the multi-resolution
        derived level of pipeline_runtime.go:624-645): frame f is the n_fft
        samples at f * frame_step, low-passed, decimated by decim and
        transformed at n_fft / decim points."""
        iq = np.ascontiguousarray(iq, np.complex64)
        din = DeviceBuffer(iq.nbytes).upload(iq)
        n_d = self.n // decim
        dout = DeviceBuffer(4 * n_d * n_frames)
        check("gpud_surv_level_spectrum_device",
              lib().gpud_surv_level_spectrum_device(self._h, din.ptr, frame_step or self.n, n_frames, decim, dout.ptr))
        self.sync()
        return dout.download(np.float32, n_d * n_frames).reshape(n_frames, n_d)

    def WelchDevice(self, iq_dev: int, frame_step: int, seg_hop: int, n_segs: int, n_frames: int, psd_dev: int):
        check("gpud_surv_welch_device", lib().gpud_surv_welch_device(self._h, iq_dev, frame_step, seg_hop, n_segs,
                                                                     n_frames, psd_dev))

    def WelchPSD(self, iq, n_segs: int, seg_hop: int, frame_step: int, n_frames: int) -> np.ndarray:
        iq = np.ascontiguousarray(iq, np.complex64)
        din = DeviceBuffer(iq.nbytes).upload(iq)
        dpsd = DeviceBuffer(4 * self.n * n_frames)
        self.WelchDevice(din.ptr, frame_step, seg_hop, n_segs, n_frames, dpsd.ptr)
        self.sync()
        return dpsd.download(np.float32, self.n * n_frames).reshape(n_frames, self.n)

    def CFARRuns(self, psd, cfar: gpud_cfar_params, max_runs: int = 2048, ambiguous: bool = False):
        """(run counts, runs per frame), plus the per-frame count of bins whose
        decision the reference's running sums could take the other way when
        ambiguous is set."""
        psd = np.ascontiguousarray(psd, np.float32).reshape(-1, self.n)
        nf = psd.shape[0]
        dpsd = DeviceBuffer(psd.nbytes).upload(psd)
        druns = DeviceBuffer(C.sizeof(gpud_surv_run) * max_runs * nf)
        dcnt = DeviceBuffer(4 * nf)
        damb = DeviceBuffer(4 * nf) if ambiguous else None
        check("gpud_surv_cfar_runs_device_ex",
              lib().gpud_surv_cfar_runs_device_ex(self._h, dpsd.ptr, nf, C.byref(cfar), druns.ptr, dcnt.ptr, max_runs,
                                                  damb.ptr if damb else None))
        self.sync()
        counts = dcnt.download(np.int32, nf)
        raw = (gpud_surv_run * (max_runs * nf))()
        check("gpud_memcpy_d2h", lib().gpud_memcpy_d2h(C.addressof(raw), druns.ptr, C.sizeof(raw)))
        out = []
        for f in range(nf):
            out.append([raw[f * max_runs + i] for i in range(min(int(counts[f]), max_runs))])
        if ambiguous:
            return counts, out, damb.download(np.int32, nf)
        return counts, out

    def DetectSpectra(self, psd, cfar, det, cap=1024, max_runs=2048):
        """CFAR runs and the device detection list (expansion, merge,
        centroid) of given dB spectra [frames][n]: (lists, signal counts)."""
        psd = np.ascontiguousarray(psd, np.float32).reshape(-1, self.n)
        nf = psd.shape[0]
        dpsd = DeviceBuffer(psd.nbytes).upload(psd)
        druns = DeviceBuffer(C.sizeof(gpud_surv_run) * max_runs * nf)
        dcnt = DeviceBuffer(4 * nf)
        dsig = DeviceBuffer(C.sizeof(gpud_surv_signal) * cap * nf)
        dout = DeviceBuffer(4 * nf)
        check("gpud_surv_cfar_runs_device", lib().gpud_surv_cfar_runs_device(self._h, dpsd.ptr, nf, C.byref(cfar),
                                                                             druns.ptr, dcnt.ptr, max_runs))
        check("gpud_surv_finish_device", lib().gpud_surv_finish_device(self._h, dpsd.ptr, nf, druns.ptr, dcnt.ptr,
                                                                       max_runs, C.byref(det), dsig.ptr, cap,
                                                                       dout.ptr))
        self.sync()
        counts = dout.download(np.int32, nf)
        raw = (gpud_surv_signal * (cap * nf))()
        check("gpud_memcpy_d2h", lib().gpud_memcpy_d2h(C.addressof(raw), dsig.ptr, C.sizeof(raw)))
        lists = [[Signal(x.first_bin, x.last_bin, x.peak_bin, x.center_hz, x.bw_hz, x.peak_db, x.snr_db, x.noise_db)
                  for x in raw[f * cap:f * cap + min(int(counts[f]), cap)]] for f in range(nf)]
        return lists, counts

    def FinishFrameHost(self, psd_row, runs, det, cap=1024):
        """The host detection list of one frame from its CFAR runs
        (gpud_surv_finish_frame): (list, signal count)."""
        row = np.ascontiguousarray(psd_row, np.float32)
        arr = (gpud_surv_run * max(len(runs), 1))(*runs)
        out = (gpud_surv_signal * cap)()
        m = lib().gpud_surv_finish_frame(row.ctypes.data, len(row), self.sample_rate, arr, len(runs), C.byref(det),
                                         out, cap)
        if m < 0:
            check("gpud_surv_finish_frame", m)
        return [Signal(x.first_bin, x.last_bin, x.peak_bin, x.center_hz, x.bw_hz, x.peak_db, x.snr_db, x.noise_db)
                for x in out[:min(m, cap)]], m

    def Detect(self, iq, n_segs, seg_hop, frame_step, n_frames, cfar, det, cap=1024, want_psd=True):
        iq = np.ascontiguousarray(iq, np.complex64)
        din = DeviceBuffer(iq.nbytes).upload(iq)
        sig = (gpud_surv_signal * (cap * n_frames))()
        counts = np.zeros(n_frames, np.int32)
        psd = np.zeros(n_frames * self.n, np.float32) if want_psd else None
        check("gpud_surv_detect", lib().gpud_surv_detect(self._h, din.ptr, frame_step, seg_hop, n_segs, n_frames,
                                                         C.byref(cfar), C.byref(det), C.addressof(sig), cap,
                                                         counts.ctypes.data,
                                                         psd.ctypes.data if psd is not None else None))
        lists = []
        for f in range(n_frames):
            lists.append([Signal(s.first_bin, s.last_bin, s.peak_bin, s.center_hz, s.bw_hz, s.peak_db, s.snr_db,
                                 s.noise_db) for s in sig[f * cap:f * cap + min(int(counts[f]), cap)]])
        return lists, (psd.reshape(n_frames, self.n) if psd is not None else None)


@dataclass
class DetectorConfig:
    """config.DetectorConfig (config.go:30-48): the fields the CFAR detection
    path reads.  Defaults are the reference's config defaults
    (config.go:468-481)."""
    ThresholdDb: float = -20.0
    EmaAlpha: float = 0.2
    HysteresisDb: float = 3.0
    CFARMode: str = "GOSCA"
    CFARGuardHz: float = 500.0
    CFARTrainHz: float = 5000.0
    CFARRank: int = 36
    CFARScaleDb: float = 6.0
    CFARWrapAround: bool = True
    EdgeMarginDb: float = 3.0
    MaxSignalBwHz: float = 150000.0
    MergeGapHz: float = 5000.0


def DetectorSettings(cfg: DetectorConfig, sample_rate: int, fft_size: int):
    """detector.New's mapping (detector.go:101-261): (cfar params, detector
    params, EMA alpha, FFT bins)."""
    c = gpud_detector_config(cfg.ThresholdDb, cfg.EmaAlpha, cfg.HysteresisDb, cfg.CFARGuardHz, cfg.CFARTrainHz,
                             cfg.CFARScaleDb, cfg.EdgeMarginDb, cfg.MaxSignalBwHz, cfg.MergeGapHz,
                             _MODES[cfg.CFARMode], cfg.CFARRank, 1 if cfg.CFARWrapAround else 0, 0)
    cf, dp = gpud_cfar_params(), gpud_detector_params()
    a, nb = C.c_double(), C.c_int()
    check("gpud_detector_settings", lib().gpud_detector_settings(C.byref(c), sample_rate, fft_size, C.byref(cf),
                                                                 C.byref(dp), C.byref(a), C.byref(nb)))
    return cf, dp, a.value, nb.value


def AdaptiveAlpha(ema_alpha: float, dt_seconds: float) -> float:
    """detector.go:263-278"""
    return float(lib().gpud_detector_adaptive_alpha(ema_alpha, dt_seconds))


class Detector:
    """The CFAR branch of detector.Process on the device for frames of one
    size: EMA smoothing with the frame-rate-adaptive alpha (smoothSpectrum,
    detector.go:806-817, state kept on the device), CFAR runs and the
    detection list (detectSignals, detector.go:408-481) per frame."""

    def __init__(self, cfg: DetectorConfig, sample_rate: int, fft_size: int, center_hz: float = 0.0):
        self.cfar, self.det, self.alpha, n = DetectorSettings(cfg, sample_rate, fft_size)
        self.det.center_hz = center_hz
        self.surv = Surveillance(n, sample_rate)
        self.n = n
        self.ema = DeviceBuffer(8 * n)
        self.first = True
        self.last_t = None

    def Smooth(self, psd, times=None) -> np.ndarray:
        """Smoothed frames of psd [frames][n] (dB, host), frame f taken at
        times[f] seconds (None: the reference's 1/15 s)."""
        psd = np.ascontiguousarray(psd, np.float32).reshape(-1, self.n)
        nf = psd.shape[0]
        alphas = np.zeros(nf, np.float64)
        for f in range(nf):
            t = None if times is None else float(times[f])
            dt = 0.0 if (t is None or self.last_t is None) else t - self.last_t
            alphas[f] = AdaptiveAlpha(self.alpha, dt)
            self.last_t = t
        dpsd = DeviceBuffer(psd.nbytes).upload(psd)
        dout = DeviceBuffer(psd.nbytes)
        check("gpud_surv_smooth_device", lib().gpud_surv_smooth_device(self.surv._h, dpsd.ptr, nf, alphas.ctypes.data,
                                                                       1 if self.first else 0, self.ema.ptr,
                                                                       dout.ptr))
        self.first = False
        self.surv.sync()
        return dout.download(np.float32, nf * self.n).reshape(nf, self.n)

    def Process(self, psd, times=None, cap=1024):
        """Smooth, then the detection list of each smoothed frame:
        (lists, smoothed frames)."""
        sm = self.Smooth(psd, times)
        lists, _ = self.surv.DetectSpectra(sm, self.cfar, self.det, cap=cap)
        return lists, sm

    def EMA(self) -> np.ndarray:
        return self.ema.download(np.float64, self.n)

    def Close(self):
        self.surv.Close()
        self.ema.free()

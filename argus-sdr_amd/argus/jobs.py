"""Job building of the reference's DSP loop (cmd/sdrd) over the C ABI.

    snapOutRate(sample_rate, target)                      helpers.go:292-310
    nearestExactDecimationRate(sample_rate, target, min)  streaming_refactor.go:105-137
    buildStreamingJobs(sample_rate, center_hz, signals, cfg)  streaming_refactor.go:16-60
    buildExtractJobs(sample_rate, center_hz, signals)     helpers.go:175-195

Signals are DetectedSignal records (the detector.Signal fields the builders
read).  buildStreamingJobs raises ValueError where the reference returns the
ExactIntegerDecimation error.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from ._lib import check, gpud_streaming_job, lib, register
from .gpudemod import ExtractJob, StreamingExtractJob, gpud_extract_job

MOD_OTHER, MOD_WFM, MOD_WFM_STEREO = 0, 1, 2
_MOD = {"WFM": MOD_WFM, "WFM_STEREO": MOD_WFM_STEREO}


class gpud_detected_signal(C.Structure):
    _fields_ = [("id", C.c_int64), ("center_hz", C.c_double), ("bw_hz", C.c_double), ("mod_type", C.c_int32),
                ("pad", C.c_int32)]


class gpud_extraction_config(C.Structure):
    _fields_ = [("fir_taps", C.c_int32), ("snap_wfm", C.c_int32), ("bw_mult", C.c_double)]


register("gpud_snap_out_rate", C.c_int, [C.c_int, C.c_int])
register("gpud_nearest_exact_decimation_rate", C.c_int, [C.c_int, C.c_int, C.c_int])
register("gpud_build_streaming_jobs", C.c_int,
         [C.c_int, C.c_double, C.POINTER(gpud_detected_signal), C.c_int, C.POINTER(gpud_extraction_config),
          C.POINTER(gpud_streaming_job), C.POINTER(C.c_int)])
register("gpud_build_extract_jobs", C.c_int,
         [C.c_int, C.c_double, C.POINTER(gpud_detected_signal), C.c_int, C.POINTER(gpud_extract_job)])


@dataclass
class DetectedSignal:
    """detector.Signal: ID, CenterHz, BWHz and Class.ModType ("" when unclassified)."""
    ID: int
    CenterHz: float
    BWHz: float
    ModType: str = ""


@dataclass
class ExtractionConfig:
    """extractionConfig (helpers.go:270-274); SnapWFM snaps the WFM out-rate."""
    FirTaps: int = 0
    BWMult: float = 0.0
    SnapWFM: bool = False


def snapOutRate(sample_rate: int, target: int) -> int:
    return int(lib().gpud_snap_out_rate(sample_rate, target))


def nearestExactDecimationRate(sample_rate: int, target_rate: int, min_rate: int) -> int:
    return int(lib().gpud_nearest_exact_decimation_rate(sample_rate, target_rate, min_rate))


def _signals(signals):
    arr = (gpud_detected_signal * max(1, len(signals)))()
    for i, s in enumerate(signals):
        arr[i].id = int(s.ID)
        arr[i].center_hz = float(s.CenterHz)
        arr[i].bw_hz = float(s.BWHz)
        arr[i].mod_type = _MOD.get(s.ModType, MOD_OTHER)
    return arr


def buildStreamingJobs(sample_rate: int, center_hz: float, signals, cfg: ExtractionConfig | None = None):
    cfg = cfg or ExtractionConfig()
    c = gpud_extraction_config(int(cfg.FirTaps), 1 if cfg.SnapWFM else 0, float(cfg.BWMult))
    jobs = (gpud_streaming_job * max(1, len(signals)))()
    bad = C.c_int(-1)
    rc = lib().gpud_build_streaming_jobs(sample_rate, center_hz, _signals(signals), len(signals), C.byref(c), jobs,
                                         C.byref(bad))
    if rc == -2:
        raise ValueError(f"streaming polyphase extractor requires integer decimation (signal index {bad.value})")
    check("gpud_build_streaming_jobs", rc)
    return [StreamingExtractJob(j.signal_id, j.offset_hz, j.bandwidth, j.out_rate, j.num_taps, j.config_hash)
            for j in jobs[:len(signals)]]


def buildExtractJobs(sample_rate: int, center_hz: float, signals):
    jobs = (gpud_extract_job * max(1, len(signals)))()
    check("gpud_build_extract_jobs", lib().gpud_build_extract_jobs(sample_rate, center_hz, _signals(signals),
                                                                   len(signals), jobs))
    return [ExtractJob(j.offset_hz, j.bw, j.out_rate, j.phase_start) for j in jobs[:len(signals)]]

"""ctypes loader for libargus_hip.so.

The library is required: there is no CPU fallback.  If it is missing or does
not load, every entry point raises ErrUnavailable (the reference's
gpudemod.ErrUnavailable, errors.go:5).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("ARGUS_HIP_LIB", Path(__file__).resolve().parent.parent / "libargus_hip.so"))


class ErrUnavailable(RuntimeError):
    """gpudemod batch runner unavailable (errors.go:5)."""


class GpudError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed with code {code}")
        self.code = code


class gpud_float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class gpud_streaming_job(C.Structure):
    _fields_ = [
        ("signal_id", C.c_int64),
        ("offset_hz", C.c_double),
        ("bandwidth", C.c_double),
        ("out_rate", C.c_int32),
        ("num_taps", C.c_int32),
        ("config_hash", C.c_uint64),
    ]


class gpud_streaming_result(C.Structure):
    _fields_ = [
        ("signal_id", C.c_int64),
        ("rate", C.c_int32),
        ("n_out", C.c_int32),
        ("phase_count", C.c_int32),
        ("history_len", C.c_int32),
        ("out_offset", C.c_int64),
        ("nco_phase", C.c_double),
    ]


_lib = None
_load_error: Exception | None = None

P = C.c_void_p
I = C.c_int
D = C.c_double
_SIGS = {
    "gpud_version": (C.c_char_p, []),
    "gpud_device_count": (I, [C.POINTER(I)]),
    "gpud_event_create": (I, [C.POINTER(P)]),
    "gpud_event_destroy": (I, [P]),
    "gpud_event_record": (I, [P, P]),
    "gpud_event_elapsed_ms": (I, [P, P, C.POINTER(C.c_float)]),
    "gpud_malloc": (I, [C.POINTER(P), C.c_size_t]),
    "gpud_free": (I, [P]),
    "gpud_host_alloc": (I, [C.POINTER(P), C.c_size_t]),
    "gpud_host_free": (I, [P]),
    "gpud_memcpy_h2d": (I, [P, P, C.c_size_t]),
    "gpud_memcpy_d2h": (I, [P, P, C.c_size_t]),
    "gpud_memcpy_d2d": (I, [P, P, C.c_size_t]),
    "gpud_memset_zero": (I, [P, C.c_size_t]),
    "gpud_device_sync": (I, []),
    "gpud_stream_create": (I, [C.POINTER(P)]),
    "gpud_stream_destroy": (I, [P]),
    "gpud_stream_sync": (I, [P]),
    "gpud_launch_freq_shift_stream_hip": (I, [P, P, I, D, D, P]),
    "gpud_launch_freq_shift_hip": (I, [P, P, I, D, D]),
    "gpud_launch_fm_discrim_hip": (I, [P, P, I]),
    "gpud_upload_fir_taps_hip": (I, [P, I]),
    "gpud_launch_fir_hip": (I, [P, P, I, I]),
    "gpud_launch_fir_stream_hip": (I, [P, P, I, I, P]),
    "gpud_launch_fir_v2_stream_hip": (I, [P, P, P, I, I, P]),
    "gpud_launch_decimate_hip": (I, [P, P, I, I]),
    "gpud_launch_decimate_stream_hip": (I, [P, P, I, I, P]),
    "gpud_launch_am_envelope_hip": (I, [P, P, I]),
    "gpud_launch_ssb_product_hip": (I, [P, P, I, D, D]),
    "gpud_launch_streaming_polyphase_prepare_hip": (
        I, [P, I, P, I, P, I, I, I, I, D, D, P, C.POINTER(I), C.POINTER(I), C.POINTER(D), P]),
    "gpud_launch_streaming_polyphase_stateful_hip": (
        I, [P, I, P, P, I, I, I, P, P, I, C.POINTER(I), C.POINTER(I), C.POINTER(D), D, P, I, C.POINTER(I)]),
    "gpud_runner_create": (I, [I, I, C.POINTER(P)]),
    "gpud_runner_destroy": (I, [P]),
    "gpud_runner_set_exec_mode": (I, [P, I]),
    "gpud_runner_set_window": (I, [P, I]),
    "gpud_runner_stream": (P, [P]),
    "gpud_runner_enable_kernel_timing": (I, [P, I]),
    "gpud_runner_kernel_time": (I, [P, C.POINTER(D), C.POINTER(I)]),
    "gpud_runner_streaming_extract": (
        I, [P, P, I, C.POINTER(gpud_streaming_job), I, P, C.c_int64, C.POINTER(gpud_streaming_result)]),
    "gpud_runner_streaming_extract_device": (
        I, [P, P, I, C.POINTER(gpud_streaming_job), I, C.POINTER(P), C.POINTER(gpud_streaming_result)]),
    "gpud_runner_get_signal_state": (
        I, [P, C.c_int64, C.POINTER(C.c_uint64), C.POINTER(D), C.POINTER(I), C.POINTER(I), C.POINTER(I), P, I]),
    "gpud_runner_num_states": (I, [P]),
    "gpud_runner_state_ids": (I, [P, P, I]),
    "gpud_runner_reset_signal_state": (I, [P, C.c_int64]),
    "gpud_runner_reset_all_signal_states": (I, [P]),
    "gpud_streaming_config_hash": (C.c_uint64, [C.c_int64, D, D, I, I, I]),
    "gpud_exact_integer_decimation": (I, [I, I]),
    "gpud_lowpass_fir": (I, [D, I, I, P]),
    "gpud_build_polyphase_taps": (I, [P, I, I, P]),
}

EXPORTS = tuple(_SIGS)


def lib():
    """The loaded library; raises ErrUnavailable if it cannot be loaded."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise ErrUnavailable(f"libargus_hip.so unavailable: {_load_error}")
    try:
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    except OSError as e:  # missing or unloadable library
        _load_error = e
        raise ErrUnavailable(f"libargus_hip.so unavailable: {e}") from e
    return _lib


def register(name: str, restype, argtypes) -> None:
    """Declare an additional entry point (used by the surveillance / serving /
    WFM modules) and include it in the export list checked by tests."""
    global EXPORTS
    _SIGS[name] = (restype, argtypes)
    EXPORTS = tuple(_SIGS)
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def check(fn: str, code: int) -> None:
    if code != 0:
        raise GpudError(fn, code)


def ptr(a) -> int | None:
    """Host address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data

"""Streaming service binding (gpud_serve_* in include/argus_gpu.h).

    s = Serve(sample_rate, block_samples, hbm_budget_bytes)
    s.Open(ChannelSpec(...))            # False if the HBM budget refuses it
    t = s.Submit(iq_block)              # async; returns a ticket
    info = s.Wait(t); iq, audio, rate = s.Output(t, channel_id)
    s.SetAudio(channel_id, 48000, 50.0); frames, rate = s.OutputAudio(t, channel_id)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, register

DEMOD_NONE, DEMOD_NFM, DEMOD_WFM, DEMOD_AM, DEMOD_USB, DEMOD_LSB, DEMOD_CW, DEMOD_WFM_STEREO = range(8)
ERR_BUDGET = -20
ERR_DUPLICATE = -21


class gpud_channel_spec(C.Structure):
    _fields_ = [("channel_id", C.c_int64), ("offset_hz", C.c_double), ("bandwidth", C.c_double),
                ("out_rate", C.c_int32), ("num_taps", C.c_int32), ("demod", C.c_int32), ("window", C.c_int32)]


class gpud_serve_block_info(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("n_channels", C.c_int32), ("n_samples", C.c_int32),
                ("iq_total", C.c_int64), ("audio_total", C.c_int64), ("gpu_ms", C.c_float),
                ("kernel_ms", C.c_float), ("extract_ms", C.c_float), ("pad", C.c_float)]


P, I = C.c_void_p, C.c_int
register("gpud_serve_create", I, [I, I, C.c_uint64, C.POINTER(P)])
register("gpud_serve_destroy", I, [P])
register("gpud_serve_open", I, [P, C.POINTER(gpud_channel_spec)])


class gpud_channel_opts(C.Structure):
    _fields_ = [("demod_rate", C.c_int32), ("pad", C.c_int32 * 7)]


register("gpud_serve_open_ex", I, [P, C.POINTER(gpud_channel_spec), C.POINTER(gpud_channel_opts)])
register("gpud_serve_channel_predemod", I, [P, C.c_int64, C.c_int64, C.POINTER(P), C.POINTER(I), C.POINTER(I)])
register("gpud_serve_close", I, [P, C.c_int64])
register("gpud_serve_info", I, [P, C.POINTER(I), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)])
register("gpud_serve_submit", I, [P, P, I, C.POINTER(C.c_int64)])
register("gpud_serve_submit_device", I, [P, P, I, C.POINTER(C.c_int64)])
register("gpud_serve_set_download", I, [P, I])
register("gpud_serve_set_download_parts", I, [P, I])
register("gpud_serve_set_download_split", I, [P, I, C.c_int64])
register("gpud_serve_poll", I, [P, C.c_int64])
register("gpud_serve_counters", I, [P, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)])
DL_IQ, DL_DEMOD, DL_WFM, DL_AUDIO, DL_PREDEMOD, DL_ALL = 1, 2, 4, 8, 16, 31
register("gpud_serve_wait", I, [P, C.c_int64, C.POINTER(gpud_serve_block_info)])
register("gpud_serve_channel_output", I, [P, C.c_int64, C.c_int64, C.POINTER(P), C.POINTER(I), C.POINTER(P),
                                          C.POINTER(I), C.POINTER(I)])
class gpud_tune(C.Structure):
    _fields_ = [("lds_entries", C.c_int32), ("narrow_outputs", C.c_int32), ("grid", C.c_int32), ("fast_grid", C.c_int32),
                ("kernel_ms", C.c_float), ("pad2", C.c_float)]


register("gpud_serve_set_graph", I, [P, I])
register("gpud_serve_get_tune", I, [P, C.POINTER(gpud_tune)])
register("gpud_serve_set_tune", I, [P, C.POINTER(gpud_tune)])
register("gpud_serve_autotune", I, [P, P, I, I, C.POINTER(gpud_tune)])
register("gpud_serve_extract_time", I, [P, P, I, I, C.POINTER(C.c_float), C.POINTER(C.c_float)])
register("gpud_runner_get_tune", I, [P, C.POINTER(gpud_tune)])
register("gpud_runner_set_tune", I, [P, C.POINTER(gpud_tune)])
register("gpud_serve_channel_wfm", I, [P, C.c_int64, C.c_int64, C.POINTER(P), C.POINTER(I), C.POINTER(I),
                                       C.POINTER(I), C.POINTER(I)])
register("gpud_serve_set_audio", I, [P, C.c_int64, I, C.c_double])
register("gpud_serve_channel_audio", I, [P, C.c_int64, C.c_int64, C.POINTER(P), C.POINTER(I), C.POINTER(I),
                                         C.POINTER(I)])
register("gpud_serve_set_rds", I, [P, C.c_int64, I])
register("gpud_serve_set_fast", I, [P, I])
register("gpud_serve_set_demod_fusion", I, [P, I])
register("gpud_fast_spectrum", I, [P, I, C.c_double, P])


def FastSpectrum(taps, phase_inc: float) -> np.ndarray:
    """The fast form's per-channel spectrum: FFT_4096(taps * e^{-j phase_inc k}) / 4096."""
    h = np.ascontiguousarray(taps, np.float32)
    out = np.zeros(4096, np.complex64)
    check("gpud_fast_spectrum", lib().gpud_fast_spectrum(h.ctypes.data, len(h), phase_inc, out.ctypes.data))
    return out
register("gpud_serve_channel_rds", I, [P, C.c_int64, C.c_int64, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_uint16),
                                       C.c_char_p, C.c_char_p, C.POINTER(P), C.POINTER(I)])


@dataclass
class ChannelSpec:
    ChannelID: int
    OffsetHz: float
    Bandwidth: float
    OutRate: int
    NumTaps: int = 101
    Demod: int = DEMOD_NONE
    Window: int = 0

    def c(self) -> gpud_channel_spec:
        return gpud_channel_spec(self.ChannelID, self.OffsetHz, self.Bandwidth, self.OutRate, self.NumTaps,
                                 self.Demod, self.Window)


class Serve:
    def __init__(self, sample_rate: int, block_samples: int, hbm_budget_bytes: int):
        h = C.c_void_p()
        check("gpud_serve_create", lib().gpud_serve_create(sample_rate, block_samples, hbm_budget_bytes, C.byref(h)))
        self._h = h
        self.sample_rate = sample_rate
        self.block_samples = block_samples

    def Close(self):
        if self._h:
            lib().gpud_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Open(self, spec: ChannelSpec, demod_rate: int = 0) -> bool:
        """True if admitted, False if the HBM budget refuses it.  demod_rate > 0
        adds the pre-demod decimating FIR to that rate (streamer.go:1098-1161)."""
        cs = spec.c()
        if demod_rate:
            op = gpud_channel_opts(int(demod_rate))
            rc = lib().gpud_serve_open_ex(self._h, C.byref(cs), C.byref(op))
        else:
            rc = lib().gpud_serve_open(self._h, C.byref(cs))
        if rc == ERR_BUDGET:
            return False
        check("gpud_serve_open", rc)
        return True

    def Retire(self, channel_id: int) -> None:
        check("gpud_serve_close", lib().gpud_serve_close(self._h, channel_id))

    def Info(self):
        n = C.c_int()
        used = C.c_uint64()
        bud = C.c_uint64()
        check("gpud_serve_info", lib().gpud_serve_info(self._h, C.byref(n), C.byref(used), C.byref(bud)))
        return n.value, used.value, bud.value

    def SetDownloadParts(self, mask: int) -> None:
        """Which outputs come to host memory (DL_* bits)."""
        check("gpud_serve_set_download_parts", lib().gpud_serve_set_download_parts(self._h, mask))

    def SetDownloadSplit(self, max_parts: int, min_part_bytes: int) -> None:
        """The early fast-form outputs in up to max_parts launches, one per
        min_part_bytes downloaded, each range downloaded as its launch ends."""
        check("gpud_serve_set_download_split",
              lib().gpud_serve_set_download_split(self._h, max_parts, min_part_bytes))

    def Poll(self, ticket: int) -> bool:
        """True once the block's outputs are on the host (non-blocking)."""
        rc = lib().gpud_serve_poll(self._h, ticket)
        if rc < 0:
            check("gpud_serve_poll", rc)
        return rc == 1

    def Counters(self):
        """(graph instantiations so far, free bytes of the channel-state arena)"""
        g, f = C.c_int64(), C.c_uint64()
        check("gpud_serve_counters", lib().gpud_serve_counters(self._h, C.byref(g), C.byref(f)))
        return g.value, f.value

    def SetDownload(self, enable: bool) -> None:
        check("gpud_serve_set_download", lib().gpud_serve_set_download(self._h, 1 if enable else 0))

    def Submit(self, iq) -> int:
        iq = np.ascontiguousarray(iq, np.complex64)
        t = C.c_int64()
        check("gpud_serve_submit", lib().gpud_serve_submit(self._h, iq.ctypes.data if len(iq) else None, len(iq),
                                                           C.byref(t)))
        return t.value

    def SubmitDevice(self, dev_ptr: int, n: int) -> int:
        t = C.c_int64()
        check("gpud_serve_submit_device", lib().gpud_serve_submit_device(self._h, dev_ptr, n, C.byref(t)))
        return t.value

    def Wait(self, ticket: int) -> gpud_serve_block_info:
        info = gpud_serve_block_info()
        check("gpud_serve_wait", lib().gpud_serve_wait(self._h, ticket, C.byref(info)))
        return info

    def Output(self, ticket: int, channel_id: int):
        """(iq complex64 copy, audio float32 copy, rate) of one channel, or None."""
        iq_p, aud_p = C.c_void_p(), C.c_void_p()
        n_iq, n_a, rate = C.c_int(), C.c_int(), C.c_int()
        rc = lib().gpud_serve_channel_output(self._h, ticket, channel_id, C.byref(iq_p), C.byref(n_iq),
                                             C.byref(aud_p), C.byref(n_a), C.byref(rate))
        if rc == -1:
            return None
        check("gpud_serve_channel_output", rc)
        iq = np.ctypeslib.as_array((C.c_float * (2 * n_iq.value)).from_address(iq_p.value)).view(np.complex64).copy() \
            if n_iq.value else np.zeros(0, np.complex64)
        au = np.ctypeslib.as_array((C.c_float * n_a.value).from_address(aud_p.value)).copy() \
            if n_a.value else np.zeros(0, np.float32)
        return iq, au, rate.value

    def OutputPreDemod(self, ticket: int, channel_id: int):
        """(pre-demod stage output complex64 copy, its rate) of one channel,
        or None when the channel has no such stage."""
        p, n, rate = C.c_void_p(), C.c_int(), C.c_int()
        rc = lib().gpud_serve_channel_predemod(self._h, ticket, channel_id, C.byref(p), C.byref(n), C.byref(rate))
        if rc == -3:
            return None
        check("gpud_serve_channel_predemod", rc)
        x = np.ctypeslib.as_array((C.c_float * (2 * n.value)).from_address(p.value)).view(np.complex64).copy() \
            if n.value else np.zeros(0, np.complex64)
        return x, rate.value

    def OutputWFM(self, ticket: int, channel_id: int):
        """(rds complex64 copy, rds_rate, locked, stereo) of a WFM stereo
        channel in a completed block, or None if it was not in the block."""
        p = C.c_void_p()
        n, rate, locked, stereo = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rc = lib().gpud_serve_channel_wfm(self._h, ticket, channel_id, C.byref(p), C.byref(n), C.byref(rate),
                                          C.byref(locked), C.byref(stereo))
        if rc == -1:
            return None
        check("gpud_serve_channel_wfm", rc)
        rds = np.ctypeslib.as_array((C.c_float * (2 * n.value)).from_address(p.value)).view(np.complex64).copy() \
            if n.value else np.zeros(0, np.complex64)
        return rds, rate.value, bool(locked.value), bool(stereo.value)

    def SetAudio(self, channel_id: int, audio_rate: int = 48000, deemphasis_us: float = 50.0) -> bool:
        """Turn a live channel's audio stage on (audio_rate > 0: low-pass for
        plain WFM, resampler to audio_rate, WFM de-emphasis and gain) or off
        (0).  False if the HBM budget refuses it."""
        rc = lib().gpud_serve_set_audio(self._h, channel_id, audio_rate, deemphasis_us)
        if rc == ERR_BUDGET:
            return False
        check("gpud_serve_set_audio", rc)
        return True

    def OutputAudio(self, ticket: int, channel_id: int):
        """(frames float32 copy shaped (n, channels), rate) of a channel's
        audio stage in a completed block; None if the channel was not in the
        block or its stage was off."""
        p = C.c_void_p()
        n, ch, rate = C.c_int(), C.c_int(), C.c_int()
        rc = lib().gpud_serve_channel_audio(self._h, ticket, channel_id, C.byref(p), C.byref(n), C.byref(ch),
                                            C.byref(rate))
        if rc in (-1, -3):
            return None
        check("gpud_serve_channel_audio", rc)
        cnt = n.value * ch.value
        a = np.ctypeslib.as_array((C.c_float * cnt).from_address(p.value)).copy() if cnt else np.zeros(0, np.float32)
        return a.reshape(n.value, ch.value), rate.value

    def SetRDS(self, channel_id: int, window_samples: int) -> bool:
        """Decode a stereo channel's RDS on the device in windows of
        window_samples baseband samples (0 turns it off).  False if the HBM
        budget refuses it."""
        rc = lib().gpud_serve_set_rds(self._h, channel_id, window_samples)
        if rc == ERR_BUDGET:
            return False
        check("gpud_serve_set_rds", rc)
        return True

    def OutputRDS(self, ticket: int, channel_id: int):
        """(decoded, found, groups (n, 4) uint16, PI, PS, RT) of a channel's
        RDS decoding in a completed block; None if the channel was not in the
        block or its decoding was off."""
        dec, found, n = C.c_int(), C.c_int(), C.c_int()
        pi = C.c_uint16()
        ps, rt = C.create_string_buffer(9), C.create_string_buffer(65)
        p = C.c_void_p()
        rc = lib().gpud_serve_channel_rds(self._h, ticket, channel_id, C.byref(dec), C.byref(found), C.byref(pi),
                                          ps, rt, C.byref(p), C.byref(n))
        if rc in (-1, -3):
            return None
        check("gpud_serve_channel_rds", rc)
        g = np.ctypeslib.as_array((C.c_uint16 * (4 * n.value)).from_address(p.value)).copy().reshape(-1, 4) \
            if n.value else np.zeros((0, 4), np.uint16)
        return (bool(dec.value), found.value, g, pi.value, ps.value.decode("latin-1"), rt.value.decode("latin-1"))

    def SetFast(self, enable: bool) -> None:
        """The fast-convolution form for eligible channels (on by default)."""
        check("gpud_serve_set_fast", lib().gpud_serve_set_fast(self._h, 1 if enable else 0))

    def SetDemodFusion(self, enable: bool) -> None:
        """Demodulation in the fast kernel's epilogue for the fast-form
        channels that demodulate their baseband directly (on by default);
        off, the separate demodulation pass gives the same outputs."""
        check("gpud_serve_set_demod_fusion", lib().gpud_serve_set_demod_fusion(self._h, 1 if enable else 0))

    def SetGraph(self, enable: bool) -> None:
        check("gpud_serve_set_graph", lib().gpud_serve_set_graph(self._h, 1 if enable else 0))

    def Tune(self) -> gpud_tune:
        t = gpud_tune()
        check("gpud_serve_get_tune", lib().gpud_serve_get_tune(self._h, C.byref(t)))
        return t

    def SetTune(self, lds_entries: int, narrow_outputs: int, grid: int, fast_grid: int = 0) -> None:
        """fast_grid 0 keeps the current fast-form grid."""
        t = gpud_tune(lds_entries, narrow_outputs, grid, fast_grid, 0.0, 0.0)
        check("gpud_serve_set_tune", lib().gpud_serve_set_tune(self._h, C.byref(t)))

    def GetTune(self) -> gpud_tune:
        t = gpud_tune()
        check("gpud_serve_get_tune", lib().gpud_serve_get_tune(self._h, C.byref(t)))
        return t

    def Autotune(self, dev_ptr: int, n: int, reps: int = 3) -> gpud_tune:
        t = gpud_tune()
        check("gpud_serve_autotune", lib().gpud_serve_autotune(self._h, dev_ptr, n, reps, C.byref(t)))
        return t

    def ExtractTime(self, dev_ptr: int, n: int, reps: int = 20):
        """(average, minimum) ms of the fused extraction launch of the live
        channel set at the current tune (scratch slot, state unchanged)."""
        avg, mn = C.c_float(), C.c_float()
        check("gpud_serve_extract_time",
              lib().gpud_serve_extract_time(self._h, dev_ptr, n, reps, C.byref(avg), C.byref(mn)))
        return avg.value, mn.value

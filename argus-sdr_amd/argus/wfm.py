"""RDS group decoder binding (gpud_rds_* in include/argus_gpu.h).

Mirrors the reference's rds.Decoder (internal/rds/rds.go:10-31, 78-264):

    d = RDSDecoder()
    d.Decode(samples, sample_rate)   # complex baseband at ~19 ksps
    d.PI, d.PS, d.RT, d.Groups()
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, register

P, I = C.c_void_p, C.c_int
register("gpud_rds_decoder_create", I, [C.POINTER(P)])
register("gpud_rds_decoder_destroy", I, [P])
register("gpud_rds_decode", I, [P, P, I, I])
register("gpud_rds_decoder_state", I, [P, C.POINTER(C.c_uint16), C.c_char_p, C.c_char_p])
register("gpud_rds_decoder_groups", I, [P, P, I])


class RDSDecoder:
    def __init__(self):
        h = C.c_void_p()
        check("gpud_rds_decoder_create", lib().gpud_rds_decoder_create(C.byref(h)))
        self._h = h

    def Close(self):
        if self._h:
            lib().gpud_rds_decoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Decode(self, samples, sample_rate: int) -> int:
        """One Decode call; returns the groups found in it."""
        x = np.ascontiguousarray(samples, np.complex64)
        rc = lib().gpud_rds_decode(self._h, x.ctypes.data if len(x) else None, len(x), sample_rate)
        if rc < 0:
            check("gpud_rds_decode", rc)
        return rc

    def _state(self):
        pi = C.c_uint16()
        ps = C.create_string_buffer(9)
        rt = C.create_string_buffer(65)
        check("gpud_rds_decoder_state", lib().gpud_rds_decoder_state(self._h, C.byref(pi), ps, rt))
        return pi.value, ps.value.decode("latin-1"), rt.value.decode("latin-1")

    @property
    def PI(self) -> int:
        return self._state()[0]

    @property
    def PS(self) -> str:
        return self._state()[1]

    @property
    def RT(self) -> str:
        return self._state()[2]

    def Groups(self) -> np.ndarray:
        """Applied groups in order, shape (n, 4) uint16."""
        total = lib().gpud_rds_decoder_groups(self._h, None, 0)
        out = np.zeros((max(total, 0), 4), np.uint16)
        if total > 0:
            lib().gpud_rds_decoder_groups(self._h, out.ctypes.data, total)
        return out


register("gpud_rds_batch_create", I, [C.POINTER(P), I])
register("gpud_rds_batch_destroy", I, [P])
register("gpud_rds_batch_reset", I, [P, I])
register("gpud_rds_batch_decode_device", I, [P, I, P, P, P, P, P, P])
register("gpud_rds_batch_state", I, [P, I, C.POINTER(C.c_uint16), C.c_char_p, C.c_char_p])
register("gpud_rds_batch_groups", I, [P, I, P, I])


class RDSBatchDecoder:
    """n_stations rds.Decoder instances whose state lives on the device.

        b = RDSBatchDecoder(20)
        b.Decode([(station, samples, rate), ...])   # one Decode call each
        b.PI(station), b.PS(station), b.RT(station), b.Groups(station)

    DecodeDevice takes baseband that is already in device memory."""

    def __init__(self, n_stations: int):
        h = C.c_void_p()
        check("gpud_rds_batch_create", lib().gpud_rds_batch_create(C.byref(h), int(n_stations)))
        self._h = h
        self.n_stations = int(n_stations)

    def Close(self):
        if self._h:
            lib().gpud_rds_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Reset(self, station: int):
        check("gpud_rds_batch_reset", lib().gpud_rds_batch_reset(self._h, int(station)))

    def DecodeDevice(self, stations, samples_dev: int, offsets, rates, stream=None):
        st = np.ascontiguousarray(stations, np.int32)
        offs = np.ascontiguousarray(offsets, np.int64)
        rt = np.ascontiguousarray(rates, np.int32)
        if len(offs) != len(st) + 1 or len(rt) != len(st):
            raise ValueError("offsets needs one entry more than stations, rates one per station")
        found = np.zeros(len(st), np.int32)
        check("gpud_rds_batch_decode_device",
              lib().gpud_rds_batch_decode_device(self._h, len(st), st.ctypes.data, samples_dev, offs.ctypes.data,
                                                 rt.ctypes.data, found.ctypes.data, stream))
        return [int(v) for v in found]

    def Decode(self, calls):
        """calls: [(station, complex samples, sample rate), ...], one Decode
        call per listed station; returns the groups found in each."""
        from .surveillance import DeviceBuffer
        xs = [np.ascontiguousarray(x, np.complex64) for _, x, _ in calls]
        offs = np.zeros(len(xs) + 1, np.int64)
        offs[1:] = np.cumsum([len(x) for x in xs])
        total = int(offs[-1])
        buf = DeviceBuffer(8 * max(total, 1))
        if total:
            buf.upload(np.concatenate(xs))
        return self.DecodeDevice([s for s, _, _ in calls], buf.ptr, offs, [r for _, _, r in calls])

    def _state(self, station):
        pi = C.c_uint16()
        ps = C.create_string_buffer(9)
        rt = C.create_string_buffer(65)
        check("gpud_rds_batch_state", lib().gpud_rds_batch_state(self._h, int(station), C.byref(pi), ps, rt))
        return pi.value, ps.value.decode("latin-1"), rt.value.decode("latin-1")

    def PI(self, station: int) -> int:
        return self._state(station)[0]

    def PS(self, station: int) -> str:
        return self._state(station)[1]

    def RT(self, station: int) -> str:
        return self._state(station)[2]

    def Groups(self, station: int) -> np.ndarray:
        total = lib().gpud_rds_batch_groups(self._h, int(station), None, 0)
        if total < 0:
            check("gpud_rds_batch_groups", total)
        out = np.zeros((total, 4), np.uint16)
        if total > 0:
            lib().gpud_rds_batch_groups(self._h, int(station), out.ctypes.data, total)
        return out


register("gpud_wfm_stereo_batch", I, [P, P, P, I, I, P, P, P])


def StereoDecodeBatch(monos, rates, mono_fallback: bool = False):
    """wfmStereoDecodeBatch (internal/recorder/wfm_batch_stereo.go:76-157) of
    several recordings in one call: returns ([L/R interleaved float32 per
    recording], [locked per recording]).  With mono_fallback an unlocked
    recording comes back as its raw discriminator on both channels, as
    demodWFMStereoBatchAudio (lines 42-49) does."""
    from .surveillance import DeviceBuffer
    monos = [np.ascontiguousarray(m, np.float32) for m in monos]
    rates = np.ascontiguousarray(rates, np.int32)
    if len(rates) != len(monos):
        raise ValueError("one rate per recording")
    offs = np.zeros(len(monos) + 1, np.int64)
    offs[1:] = np.cumsum([len(m) for m in monos])
    total = int(offs[-1])
    flat = np.concatenate(monos) if total else np.zeros(1, np.float32)
    din = DeviceBuffer(4 * max(total, 1)).upload(flat)
    dout = DeviceBuffer(8 * max(total, 1))
    locked = np.zeros(len(monos), np.int32)
    check("gpud_wfm_stereo_batch",
          lib().gpud_wfm_stereo_batch(din.ptr, offs.ctypes.data, rates.ctypes.data, len(monos),
                                      1 if mono_fallback else 0, dout.ptr, locked.ctypes.data, None))
    out = dout.download(np.float32, 2 * total) if total else np.zeros(0, np.float32)
    audio = []
    for i, m in enumerate(monos):
        if len(m) == 0 or rates[i] <= 0:
            audio.append(np.zeros(0, np.float32))
        else:
            audio.append(out[2 * offs[i]:2 * offs[i + 1]].copy())
    return audio, [bool(x) for x in locked]


register("gpud_stereo_pilot_present", I, [P, P, P, I, P, P, P, P, P])
register("gpud_stereo_pilot_present_noalloc", I, [P, P, P, I, P, P, P, P, P])


def StereoPilotPresent(iqs, rates):
    """StereoPilotPresent / estimateWFMPilot (internal/classifier/pll.go:82-131)
    of several baseband IQ slices in one call: returns (present, ratio,
    offset_hz, mags[n, 201]) per slice."""
    from .surveillance import DeviceBuffer
    iqs = [np.ascontiguousarray(x, np.complex64) for x in iqs]
    rates = np.ascontiguousarray(rates, np.int32)
    if len(rates) != len(iqs):
        raise ValueError("one rate per slice")
    offs = np.zeros(len(iqs) + 1, np.int64)
    offs[1:] = np.cumsum([len(x) for x in iqs])
    total = int(offs[-1])
    flat = np.concatenate(iqs) if total else np.zeros(1, np.complex64)
    din = DeviceBuffer(8 * max(total, 1)).upload(flat)
    k = len(iqs)
    present = np.zeros(k, np.int32)
    ratio = np.zeros(k, np.float64)
    off = np.zeros(k, np.float64)
    mags = np.zeros((k, 201), np.float64)
    check("gpud_stereo_pilot_present",
          lib().gpud_stereo_pilot_present(din.ptr, offs.ctypes.data, rates.ctypes.data, k, present.ctypes.data,
                                          ratio.ctypes.data, off.ctypes.data, mags.ctypes.data, None))
    return present.astype(bool), ratio, off, mags

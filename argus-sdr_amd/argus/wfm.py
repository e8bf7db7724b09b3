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

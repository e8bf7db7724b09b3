"""The daemon's RDS caller flow over the C ABI (gpud_rds_runtime_* in
include/argus_gpu.h): updateRDS (cmd/sdrd/pipeline_runtime.go:1044-1156)
and the ID prune of maintenance (1158-1176), every signal due in one tick
decoded in the same batched device passes.

    rt = RDSRuntime(max_stations=1024)
    out = rt.Update(now_s, [(signal_id, center_hz), ...], window, window_rate, window_center_hz)
    for stereo, station, turn in out: ...
    rt.Maintain(active_ids)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, register

P, I, L, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
register("gpud_rds_runtime_create", I, [I, C.POINTER(P)])
register("gpud_rds_runtime_destroy", I, [P])
register("gpud_rds_runtime_update", I, [P, D, P, P, I, P, L, I, I, D, P, P, P])
register("gpud_rds_runtime_maintain", I, [P, P, I])
register("gpud_rds_runtime_num_states", I, [P])
register("gpud_rds_runtime_state", I, [P, L, C.POINTER(D), C.POINTER(D), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_uint16), C.c_char_p, C.c_char_p])
register("gpud_rds_runtime_last", I, [P, L, P, L, C.POINTER(L), C.POINTER(C.c_int32), P, L, C.POINTER(L),
                                      C.POINTER(C.c_int32)])
register("gpud_rds_runtime_groups", I, [P, L, P, I])


class RDSRuntime:
    def __init__(self, max_stations: int = 1024):
        h = C.c_void_p()
        check("gpud_rds_runtime_create", lib().gpud_rds_runtime_create(int(max_stations), C.byref(h)))
        self._h = h

    def Close(self) -> None:
        if self._h:
            lib().gpud_rds_runtime_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass

    def Update(self, now_s: float, signals, window=None, window_rate: int = 0, window_center_hz: float = 0.0,
               window_dev: int = 0, n_window: int = 0):
        """updateRDS for each (signal_id, center_hz) in order.  window: host
        complex64 IQ, or pass window_dev (a device address) with n_window.
        Returns [(stereo, station, turn)] per signal (turn: 1 decode turn
        taken, 0 not due, -1 ID 0)."""
        n = len(signals)
        ids = np.ascontiguousarray([int(s[0]) for s in signals], np.int64)
        cen = np.ascontiguousarray([float(s[1]) for s in signals], np.float64)
        if window_dev:
            wptr, nw, on_dev = window_dev, int(n_window), 1
        else:
            w = np.ascontiguousarray(window if window is not None else np.zeros(0, np.complex64), np.complex64)
            wptr, nw, on_dev = (w.ctypes.data if len(w) else None), len(w), 0
        stereo = np.zeros(max(n, 1), np.int32)
        turn = np.zeros(max(n, 1), np.int32)
        station = C.create_string_buffer(9 * max(n, 1))
        check("gpud_rds_runtime_update",
              lib().gpud_rds_runtime_update(self._h, float(now_s), ids.ctypes.data if n else None,
                                            cen.ctypes.data if n else None, n, wptr, nw, on_dev, int(window_rate),
                                            float(window_center_hz), stereo.ctypes.data, station, turn.ctypes.data))
        raw = station.raw
        out = []
        for i in range(n):
            name = raw[9 * i:9 * i + 9].split(b"\0", 1)[0].decode("latin-1")
            out.append((int(stereo[i]), name, int(turn[i])))
        return out

    def Maintain(self, active_ids) -> None:
        a = np.ascontiguousarray(list(active_ids), np.int64)
        check("gpud_rds_runtime_maintain", lib().gpud_rds_runtime_maintain(self._h, a.ctypes.data if len(a) else None,
                                                                            len(a)))

    def NumStates(self) -> int:
        return int(lib().gpud_rds_runtime_num_states(self._h))

    def State(self, signal_id: int):
        """dict of the ID's state, or None."""
        ls, ld, st, pi = C.c_double(), C.c_double(), C.c_int32(), C.c_uint16()
        ps, rt = C.create_string_buffer(9), C.create_string_buffer(65)
        rc = lib().gpud_rds_runtime_state(self._h, int(signal_id), C.byref(ls), C.byref(ld), C.byref(st),
                                          C.byref(pi), ps, rt)
        if rc < 0:
            return None
        return {"last_seen": ls.value, "last_decode": ld.value, "stereo": st.value, "has_result": rc == 1,
                "pi": pi.value, "ps": ps.value.decode("latin-1"), "rt": rt.value.decode("latin-1")}

    def Last(self, signal_id: int):
        """(baseband, rate, rds baseband, rds rate) of the last decoding update,
        or None if that update did not decode the ID."""
        nb, br, nr, rr = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int32()
        rc = lib().gpud_rds_runtime_last(self._h, int(signal_id), None, 0, C.byref(nb), C.byref(br), None, 0,
                                         C.byref(nr), C.byref(rr))
        if rc != 0:
            return None
        base = np.zeros(max(nb.value, 1), np.complex64)
        rds = np.zeros(max(nr.value, 1), np.complex64)
        check("gpud_rds_runtime_last",
              lib().gpud_rds_runtime_last(self._h, int(signal_id), base.ctypes.data, nb.value, C.byref(nb),
                                          C.byref(br), rds.ctypes.data, nr.value, C.byref(nr), C.byref(rr)))
        return base[:nb.value], br.value, rds[:nr.value], rr.value

    def Groups(self, signal_id: int) -> np.ndarray:
        total = lib().gpud_rds_runtime_groups(self._h, int(signal_id), None, 0)
        if total < 0:
            check("gpud_rds_runtime_groups", total)
        out = np.zeros((total, 4), np.uint16)
        if total > 0:
            lib().gpud_rds_runtime_groups(self._h, int(signal_id), out.ctypes.data, total)
        return out

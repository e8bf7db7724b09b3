"""Capture files and the ingest ring (gpud_iqfile_* / gpud_ingest_* in
include/argus_gpu.h).

Mirrors internal/iqfile/iqfile.go:

    WriteIQFile(path, iq, sample_rate, center_hz)        # iqfile.Write
    iq, meta = ReadIQFile(path)                           # iqfile.Read
    src = IQFile(path); src.ReadIQ(n)                      # Source.ReadIQ (loops)

and the ring that moves a source's blocks to the device ahead of the
consumer:

    ing = Ingest(block_samples, depth)
    ing.StartFile(src)            # or ing.Push(block) per block
    dev, n, seq = ing.Acquire()   # device pointer, valid until Release(seq)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import GpudError, check, lib, register

P, I = C.c_void_p, C.c_int
I64 = C.c_int64
register("gpud_iqfile_write", I, [C.c_char_p, P, I64, I64, C.c_double])
register("gpud_iqfile_open", I, [C.c_char_p, C.POINTER(P), C.POINTER(I64), C.POINTER(C.c_double), C.POINTER(I64)])
register("gpud_iqfile_read", I, [P, P, I])
register("gpud_iqfile_seek", I, [P, I64])
register("gpud_iqfile_close", I, [P])
register("gpud_ingest_create", I, [I, I, C.POINTER(P)])
register("gpud_ingest_destroy", I, [P])
register("gpud_ingest_start_file", I, [P, P])
register("gpud_ingest_push", I, [P, P, I])
register("gpud_ingest_acquire", I, [P, C.POINTER(P), C.POINTER(I), C.POINTER(I64), I])
register("gpud_ingest_release", I, [P, I64])
register("gpud_ingest_stats", I, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I)])


@dataclass
class Meta:
    """iqfile.Meta (iqfile.go:30-34)."""
    SampleRate: int
    CenterHz: float
    Samples: int


def WriteIQFile(path: str, iq, sample_rate: int, center_hz: float) -> None:
    x = np.ascontiguousarray(iq, np.complex64)
    check("gpud_iqfile_write", lib().gpud_iqfile_write(str(path).encode(), x.ctypes.data if len(x) else None,
                                                        len(x), int(sample_rate), float(center_hz)))


class IQFile:
    """A capture opened for replay: Source.ReadIQ loops at the end."""

    def __init__(self, path: str):
        h = C.c_void_p()
        sr, n = I64(), I64()
        ch = C.c_double()
        check("gpud_iqfile_open", lib().gpud_iqfile_open(str(path).encode(), C.byref(h), C.byref(sr), C.byref(ch),
                                                          C.byref(n)))
        self._h = h
        self.meta = Meta(sr.value, ch.value, n.value)

    @property
    def handle(self):
        return self._h

    def ReadIQ(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.complex64)
        check("gpud_iqfile_read", lib().gpud_iqfile_read(self._h, out.ctypes.data if n else None, n))
        return out

    def Seek(self, pos: int) -> None:
        check("gpud_iqfile_seek", lib().gpud_iqfile_seek(self._h, pos))

    def Close(self):
        if self._h:
            lib().gpud_iqfile_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass


def ReadIQFile(path: str):
    """iqfile.Read: (samples, Meta)."""
    f = IQFile(path)
    try:
        return f.ReadIQ(f.meta.Samples), f.meta
    finally:
        f.Close()


class Ingest:
    def __init__(self, block_samples: int, depth: int = 4):
        h = C.c_void_p()
        check("gpud_ingest_create", lib().gpud_ingest_create(int(block_samples), int(depth), C.byref(h)))
        self._h = h
        self._file = None

    def StartFile(self, f: IQFile) -> None:
        self._file = f  # kept open while the producer reads it
        check("gpud_ingest_start_file", lib().gpud_ingest_start_file(self._h, f.handle))

    def Push(self, iq) -> None:
        x = np.ascontiguousarray(iq, np.complex64)
        check("gpud_ingest_push", lib().gpud_ingest_push(self._h, x.ctypes.data if len(x) else None, len(x)))

    def Acquire(self, timeout_ms: int = -1):
        """(device pointer, samples, sequence number) of the next block, or
        None if timeout_ms passed first."""
        p, n, seq = C.c_void_p(), I(), I64()
        rc = lib().gpud_ingest_acquire(self._h, C.byref(p), C.byref(n), C.byref(seq), int(timeout_ms))
        if rc == 1:
            return None
        check("gpud_ingest_acquire", rc)
        return p.value, n.value, seq.value

    def Release(self, seq: int) -> None:
        check("gpud_ingest_release", lib().gpud_ingest_release(self._h, int(seq)))

    def Stats(self):
        a, b, q = I64(), I64(), I()
        check("gpud_ingest_stats", lib().gpud_ingest_stats(self._h, C.byref(a), C.byref(b), C.byref(q)))
        return a.value, b.value, q.value

    def Close(self):
        if self._h:
            lib().gpud_ingest_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Close()
        except Exception:
            pass


__all__ = ["Meta", "WriteIQFile", "ReadIQFile", "IQFile", "Ingest", "GpudError"]

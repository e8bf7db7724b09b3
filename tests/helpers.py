"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_rms(got, want):
    got = np.asarray(got, np.complex128)
    want = np.asarray(want, np.complex128)
    assert got.shape == want.shape, (got.shape, want.shape)
    if want.size == 0:
        return 0.0
    den = np.sqrt(np.mean(np.abs(want) ** 2))
    num = np.sqrt(np.mean(np.abs(got - want) ** 2))
    return float(num / den) if den > 0 else float(num)


def phase_diff(a, b):
    d = a - b
    while d > np.pi:
        d -= 2 * np.pi
    while d < -np.pi:
        d += 2 * np.pi
    return abs(d)

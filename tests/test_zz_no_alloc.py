"""The no-allocation pilot-presence entry point through the C ABI: its
results equal gpud_stereo_pilot_present's (same kernels, workspace in the
library's file-scope scratch array instead of the stream-ordered pool) on
two calls in a row, the second with a smaller batch, and a batch larger than
the scratch array is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, din, offs, rates):
    from argus._lib import lib
    k = len(rates)
    out = (np.zeros(k, np.int32), np.zeros(k), np.zeros(k), np.zeros((k, 201)))
    rc = getattr(lib(), fn)(din, offs.ctypes.data, rates.ctypes.data, k, out[0].ctypes.data, out[1].ctypes.data,
                            out[2].ctypes.data, out[3].ctypes.data, None)
    return rc, out


def _batch(rng, n_slices, fs=250_000):
    from argus.synth import WFMStation, rds_symbols
    xs, rates = [], []
    for i in range(n_slices):
        st = WFMStation(fs, 0.0, rds_symbols(np.zeros(800, np.uint8)), seed=i).block(int(fs * rng.uniform(0.1, 0.6)))
        noise = (0.05 * (rng.standard_normal(len(st)) + 1j * rng.standard_normal(len(st)))).astype(np.complex64)
        xs.append(st + noise if i % 3 else noise)
        rates.append(fs)
    offs = np.zeros(n_slices + 1, np.int64)
    offs[1:] = np.cumsum([len(x) for x in xs])
    return np.concatenate(xs), offs, np.asarray(rates, np.int32)


def test_noalloc_equals_pool_version_twice(gpu):
    from argus.surveillance import DeviceBuffer
    rng = np.random.default_rng(31)
    for n_slices in (24, 5):  # the second call smaller than the first
        flat, offs, rates = _batch(rng, n_slices)
        din = DeviceBuffer(flat.nbytes).upload(flat)
        print(f"pilot batch of {n_slices} slices", flush=True)
        rc_a, a = _call("gpud_stereo_pilot_present", din.ptr, offs, rates)
        rc_b, b = _call("gpud_stereo_pilot_present_noalloc", din.ptr, offs, rates)
        assert rc_a == 0 and rc_b == 0
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert a[0].any() and not a[0].all()
        din.free()


def test_noalloc_refuses_oversize_batch(gpu):
    from argus.surveillance import DeviceBuffer
    fs = 250_000
    n = fs // 2
    k = 64  # 64 slices of 0.5 s need about 25 MiB of workspace, beyond the 16 MiB array
    flat = np.ones(n * k, np.complex64)
    offs = np.arange(k + 1, dtype=np.int64) * n
    rates = np.full(k, fs, np.int32)
    din = DeviceBuffer(flat.nbytes).upload(flat)
    rc, _ = _call("gpud_stereo_pilot_present_noalloc", din.ptr, offs, rates)
    assert rc == -7
    rc, _ = _call("gpud_stereo_pilot_present", din.ptr, offs, rates)
    assert rc == 0  # the pool version takes it
    din.free()

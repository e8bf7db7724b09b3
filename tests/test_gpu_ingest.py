"""The ingest ring on the GPU (gpud_ingest_*): blocks replayed from a
capture file or pushed by the caller arrive on the device in order and
byte-identical to the source (the replay looping at the end of the
capture, as iqfile's Source does), and the streaming service fed from the
ring's device blocks gives bit-identical outputs to the service fed the
same blocks from host memory."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOCK = 50_000


@pytest.fixture()
def ing(gpu):
    from argus import ingest
    return ingest


def _download(ptr, n):
    from argus._lib import check, lib
    out = np.zeros(n, np.complex64)
    if n:
        check("gpud_memcpy_d2h", lib().gpud_memcpy_d2h(out.ctypes.data, C.c_void_p(ptr), out.nbytes))
    return out


def _capture(tmp_path, n, seed=0, rate=2_500_000):
    from argus import ingest
    rng = np.random.default_rng(seed)
    iq = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    path = tmp_path / "cap.cf32"
    ingest.WriteIQFile(path, iq, rate, 100.1e6)
    return path, iq


def test_file_replay_reaches_the_device_in_order(ing, tmp_path):
    path, iq = _capture(tmp_path, 3 * BLOCK + 12_345)
    src = ing.IQFile(path)
    g = ing.Ingest(BLOCK, depth=3)
    g.StartFile(src)
    looped = np.concatenate([iq] * 4)
    for k in range(9):
        print(f"block {k}", flush=True)
        dev, n, seq = g.Acquire()
        assert (n, seq) == (BLOCK, k)
        np.testing.assert_array_equal(_download(dev, n), looped[k * BLOCK:(k + 1) * BLOCK])
        g.Release(seq)
    produced, consumed, queued = g.Stats()
    assert consumed == 9 and produced >= 9 and 0 <= queued <= 3
    g.Close()
    src.Close()


def test_push_acquire_release(ing):
    from argus._lib import GpudError
    rng = np.random.default_rng(2)
    g = ing.Ingest(BLOCK, depth=2)
    assert g.Acquire(timeout_ms=50) is None  # nothing pushed yet
    sizes = [BLOCK, 17, 0, BLOCK - 1, 4096]
    blocks = [(rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64) for n in sizes]
    g.Push(blocks[0])
    g.Push(blocks[1])  # the ring (depth 2) is now full
    for k, b in enumerate(blocks):
        dev, n, seq = g.Acquire()
        assert (n, seq) == (len(b), k)
        np.testing.assert_array_equal(_download(dev, n), b)
        with pytest.raises(GpudError):
            g.Release(seq + 1)  # not acquired
        g.Release(seq)
        if k + 2 < len(blocks):
            g.Push(blocks[k + 2])
    with pytest.raises(GpudError):
        g.Push(np.zeros(BLOCK + 1, np.complex64))  # longer than a slot
    g.Close()


def test_service_fed_from_the_ring_equals_host_submit(ing, tmp_path):
    from argus import serve as S
    fs = 2_500_000
    path, iq = _capture(tmp_path, 6 * BLOCK, seed=4, rate=fs)
    chans = [S.ChannelSpec(1, 200_000.0, 12_500.0, 50_000, 127, S.DEMOD_NFM, 0),
             S.ChannelSpec(2, -600_000.0, 200_000.0, 250_000, 101, S.DEMOD_WFM, 0),
             S.ChannelSpec(3, 31_250.0, 10_000.0, 12_500, 255, S.DEMOD_AM, 0)]
    host = S.Serve(fs, BLOCK, 1 << 28)
    dev = S.Serve(fs, BLOCK, 1 << 28)
    for c in chans:
        assert host.Open(c) and dev.Open(c)
    src = ing.IQFile(path)
    g = ing.Ingest(BLOCK, depth=4)
    g.StartFile(src)
    for k in range(8):  # wraps around the 6-block capture
        print(f"block {k}", flush=True)
        p, n, seq = g.Acquire()
        td = dev.SubmitDevice(p, n)
        th = host.Submit(np.concatenate([iq] * 2)[k * BLOCK:(k + 1) * BLOCK])
        dev.Wait(td)
        host.Wait(th)
        g.Release(seq)
        for c in chans:
            a, b = dev.Output(td, c.ChannelID), host.Output(th, c.ChannelID)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
    g.Close()
    src.Close()
    host.Close()
    dev.Close()

"""SDRIQ001 capture files through the library (gpud_iqfile_*), against the
reference's own iqfile tests (internal/iqfile/iqfile_test.go) and the format
its header comment fixes (iqfile.go:5-11).  Host code only."""
import struct

import numpy as np
import pytest

from argus._lib import GpudError


@pytest.fixture()
def iqf():
    from argus import ingest
    return ingest


def test_round_trip(iqf, tmp_path):
    """iqfile_test.go:9-30 TestRoundTrip, same values."""
    i = np.arange(1000, dtype=np.float32)
    iq = (i * np.float32(0.001) + 1j * (-i * np.float32(0.002))).astype(np.complex64)
    path = tmp_path / "snap.cf32"
    iqf.WriteIQFile(path, iq, 4_096_000, 101.5e6)
    got, meta = iqf.ReadIQFile(path)
    assert (meta.SampleRate, meta.CenterHz, meta.Samples) == (4_096_000, 101.5e6, 1000)
    np.testing.assert_array_equal(got, iq)


def test_layout_is_the_documented_format(iqf, tmp_path):
    iq = np.array([1 + 2j, -3.5 + 0.25j], np.complex64)
    path = tmp_path / "two.cf32"
    iqf.WriteIQFile(path, iq, 20_000_000, -1.25e3)
    raw = path.read_bytes()
    assert raw[:8] == b"SDRIQ001" and len(raw) == 32 + 16
    assert struct.unpack("<qdq", raw[8:32]) == (20_000_000, -1.25e3, 2)
    assert struct.unpack("<4f", raw[32:]) == (1.0, 2.0, -3.5, 0.25)


def test_replay_loops(iqf, tmp_path):
    """iqfile_test.go:32-50 TestReplaySourceLoops: 3 samples read as 7."""
    path = tmp_path / "s.cf32"
    iqf.WriteIQFile(path, np.array([1, 2, 3], np.complex64), 1000, 0)
    src = iqf.IQFile(path)
    np.testing.assert_array_equal(src.ReadIQ(7), np.array([1, 2, 3, 1, 2, 3, 1], np.complex64))
    np.testing.assert_array_equal(src.ReadIQ(4), np.array([2, 3, 1, 2], np.complex64))
    src.Seek(2)
    np.testing.assert_array_equal(src.ReadIQ(2), np.array([3, 1], np.complex64))
    src.Close()


def test_large_reads_stream_in_pieces(iqf, tmp_path):
    rng = np.random.default_rng(1)
    iq = (rng.standard_normal(100_003) + 1j * rng.standard_normal(100_003)).astype(np.complex64)
    path = tmp_path / "big.cf32"
    iqf.WriteIQFile(path, iq, 2_048_000, 98.1e6)
    src = iqf.IQFile(path)
    want = np.concatenate([iq, iq, iq])
    got = np.concatenate([src.ReadIQ(65_536) for _ in range(4)])
    np.testing.assert_array_equal(got, want[:len(got)])


def test_bad_files(iqf, tmp_path):
    bad = tmp_path / "bad.cf32"
    bad.write_bytes(b"NOTIQ001" + bytes(24))
    with pytest.raises(GpudError) as e:
        iqf.IQFile(bad)
    assert e.value.code == -3
    short = tmp_path / "short.cf32"
    short.write_bytes(b"SDRIQ001" + struct.pack("<qdq", 1000, 0.0, 10) + bytes(8 * 9))
    with pytest.raises(GpudError) as e:
        iqf.IQFile(short)
    assert e.value.code == -4
    with pytest.raises(GpudError) as e:
        iqf.IQFile(tmp_path / "missing.cf32")
    assert e.value.code == -2


def test_empty_capture_reads_zeros(iqf, tmp_path):
    path = tmp_path / "empty.cf32"
    iqf.WriteIQFile(path, np.zeros(0, np.complex64), 1000, 0)
    src = iqf.IQFile(path)
    assert src.meta.Samples == 0
    np.testing.assert_array_equal(src.ReadIQ(5), np.zeros(5, np.complex64))

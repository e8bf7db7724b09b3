"""Graph replay and autotuning of the streaming service.

Graph mode replays a captured per-slot hipGraph of each block's launches
(extraction, demodulation, stereo stage).  It must give bit-identical
outputs to the same launches issued one by one, replay after replay, with
eager work and host synchronisation between replays, and through channel
churn (which recaptures).  The autotuner must not advance channel state:
blocks after it continue the oracle's stream."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import AM, NFM, USB, WFM, OracleChannel, OracleStereoChannel, synth_scene

pytestmark = pytest.mark.gpu

SR = 2_500_000
SPECS = [  # id, kind, offset, bw, rate, taps, demod
    (1, "fm", -400_000.0, 12_500, 62_500, 127, NFM),
    (2, "wfm", 300_000.0, 200_000, 250_000, 101, 7),  # stereo
    (3, "am", 100_000.0, 10_000, 31_250, 201, AM),
    (4, "ssb", -150_000.0, 3_000, 12_500, 255, USB),
    (5, "wfm", 800_000.0, 200_000, 250_000, 63, WFM),
]


@pytest.fixture()
def S(gpu):
    from argus import serve
    return serve


def scene(n, seed):
    return synth_scene(n, SR, [(k, off, 0.5) for _, k, off, *_ in SPECS], seed=seed)


def test_graph_replays_equal_eager(S):
    x = scene(50_000 * 12, seed=1)
    a = S.Serve(SR, 50_000, 1 << 30)
    b = S.Serve(SR, 50_000, 1 << 30)
    a.SetGraph(True)
    b.SetGraph(False)
    for srv in (a, b):
        for cid, _, off, bw, rate, taps, dm in SPECS[:3]:
            assert srv.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0))
    sizes = [50_000, 50_000, 777, 50_000, 0, 50_000, 49_999, 50_000, 50_000, 50_000, 50_000, 50_000]
    pos = 0
    for k, n in enumerate(sizes):
        if k == 4:  # churn: recapture
            for srv in (a, b):
                srv.Retire(1)
                for cid, _, off, bw, rate, taps, dm in SPECS[3:]:
                    assert srv.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0))
        blk = x[pos:pos + n]
        pos += n
        print(f"block {k} n={n}", flush=True)
        ta = a.Submit(blk)
        a.Wait(ta)
        tb = b.Submit(blk)  # eager work between replays
        b.Wait(tb)
        live = [1, 2, 3] if k < 4 else [2, 3, 4, 5]
        for cid in live:
            ga, gb = a.Output(ta, cid), b.Output(tb, cid)
            assert ga is not None and gb is not None, (k, cid)
            np.testing.assert_array_equal(ga[0], gb[0])
            np.testing.assert_array_equal(ga[1], gb[1])
            if cid == 2:
                ra, rb = a.OutputWFM(ta, cid), b.OutputWFM(tb, cid)
                np.testing.assert_array_equal(ra[0], rb[0])
                assert ra[1:] == rb[1:]
    a.Close()
    b.Close()


def test_autotune_keeps_state_and_parity(S, gpu):
    """Autotuning and timing the extraction launch between blocks leave the
    channels' streams exactly where the oracle expects them; each block's
    wait reports its extraction time within its kernel time."""
    from argus._lib import check, lib
    L = lib()
    x = scene(50_000 * 8, seed=2)
    s = S.Serve(SR, 50_000, 1 << 30)
    orc = {}
    for cid, _, off, bw, rate, taps, dm in SPECS:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0))
        orc[cid] = (OracleStereoChannel(SR, off, bw, rate, taps) if dm == 7
                    else OracleChannel(SR, off, bw, rate, taps, dm))
    d = C.c_void_p()
    check("malloc", L.gpud_malloc(C.byref(d), 8 * 50_000))
    check("h2d", L.gpud_memcpy_h2d(d, x[:50_000].ctypes.data, 8 * 50_000))
    acc = {c: ([], []) for c in orc}
    for k in range(8):
        if k == 3:
            before = s.Tune()
            t = s.Autotune(d.value, 50_000, reps=2)
            print("tuned", t.lds_entries, t.narrow_outputs, t.grid, t.kernel_ms, "from", before.lds_entries,
                  before.narrow_outputs, before.grid, flush=True)
            assert t.grid % 8 == 0 and t.kernel_ms > 0
        if k == 5:  # timing the extraction launch also leaves the stream untouched
            avg, mn = s.ExtractTime(d.value, 50_000, reps=3)
            assert 0 < mn <= avg
        blk = x[50_000 * k:50_000 * (k + 1)]
        tk = s.Submit(blk)
        info = s.Wait(tk)
        assert 0 < info.extract_ms <= info.kernel_ms
        for cid, oc in orc.items():
            want = oc.block(blk)
            got = s.Output(tk, cid)
            assert len(got[0]) == len(want[0]) and len(got[1]) == len(want[1]), (k, cid)
            acc[cid][0].append(got[1])
            acc[cid][1].append(want[1])
    for cid, (g, w) in acc.items():
        e = rel_rms(np.concatenate(g), np.concatenate(w))
        assert e <= 1e-4, (cid, e)
    L.gpud_free(d)
    s.Close()


def test_set_tune_rejects_bad_shapes_and_changes_nothing_numerically(S):
    from argus._lib import GpudError
    s = S.Serve(SR, 50_000, 1 << 30)
    with pytest.raises(GpudError):
        s.SetTune(100, 32, 2048)
    with pytest.raises(GpudError):
        s.SetTune(6144, 32, 1001)
    s.SetTune(4096, 16, 1024)
    t = s.Tune()
    assert (t.lds_entries, t.narrow_outputs, t.grid) == (4096, 16, 1024)
    s.Close()

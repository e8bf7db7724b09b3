"""StreamingExtractGPU with the reference's output ownership
(streaming_gpu_native_prepare.go:78-104, streaming_out_ring.go:22-112):
the ring variant returns the same outputs as the plain call (same kernels,
bit-exact), every output buffer stays untouched for the next three calls,
and the runner's rings follow the oracle model (prune with each call's job
list, then one ring buffer per job) through a signal leaving and another
arriving: live rings, pooled buffers and allocations equal after every
call."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FS = 2_048_000


def _jobs(ids):
    from argus.gpudemod import StreamingConfigHash, StreamingExtractJob
    specs = {1: (-300e3, 20e3, 32000, 101), 2: (150e3, 12e3, 16000, 63), 3: (400e3, 150e3, 256000, 127),
             4: (-700e3, 5e3, 8000, 255), 5: (20e3, 25e3, 64000, 101), 6: (600e3, 5e3, 8000, 95)}
    return [StreamingExtractJob(i, *specs[i], StreamingConfigHash(i, *specs[i], FS)) for i in ids]


def test_ring_outputs_and_ownership(gpu):
    from argus.gpudemod import OUT_RING_DEPTH, NewBatchRunner
    rng = np.random.default_rng(4)
    a = NewBatchRunner(1 << 18, FS)
    b = NewBatchRunner(1 << 18, FS)
    held = []  # (call, signal, view, copy)
    model = O.OutRings()
    ids = [1, 2, 3, 4, 5]
    for call in range(10):
        if call == 6:
            ids = [1, 2, 3, 5, 6]  # 4 leaves, 6 (same rate) arrives: 4's pooled buffers serve 6
        n = int(rng.integers(30_000, 200_000))
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        print(f"ring call {call}: n={n} signals {ids}", flush=True)
        ra = a.StreamingExtractGPURing(x, _jobs(ids))
        rb = b.StreamingExtractGPU(x, _jobs(ids))
        model.prune(ids)
        for p, q in zip(ra, rb):
            assert p.SignalID == q.SignalID and p.NOut == q.NOut
            assert np.array_equal(p.IQ, q.IQ), (call, p.SignalID)
            held.append((call, p.SignalID, p.IQ, p.IQ.copy()))
            model.next(p.SignalID, p.NOut)
        st = a.OutRings().Stats()
        assert (st["rings"], st["free"], st["allocated"]) == (len(model.rings), len(model.pool.free),
                                                            model.pool.allocated), call
        # every buffer handed out in the last OUT_RING_DEPTH calls still holds its data
        for c, sid, view, ref in held:
            if call - c < OUT_RING_DEPTH and (sid != 4 or call < 6):
                assert np.array_equal(view, ref), (call, c, sid)
    a.Close()
    b.Close()

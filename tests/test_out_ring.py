"""Output rings (streamOutRing / outBufPool, streaming_out_ring.go:22-112)
through the C ABI, on the host only (no GPU):

* the reference's two tests (streaming_out_ring_test.go:5-36) restated;
* closed-form pins of the oracle model: a signal's fifth call takes the
  first call's buffer again, a slot grows by drawing from the pool, the
  pool keeps at most 128 buffers;
* seeded random sequences of next / prune / pool get / pool put against the
  oracle model: live rings, pooled buffers, allocations, and buffer
  identity (the C side hands back the very buffer the model reuses)."""
import numpy as np
import pytest

import oracle as O


@pytest.fixture
def rings():
    from argus.gpudemod import OutRings
    r = OutRings()
    yield r
    r.Close()


def test_pool_recycles_across_signals(rings):
    """streaming_out_ring_test.go:5-26"""
    for _ in range(O.OUT_RING_DEPTH):
        rings.Next(1, 1000)
    rings.Prune([])  # signal 1 inactive: its buffers go to the pool
    st = rings.Stats()
    assert st["free"] > 0 and st["rings"] == 0
    before = st["free"]
    b = rings.Next(2, 1000)
    assert b
    assert rings.Slot(2)[1][0] >= 1000
    assert rings.Stats()["free"] == before - 1
    assert rings.Stats()["allocated"] == O.OUT_RING_DEPTH  # nothing new was allocated


def test_pool_grows_when_too_small(rings):
    """streaming_out_ring_test.go:28-36"""
    small, cap = rings.PoolGet(100)
    assert cap == 100
    rings.PoolPut(small, cap)
    buf, cap = rings.PoolGet(5000)
    assert buf and cap >= 5000
    st = rings.Stats()
    assert st["free"] == 0 and st["released"] == 1  # the too-small buffer was dropped
    rings.PoolPut(buf, cap)


def test_ring_depth_and_growth_closed_form(rings):
    seen = [rings.Next(5, 64) for _ in range(O.OUT_RING_DEPTH)]
    assert len(set(seen)) == O.OUT_RING_DEPTH
    assert rings.Next(5, 64) == seen[0]          # call 5 reuses call 1's buffer
    assert rings.Next(5, 32) == seen[1]          # a smaller call keeps the slot's buffer
    assert rings.Slot(5) == (2, [64, 64, 64, 64])
    rings.Next(5, 1000)                          # slot 2 grows: old buffer released, a new one made
    st = rings.Stats()
    assert st["allocated"] == 5 and st["released"] == 1
    assert rings.Slot(5)[1][2] == 1000
    assert rings.Next(9, 0) == 0                 # an empty slot serves zero samples without a buffer


def test_pool_keeps_at_most_128(rings):
    n_sig = O.OUT_POOL_MAX // O.OUT_RING_DEPTH + 3
    for s in range(n_sig):
        for _ in range(O.OUT_RING_DEPTH):
            rings.Next(s, 16)
    rings.Prune([])
    st = rings.Stats()
    assert st["free"] == O.OUT_POOL_MAX
    assert st["released"] == n_sig * O.OUT_RING_DEPTH - O.OUT_POOL_MAX


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_sequences_match_model(rings, seed):
    rng = np.random.default_rng(seed)
    model = O.OutRings()
    uid_ptr = {}
    pooled_out = []  # (model buffer, C pointer, cap) taken by pool get and not yet put back
    for step in range(3000):
        op = rng.random()
        if op < 0.75:
            sid = int(rng.integers(1, 40))
            n = int(rng.choice([0, 1, 50, 200, 1000, 4096, int(rng.integers(1, 5000))]))
            mb = model.next(sid, n)
            p = rings.Next(sid, n)
            if mb is None:
                assert p == 0, step
            elif mb.uid in uid_ptr:
                assert p == uid_ptr[mb.uid], step
            else:
                uid_ptr[mb.uid] = p
        elif op < 0.85:
            active = [int(v) for v in rng.integers(1, 40, int(rng.integers(0, 30)))]
            model.prune(active)
            rings.Prune(active)
        elif op < 0.93:
            n = int(rng.integers(0, 6000))
            mb = model.pool.get(n)
            p, cap = rings.PoolGet(n)
            assert cap == mb.cap, step
            if mb.uid in uid_ptr:
                assert p == uid_ptr[mb.uid], step
            uid_ptr[mb.uid] = p
            pooled_out.append((mb, p, cap))
        elif pooled_out:
            mb, p, cap = pooled_out.pop(int(rng.integers(0, len(pooled_out))))
            model.pool.put([mb])
            rings.PoolPut(p, cap)
        st = rings.Stats()
        assert st["rings"] == len(model.rings), step
        assert st["free"] == len(model.pool.free), step
        assert st["allocated"] == model.pool.allocated, step
    for mb, p, cap in pooled_out:
        rings.PoolPut(p, cap)

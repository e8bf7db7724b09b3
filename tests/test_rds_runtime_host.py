"""Bookkeeping of the RDS caller flow (updateRDS, pipeline_runtime.go:
1044-1080, and maintenance's prune, 1158-1176) through the C ABI with no
window to decode, so no device work runs (no GPU needed):

* closed-form pins: decode turns every 4 s of the caller's clock, the first
  sight due at once, ID 0 skipped, a state dropped only when unseen for
  MORE than 30 s (and then recreated fresh), maintain() dropping absent IDs,
  -8 past the decoder slots;
* seeded random tick sequences against the oracle's RDSRuntime: turns, the
  set of live states and their clock stamps."""
import numpy as np
import pytest

import oracle as O


@pytest.fixture
def rt():
    from argus.rds_runtime import RDSRuntime
    r = RDSRuntime(64)
    yield r
    r.Close()


def _turns(rt, now, sigs):
    return [t for _, _, t in rt.Update(now, sigs)]


def test_cadence_closed_form(rt):
    sig = [(7, 98.1e6)]
    turns = [_turns(rt, t * 0.5, sig)[0] for t in range(26)]  # 0 .. 12.5 s
    assert [i * 0.5 for i, t in enumerate(turns) if t == 1] == [0.0, 4.0, 8.0, 12.0]
    st = rt.State(7)
    assert st["last_decode"] == 12.0 and st["last_seen"] == 12.5 and not st["has_result"]
    assert _turns(rt, 13.0, [(0, 1e6), (7, 98.1e6)]) == [-1, 0]
    assert rt.NumStates() == 1


def test_prune_after_more_than_30s(rt):
    rt.Update(0.0, [(1, 0.0), (2, 0.0)])
    rt.Update(29.0, [(2, 0.0)])
    rt.Update(30.0, [(2, 0.0)])
    assert rt.State(1) is not None          # unseen for exactly 30 s: kept
    rt.Update(30.25, [(2, 0.0)])
    assert rt.State(1) is None              # more than 30 s: dropped
    assert rt.State(2)["last_decode"] == 29.0  # turns at 0 and 29 (seen at 0, 29, 30, 30.25)
    assert _turns(rt, 31.0, [(1, 0.0)]) == [1]  # recreated fresh: due at once
    rt.Maintain([1])
    assert rt.State(2) is None and rt.NumStates() == 1
    rt.Maintain([])
    assert rt.NumStates() == 0


def test_slots_bounded():
    from argus.rds_runtime import RDSRuntime
    from argus import GpudError
    r = RDSRuntime(3)
    r.Update(0.0, [(1, 0.0), (2, 0.0), (3, 0.0)])
    with pytest.raises(GpudError) as e:
        r.Update(1.0, [(4, 0.0)])
    assert e.value.code == -8
    r.Maintain([1, 2])                      # a slot comes free
    assert _turns(r, 2.0, [(4, 0.0)]) == [1]
    r.Close()


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_random_ticks_match_oracle(rt, seed):
    rng = np.random.default_rng(seed)
    orc = O.RDSRuntime()
    now = 0.0
    for tick in range(400):
        now += float(rng.choice([0.05, 0.5, 1.0, 3.99, 4.0, 7.0, 31.0], p=[.3, .25, .2, .1, .1, .04, .01]))
        ids = [int(v) for v in rng.integers(0, 24, int(rng.integers(0, 10)))]
        sigs = [(i, 88e6 + i * 1e5) for i in ids]
        got = _turns(rt, now, sigs)
        want = [t for _, _, t in orc.Update(now, sigs, np.zeros(0, np.complex64), 0, 0.0)]
        assert got == want, (tick, sigs)
        if rng.random() < 0.05:
            active = [int(v) for v in rng.integers(0, 24, 8)]
            rt.Maintain(active)
            orc.Maintain(active)
        assert rt.NumStates() == len(orc.states), tick
        for sid, st in orc.states.items():
            g = rt.State(sid)
            assert g is not None and g["last_seen"] == st["last_seen"], (tick, sid)
            assert g["last_decode"] == st["last_decode"], (tick, sid)

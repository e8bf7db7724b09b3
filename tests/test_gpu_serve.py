"""GPU tests of the streaming service: parity of every demod mode with the
oracle over ragged blocks, identity of a channel's results alone and in a
mix, admission against the HBM budget, and births/deaths while blocks are in
flight.  Baseband and audio within 1e-4 relative RMS of the oracle."""
import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import AM, CW, LSB, NFM, NONE, USB, WFM, OracleChannel, synth_scene

pytestmark = pytest.mark.gpu

TOL = 1e-4
SR = 2_048_000


@pytest.fixture()
def serve_mod(gpu):
    from argus import serve
    return serve


SPECS = [  # id, kind, offset, bw, out_rate, taps, demod, window
    (1, "fm", -400_000.0, 12_500, 64_000, 127, NFM, 0),
    (2, "wfm", 300_000.0, 200_000, 256_000, 127, WFM, 1),
    (3, "am", 100_000.0, 10_000, 32_000, 201, AM, 0),
    (4, "ssb", -150_000.0, 3_000, 16_000, 255, USB, 0),
    (5, "ssb", 600_000.0, 3_000, 16_000, 63, LSB, 1),
    (6, "tone", -700_000.0, 2_000, 8_000, 511, CW, 0),
    (7, "fm", 800_000.0, 25_000, 128_000, 65, NONE, 0),
]


def scene(n, seed=0):
    return synth_scene(n, SR, [(k, off, 0.5) for _, k, off, *_ in SPECS], seed=seed)


def open_all(S, s, specs):
    orc = {}
    for cid, _, off, bw, rate, taps, dm, win in specs:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, win))
        orc[cid] = OracleChannel(SR, off, bw, rate, taps, dm, win)
    return orc


def test_every_mode_matches_oracle_over_ragged_blocks(serve_mod):
    S = serve_mod
    s = S.Serve(SR, 65536, 1 << 30)
    orc = open_all(S, s, SPECS)
    x = scene(65536 * 4)
    sizes = [65536, 1, 0, 4097, 511, 65536, 30000, 65536]
    pos = 0
    acc = {c: ([], [], [], []) for c in orc}
    for k, n in enumerate(sizes):
        blk = x[pos:pos + n]
        pos += n
        print(f"block {k} n={n}", flush=True)
        t = s.Submit(blk)
        s.Wait(t)
        for cid, oc in orc.items():
            want_iq, want_au = oc.block(blk)
            got_iq, got_au, rate = s.Output(t, cid)
            assert len(got_iq) == len(want_iq) and len(got_au) == len(want_au), (cid, k)
            acc[cid][0].append(got_iq)
            acc[cid][1].append(want_iq)
            acc[cid][2].append(got_au)
            acc[cid][3].append(want_au)
    for cid, (gi, wi, ga, wa) in acc.items():
        e_iq = rel_rms(np.concatenate(gi), np.concatenate(wi))
        assert e_iq <= TOL, (cid, e_iq)
        if len(np.concatenate(wa)):
            e_au = rel_rms(np.concatenate(ga), np.concatenate(wa))
            assert e_au <= TOL, (cid, e_au)
    s.Close()


def test_alone_equals_in_mix(serve_mod):
    """Continuous batching must not change any channel's result."""
    S = serve_mod
    x = scene(65536 * 3, seed=4)
    mix = S.Serve(SR, 65536, 1 << 30)
    open_all(S, mix, SPECS)
    outs_mix = {c[0]: [] for c in SPECS}
    for b in range(3):
        t = mix.Submit(x[b * 65536:(b + 1) * 65536])
        mix.Wait(t)
        for c in SPECS:
            outs_mix[c[0]].append(mix.Output(t, c[0]))
    mix.Close()
    for spec in SPECS:
        alone = S.Serve(SR, 65536, 1 << 30)
        open_all(S, alone, [spec])
        for b in range(3):
            t = alone.Submit(x[b * 65536:(b + 1) * 65536])
            alone.Wait(t)
            iq, au, _ = alone.Output(t, spec[0])
            miq, mau, _ = outs_mix[spec[0]][b]
            assert np.array_equal(iq, miq) and np.array_equal(au, mau), spec
        alone.Close()


def test_admission_against_hbm_budget(serve_mod):
    S = serve_mod
    spec = lambda i: S.ChannelSpec(100 + i, 1000.0 * i, 12_500, 64_000, 255, NFM, 0)
    probe = S.Serve(SR, 65536, 1 << 30)
    assert probe.Open(spec(0))
    one = probe.Info()[1]
    probe.Close()
    s = S.Serve(SR, 65536, 3 * one + one // 2)
    assert s.Open(spec(0)) and s.Open(spec(1)) and s.Open(spec(2))
    assert not s.Open(spec(3))                       # refused: budget
    with pytest.raises(Exception):
        s.Open(spec(1))                              # duplicate id
    t = s.Submit(scene(4096))
    s.Retire(101)                                    # retired while its block is in flight
    s.Wait(t)
    assert s.Output(t, 101) is not None              # it still produced that block
    t2 = s.Submit(scene(4096))
    s.Wait(t2)
    assert s.Output(t2, 101) is None                 # and left the next one
    assert s.Open(spec(3))                           # its memory was returned
    assert s.Info()[0] == 3
    s.Close()


def test_births_and_deaths_in_flight(serve_mod):
    """Channels join and leave between submits with blocks in flight; each
    channel's stream matches an oracle started at its first block."""
    S = serve_mod
    rng = np.random.default_rng(9)
    s = S.Serve(SR, 16384, 1 << 30)
    x = scene(16384 * 12, seed=2)
    live, orc, tickets = {}, {}, []
    acc = {}
    for b in range(12):
        if b % 3 == 0:  # birth
            cid = 1000 + b
            spec = SPECS[b % len(SPECS)]
            assert s.Open(S.ChannelSpec(cid, spec[2], spec[3], spec[4], spec[5], spec[6], spec[7]))
            orc[cid] = OracleChannel(SR, spec[2], spec[3], spec[4], spec[5], spec[6], spec[7])
            acc[cid] = ([], [], [], [])
            live[cid] = True
        if b % 4 == 3 and len(live) > 1:  # death
            cid = sorted(live)[0]
            s.Retire(cid)
            del live[cid]
        blk = x[b * 16384:(b + 1) * 16384]
        t = s.Submit(blk)
        tickets.append((t, blk, list(live)))
        if len(tickets) == 3 or b == 11:  # wait with up to 3 blocks in flight
            for tt, bb, ids in tickets:
                s.Wait(tt)
                for cid in ids:
                    wi, wa = orc[cid].block(bb)
                    gi, ga, _ = s.Output(tt, cid)
                    acc[cid][0].append(gi)
                    acc[cid][1].append(wi)
                    acc[cid][2].append(ga)
                    acc[cid][3].append(wa)
            tickets = []
    for cid, (gi, wi, ga, wa) in acc.items():
        assert rel_rms(np.concatenate(gi), np.concatenate(wi)) <= TOL, cid
        if len(np.concatenate(wa)):
            assert rel_rms(np.concatenate(ga), np.concatenate(wa)) <= TOL, cid
    s.Close()

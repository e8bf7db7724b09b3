"""GPU tests of the streaming service's post-demodulation audio stage
(gpud_serve_set_audio): the low-pass for plain WFM, the rational polyphase
resampler to 48 kHz, WFM de-emphasis and gain, against the oracle chain
(serve_oracle.OracleAudioStage after the oracle channel) over ragged blocks:
frame counts exact, samples within 1e-4 relative RMS."""
import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import (AM, CW, LSB, NFM, NONE, USB, WFM, WFM_STEREO, OracleAudioStage, OracleChannel,
                          OracleStereoChannel, synth_scene)

pytestmark = pytest.mark.gpu

TOL = 1e-4
SR = 2_048_000


@pytest.fixture()
def serve_mod(gpu):
    from argus import serve
    return serve


SPECS = [  # id, kind, offset, bw, out_rate, taps, demod, window, deemphasis us
    (1, "fm", -400_000.0, 12_500, 64_000, 127, NFM, 0, 50.0),     # L/M 3/4
    (2, "wfm", 300_000.0, 200_000, 256_000, 127, WFM, 1, 75.0),   # 3/16, low-pass first
    (3, "am", 100_000.0, 10_000, 32_000, 201, AM, 0, 50.0),       # 3/2
    (4, "ssb", -150_000.0, 3_000, 16_000, 255, USB, 0, 50.0),     # 3/1
    (5, "ssb", 600_000.0, 3_000, 16_000, 63, LSB, 1, 0.0),
    (6, "tone", -700_000.0, 2_000, 8_000, 511, CW, 0, 50.0),      # 6/1
    (7, "wfm", 800_000.0, 200_000, 128_000, 65, WFM, 0, 0.0),     # 3/8, no de-emphasis
    (8, "fm", -900_000.0, 25_000, 128_000, 65, NONE, 0, 50.0),    # no demodulator: no audio
]


def _compare(acc):
    for cid, (g, w) in acc.items():
        g, w = np.concatenate(g), np.concatenate(w)
        assert g.shape == w.shape, (cid, g.shape, w.shape)
        if w.size:
            e = rel_rms(g, w)
            assert e <= TOL, (cid, e)


def test_every_mode_resampled_over_ragged_blocks(serve_mod):
    S = serve_mod
    x = synth_scene(65536 * 5, SR, [(k, off, 0.5) for _, k, off, *_ in SPECS], seed=5)
    s = S.Serve(SR, 65536, 1 << 30)
    orc, aud = {}, {}
    for cid, _, off, bw, rate, taps, dm, win, us in SPECS:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, win))
        orc[cid] = OracleChannel(SR, off, bw, rate, taps, dm, win)
    # equal sizes in a row take the tables prepared while the previous block ran
    sizes = [65536, 1, 0, 4097, 511, 65536, 30000, 30000, 30000, 65536, 3]
    acc = {cid: ([], []) for cid in orc}
    pos = 0
    for k, n in enumerate(sizes):
        if k == 1:  # the stage starts from fresh state at the next block
            for cid, *_, rate, taps, dm, win, us in SPECS:
                assert s.SetAudio(cid, 48000, us)
                aud[cid] = OracleAudioStage(dm, rate, 48000, us)
        blk = x[pos:pos + n]
        pos += n
        print(f"block {k} n={n}", flush=True)
        t = s.Submit(blk)
        s.Wait(t)
        for cid, oc in orc.items():
            _, want_au = oc.block(blk)
            got = s.OutputAudio(t, cid)
            if cid not in aud:
                assert got is None
                continue
            frames, rate = got
            want = aud[cid].process(want_au)
            assert rate == 48000 and frames.shape == want.shape, (cid, k, frames.shape, want.shape)
            acc[cid][0].append(frames)
            acc[cid][1].append(want)
    s.Close()
    _compare(acc)
    assert sum(len(g) for g in acc[2][0]) > 1000


def test_graph_and_eager_agree(serve_mod):
    """Captured-graph replays and eager launches give the same audio."""
    S = serve_mod
    x = synth_scene(65536 * 3, SR, [(k, off, 0.5) for _, k, off, *_ in SPECS], seed=6)
    outs = []
    for graph in (True, False):
        s = S.Serve(SR, 65536, 1 << 30)
        s.SetGraph(graph)
        for cid, _, off, bw, rate, taps, dm, win, us in SPECS[:4]:
            assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, win))
            assert s.SetAudio(cid, 48000, us)
        got = []
        for k in range(3):
            t = s.Submit(x[k * 65536:(k + 1) * 65536])
            s.Wait(t)
            got.append([s.OutputAudio(t, cid)[0] for cid, *_ in SPECS[:4]])
        s.Close()
        outs.append(got)
    for a, b in zip(*outs):
        for fa, fb in zip(a, b):
            np.testing.assert_array_equal(fa, fb)


def test_identity_rate_copies_the_demodulator(serve_mod):
    """Demod rate 48 kHz: the reference skips resampling, so NFM audio is the
    demodulator output itself and WFM its low-pass, de-emphasis and gain."""
    S = serve_mod
    sr = 1_920_000
    x = synth_scene(96000 * 2, sr, [("fm", 200_000.0, 0.5), ("wfm", -500_000.0, 0.5)], seed=7)
    s = S.Serve(sr, 96000, 1 << 30)
    assert s.Open(S.ChannelSpec(1, 200_000.0, 12_500, 48_000, 127, S.DEMOD_NFM, 0))
    assert s.Open(S.ChannelSpec(2, -500_000.0, 40_000, 48_000, 127, S.DEMOD_WFM, 0))
    assert s.SetAudio(1, 48000, 50.0) and s.SetAudio(2, 48000, 50.0)
    aud2 = OracleAudioStage(WFM, 48000, 48000, 50.0)
    o2 = OracleChannel(sr, -500_000.0, 40_000, 48_000, 127, WFM, 0)
    g2, w2 = [], []
    pos = 0
    for n in (96000, 777, 95223):
        blk = x[pos:pos + n]
        pos += n
        t = s.Submit(blk)
        s.Wait(t)
        _, au1, _ = s.Output(t, 1)
        frames1, _ = s.OutputAudio(t, 1)
        np.testing.assert_array_equal(frames1[:, 0], au1)
        _, want = o2.block(blk)
        g2.append(s.OutputAudio(t, 2)[0])
        w2.append(aud2.process(want))
    s.Close()
    _compare({2: (g2, w2)})


def test_stereo_frames_resampled(serve_mod):
    """WFM stereo at 250 ksps to 48 kHz (L/M 24/125): two resamplers over the
    L/R frames of the stereo stage (mono fallback before lock, the matrix
    after), de-emphasis per side, gain."""
    S = serve_mod
    import oracle as O
    from argus.synth import WFMStation, rds_symbols
    fs, block = 2_500_000, 50_000
    bits = O.RDSEncodeGroups(0x1A2B, "ALPHA FM", "radiotext", 12)
    st = WFMStation(fs, -300_000.0, rds_symbols(bits), seed=3, amp=1.0)
    rng = np.random.default_rng(8)
    s = S.Serve(fs, block, 1 << 30)
    assert s.Open(S.ChannelSpec(1, -300_000.0, 200_000.0, 250_000, 101, S.DEMOD_WFM_STEREO, 0))
    assert s.SetAudio(1, 48000, 50.0)
    oc = OracleStereoChannel(fs, -300_000.0, 200_000.0, 250_000, 101, 0)
    aud = OracleAudioStage(WFM_STEREO, 250_000, 48000, 50.0)
    acc = {1: ([], [])}
    states = []
    for k, n in enumerate([7, block - 7, 0, 1234, block - 1234] + [block] * 10):
        print(f"block {k} n={n}", flush=True)
        blk = (st.block(n) + 0.02 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))).astype(np.complex64)
        t = s.Submit(blk)
        s.Wait(t)
        _, w_au, _, _, w_st = oc.block(blk)
        frames, rate = s.OutputAudio(t, 1)
        want = aud.process(w_au)
        assert rate == 48000 and frames.shape == want.shape, (k, frames.shape, want.shape)
        acc[1][0].append(frames)
        acc[1][1].append(want)
        states.append(w_st)
    s.Close()
    assert states[-1] and not states[0]  # both the fallback and the matrix were resampled
    _compare(acc)


def test_stage_off_and_budget(serve_mod):
    S = serve_mod
    s = S.Serve(SR, 65536, 1 << 30)
    assert s.Open(S.ChannelSpec(1, 0.0, 12_500, 64_000, 127, S.DEMOD_NFM, 0))
    _, used0, _ = s.Info()
    assert s.SetAudio(1, 48000, 0.0)
    _, used1, _ = s.Info()
    assert used1 > used0
    t = s.Submit(np.zeros(4096, np.complex64))
    s.Wait(t)
    assert s.OutputAudio(t, 1)[0].shape[1] == 1
    assert s.SetAudio(1, 0, 0.0)
    _, used2, _ = s.Info()
    assert used2 == used0
    t = s.Submit(np.zeros(4096, np.complex64))
    s.Wait(t)
    assert s.OutputAudio(t, 1) is None
    s.Close()
    # a budget that holds the channel but not its audio stage
    tight = S.Serve(SR, 65536, used0)
    assert tight.Open(S.ChannelSpec(1, 0.0, 12_500, 64_000, 127, S.DEMOD_NFM, 0))
    assert not tight.SetAudio(1, 48000, 0.0)
    t = tight.Submit(np.zeros(4096, np.complex64))
    tight.Wait(t)
    assert tight.OutputAudio(t, 1) is None
    tight.Close()

"""The fast-convolution form of the service's extraction (fastconv.hip)
against the oracle and against the direct form: channels of decimation 8,
16, 32, 64, 40 and 80 with up to 511 taps over ragged blocks (shorter than the
511-sample tail, one segment exactly, segment boundaries, empty), a channel
opened mid-stream (direct form until its history is full, then fast), every
demod mode, and ineligible channels beside them.  Baseband and audio within
1e-4 relative RMS of the oracle; fast and direct forms agree to FP32
rounding but not bit for bit (so the fast form did run).

The demodulation fused into the fast kernel's epilogue and into the direct
form's narrow items (the default) gives the same bits as the separate
demodulation pass over the same baseband: FM with its segment and item
seams and carry, SSB, CW, on the same ragged blocks.
Split into 2-4 launches over channel ranges (each range downloaded as its
launch ends), the outputs are the same bits as in one launch."""
import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import AM, CW, LSB, NFM, NONE, USB, WFM, OracleChannel, synth_scene

pytestmark = pytest.mark.gpu

TOL = 1e-4
SR = 2_560_000
BLOCK = 51_200  # 20 ms: 15 segments of 3584

SPECS = [  # id, kind, offset, bw, out_rate, taps, demod, window
    (1, "wfm", 300_000.0, 200_000, 320_000, 511, WFM, 0),   # D 8, longest taps
    (2, "fm", -400_000.0, 25_000, 160_000, 127, NFM, 1),    # D 16
    (3, "am", 100_000.0, 10_000, 80_000, 65, AM, 0),        # D 32
    (4, "ssb", -150_000.0, 3_000, 40_000, 255, USB, 0),     # D 64
    (5, "fm", 700_000.0, 12_500, 256_000, 127, NFM, 0),     # D 10: direct form
    (6, "wfm", -800_000.0, 200_000, 320_000, 601, WFM, 0),  # 601 taps: direct form
    (7, "tone", 950_000.0, 5_000, 160_000, 3, NONE, 0),     # D 16 >= T: direct form
    (9, "fm", -250_000.0, 12_500, 64_000, 255, NFM, 0),     # D 40: fold by 8, every 5th
    (10, "am", 450_000.0, 10_000, 32_000, 301, AM, 1),      # D 80: fold by 16, every 5th
]
LATE = (8, "fm", 500_000.0, 12_500, 80_000, 201, NFM, 0)    # opened at block 3


@pytest.fixture()
def serve_mod(gpu):
    from argus import serve
    return serve


def _run(S, fast, x, sizes, fusion=True):
    s = S.Serve(SR, BLOCK, 1 << 30)
    s.SetFast(fast)
    s.SetDemodFusion(fusion)
    orc = {}
    specs = list(SPECS)
    for cid, _, off, bw, rate, taps, dm, win in specs:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, win))
        orc[cid] = OracleChannel(SR, off, bw, rate, taps, dm, win)
    got = {}
    pos = 0
    for k, n in enumerate(sizes):
        if k == 3:
            cid, _, off, bw, rate, taps, dm, win = LATE
            assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, win))
            orc[cid] = OracleChannel(SR, off, bw, rate, taps, dm, win)
        blk = x[pos:pos + n]
        pos += n
        print(f"fast={fast} block {k} n={n}", flush=True)
        t = s.Submit(blk)
        s.Wait(t)
        for cid, oc in orc.items():
            w_iq, w_au = oc.block(blk)
            g_iq, g_au, _ = s.Output(t, cid)
            assert len(g_iq) == len(w_iq) and len(g_au) == len(w_au), (cid, k)
            a = got.setdefault(cid, ([], [], [], []))
            a[0].append(g_iq)
            a[1].append(w_iq)
            a[2].append(g_au)
            a[3].append(w_au)
    s.Close()
    return {c: tuple(np.concatenate(v) for v in a) for c, a in got.items()}


def test_fast_form_matches_oracle_and_direct_form(serve_mod):
    S = serve_mod
    kinds = [(k, off, 0.5) for _, k, off, *_ in SPECS + [LATE]]
    sizes = [BLOCK, 7, 300, BLOCK, 3584, 3585, 0, 12_345, BLOCK, BLOCK, 7168, BLOCK]
    x = synth_scene(sum(sizes), SR, kinds, seed=4)
    fast = _run(S, True, x, sizes)
    direct = _run(S, False, x, sizes)
    differ = 0
    for cid in fast:
        g_iq, w_iq, g_au, w_au = fast[cid]
        e = rel_rms(g_iq, w_iq)
        print(f"channel {cid}: baseband rel RMS {e:.2e}", flush=True)
        assert e <= TOL, (cid, e)
        if len(w_au):
            assert rel_rms(g_au, w_au) <= TOL, cid
        d_iq = direct[cid][0]
        assert rel_rms(g_iq, d_iq) <= 1e-5, cid
        differ += int(not np.array_equal(g_iq, d_iq))
    assert differ >= 6  # the eligible channels took the other form


FUSED = [  # id, offset, bw, out_rate, taps, demod: every mode the epilogue demodulates, and AM beside them
    (21, 300_000.0, 200_000, 320_000, 511, WFM),   # D 8: 448 outputs per segment
    (22, -400_000.0, 25_000, 160_000, 127, NFM),   # D 16
    (23, -150_000.0, 3_000, 40_000, 255, USB),     # D 64: fold by 64
    (24, 250_000.0, 3_000, 80_000, 129, LSB),      # D 32
    (25, 600_000.0, 1_000, 320_000, 63, CW),       # D 8
    (26, -250_000.0, 12_500, 64_000, 255, NFM),    # D 40: every 5th fold position
    (27, 100_000.0, 10_000, 80_000, 65, AM),       # D 32: the separate pass
    # the direct form's narrow items (D >= T) demodulate in their epilogue too
    (31, 420_000.0, 12_500, 80_000, 31, NFM),      # D 32, 31 taps
    (32, -520_000.0, 3_000, 40_000, 63, USB),      # D 64
    (33, 820_000.0, 3_000, 64_000, 39, LSB),       # D 40
    (34, -820_000.0, 1_000, 80_000, 17, CW),       # D 32
    (35, 150_000.0, 200_000, 320_000, 7, WFM),     # D 8, 7 taps
    (36, -60_000.0, 12_500, 256_000, 127, NFM),    # D 10 < T: a dense tile, the separate pass
]


def _run_fusion(S, fusion, x, sizes, split=None):
    s = S.Serve(SR, BLOCK, 1 << 30)
    s.SetDemodFusion(fusion)
    if split is not None:
        s.SetDownloadSplit(*split)
    for cid, off, bw, rate, taps, dm in FUSED:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0))
    out = []
    pos = 0
    for k, n in enumerate(sizes):
        if k == 2:  # opened mid-stream: direct form first, then fast with the carry
            assert s.Open(S.ChannelSpec(28, -700_000.0, 12_500, 160_000, 201, NFM, 0))
        blk = x[pos:pos + n]
        pos += n
        print(f"fusion={fusion} split={split} block {k} n={n}", flush=True)
        t = s.Submit(blk)
        s.Wait(t)
        cids = [c[0] for c in FUSED] + ([28] if k >= 2 else [])
        out.append({cid: tuple(np.copy(a) for a in s.Output(t, cid)[:2]) for cid in cids})
    s.Close()
    return out


def test_fused_demod_same_bits_as_separate_pass(serve_mod):
    S = serve_mod
    kinds = [("wfm", 300_000.0, 0.5), ("fm", -400_000.0, 0.5), ("ssb", -150_000.0, 0.5),
             ("ssb", 250_000.0, 0.5), ("tone", 600_000.0, 0.5), ("fm", -250_000.0, 0.5),
             ("am", 100_000.0, 0.5), ("fm", -700_000.0, 0.5), ("fm", 420_000.0, 0.5),
             ("ssb", -520_000.0, 0.5), ("ssb", 820_000.0, 0.5), ("tone", -820_000.0, 0.5),
             ("wfm", 150_000.0, 0.5), ("fm", -60_000.0, 0.5)]
    sizes = [BLOCK, BLOCK, 3584, 3585, 7, 0, 12_345, BLOCK, 7168, BLOCK]
    x = synth_scene(sum(sizes), SR, kinds, seed=9)
    fused = _run_fusion(S, True, x, sizes)
    sep = _run_fusion(S, False, x, sizes)
    n_audio = 0
    for k, (bf, bs) in enumerate(zip(fused, sep)):
        assert bf.keys() == bs.keys()
        for cid in bf:
            (fi, fa), (si, sa) = bf[cid], bs[cid]
            assert np.array_equal(fi, si), (k, cid)
            assert len(fa) == len(sa), (k, cid)
            assert np.array_equal(fa, sa), (k, cid, np.flatnonzero(fa != sa)[:8])
            n_audio += len(fa)
    assert n_audio > 100_000


@pytest.mark.parametrize("parts", [2, 3, 4])
def test_download_split_same_outputs(serve_mod, parts):
    """The early channels' outputs computed and downloaded in 2-4 parts
    (consecutive channel ranges, each downloaded as its launch ends) are
    the same bits as in one part, with and without the fused demodulation,
    over ragged blocks and a channel opened mid-stream."""
    S = serve_mod
    kinds = [("wfm", 300_000.0, 0.5), ("fm", -400_000.0, 0.5), ("ssb", -150_000.0, 0.5),
             ("ssb", 250_000.0, 0.5), ("tone", 600_000.0, 0.5), ("fm", -250_000.0, 0.5),
             ("am", 100_000.0, 0.5), ("fm", -700_000.0, 0.5)]
    sizes = [BLOCK, BLOCK, 3585, 0, 12_345, BLOCK]
    x = synth_scene(sum(sizes), SR, kinds, seed=10 + parts)
    for fusion in (True, False):
        one = _run_fusion(S, fusion, x, sizes, split=(1, 0))
        many = _run_fusion(S, fusion, x, sizes, split=(parts, 0))
        for k, (b1, bm) in enumerate(zip(one, many)):
            assert b1.keys() == bm.keys()
            for cid in b1:
                for a1, am in zip(b1[cid], bm[cid]):
                    assert len(a1) == len(am) and np.array_equal(a1, am), (fusion, k, cid)



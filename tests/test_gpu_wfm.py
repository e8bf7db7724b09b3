"""GPU tests of the WFM multiplex stage of the streaming service: stereo
audio, stereo lock and the RDS baseband of each station within 1e-4
relative RMS of the oracle chain, and the RDS groups decoded from the GPU
baseband bit-identical to those the oracle decoder finds in the oracle
baseband (and equal to the transmitted PI/PS)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_rms
from serve_oracle import OracleStereoChannel

pytestmark = pytest.mark.gpu

TOL = 1e-4
FS = 2_500_000     # decimation 10 to 250 ksps
BLOCK = 50_000     # 20 ms
STATIONS = [       # id, offset, PI, PS, RT, seed, level
    (1, -700_000.0, 0x1A2B, "ALPHA FM", "first station radiotext", 21, 1.0),
    (2, 150_000.0, 0x3C4D, "BRAVO 99", "second station radiotext", 22, 0.6),
    (3, 800_000.0, 0x5E6F, "CHARLIE ", "third", 23, 0.8),
]


def _stations(seconds):
    from argus.synth import WFMStation, rds_symbols
    out = []
    for cid, off, pi, ps, rt, seed, lvl in STATIONS:
        bits = O.RDSEncodeGroups(pi, ps, rt, int(seconds * 1187.5 / 104) + 4)
        out.append(WFMStation(FS, off, rds_symbols(bits), seed=seed, amp=lvl))
    return out


def _mix(stations, n, rng):
    x = 0.02 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    for st in stations:
        x = x + st.block(n)
    return x.astype(np.complex64)


@pytest.fixture()
def serve_mod(gpu):
    from argus import serve
    return serve


def test_stereo_and_rds_match_oracle(serve_mod):
    S = serve_mod
    from argus.wfm import RDSDecoder
    seconds = 1.6
    sizes = [7, BLOCK - 7, 0, 1234, BLOCK - 1234] + [BLOCK] * (int(seconds * FS) // BLOCK - 2)
    st = _stations(seconds)
    rng = np.random.default_rng(9)
    s = S.Serve(FS, BLOCK, 1 << 30)
    orc = {}
    for cid, off, *_ in STATIONS:
        assert s.Open(S.ChannelSpec(cid, off, 200_000.0, 250_000, 101, S.DEMOD_WFM_STEREO, 0))
        orc[cid] = OracleStereoChannel(FS, off, 200_000.0, 250_000, 101, 0)
    acc = {cid: {k: [] for k in ("gi", "wi", "ga", "wa", "gr", "wr")} for cid in orc}
    locks = {cid: [] for cid in orc}
    for k, n in enumerate(sizes):
        print(f"block {k} n={n}", flush=True)
        blk = _mix(st, n, rng)
        t = s.Submit(blk)
        s.Wait(t)
        for cid, oc in orc.items():
            w_iq, w_au, w_rds, w_lock, w_st = oc.block(blk)
            g_iq, g_au, rate = s.Output(t, cid)
            g_rds, rds_rate, g_lock, g_st = s.OutputWFM(t, cid)
            assert rate == 250_000 and rds_rate == oc.rds.rate
            assert (len(g_iq), len(g_au), len(g_rds)) == (len(w_iq), len(w_au), len(w_rds)), (cid, k)
            assert (g_lock, g_st) == (w_lock, w_st), (cid, k)
            locks[cid].append(g_st)
            a = acc[cid]
            a["gi"].append(g_iq); a["wi"].append(w_iq)
            a["ga"].append(g_au); a["wa"].append(w_au)
            a["gr"].append(g_rds); a["wr"].append(w_rds)
    s.Close()
    for cid, off, pi, ps, rt, *_ in STATIONS:
        a = {k: np.concatenate(v) for k, v in acc[cid].items()}
        for g, w in (("gi", "wi"), ("ga", "wa"), ("gr", "wr")):
            e = rel_rms(a[g], a[w])
            print(cid, g, e, flush=True)
            assert e <= TOL, (cid, g, e)
        assert locks[cid][-1], cid  # stereo enabled by the end
        # RDS: decode both basebands in the same 0.5 s windows
        rate = orc[cid].rds.rate
        dg, dw = RDSDecoder(), O.RDSDecoder()
        win = rate // 2
        for i in range(0, len(a["gr"]), win):
            dg.Decode(a["gr"][i:i + win], rate)
            dw.Decode(a["wr"][i:i + win], rate)
        gg = dg.Groups()
        np.testing.assert_array_equal(gg, np.asarray(dw.Groups(), np.uint16).reshape(-1, 4))
        assert len(gg) >= 4 and dg.PI == pi, (cid, len(gg), hex(dg.PI))


def test_rds_decoded_on_device_in_the_service(serve_mod):
    """Device-side RDS decoding inside the service: each window the service
    decodes equals one Decode call of the oracle decoder on the same window
    of the service's own RDS baseband (groups found, group bits, PI/PS/RT),
    in the block where the window completes; the groups also equal the
    oracle decoder's on the oracle chain's baseband in the same windows."""
    S = serve_mod
    from argus._lib import GpudError
    seconds = 2.0
    sizes = [7, BLOCK - 7, 0, 1234, BLOCK - 1234] + [BLOCK] * (int(seconds * FS) // BLOCK - 2)
    st = _stations(seconds)
    rng = np.random.default_rng(10)
    s = S.Serve(FS, BLOCK, 1 << 30)
    wins = {1: 9000, 2: 4000}  # station 3 leaves decoding off
    orc = {}
    for cid, off, *_ in STATIONS:
        assert s.Open(S.ChannelSpec(cid, off, 200_000.0, 250_000, 101, S.DEMOD_WFM_STEREO, 0))
        orc[cid] = OracleStereoChannel(FS, off, 200_000.0, 250_000, 101, 0)
    assert s.Open(S.ChannelSpec(9, 0.0, 12_500.0, 50_000, 101, S.DEMOD_NFM, 0))
    with pytest.raises(GpudError):
        s.SetRDS(9, 4000)       # not a stereo channel
    with pytest.raises(GpudError):
        s.SetRDS(1, 100)        # shorter than a group
    for cid, w in wins.items():
        assert s.SetRDS(cid, w)
    s.SetDownload(True)
    pend = {cid: np.zeros(0, np.complex64) for cid in wins}
    pend_w = {cid: np.zeros(0, np.complex64) for cid in wins}
    dec_g = {cid: O.RDSDecoder() for cid in wins}   # on the service's baseband
    dec_w = {cid: O.RDSDecoder() for cid in wins}   # on the oracle chain's baseband
    n_dec = {cid: 0 for cid in wins}
    for k, n in enumerate(sizes):
        print(f"block {k} n={n}", flush=True)
        blk = _mix(st, n, rng)
        t = s.Submit(blk)
        s.Wait(t)
        assert s.OutputRDS(t, 3) is None and s.OutputRDS(t, 9) is None
        for cid, oc in orc.items():
            _, _, w_rds, _, _ = oc.block(blk)
            if cid not in wins:
                continue
            g_rds, rate, _, _ = s.OutputWFM(t, cid)
            pend[cid] = np.concatenate([pend[cid], g_rds])
            pend_w[cid] = np.concatenate([pend_w[cid], w_rds])
            decoded, found, groups, pi, ps, rt = s.OutputRDS(t, cid)
            W = wins[cid]
            want_dec = len(pend[cid]) >= W
            assert decoded == want_dec, (cid, k, decoded, len(pend[cid]))
            if not decoded:
                assert found == 0 and groups.shape == (0, 4)
                continue
            before = len(dec_g[cid].Groups())
            f = dec_g[cid].Decode(pend[cid][:W], rate)
            dec_w[cid].Decode(pend_w[cid][:W], rate)
            pend[cid] = pend[cid][W:]
            pend_w[cid] = pend_w[cid][W:]
            want = np.asarray(dec_g[cid].Groups(), np.uint16).reshape(-1, 4)[before:]
            assert found == f, (cid, k, found, f)
            np.testing.assert_array_equal(groups, want)
            assert (pi, ps, rt) == (dec_g[cid].PI, dec_g[cid].PS, dec_g[cid].RT), (cid, k)
            n_dec[cid] += 1
    s.Close()
    for cid, off, pi, ps, *_ in STATIONS:
        if cid not in wins:
            continue
        gg = np.asarray(dec_g[cid].Groups(), np.uint16).reshape(-1, 4)
        np.testing.assert_array_equal(gg, np.asarray(dec_w[cid].Groups(), np.uint16).reshape(-1, 4))
        assert n_dec[cid] >= 3 and len(gg) >= 4 and dec_g[cid].PI == pi, (cid, n_dec[cid], len(gg))

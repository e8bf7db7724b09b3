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

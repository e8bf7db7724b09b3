"""The live recorder's pre-demod stage in the serving loop (D4:
streamer.go:1098-1161, dsp.StatefulDecimatingFIRComplex
decimating_fir.go:51-95), against the oracle chain (streaming extractor,
the oracle's StatefulDecimatingFIRComplex, the demodulators and the stereo
decoder at the demod rate) over ragged blocks at 20 Msps with the rates the
reference reaches there: WFM snapped to 500 ksps (snapOutRate) and forced
to /2 -> 250 ksps, NFM at 200 ksps to a 48 kHz demod (/4 -> 50 ksps), and
a channel with no stage beside them.  Baseband, pre-demod output and audio
within 1e-4 relative RMS; counts, rates and stereo lock decisions exact."""
import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import NFM, NONE, WFM, WFM_STEREO, OracleChannel, OracleStereoChannel, synth_scene

pytestmark = pytest.mark.gpu

SR = 20_000_000
TOL = 1e-4


def test_predemod_stage_matches_oracle(gpu):
    from argus import serve as S
    from argus.jobs import snapOutRate
    wfm_rate = snapOutRate(SR, 512000)
    assert wfm_rate == 500000
    specs = [  # id, offset, bw, out_rate, taps, demod, demod_rate
        (1, 2_000_000.0, 250_000.0, wfm_rate, 101, WFM, 192000),        # /round(2.6) = 3 -> forced 2
        (2, -3_000_000.0, 20_000.0, 200_000, 101, NFM, 48000),          # /round(4.17) = 4
        (3, 5_000_000.0, 250_000.0, wfm_rate, 127, WFM_STEREO, 192000),  # stereo at 250 ksps
        (4, -7_000_000.0, 20_000.0, 200_000, 63, NFM, 0),               # no stage
        (5, 1_000_000.0, 20_000.0, 200_000, 101, NFM, 250000),          # round(0.8) = 1: no stage
    ]
    kinds = [("wfm", 2e6, 0.5), ("fm", -3e6, 0.5), ("wfm", 5e6, 0.5), ("fm", -7e6, 0.5), ("fm", 1e6, 0.5)]
    sizes = [400_000, 1, 0, 39_999, 400_000, 123_457, 400_000]
    x = synth_scene(sum(sizes), SR, kinds, seed=21)
    s = S.Serve(SR, 400_000, 1 << 30)
    orc = {}
    for cid, off, bw, rate, taps, dm, drate in specs:
        assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0), demod_rate=drate)
        orc[cid] = (OracleStereoChannel(SR, off, bw, rate, taps, 0, drate) if dm == WFM_STEREO
                    else OracleChannel(SR, off, bw, rate, taps, dm, 0, drate))
    acc = {cid: ([], [], [], [], [], []) for cid in orc}
    pos = 0
    for k, n in enumerate(sizes):
        blk = x[pos:pos + n]
        pos += n
        print(f"predemod block {k} n={n}", flush=True)
        t = s.Submit(blk)
        s.Wait(t)
        for cid, oc in orc.items():
            if isinstance(oc, OracleStereoChannel):
                w_iq, w_au, _, w_lock, _ = oc.block(blk)
                pre = oc.ch.last_pre
                g_lock = s.OutputWFM(t, cid)[2]
                assert g_lock == int(w_lock) or len(w_au) == 0, (cid, k)
            else:
                w_iq, w_au = oc.block(blk)
                pre = oc.last_pre
            g_iq, g_au, rate = s.Output(t, cid)
            assert len(g_iq) == len(w_iq) and len(g_au) == len(w_au), (cid, k, len(g_au), len(w_au))
            gp = s.OutputPreDemod(t, cid)
            if pre is None:
                assert gp is None, cid
            else:
                assert gp[1] == oc.ch.out_rate if isinstance(oc, OracleStereoChannel) else gp[1] == oc.out_rate
                assert len(gp[0]) == len(pre), (cid, k, len(gp[0]), len(pre))
                acc[cid][4].append(gp[0])
                acc[cid][5].append(pre)
            for i, v in enumerate((g_iq, w_iq, g_au, w_au)):
                acc[cid][i].append(v)
    for cid, a in acc.items():
        g_iq, w_iq, g_au, w_au = (np.concatenate(v) for v in a[:4])
        assert rel_rms(g_iq, w_iq) <= TOL, (cid, rel_rms(g_iq, w_iq))
        assert rel_rms(g_au, w_au) <= TOL, (cid, rel_rms(g_au, w_au))
        if a[4]:
            e = rel_rms(np.concatenate(a[4]), np.concatenate(a[5]))
            print(f"channel {cid}: pre-demod rel RMS {e:.2e}", flush=True)
            assert e <= TOL, (cid, e)
    assert [len(acc[c][4]) > 0 for c in (1, 2, 3, 4, 5)] == [True, True, True, False, False]
    s.Close()

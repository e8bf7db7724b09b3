"""Parity at the benchmarked shapes (20 Msps, 20 ms = 400 000-sample blocks).

* Serving mix: 8192 channels of the bench's uneven mode mix (NFM/AM/USB/WFM,
  decimations 8..1600, taps 63..511, their demodulators), 2230 of them
  eligible for the fast-convolution form -- two work items each, more than
  the fast kernel's persistent grid, so its grid-stride loop runs.  Every channel of the
  served form (fast form on) against the direct form (fast form off) over
  four blocks, and a sampled subset of every decimation class against the
  oracle chain (streaming host core + demodulators).  The input is noise, so
  an FM discriminator output near +-pi can come out on either side of the
  branch cut for a last-bit difference of its inputs: FM outputs are
  compared as phase steps, their difference taken modulo 2 pi.
* WFM multiplex: 20 stereo stations with RDS at SNR 15-45 dB, 200 blocks
  (4 s), decimation 80 to 250 ksps, RDS decoded on the device in 0.5 s
  windows.  Against the oracle chain run on the same blocks (extraction,
  discriminator, stereo decoder, streaming RDS baseband, RDS decoder on the
  same windows): RDS baseband within 1e-4 relative RMS, decoded group bits
  identical per station, every PI decoded equal to the transmitted one, and
  PS recovered in full where the station's SNR allows (>= 24 dB)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from helpers import rel_rms
from serve_oracle import OracleChannel, OracleStereoChannel

pytestmark = pytest.mark.gpu

FS = 20_000_000
BLOCK = 400_000
THREADS = min(16, os.cpu_count() or 4)


def audio_rel(mode, got, want):
    """Relative RMS of demodulator outputs; FM (NFM/WFM) phase steps modulo 2 pi."""
    from argus import serve as S
    if mode in (S.DEMOD_NFM, S.DEMOD_WFM):
        d = np.angle(np.exp(1j * (np.asarray(got, np.float64) - np.asarray(want, np.float64))))
        den = np.sqrt(np.mean(np.asarray(want, np.float64) ** 2))
        return float(np.sqrt(np.mean(d ** 2)) / den) if len(d) else 0.0
    return rel_rms(got, want)


def _sample(chans):
    """A few channels of every decimation, including the longest and shortest taps."""
    by_d = {}
    for c in chans:
        by_d.setdefault(FS // c[3], []).append(c)
    pick = []
    for d, cs in sorted(by_d.items()):
        cs = sorted(cs, key=lambda c: c[4])
        pick += [cs[0], cs[-1]]
    return pick


def test_serving_mix_fast_and_direct_forms_and_oracle(gpu):
    import bench
    from argus import serve as S
    chans = bench.make_channels(8192)
    fast_eligible = sum(1 for c in chans if (FS // c[3]) % 8 == 0 and FS // c[3] < c[4] <= 512)
    assert fast_eligible > 2048
    svc = {}
    for fast in (True, False):
        s = S.Serve(FS, BLOCK, 32 << 30)
        s.SetFast(fast)
        for (cid, off, bw, rate, taps, dm) in chans:
            assert s.Open(S.ChannelSpec(cid, off, bw, rate, taps, dm, 0))
        svc[fast] = s
    sample = _sample(chans)
    orc = {c[0]: OracleChannel(FS, c[1], c[2], c[3], c[4], c[5], 0) for c in sample}
    acc = {cid: ([], [], [], []) for cid in orc}
    differ = 0
    pool = ThreadPoolExecutor(THREADS)
    for k in range(4):
        blk = bench.make_block(100 + k)
        print(f"mix block {k}: {len(chans)} channels", flush=True)
        t = {f: s.Submit(blk) for f, s in svc.items()}
        want = dict(zip(orc, pool.map(lambda cid: orc[cid].block(blk), orc)))
        for f, s in svc.items():
            s.Wait(t[f])
        worst = 0.0
        for (cid, _off, _bw, _rate, _taps, mode) in chans:
            a_iq, a_au, _ = svc[True].Output(t[True], cid)
            b_iq, b_au, _ = svc[False].Output(t[False], cid)
            assert len(a_iq) == len(b_iq) and len(a_au) == len(b_au), cid
            e = rel_rms(a_iq, b_iq)
            worst = max(worst, e)
            assert e <= 1e-5, (cid, k, e)
            if len(a_au):
                assert audio_rel(mode, a_au, b_au) <= 1e-4, (cid, k, audio_rel(mode, a_au, b_au))
            differ += int(not np.array_equal(a_iq, b_iq))
            if cid in want:
                w_iq, w_au = want[cid]
                assert len(a_iq) == len(w_iq) and len(a_au) == len(w_au), (cid, k)
                for i, v in enumerate((a_iq, w_iq, a_au, w_au)):
                    acc[cid][i].append(v)
        print(f"  fast vs direct worst rel RMS {worst:.2e}", flush=True)
    assert differ >= fast_eligible  # the eligible channels took the fast form after their first block
    modes = {c[0]: c[5] for c in chans}
    for cid, (gi, wi, ga, wa) in acc.items():
        e = rel_rms(np.concatenate(gi), np.concatenate(wi))
        assert e <= 1e-4, (cid, e)
        if sum(len(x) for x in wa):
            ea = audio_rel(modes[cid], np.concatenate(ga), np.concatenate(wa))
            assert ea <= 1e-4, (cid, ea)
    for s in svc.values():
        s.Close()


def test_wfm_multiplex_rds_200_blocks(gpu):
    from argus import serve as S
    from argus.synth import StereoScene, rds_group_bits
    rng = np.random.default_rng(5)
    n_st, n_blocks = 20, 200
    offs = [-9.0e6 + k * 0.9e6 + float(rng.uniform(-100e3, 100e3)) for k in range(n_st)]
    snr = rng.uniform(15.0, 45.0, n_st)
    pis = [int(v) for v in rng.integers(0x1000, 0xFFFF, n_st)]
    ps_tx = [f"ST{k:02d} FM" for k in range(n_st)]
    bits = [rds_group_bits(pis[k], ps_tx[k], f"radiotext of station {k}", 64) for k in range(n_st)]
    scene = StereoScene(FS, offs, snr, bits, seed=5, low_rate=1_000_000.0)
    rate = 250_000
    win = (rate // 13) // 2
    s = S.Serve(FS, BLOCK, 1 << 30)
    for k in range(n_st):
        assert s.Open(S.ChannelSpec(k + 1, offs[k], 200_000.0, rate, 101, S.DEMOD_WFM_STEREO, 0))
        assert s.SetRDS(k + 1, win)
    orc = [OracleStereoChannel(FS, offs[k], 200_000.0, rate, 101) for k in range(n_st)]
    g_rds = [[] for _ in range(n_st)]
    w_rds = [[] for _ in range(n_st)]
    g_groups = [[] for _ in range(n_st)]
    g_state = [None] * n_st
    pool = ThreadPoolExecutor(THREADS)
    for b in range(n_blocks):
        blk = scene.block(BLOCK)
        if b % 20 == 0:
            print(f"multiplex block {b}", flush=True)
        t = s.Submit(blk)
        want = list(pool.map(lambda k: orc[k].block(blk), range(n_st)))
        s.Wait(t)
        for k in range(n_st):
            x, _, _, _ = s.OutputWFM(t, k + 1)
            g_rds[k].append(x)
            w_rds[k].append(want[k][2])
            dec, _, g, pi, ps, _ = s.OutputRDS(t, k + 1)
            if dec:
                g_groups[k].append(g)
            g_state[k] = (pi, ps)
    s.Close()
    decoded = 0
    for k in range(n_st):
        gx, wx = np.concatenate(g_rds[k]), np.concatenate(w_rds[k])
        assert len(gx) == len(wx)
        assert rel_rms(gx, wx) <= 1e-4, (k, rel_rms(gx, wx))
        d = O.RDSDecoder()
        for i in range(0, len(wx) - win + 1, win):  # the windows the device decoded
            d.Decode(wx[i:i + win], rate // 13)
        dev = np.concatenate(g_groups[k]) if g_groups[k] else np.zeros((0, 4), np.uint16)
        assert [tuple(int(v) for v in r) for r in dev] == d.Groups(), k
        pi, ps = g_state[k]
        print(f"station {k}: SNR {snr[k]:.1f} dB, {len(dev)} groups, PI {pi:#06x}, PS {ps!r}", flush=True)
        assert pi == d.PI
        if pi:
            decoded += 1
            assert pi == pis[k], k
        if snr[k] >= 24.0:
            assert pi == pis[k] and ps == ps_tx[k], (k, pi, ps)
    assert decoded >= 15

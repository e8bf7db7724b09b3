"""The batch runner's fast-convolution form (StreamingExtractGPU,
streaming_gpu_stub.go:21-26, with eligible signals filtered through shared
segment transforms) against the oracle runner and the runner's own direct
form on the same chunk sequence: ragged chunks (below the 511-sample tail,
empty, a single sample), signals of D 8/16/32/40/64 with D < T and
ineligible ones beside them, a signal added mid-stream, a configuration
change (state reset, direct form until the history refills), a signal that
leaves and returns.  Samples within 1e-4 of the oracle (the reference's
native tolerance), phase 1e-8 rad, counts and histories as the oracle's;
fast and direct forms within 1e-5 relative RMS and not bit-identical (so
the fast form ran)."""
import numpy as np
import pytest

import oracle as O
from helpers import phase_diff, rel_rms

pytestmark = pytest.mark.gpu

SR = 2_560_000


def _job(sid, off, bw, rate, taps, hash_salt=0):
    from argus.gpudemod import StreamingConfigHash, StreamingExtractJob
    return StreamingExtractJob(sid, off, bw, rate, taps, StreamingConfigHash(sid, off, bw + hash_salt, rate, taps, SR))


BASE = [  # id, offset, bw, out rate, taps
    (1, 300_000.0, 200_000.0, 320_000, 101),   # D 8
    (2, -400_000.0, 25_000.0, 160_000, 127),   # D 16
    (3, 100_000.0, 10_000.0, 80_000, 65),      # D 32
    (4, -150_000.0, 3_000.0, 40_000, 255),     # D 64
    (5, 250_000.0, 12_500.0, 64_000, 255),     # D 40: fold by 8, every 5th
    (6, 700_000.0, 12_500.0, 256_000, 127),    # D 10: direct form
    (7, -800_000.0, 5_000.0, 6_400, 63),       # D 400 >= T: direct form
]
CHUNKS = [65_536, 1, 300, 0, 51_200, 7, 65_536, 40_000, 65_536, 3_584, 65_536]


def test_fast_form_against_oracle_and_direct(gpu):
    rng = np.random.default_rng(9)
    n_tot = sum(CHUNKS)
    t = np.arange(n_tot) / SR
    x = (0.5 * (rng.standard_normal(n_tot) + 1j * rng.standard_normal(n_tot))
         + np.exp(2j * np.pi * 300_000.0 * t) + 0.5 * np.exp(2j * np.pi * -150_010.0 * t)).astype(np.complex64)
    fast = gpu.NewBatchRunner(1 << 17, SR)
    direct = gpu.NewBatchRunner(1 << 17, SR)
    direct.SetFast(False)
    orc = O.CPUOracleRunner(SR)
    acc = {}
    pos = 0
    for k, n in enumerate(CHUNKS):
        specs = [b for b in BASE if not (b[0] == 6 and 4 <= k <= 5)]        # 6 leaves, then returns
        if k >= 3:
            specs = specs + [(8, -300_000.0, 25_000.0, 160_000, 201)]         # D 16, added mid-stream
        salt = 1.0 if k >= 6 else 0.0                                         # 2 changes configuration
        jobs = [_job(*b, hash_salt=salt if b[0] == 2 else 0.0) for b in specs]
        iq = x[pos:pos + n]
        pos += n
        print(f"runner fast chunk {k}: n={n} signals {[j.SignalID for j in jobs]}", flush=True)
        gf = fast.StreamingExtractGPU(iq, jobs)
        gd = direct.StreamingExtractGPU(iq, jobs)
        want = orc.StreamingExtract(iq, jobs)
        st = orc.States
        for a, b, w in zip(gf, gd, want):
            assert a.NOut == b.NOut == w.NOut and a.PhaseCount == w.PhaseCount, (k, a.SignalID)
            if w.NOut:
                assert np.max(np.abs(a.IQ - w.IQ)) <= 1e-4, (k, a.SignalID, np.max(np.abs(a.IQ - w.IQ)))
            acc.setdefault(a.SignalID, ([], []))
            acc[a.SignalID][0].append(a.IQ)
            acc[a.SignalID][1].append(b.IQ)
            gs, os_ = fast.SignalState(a.SignalID), st[a.SignalID]
            assert gs.PhaseCount == os_.PhaseCount and len(gs.ShiftedHistory) == len(os_.ShiftedHistory)
            assert phase_diff(gs.NCOPhase, os_.NCOPhase) <= 1e-8
            if len(os_.ShiftedHistory):
                assert np.max(np.abs(gs.ShiftedHistory - os_.ShiftedHistory)) <= 1e-4
    differ = 0
    for sid, (fa, da) in acc.items():
        f, d = np.concatenate(fa), np.concatenate(da)
        e = rel_rms(f, d)
        print(f"signal {sid}: {len(f)} outputs, fast vs direct rel RMS {e:.2e}", flush=True)
        assert e <= 1e-5, (sid, e)
        differ += int(not np.array_equal(f, d))
    assert differ >= 5  # the eligible signals (1-5, 8) took the fast form
    fast.Close()
    direct.Close()

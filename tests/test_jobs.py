"""Job building of the DSP loop (cmd/sdrd): snapOutRate, nearestExact-
DecimationRate, buildStreamingJobs and the legacy ExtractJob builder.

The oracle (oracle/__init__.py) is pinned by the reference's own cases and
stated answers -- snap_test.go:5-30 (divisor within 0.7..1.4 of the target),
helpers.go:283-291 and streaming_refactor.go:101-104 (4.096 MS/s with a
200 kHz target gives 204 800), the 20 Msps WFM finding (512 k does not
divide 20 M; snapped it is 500 k) -- and by exhaustive search over all
divisors on small rates.  The library (through the C ABI, host code only:
no GPU needed) must then equal the oracle exactly on random cases."""
import numpy as np
import pytest

import oracle as O

SNAP_CASES = [(4096000, 200000), (4096000, 512000), (2000000, 200000), (2048000, 200000), (4000000, 200000)]


@pytest.mark.parametrize("sr,target", SNAP_CASES)
def test_oracle_snap_reference_cases(sr, target):
    got = O.snapOutRate(sr, target)  # snap_test.go:5-30
    assert got > 0 and sr % got == 0
    assert 0.7 <= got / target <= 1.4


def test_oracle_snap_known_answers():
    assert O.snapOutRate(4096000, 200000) == 204800   # helpers.go:289-290
    assert O.snapOutRate(4096000, 512000) == 512000   # snap_test.go:9: 8.0 exact
    assert O.snapOutRate(20000000, 512000) == 500000  # 39.06 -> divisor 40
    assert O.snapOutRate(2048000, 200000) == 204800   # 10.24 -> 10
    assert O.snapOutRate(0, 200000) == 200000
    assert O.snapOutRate(48000, -1) == -1
    assert O.nearestExactDecimationRate(4096000, 200000, 128000) == 204800  # streaming_refactor.go:101-104
    assert O.nearestExactDecimationRate(20000000, 200000, 128000) == 200000  # exact already
    assert O.nearestExactDecimationRate(2048000, 200000, 250000) == 256000  # d 8..12: only 8 reaches min
    assert O.nearestExactDecimationRate(2048000, 200000, 300000) == 200000  # nothing >= min: target back


def test_oracle_snap_exhaustive_small_rates():
    """Against a search over every divisor (ties to the smaller divisor)."""
    rng = np.random.default_rng(1)
    for sr in list(range(1, 400)) + [int(v) for v in rng.integers(1000, 5_000_000, 200)]:
        for target in (1, 7, 48000, 200000, 512000, sr, sr + 1):
            divs = [d for d in range(1, min(256, sr) + 1) if sr % d == 0]
            ideal = sr / target
            best = min(divs, key=lambda d: (abs(d - ideal), d))
            assert O.snapOutRate(sr, target) == sr // best


def test_oracle_nearest_exact_properties():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        sr = int(rng.integers(100_000, 40_000_000))
        target, lo = int(rng.integers(50_000, 600_000)), int(rng.integers(1, 300_000))
        got = O.nearestExactDecimationRate(sr, target, lo)
        if sr % target == 0:
            assert got == target
            continue
        cands = [sr // d for d in range(max(1, sr // target - 2), sr // target + 3)
                 if sr // d >= lo and sr % (sr // d) == 0]
        cands = [c for c in cands if abs(target - c) < sr]  # the search starts at distance sr
        if not cands:
            assert got == target
        else:
            assert got in cands and abs(target - got) == min(abs(target - c) for c in cands)


def test_oracle_streaming_jobs_rules():
    sigs = [(1, 100.0e6, 180e3, ""),        # FM band: WFM floor 250 k
            (2, 87.5e6, 300e3, ""),         # band edges are inclusive
            (3, 108.0e6, 10e3, ""),
            (4, 145.5e6, 5e3, ""),          # NFM floor 20 k
            (5, 145.6e6, 30e3, "WFM"),      # classified WFM outside the band
            (6, 145.7e6, 30e3, "NFM")]
    jobs, bad = O.buildStreamingJobs(2_400_000, 100.0e6, sigs)
    assert bad == 0 and jobs is None        # 512 k does not divide 2.4 M
    jobs, bad = O.buildStreamingJobs(4_096_000, 100.0e6, sigs, bw_mult=1.2)
    assert bad == -1
    want_bw = [250e3, 360e3, 250e3, 20e3, 250e3, 36e3]
    for j, (sid, cen, *_), bw in zip(jobs, sigs, want_bw):
        assert j[0] == sid and j[1] == cen - 100.0e6 and j[2] == pytest.approx(bw)
        assert j[4] == 101
    assert [j[3] for j in jobs] == [512000, 512000, 512000, 204800, 512000, 204800]
    assert jobs[0][5] == O.StreamingConfigHash(1, 0.0, 250e3, 512000, 101, 4_096_000)
    # 20 Msps: the reference rejects the batch for WFM; snapped it extracts at 500 k
    jobs, bad = O.buildStreamingJobs(20_000_000, 100.0e6, sigs[:1])
    assert jobs is None and bad == 0
    jobs, bad = O.buildStreamingJobs(20_000_000, 100.0e6, sigs, fir_taps=63, snap_wfm=True)
    assert [j[3] for j in jobs] == [500000, 500000, 500000, 200000, 500000, 200000]
    assert all(j[4] == 63 for j in jobs)


def _random_signals(rng, n):
    mods = ["", "", "WFM", "WFM_STEREO", "NFM", "AM"]
    out = []
    for i in range(n):
        cen = float(rng.choice([rng.uniform(80e6, 115e6), rng.uniform(87.5e6, 108e6), rng.uniform(1e6, 500e6),
                                87.5e6, 108.0e6]))
        out.append((int(rng.integers(1, 1 << 40)), cen, float(rng.uniform(100, 400e3)), str(rng.choice(mods))))
    return out


def test_library_matches_oracle():
    from argus import jobs as J
    rng = np.random.default_rng(3)
    rates = [2_048_000, 2_000_000, 2_400_000, 4_096_000, 8_000_000, 10_000_000, 20_000_000, 3_200_000, 2_560_000]
    for sr in rates + [int(v) for v in rng.integers(1000, 50_000_000, 300)]:
        for target in (48000, 200000, 512000, 192000, 1, sr, 3 * sr):
            assert J.snapOutRate(sr, target) == O.snapOutRate(sr, target), (sr, target)
            assert (J.nearestExactDecimationRate(sr, target, 128000)
                    == O.nearestExactDecimationRate(sr, target, 128000)), (sr, target)
    for sr in rates:
        for snap in (False, True):
            for _ in range(20):
                sigs = _random_signals(rng, int(rng.integers(0, 12)))
                taps, mult = int(rng.choice([0, 63, 101, 127])), float(rng.choice([0.0, 1.0, 1.2]))
                cen = float(rng.uniform(80e6, 120e6))
                want, bad = O.buildStreamingJobs(sr, cen, sigs, fir_taps=taps, bw_mult=mult, snap_wfm=snap)
                ds = [J.DetectedSignal(*s) for s in sigs]
                cfg = J.ExtractionConfig(taps, mult, snap)
                if want is None:
                    with pytest.raises(ValueError, match=f"index {bad}"):
                        J.buildStreamingJobs(sr, cen, ds, cfg)
                    continue
                got = J.buildStreamingJobs(sr, cen, ds, cfg)
                assert [(g.SignalID, g.OffsetHz, g.Bandwidth, g.OutRate, g.NumTaps, g.ConfigHash) for g in got] \
                    == want
                got = J.buildExtractJobs(sr, cen, ds)
                assert [(g.OffsetHz, g.BW, g.OutRate, g.PhaseStart) for g in got] \
                    == O.buildExtractJobs(sr, cen, sigs)

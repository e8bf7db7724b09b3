"""GPU tests of gpud_wfm_stereo_batch, the one-shot stereo decoder of whole
recordings (wfmStereoDecodeBatch, internal/recorder/wfm_batch_stereo.go:
76-157): L/R audio within 1e-4 relative RMS of the oracle and the lock
decision identical, for several recordings of different lengths and rates
decoded in one call; ragged inputs (empty, rate 0, shorter than the filters,
not a whole number of tiles); the dual-mono fallback of
demodWFMStereoBatchAudio (lines 42-49)."""
import numpy as np
import pytest

import oracle as O
from helpers import rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _mpx(fs, seconds, pilot=0.09, seed=0, left_only=True):
    n = int(seconds * fs)
    t = np.arange(n) / fs
    rng = np.random.default_rng(seed)
    left = 0.5 * np.sin(2 * np.pi * 1000 * t)
    right = 0.0 * t if left_only else 0.4 * np.sin(2 * np.pi * 2500 * t)
    x = (0.45 * (left + right) + 0.45 * (left - right) * np.sin(2 * np.pi * 38000 * t)
         + pilot * np.sin(2 * np.pi * 19000 * t) + 0.02 * np.sin(2 * np.pi * 57000 * t)
         + 0.005 * rng.standard_normal(n))
    return x.astype(np.float32)


@pytest.fixture()
def wfm(gpu):
    from argus import wfm
    return wfm


def _check(wfm, monos, rates, fallback):
    print(f"batch of {len(monos)} lengths={[len(m) for m in monos]} rates={list(rates)} "
          f"fallback={fallback}", flush=True)
    audio, locked = wfm.StereoDecodeBatch(monos, rates, mono_fallback=fallback)
    for i, (m, r) in enumerate(zip(monos, rates)):
        if len(m) == 0 or r <= 0:
            assert len(audio[i]) == 0 and not locked[i]
            continue
        want, want_locked = O.StereoDecodeBatch(m, r)
        assert locked[i] == want_locked, i
        if fallback and not want_locked:
            want = np.repeat(m, 2)
        assert len(audio[i]) == len(want)
        assert rel_rms(audio[i], want) < TOL, (i, rel_rms(audio[i], want))
    return locked


def test_batch_matches_oracle_mixed_rates(wfm):
    rng = np.random.default_rng(3)
    monos = [
        _mpx(250_000, 0.4, seed=1),                       # locked, L only
        _mpx(200_000, 0.3, seed=2, left_only=False),      # locked, both channels
        _mpx(250_000, 0.2, pilot=0.005, seed=3),          # PLL tracks, pilot ratio below 0.003
        (0.3 * rng.standard_normal(60_000)).astype(np.float32),  # noise
        _mpx(192_000, 0.25, seed=4),
    ]
    rates = [250_000, 200_000, 250_000, 250_000, 192_000]
    locked = _check(wfm, monos, rates, False)
    assert locked[0] and locked[1] and locked[4]
    assert not locked[2] and not locked[3]


def test_batch_fallback_and_ragged(wfm):
    monos = [
        _mpx(250_000, 0.3, seed=5),
        np.zeros(0, np.float32),                          # empty recording
        _mpx(250_000, 0.01, seed=6)[:77],                 # shorter than the 101-tap filters
        _mpx(250_000, 0.05, seed=7)[:1024 * 5 + 333],     # not a whole number of tiles
        _mpx(250_000, 0.1, seed=8),                       # rate 0: skipped
        (0.2 * np.random.default_rng(9).standard_normal(30_001)).astype(np.float32),
    ]
    rates = [250_000, 250_000, 250_000, 250_000, 0, 250_000]
    locked = _check(wfm, monos, rates, True)
    assert locked[0] and not locked[5]
    _check(wfm, monos, rates, False)


def test_batch_single_and_repeated(wfm):
    """The same recording alone and inside a batch decodes identically."""
    m = _mpx(250_000, 0.2, seed=11, left_only=False)
    a1, l1 = wfm.StereoDecodeBatch([m], [250_000])
    a2, l2 = wfm.StereoDecodeBatch([_mpx(250_000, 0.1, seed=12), m], [250_000, 250_000])
    assert l1[0] == l2[1]
    np.testing.assert_array_equal(a1[0], a2[1])

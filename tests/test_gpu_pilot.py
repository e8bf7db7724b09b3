"""GPU tests of gpud_stereo_pilot_present (StereoPilotPresent /
estimateWFMPilot, internal/classifier/pll.go:82-131) against the oracle: the
201 bin magnitudes to FP64 rounding, the ratio, the offset of the strongest
bin and the present decision identical, for a batch of slices that covers a
stereo station, a pilot off 19 kHz, mono FM, noise, the 0.5 s window, a rate
below 40 kHz and slices too short to analyse."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _fm(mpx, fs, dev=75000.0):
    return np.exp(1j * 2 * np.pi * dev * np.cumsum(mpx) / fs).astype(np.complex64)


@pytest.fixture()
def wfm(gpu):
    from argus import wfm
    return wfm


def test_pilot_presence_matches_oracle(wfm):
    from argus.synth import WFMStation, rds_symbols
    fs = 250_000
    rng = np.random.default_rng(21)
    station = WFMStation(fs, 0.0, rds_symbols(np.zeros(800, np.uint8)), seed=3).block(fs // 2)
    t = np.arange(fs) / fs
    off_pilot = _fm(0.3 * np.sin(2 * np.pi * 900 * t) + 0.08 * np.sin(2 * np.pi * 19_300 * t)
                    + 0.01 * rng.standard_normal(fs), fs)
    # mono programme with a noise floor (an exact tone alone leaves every bin at rounding level)
    mono = _fm(0.5 * np.sin(2 * np.pi * 1000 * t[:fs // 2]) + 0.01 * rng.standard_normal(fs // 2), fs)
    noise = ((rng.standard_normal(fs // 3) + 1j * rng.standard_normal(fs // 3)) / 2).astype(np.complex64)
    long_then_mono = np.concatenate([WFMStation(fs, 0.0, rds_symbols(np.zeros(2000, np.uint8)), seed=5).block(fs),
                                     mono])
    fs2 = 192_000
    t2 = np.arange(fs2 // 4) / fs2
    short_station = _fm(0.2 * np.sin(2 * np.pi * 700 * t2) + 0.1 * np.sin(2 * np.pi * 18_750 * t2), fs2)
    slices = [station, off_pilot, mono, noise, long_then_mono, short_station,
              station[:40_000], station[:1], np.zeros(0, np.complex64), station[:5]]
    rates = [fs, fs, fs, fs, fs, fs2, 30_000, fs, fs, fs]
    print("pilot batch", [len(s) for s in slices], rates, flush=True)
    present, ratio, off, mags = wfm.StereoPilotPresent(slices, rates)
    for i, (x, r) in enumerate(zip(slices, rates)):
        ok, w_ratio, w_off, w_mags = O.StereoPilotEstimate(x, r)
        assert present[i] == ok, i
        assert off[i] == w_off, (i, off[i], w_off)
        top = max(w_mags.max(), 1e-300)
        np.testing.assert_allclose(mags[i], w_mags, rtol=1e-7, atol=1e-9 * top, err_msg=str(i))
        assert ratio[i] == pytest.approx(w_ratio, rel=1e-7, abs=1e-12), i
    assert list(present[:6]) == [True, True, False, False, False, True]
    assert off[1] == 300.0 and off[5] == -250.0
    assert not present[6:9].any() and not mags[6:9].any()  # below 40 kHz, one sample, empty

"""The product RDS group decoder (gpud_rds_*, host code in libargus_hip.so)
against the oracle decoder: identical group logs, PI, PS and RT on the same
baseband, fed in the same windows.  Host code only: runs without a GPU."""
import numpy as np
import pytest

import oracle as O


def _baseband(fs, seconds, seed, pi, ps, rt, n_groups, noise=0.0):
    from argus.synth import WFMStation, rds_symbols
    bits = O.RDSEncodeGroups(pi, ps, rt, n_groups)
    st = WFMStation(fs, 0.0, rds_symbols(bits), seed=seed)
    fm = O.BridgedFMDiscrim()
    s = O.RDSBasebandStream(fs)
    rng = np.random.default_rng(seed)
    out = []
    n = int(seconds * fs)
    for i in range(0, n, 5000):
        iq = st.block(min(5000, n - i))
        if noise:
            iq = (iq + noise * (rng.standard_normal(len(iq)) + 1j * rng.standard_normal(len(iq)))).astype(np.complex64)
        out.append(s.Process(fm.Demod(iq)))
    return np.concatenate(out), s.rate


@pytest.mark.parametrize("noise,window_s", [(0.0, 1.0), (0.1, 0.5), (0.0, 0.1), (0.0, 0.25)])
def test_decoder_matches_oracle(noise, window_s):
    from argus.wfm import RDSDecoder
    x, rate = _baseband(250_000, 3.0, 11, 0x54A8, "KXYZ FM ", "Now playing: a synthetic song", 36, noise)
    win = max(1, int(window_s * rate))
    got, want = RDSDecoder(), O.RDSDecoder()
    for i in range(0, len(x), win):
        g = got.Decode(x[i:i + win], rate)
        w = want.Decode(x[i:i + win], rate)
        assert g == w, (i, g, w)
    np.testing.assert_array_equal(got.Groups(), np.asarray(want.Groups(), np.uint16).reshape(-1, 4))
    assert (got.PI, got.PS, got.RT) == (want.PI, want.PS, want.RT)
    assert len(got.Groups()) > 0
    if window_s >= 0.5 and noise == 0.0:
        assert got.PI == 0x54A8 and got.PS == "KXYZ FM"


def test_decoder_rejects_short_and_bad_input():
    from argus.wfm import RDSDecoder
    d = RDSDecoder()
    assert d.Decode(np.zeros(103, np.complex64), 19230) == 0  # fewer than 104 samples
    assert d.Decode(np.zeros(500, np.complex64), 0) == 0      # no rate
    assert d.Groups().shape == (0, 4) and d.PI == 0 and d.PS == "" and d.RT == ""


def test_decoder_recovers_groups_from_clean_bits():
    """Known answer: NRZ symbols of encoded groups at 16 samples per bit decode
    to exactly those groups (direct polarity)."""
    from argus.wfm import RDSDecoder
    bits = O.RDSEncodeGroups(0x1234, "ABCDEFGH", "x" * 64, 8)
    sym = np.where(np.bitwise_xor.accumulate(np.concatenate([[0], bits]).astype(np.uint8)) > 0, 1.0, -1.0)
    rate = 19000  # 16 samples per symbol
    x = np.repeat(sym, 16).astype(np.complex64)
    d = RDSDecoder()
    assert d.Decode(x, rate) >= 7
    g = d.Groups()
    assert np.all(g[:, 0] == 0x1234) and d.PS == "ABCDEFGH"

"""The device-side RDS group decoder (gpud_rds_batch_*) against the oracle
decoder and the host decoder: per call the same number of groups found, and
at the end the same group log, PI, PS and RT for every station, with the
stations' windows of uneven and ragged sizes decoded together in each call."""
import numpy as np
import pytest

import oracle as O
from test_rds_decoder import _baseband

pytestmark = pytest.mark.gpu

STATIONS = [  # PI, PS, RT, seed, noise, window (s; 0 = ragged)
    (0x54A8, "KXYZ FM ", "Now playing: a synthetic song", 11, 0.0, 0.5),
    (0x1A2B, "ALPHA FM", "first station radiotext", 12, 0.1, 0.0),
    (0x3C4D, "BRAVO 99", "second station radiotext", 13, 0.0, 4.0),
    (0x5E6F, "CHARLIE ", "third", 14, 0.0, 0.1),
    (0x7001, "DELTA   ", "", 15, 0.05, 1.0),
]


@pytest.fixture()
def wfm_mod(gpu):
    from argus import wfm
    return wfm


def _windows(n, win, rng):
    if win > 0:
        return [(i, min(n, i + win)) for i in range(0, n, win)]
    out, i = [], 0
    while i < n:  # ragged: below one group, a few groups, and odd sizes
        w = int(rng.choice([7, 103, 104, 105, 999, 2048, 5000, 12345]))
        out.append((i, min(n, i + w)))
        i += w
    return out


def test_batch_matches_oracle_and_host(wfm_mod):
    rng = np.random.default_rng(3)
    sig = []
    for k, (pi, ps, rt, seed, noise, win_s) in enumerate(STATIONS):
        x, rate = _baseband(250_000, 4.5, seed, pi, ps, rt, 56, noise)
        wins = _windows(len(x), int(win_s * rate), rng)
        sig.append((x, rate, wins))
    dev = wfm_mod.RDSBatchDecoder(len(STATIONS) + 3)  # spare stations stay untouched
    host = [wfm_mod.RDSDecoder() for _ in STATIONS]
    orc = [O.RDSDecoder() for _ in STATIONS]
    step = 0
    while True:
        calls = [(k, sig[k][0][a:b], sig[k][1]) for k in range(len(STATIONS))
                 if step < len(sig[k][2]) for a, b in [sig[k][2][step]]]
        if not calls:
            break
        print(f"call {step}: stations {[c[0] for c in calls]}", flush=True)
        got = dev.Decode(calls)
        for (k, x, rate), g in zip(calls, got):
            w = orc[k].Decode(x, rate)
            h = host[k].Decode(x, rate)
            assert g == w == h, (step, k, g, w, h)
        step += 1
    for k, (pi, ps, *_rest) in enumerate(STATIONS):
        want = np.asarray(orc[k].Groups(), np.uint16).reshape(-1, 4)
        np.testing.assert_array_equal(dev.Groups(k), want)
        assert (dev.PI(k), dev.PS(k), dev.RT(k)) == (orc[k].PI, orc[k].PS, orc[k].RT)
        assert len(want) > 0, k
    assert dev.PI(0) == 0x54A8 and dev.PS(0) == "KXYZ FM"
    for spare in range(len(STATIONS), len(STATIONS) + 3):
        assert dev.Groups(spare).shape == (0, 4) and dev.PI(spare) == 0


@pytest.mark.parametrize("inverted", [False, True])
def test_batch_known_answer_clean_bits(wfm_mod, inverted):
    """NRZ symbols of encoded groups, 16 samples per bit: the direct stream
    decodes to exactly those groups; with every data bit inverted before the
    differential encoding the inverted-stream scan finds them instead."""
    bits = O.RDSEncodeGroups(0x1234, "ABCDEFGH", "x" * 64, 8).astype(np.uint8)
    if inverted:
        bits = 1 - bits
    sym = np.where(np.bitwise_xor.accumulate(np.concatenate([[0], bits]).astype(np.uint8)) > 0, 1.0, -1.0)
    x = np.repeat(sym, 16).astype(np.complex64)
    dev = wfm_mod.RDSBatchDecoder(2)
    want = O.RDSDecoder()
    (g,) = dev.Decode([(1, x, 19000)])
    assert g == want.Decode(x, 19000) and g >= 7
    got = dev.Groups(1)
    np.testing.assert_array_equal(got, np.asarray(want.Groups(), np.uint16).reshape(-1, 4))
    assert np.all(got[:, 0] == 0x1234) and dev.PS(1) == "ABCDEFGH"


def test_batch_short_bad_and_reset(wfm_mod):
    from argus._lib import GpudError
    dev = wfm_mod.RDSBatchDecoder(3)
    assert dev.Decode([(0, np.zeros(103, np.complex64), 19230), (1, np.zeros(500, np.complex64), 0),
                       (2, np.zeros(0, np.complex64), 19230)]) == [0, 0, 0]
    with pytest.raises(GpudError):
        dev.Decode([(0, np.zeros(200, np.complex64), 19230), (0, np.zeros(200, np.complex64), 19230)])
    with pytest.raises(GpudError):
        dev.Decode([(3, np.zeros(200, np.complex64), 19230)])
    x, rate = _baseband(250_000, 1.0, 11, 0x54A8, "KXYZ FM ", "radiotext", 12)
    (g,) = dev.Decode([(2, x, rate)])
    assert g > 0 and dev.PI(2) == 0x54A8
    dev.Reset(2)
    assert dev.PI(2) == 0 and dev.Groups(2).shape == (0, 4) and dev.PS(2) == ""
    # a fresh decoder again: the same call finds the same groups
    assert dev.Decode([(2, x, rate)]) == [g]


@pytest.mark.parametrize("n_st", [257, 301, 771])
def test_batch_more_stations_than_cus(wfm_mod, n_st):
    """More calls than compute units: the symbol kernel packs two to four
    calls per workgroup, the last workgroup partly filled, and the waves of
    one workgroup run windows of different lengths.  Every station's groups
    equal the host decoder's, call by call."""
    base = [_baseband(250_000, 1.0, seed, pi, ps, rt, 56, noise)
            for pi, ps, rt, seed, noise, _ in STATIONS]
    dev = wfm_mod.RDSBatchDecoder(n_st)
    host = [wfm_mod.RDSDecoder() for _ in range(n_st)]
    for step in range(2):
        calls = []
        for k in range(n_st):
            x, rate = base[k % len(base)]
            cut = len(x) // 2 - (k % 7) * 13
            calls.append((k, x[:cut] if step == 0 else x[cut:], rate))
        print(f"{n_st} stations, call {step}", flush=True)
        got = dev.Decode(calls)
        for (k, x, rate), g in zip(calls, got):
            assert g == host[k].Decode(x, rate), (n_st, step, k)
    for k in range(n_st):
        want = np.asarray(host[k].Groups(), np.uint16).reshape(-1, 4)
        np.testing.assert_array_equal(dev.Groups(k), want)
        assert dev.PI(k) == host[k].PI, k
    assert len(dev.Groups(0)) > 0

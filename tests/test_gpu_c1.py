"""The plain CPU-runnable configuration at full size through the GPU service:
2.048 Msps, 1 s, 4 WFM stations at seeded offsets within +-900 kHz (>= 200
kHz apart), 75 kHz deviation, 1 kHz mono tone, AWGN at 30 dB in-band SNR;
per station 127-tap Hamming low-pass at 100 kHz, /8 to 256 ksps, FM
discriminator; 16384-sample chunks.  Baseband and audio within 1e-4 relative
RMS of the oracle chain; the recovered audio is the 1 kHz tone."""
import numpy as np
import pytest

from helpers import rel_rms
from serve_oracle import WFM, OracleChannel

pytestmark = pytest.mark.gpu

FS, CHUNK = 2_048_000, 16384


def scene():
    rng = np.random.default_rng(0)
    offs = []
    while len(offs) < 4:
        f = float(rng.uniform(-900e3, 900e3))
        if all(abs(f - g) >= 200e3 for g in offs):
            offs.append(f)
    t = np.arange(FS) / FS
    x = np.zeros(FS, np.complex128)
    for f in offs:
        x += np.exp(1j * (2 * np.pi * f * t + (75000 / 1000) * np.sin(2 * np.pi * 1000 * t)))
    sig_p = 1.0  # each station has unit power
    sigma = np.sqrt(sig_p * 200e3 / FS / 10 ** 3.0 / 2)
    x += sigma * (rng.standard_normal(FS) + 1j * rng.standard_normal(FS))
    return x.astype(np.complex64), offs


def test_c1_four_wfm_stations(gpu):
    from argus import serve as S
    x, offs = scene()
    s = S.Serve(FS, CHUNK, 1 << 30)
    orc = {}
    for k, f in enumerate(offs):
        assert s.Open(S.ChannelSpec(k + 1, f, 200e3, 256_000, 127, S.DEMOD_WFM, 1))
        orc[k + 1] = OracleChannel(FS, f, 200e3, 256_000, 127, WFM, 1)
    acc = {c: ([], [], [], []) for c in orc}
    for i in range(0, FS, CHUNK):
        blk = x[i:i + CHUNK]
        t = s.Submit(blk)
        s.Wait(t)
        for c, oc in orc.items():
            wi, wa = oc.block(blk)
            gi, ga, rate = s.Output(t, c)
            assert rate == 256_000 and len(gi) == len(wi) and len(ga) == len(wa)
            for lst, v in zip(acc[c], (gi, wi, ga, wa)):
                lst.append(v)
    s.Close()
    for c, (gi, wi, ga, wa) in acc.items():
        gi, wi, ga, wa = map(np.concatenate, (gi, wi, ga, wa))
        assert len(wi) == 256_000
        e_iq, e_au = rel_rms(gi, wi), rel_rms(ga, wa)
        print(c, e_iq, e_au, flush=True)
        assert e_iq <= 1e-4 and e_au <= 1e-4, (c, e_iq, e_au)
        # the discriminator of a 1 kHz tone at 75 kHz deviation: 2 pi 75e3/256e3 peak
        spec = np.abs(np.fft.rfft(ga[1000:1000 + 256_00]))
        assert abs(np.argmax(spec[1:]) + 1 - 100) <= 1  # 1 kHz at 10 Hz bins

"""Oracle model of one served channel: the reference's streaming extractor
(streaming_host_core.go) followed by the reference demodulators, with the
block-to-block state the service carries.  Test infrastructure only."""
import numpy as np

import oracle as O

NONE, NFM, WFM, AM, USB, LSB, CW, WFM_STEREO = range(8)


def pre_demod_stage(out_rate, demod_rate, wfm):
    """The live recorder's pre-demod decimating FIR (streamer.go:1098-1124):
    factor round(snipRate / demodRate), forced to 2 for WFM when that keeps
    >= 200 kHz; LowpassFIR at 90 kHz (WFM) or 0.8 of the demod Nyquist, 101
    taps; None when the factor is 1."""
    if not demod_rate:
        return None, 1
    f = max(1, O.go_round(out_rate / demod_rate))
    if wfm and f > 2 and out_rate // 2 >= 200000:
        f = 2
    if f <= 1:
        return None, 1
    cutoff = 90000.0 if wfm else (out_rate // f) / 2.0 * 0.8
    return O.StatefulDecimatingFIRComplex(O.LowpassFIR(cutoff, out_rate, 101), f), f


class OracleChannel:
    def __init__(self, sample_rate, offset_hz, bandwidth, out_rate, num_taps=101, demod=NONE, window=0,
                 demod_rate=0):
        self.sr = sample_rate
        self.off = offset_hz
        self.decim = O.ExactIntegerDecimation(sample_rate, out_rate)
        self.T = num_taps if num_taps > 0 else 101
        cutoff = max(bandwidth / 2, 200)
        design = O.LowpassFIRHamming if window == 1 else O.LowpassFIR
        base = design(cutoff, sample_rate, self.T).astype(np.float32)[:self.T]
        self.poly = O.BuildPolyphaseTapsPhaseMajor(base, self.decim)
        self.phase, self.pc, self.hist = 0.0, 0, np.zeros(0, np.complex64)
        self.demod = demod
        self.pre, f = pre_demod_stage(out_rate, demod_rate, demod in (WFM, WFM_STEREO))
        self.out_rate = out_rate // f   # the demodulator's rate
        self.fm = O.BridgedFMDiscrim()
        self.bfo = 0.0
        self.last_pre = None

    def block(self, iq):
        bb, self.phase, self.pc, self.hist = O.runStreamingPolyphaseHostCore(
            iq, self.sr, self.off, self.phase, self.pc, self.T, self.decim, self.hist, self.poly)
        out = bb
        if self.pre is not None:
            out = self.pre.Process(bb) if len(bb) else np.zeros(0, np.complex64)
            self.last_pre = out
        if self.demod in (NFM, WFM):
            audio = self.fm.Demod(out)
        elif self.demod == AM:
            audio = O.AMDemod(out)
        elif self.demod in (USB, LSB):
            audio, self.bfo = O.SSBDemod(out, self.out_rate, self.demod == USB, self.bfo)
        elif self.demod == CW:
            audio, self.bfo = O.CWDemod(out, self.out_rate, self.bfo)
        else:
            audio = np.zeros(0, np.float32)
        return bb, audio


class OracleStereoChannel:
    """A WFM stereo channel: extraction and the bridged discriminator
    (OracleChannel), then stereoDecodeStateful with the lock hysteresis and
    the mono fallback filter (streamer.go:1367-1395, 1511-1643), and the
    streaming RDS baseband of the discriminator output (fm.go:234-271)."""

    def __init__(self, sample_rate, offset_hz, bandwidth, out_rate, num_taps=101, window=0, demod_rate=0):
        self.ch = OracleChannel(sample_rate, offset_hz, bandwidth, out_rate, num_taps, WFM, window, demod_rate)
        rate = self.ch.out_rate
        self.stereo = O.StereoDecoder(rate)
        self.fallback = O.StatefulFIRReal(O.LowpassFIR(15000, rate, 101))
        self.rds = O.RDSBasebandStream(rate)

    def block(self, iq):
        """(baseband, audio L/R interleaved, rds baseband, locked, stereo)."""
        out, mono = self.ch.block(iq)
        locked, stereo = False, self.stereo.Enabled
        audio = np.zeros(0, np.float32)
        if len(mono):
            lr, locked = self.stereo.Decode(mono)
            stereo = self.stereo.Enabled
            audio = lr if stereo else np.repeat(self.fallback.Process(mono), 2)
        return out, audio, self.rds.Process(mono), locked, stereo


def synth_scene(n, sr, carriers, seed=0, noise=0.05):
    """Sum of modulated carriers + complex AWGN.  carriers: (kind, offset_hz, level)."""
    rng = np.random.default_rng(seed)
    t = np.arange(n) / sr
    x = noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
    for kind, f0, lvl in carriers:
        if kind == "fm":
            dev, fm = 5000.0, 1000.0
            ph = 2 * np.pi * f0 * t + dev / fm * np.sin(2 * np.pi * fm * t)
            x += lvl * np.exp(1j * ph)
        elif kind == "wfm":
            dev, fm = 75000.0, 1000.0
            ph = 2 * np.pi * f0 * t + dev / fm * np.sin(2 * np.pi * fm * t)
            x += lvl * np.exp(1j * ph)
        elif kind == "am":
            x += lvl * (1 + 0.5 * np.cos(2 * np.pi * 700 * t)) * np.exp(2j * np.pi * f0 * t)
        elif kind == "ssb":
            x += lvl * np.exp(2j * np.pi * (f0 + 900) * t)
        else:
            x += lvl * np.exp(2j * np.pi * f0 * t)
    return x.astype(np.complex64)


class OracleAudioStage:
    """The post-demodulation audio stage of one channel
    (streamer.go:1382-1470): plain WFM low-passed at 15 kHz
    (wfmAudioFilter, streamer.go:1511-1520), resampled from the demod rate to
    audio_rate (NewResampler / NewStereoResampler with 32 taps per phase,
    skipped when the rates are equal), then for WFM and WFM stereo the
    de-emphasis y = a y + (1 - a) x with a = exp(-1 / (audio_rate tau)) and
    the gain 0.35.  process() takes the demodulator's block output (L/R
    interleaved for stereo) and returns (frames, channels)."""

    def __init__(self, demod, in_rate, audio_rate=48000, deemph_us=50.0):
        self.wfm = demod in (WFM, WFM_STEREO)
        self.ch = 2 if demod == WFM_STEREO else 1
        self.lpf = O.StatefulFIRReal(O.LowpassFIR(15000, in_rate, 101)) if demod == WFM else None
        self.same = in_rate == audio_rate
        self.rs = O.StereoResampler(in_rate, audio_rate, 32) if self.ch == 2 else O.Resampler(in_rate, audio_rate, 32)
        self.deemph = self.wfm and deemph_us > 0
        self.alpha = np.exp(-1.0 / (audio_rate * deemph_us * 1e-6)) if self.deemph else 0.0
        self.y = [0.0, 0.0]

    def process(self, audio):
        a = np.ascontiguousarray(audio, np.float32)
        if self.lpf is not None and len(a):
            a = self.lpf.Process(a)
        if not self.same:
            a = self.rs.Process(a)
        if self.deemph:
            for s in range(self.ch):
                a, self.y[s] = O.Deemphasis(a, self.alpha, self.y[s], stride=self.ch, offset=s)
        if self.wfm:
            a = a * np.float32(0.35)
        return a.reshape(-1, self.ch)

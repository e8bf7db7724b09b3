"""The RDS caller flow on the device (updateRDS, pipeline_runtime.go:
1044-1156, with extractRDS helpers.go:117-160 and the fallback chain of
1104-1124) against the oracle's RDSRuntime on the same ticks and windows:

* route A (ShiftFilterDecimateBatch) at 2.5 Msps, D = 10 -> 250 ksps, three
  stereo stations with RDS, ticks every half second up to 9.5 s: decode
  turns at 0.5 s (a window shorter than one second: turn taken, nothing
  decoded), 4.5 and 8.5 s on the last 4 s of IQ;
* route B (the fallback chain) at 2.4 Msps, where 2.4 M / 9 does not divide
  2.4 M: 51-tap /2 to 1.2 Msps, then 101-tap /4 to 300 ksps;
* the daemon's shape, 20 Msps (D = 80): 20 stations decoded in one tick
  from a window resident in device memory, two checked against the oracle
  chain and every station's PI against the transmitted one.
Per decoded station: baseband and RDS baseband within 1e-4 relative RMS,
rates equal, the stereo flag equal, decoded groups identical, the station
name equal to the oracle's.

Route B's first stage is dsp.FreqShift (freqshift.go:6-22), which adds the
phase increment to an FP64 accumulator once per sample; over a 4 s window
at 2.4 Msps the accumulator drifts from the exact phase (i + 1) inc by up to
3e-3 rad (RMS 1.3e-3 at -600 kHz).  The device rotates by the exact phase,
so the oracle's baseband is compared after removing that drift, computed
independently here by sequential FP64 accumulation (numpy's accumulate);
the drift shifts the carrier phase only, so the discriminator, the RDS
baseband and everything decoded from it are compared as they are."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import rel_rms

pytestmark = pytest.mark.gpu

CENTER = 98.0e6


def _scene(fs, offs, snr, seed, low_rate):
    from argus.synth import StereoScene, rds_group_bits
    pis = [0x2000 + 0x111 * k for k in range(len(offs))]
    names = [f"ST{k:02d} FM" for k in range(len(offs))]
    bits = [rds_group_bits(pis[k], names[k], f"radiotext {k}", 64) for k in range(len(offs))]
    return StereoScene(fs, offs, snr, bits, seed=seed, low_rate=low_rate), pis, names


def _freqshift_drift(fs, offset, n):
    """FP64 accumulated phase minus the exact (i + 1) inc, i < n."""
    inc = -2 * np.pi * offset / fs
    return np.add.accumulate(np.full(n, inc)) - inc * np.arange(1, n + 1, dtype=np.float64)


def _check_tick(rt, orc, sigs, got, want, tag, center=None, fs=None, route="A", n_win=0):
    assert [g[2] for g in got] == [w[2] for w in want], tag
    for (sid, _), g, w in zip(sigs, got, want):
        assert g[0] == w[0], (tag, sid, "stereo")
        assert g[1] == w[1], (tag, sid, "station", g[1], w[1])
        if g[2] != 1 or sid not in orc.last:
            continue
        last = rt.Last(sid)
        assert last is not None, (tag, sid)
        base, brate, rds, rrate = last
        wb, wbr, wr, wrr = orc.last[sid]
        assert (brate, rrate) == (wbr, wrr), (tag, sid)
        assert len(base) == len(wb) and len(rds) == len(wr), (tag, sid, len(base), len(wb), len(rds), len(wr))
        if route == "B":
            d1 = max(1, fs // 1_000_000)
            d2 = max(1, (fs // d1) // 250_000)
            err = _freqshift_drift(fs, dict(sigs)[sid] - center, n_win)
            wb = wb * np.exp(-1j * err[np.arange(len(wb)) * d1 * d2])
        eb, er = rel_rms(base, wb), rel_rms(rds, wr)
        print(f"{tag} station {sid}: baseband {len(base)} @ {brate}, rel RMS {eb:.2e}; "
              f"RDS {len(rds)} @ {rrate}, rel RMS {er:.2e}; stereo {g[0]} station {g[1]!r}", flush=True)
        assert eb <= 1e-4 and er <= 1e-4, (tag, sid, eb, er)
        dev = [tuple(int(v) for v in r) for r in rt.Groups(sid)]
        assert dev == orc.states[sid]["dec"].Groups(), (tag, sid)


@pytest.mark.parametrize("fs,low,offs,route", [
    (2_500_000, 500_000, [-800e3, 50e3, 700e3], "A"),
    (2_400_000, 600_000, [-600e3, 500e3], "B"),
])
def test_ticks_against_oracle(gpu, fs, low, offs, route):
    from argus.rds_runtime import RDSRuntime
    d = max(1, fs // 250000)
    assert (fs % (fs // d) == 0) == (route == "A")
    scene, pis, names = _scene(fs, offs, [40.0] * len(offs), 3, low)
    rt = RDSRuntime(16)
    orc = O.RDSRuntime()
    sigs = [(k + 1, CENTER + off) for k, off in enumerate(offs)]
    hist = np.zeros(0, np.complex64)
    blk = fs // 2
    for t in range(19):  # ticks at 0.5, 1.0, ..., 9.5 s
        hist = np.concatenate([hist, scene.block(blk)])[-4 * fs:]
        now = 0.5 * (t + 1)
        print(f"route {route} tick {now:.1f} s: window {len(hist)}", flush=True)
        got = rt.Update(now, sigs, hist, fs, CENTER)
        want = orc.Update(now, sigs, hist, fs, CENTER)
        _check_tick(rt, orc, sigs, got, want, f"{route}@{now:.1f}", CENTER, fs, route, len(hist))
    for k, (sid, _) in enumerate(sigs):
        st = rt.State(sid)
        assert st["pi"] == pis[k] and st["ps"] == names[k], (sid, st)
        assert st["stereo"] == 1
    rt.Close()


def test_daemon_shape_20msps_device_window(gpu):
    from argus import lib
    from argus.rds_runtime import RDSRuntime
    fs, n_st = 20_000_000, 20
    rng = np.random.default_rng(8)
    offs = [-9.0e6 + k * 0.9e6 + float(rng.uniform(-100e3, 100e3)) for k in range(n_st)]
    snr = rng.uniform(24.0, 45.0, n_st)
    scene, pis, names = _scene(fs, offs, snr, 8, 1_000_000)
    win = np.concatenate([scene.block(400_000) for _ in range(200)])  # 4 s
    dev = C.c_void_p()
    assert lib().gpud_malloc(C.byref(dev), win.nbytes) == 0
    try:
        assert lib().gpud_memcpy_h2d(dev, win.ctypes.data, win.nbytes) == 0
        rt = RDSRuntime(64)
        sigs = [(100 + k, CENTER + off) for k, off in enumerate(offs)]
        got = rt.Update(4.0, sigs, window_rate=fs, window_center_hz=CENTER, window_dev=dev.value,
                        n_window=len(win))
        assert all(g[2] == 1 for g in got)
        for k, (sid, _) in enumerate(sigs):
            st = rt.State(sid)
            print(f"station {sid}: SNR {snr[k]:.1f} dB, PI {st['pi']:#06x}, PS {st['ps']!r}, stereo {st['stereo']}",
                  flush=True)
            assert st["pi"] == pis[k] and st["stereo"] == 1, (sid, st)
        orc = O.RDSRuntime()
        pick = [sigs[0], sigs[13]]
        want = orc.Update(4.0, pick, win, fs, CENTER)
        _check_tick(rt, orc, pick, [got[0], got[13]], want, "20M")
        rt.Close()
    finally:
        lib().gpud_free(dev)

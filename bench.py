#!/usr/bin/env python3
"""Benchmark: real-time channels sustained at a 20 Msps span on one MI355X.

Workload (the serving shape at 20 Msps): complex64 IQ at 20 Msps in 20 ms
blocks (400 000 samples); CH channels drawn from the uneven mode mix
{NFM 12.5 kHz, AM 10 kHz, USB 3 kHz, WFM 200 kHz} with decimations 8..1600
and odd FIR lengths 63..511 (seed 0).  One step = one block through the fused
batched extractor (NCO shift + polyphase FIR/decimate for every channel, one
launch) with the block already resident in HBM.

    value = CH * block_duration / ms_per_step   (real-time channels)

Reported beside it, in the same run: per-block latency p50/p99 (event-timed
device work, and host-to-host with the block's upload and the outputs'
download), the plain one-kernel-per-signal GPU form (the reference's call
pattern: per-signal upload, launch, synchronous copies), the roofline of the
fused kernel, and the CPU oracle timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "argus-sdr_amd"))

FS = 20_000_000
BLOCK = 400_000            # 20 ms at 20 Msps
BLOCK_S = BLOCK / FS
MODES = [("NFM", 12_500.0, [400, 800]), ("AM", 10_000.0, [500, 1000]), ("USB", 3_000.0, [1250, 1600]),
         ("WFM", 200_000.0, [8, 16, 40, 80])]
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "real-time channels sustained @20 Msps span (1 GPU); p99 block latency; HBM GB/s"


def make_channels(n: int, seed: int = 0):
    from argus.gpudemod import StreamingExtractJob
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        name, bw, ds = MODES[int(rng.integers(0, 4))]
        d = int(rng.choice(ds))
        t = 2 * int(rng.integers(31, 256)) + 1
        off = float(rng.uniform(-9.5e6, 9.5e6))
        jobs.append(StreamingExtractJob(SignalID=i + 1, OffsetHz=off, Bandwidth=bw, OutRate=FS // d, NumTaps=t,
                                        ConfigHash=0x9E3779B97F4A7C15 ^ (i + 1)))
    return jobs


def make_block(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(2 * BLOCK, dtype=np.float32) * np.float32(0.5)).view(np.complex64)
    return x


def alg_figures(jobs, n_in):
    """Algorithmic bytes and flops of one fused launch (unique bytes only)."""
    b = 8 * n_in
    f = 0
    for j in jobs:
        d = FS // j.OutRate
        n_out = n_in // d
        b += n_out * 8 + 16 * (j.NumTaps - 1) + 4 * j.NumTaps
        f += 6 * n_in + 4 * j.NumTaps * n_out
    return b, f


def cpu_baseline(jobs, block, threads):
    import oracle as O
    r = O.CPUOracleRunner(FS, threads=threads)
    t0 = time.perf_counter()
    r.StreamingExtract(block, jobs)
    dt = time.perf_counter() - t0
    return dt


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--channels", type=int, default=256)
    ap.add_argument("--plain-blocks", type=int, default=2)
    ap.add_argument("--cpu-channels", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the extraction path shards by channel and this build runs on a "
              "single MI355X; run with --gpus 1.", file=sys.stderr)
        return 2

    import argus
    from argus import gpudemod as G
    from argus._lib import lib, check
    L = lib()
    if not G.Available():
        print("no HIP device", file=sys.stderr)
        return 1

    jobs = make_channels(a.channels)
    blocks = [make_block(s) for s in range(4)]
    print(f"[bench] {a.channels} channels, block {BLOCK} samples @ {FS/1e6:.0f} Msps", file=sys.stderr, flush=True)

    # ---- fused batched form, block resident in HBM --------------------------
    r = G.NewBatchRunner(BLOCK, FS)
    d_blocks = []
    for blk in blocks:
        p = C.c_void_p()
        check("gpud_malloc", L.gpud_malloc(C.byref(p), blk.nbytes))
        check("gpud_memcpy_h2d", L.gpud_memcpy_h2d(p, blk.ctypes.data, blk.nbytes))
        d_blocks.append(p.value)
    stream = r.stream()
    for i in range(a.warmup):
        r.StreamingExtractDevice(d_blocks[i % 4], BLOCK, jobs)
    check("gpud_stream_sync", L.gpud_stream_sync(stream))
    evs = []
    for _ in range(2 * a.steps):
        e = C.c_void_p()
        check("gpud_event_create", L.gpud_event_create(C.byref(e)))
        evs.append(e.value)
    check("enable timing", L.gpud_runner_enable_kernel_timing(r._h, 1))
    check("gpud_device_sync", L.gpud_device_sync())
    t0 = time.perf_counter()
    for i in range(a.steps):
        L.gpud_event_record(evs[2 * i], stream)
        r.StreamingExtractDevice(d_blocks[i % 4], BLOCK, jobs)
        L.gpud_event_record(evs[2 * i + 1], stream)
    check("gpud_stream_sync", L.gpud_stream_sync(stream))
    check("gpud_device_sync", L.gpud_device_sync())
    wall = time.perf_counter() - t0
    tot = C.c_double()
    nl = C.c_int()
    check("kernel time", L.gpud_runner_kernel_time(r._h, C.byref(tot), C.byref(nl)))
    L.gpud_runner_enable_kernel_timing(r._h, 0)
    step_ms = []
    for i in range(a.steps):
        ms = C.c_float()
        check("elapsed", L.gpud_event_elapsed_ms(evs[2 * i], evs[2 * i + 1], C.byref(ms)))
        step_ms.append(ms.value)
    ms_per_step = wall * 1e3 / a.steps
    kern_ms = tot.value / max(nl.value, 1)
    value = a.channels * BLOCK_S * 1e3 / ms_per_step

    # ---- host-to-host per-block latency (upload + extract + download) ------
    lat = []
    for i in range(min(50, a.steps)):
        t1 = time.perf_counter()
        r.StreamingExtractGPU(blocks[i % 4], jobs)
        lat.append((time.perf_counter() - t1) * 1e3)
    r.Close()

    # ---- plain per-signal GPU form ----------------------------------------
    rp = G.NewBatchRunner(BLOCK, FS, exec_mode=G.EXEC_PER_SIGNAL)
    rp.StreamingExtractGPU(blocks[0], jobs)
    t1 = time.perf_counter()
    for i in range(a.plain_blocks):
        rp.StreamingExtractGPU(blocks[(i + 1) % 4], jobs)
    plain_ms = (time.perf_counter() - t1) * 1e3 / a.plain_blocks
    rp.Close()
    plain_channels = a.channels * BLOCK_S * 1e3 / plain_ms

    for p in d_blocks:
        L.gpud_free(p)
    for e in evs:
        L.gpud_event_destroy(e)

    b_alg, f_alg = alg_figures(jobs, BLOCK)
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    valu = f_alg / (kern_ms * 1e-3) / 1e12

    cpu = None
    if not a.no_cpu:
        sub = jobs[:a.cpu_channels]
        dt = cpu_baseline(sub, blocks[0], 1)
        cpu = {"value": round(len(sub) * BLOCK_S / dt, 3), "unit": "real-time channels", "cores": 1,
               "kind": "port",
               "sample": f"oracle streaming extractor, {len(sub)} channels of the same mix x one 20 ms block "
                         f"({dt:.2f} s of CPU work, single thread)"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "real-time channels",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uneven-mode extraction at 20 Msps, 20 ms blocks", "sample_rate": FS,
                   "block_samples": BLOCK, "channels": a.channels, "modes": "NFM/AM/USB/WFM",
                   "decim": "8-1600", "taps": "63-511", "seed": 0},
        "p50_block_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_block_ms": round(float(np.percentile(step_ms, 99)), 4),
        "p50_block_ms_host": round(float(np.percentile(lat, 50)), 3),
        "p99_block_ms_host": round(float(np.percentile(lat, 99)), 3),
        "hbm_gbps": round(achieved, 1),
        "plain_per_signal": {"ms_per_block": round(plain_ms, 2), "real_time_channels": round(plain_channels, 1),
                             "fused_over_plain": round(value / plain_channels, 1)},
        "roofline": {"kernel": "argus_extract_batch_kernel", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel_ms": round(kern_ms, 4), "alg_bytes": b_alg, "alg_flops": f_alg,
                     "valu_tflops": round(valu, 2), "valu_frac": round(valu / FP32_PEAK_TFLOPS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

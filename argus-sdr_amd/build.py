"""Build recipe for libargus_hip.so (gfx950) and the CPU oracle.

    python argus-sdr_amd/build.py            # both
    python argus-sdr_amd/build.py --lib      # HIP library only

The HIP sources are compiled in one hipcc invocation into a shared library
that sits next to this file (argus-sdr_amd/libargus_hip.so), so it travels
with the source tree.  Packed FP32 instructions are disabled for all device
code (-packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libargus_hip.so"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ARGUS_OFFLOAD_ARCH", "gfx950")


def hip_sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build_lib(verbose: bool = False) -> Path:
    srcs = hip_sources()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        HIPCC,
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-fvisibility=hidden",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        # release paths (hipFree/hipHostFree on teardown) are best effort
        "-Wno-unused-value", "-Wno-unused-result",
        f"-I{ROOT / 'include'}",
        "-o", str(tmp),
        *map(str, srcs),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; drop that noise
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l and "(ignoring feature)" not in l)
    if res.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    if err.strip() and verbose:
        sys.stderr.write(err + "\n")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    res = subprocess.run(["make", "-s", "-C", str(ORACLE), "-B", "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    return ORACLE / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    both = not (a.lib or a.oracle)
    if a.lib or both:
        print(build_lib(a.verbose))
    if a.oracle or both:
        print(build_oracle(a.verbose))


if __name__ == "__main__":
    main()

"""The FP64 sine/cosine on the RDS Costas recurrence
(argus-sdr_amd/csrc/fp64_sincos.hpp) against the C library on the host:
over [-pi, pi] its results round to the same float32 as libm's (the
decoder uses float32(cos(phase)) and float32(sin(phase)), rds.go:160-161),
and they are within a few FP64 ulps.  Host code only."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

HARNESS = r"""
#include "fp64_sincos.hpp"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
static int64_t ulps(double a, double b) {
    int64_t x, y; std::memcpy(&x, &a, 8); std::memcpy(&y, &b, 8); return x > y ? x - y : y - x; }
int main(int argc, char** argv) {
    const long n = std::atol(argv[1]);
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> u(-M_PI, M_PI);
    const double fixed[] = {0.0, M_PI, -M_PI, M_PI / 2, -M_PI / 2, M_PI / 4, -M_PI / 4, 3 * M_PI / 4,
                            -3 * M_PI / 4, 1e-300, -1e-300, 1e-8, std::nextafter(M_PI / 4, 0.0), 3.5, -3.9};
    long mism = 0; int64_t worst = 0;
    for (long i = 0; i < n; i++) {
        const double x = i < 15 ? fixed[i] : u(g);
        double s, c;
        argus_sincos_wrapped(x, &s, &c);
        const double S = std::sin(x), C = std::cos(x);
        mism += ((float)s != (float)S) + ((float)c != (float)C);
        if (std::fabs(S) > 1e-6 && ulps(s, S) > worst) worst = ulps(s, S);
        if (std::fabs(C) > 1e-6 && ulps(c, C) > worst) worst = ulps(c, C);
    }
    std::printf("%ld %lld\n", mism, (long long)worst);
    return 0;
}
"""


def test_sincos_rounds_to_libm_float(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "h"
    r = subprocess.run(["g++", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas",
                        f"-I{ROOT / 'argus-sdr_amd' / 'csrc'}", str(src), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe), "5000000"], capture_output=True, text=True, check=True).stdout.split()
    mism, worst = int(out[0]), int(out[1])
    assert mism == 0, mism
    assert worst <= 4, worst

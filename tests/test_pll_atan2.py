"""The FP32 atan2 of the stereo PLL's phase detector
(argus-sdr_amd/csrc/pll_atan2.hpp) against the C library's FP64 atan2 on the
host, with an exact reciprocal in place of the hardware one (1 ulp): the
error bound over every quadrant and magnitude, the signed zeros and axes,
and the octant boundary.  Host code only."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

HARNESS = r"""
#include "pll_atan2.hpp"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
static float at(float y, float x) { return argus_pll_atan2f(y, x, [](float v) { return 1.0f / v; }); }
int main(int argc, char** argv) {
    const long n = std::atol(argv[1]);
    std::mt19937_64 g(7);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    std::uniform_real_distribution<float> ex(-20.0f, 5.0f);
    double worst = 0.0;
    for (long i = 0; i < n; i++) {
        const float y = nd(g) * std::pow(10.0f, ex(g)), x = nd(g) * std::pow(10.0f, ex(g));
        const double e = std::fabs((double)at(y, x) - std::atan2((double)y, (double)x));
        if (e > worst) worst = e;
    }
    // near the octant boundary |y| = |x| and on the axes
    for (int k = -1000; k <= 1000; k++) {
        const float y = 1.0f + k * 1e-6f;
        for (float sx : {1.0f, -1.0f})
            for (float sy : {1.0f, -1.0f}) {
                const double e = std::fabs((double)at(sy * y, sx) - std::atan2((double)(sy * y), (double)sx));
                if (e > worst) worst = e;
            }
    }
    int special = 0;
    const float z[][2] = {{0.0f, 0.0f}, {-0.0f, 0.0f}, {0.0f, -0.0f}, {-0.0f, -0.0f}, {1.0f, 0.0f}, {-1.0f, 0.0f},
                          {0.0f, 1.0f}, {0.0f, -1.0f}, {-0.0f, -1.0f}, {3.0f, -3.0f}, {-3.0f, 3.0f}};
    for (auto& v : z) {
        const float r = at(v[0], v[1]), w = (float)std::atan2((double)v[0], (double)v[1]);
        special += (std::fabs(r - w) > 4e-7f) || (std::signbit(r) != std::signbit(w));
    }
    std::printf("%.9e %d\n", worst, special);
    return 0;
}
"""


def test_pll_atan2_error_and_special_cases(tmp_path):
    src = tmp_path / "a.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "a"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                        f"-I{ROOT / 'argus-sdr_amd' / 'csrc'}", str(src), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    out = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, check=True).stdout.split()
    worst, special = float(out[0]), int(out[1])
    assert worst <= 3.5e-7, worst
    assert special == 0, special

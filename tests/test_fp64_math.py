"""The binary64 sin/cos and log of the FP64 particle update and noise
(ss_device.h: sincos_angle_f64, log_unit_f64), compiled for the host with
hipcc from the same header the kernels include and checked against glibc's
long-double sinl, cosl and logl (64-bit significands, so the reference adds
under 0.001 ulp of its own) over the ranges the path uses: particle and axis
angles, half rotation angles, 2 pi u of the noise, and the noise uniforms
down to 2^-53 (measured: sin and cos 0.73 ulp, log 0.95 ulp).  Each of
these mistakes fails it: the highest Taylor term of sin, cos or log dropped
(1.03 to 9.5 ulp), the third part of pi/2 or the first-order term of the
reduction's low part left out (39 ulp), the low part of ln 2 left out (991
ulp).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "stochasticsampling_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HARNESS = r"""
#include "ss_device.h"
#include <cmath>
#include <cstdio>
#include <random>

static double ulps(double got, long double ref) {
    // the spacing of doubles above |ref| (at a power of two, the larger one)
    const double r = std::fabs((double)ref);
    const long double u = (long double)std::nextafter(r, INFINITY) - (long double)r;
    return (double)(std::fabs((long double)got - ref) / u);
}

int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> ang(-7.0, 14.0), small(-1e-3, 1e-3), un(0.0, 1.0);
    double es = 0, ec = 0, el = 0;
    for (int i = 0; i < 2000000; ++i) {
        const double x = (i % 3 == 0) ? small(g) : ang(g);
        double s, c;
        ss::sincos_angle_f64(x, &s, &c);
        es = std::fmax(es, ulps(s, sinl((long double)x)));
        ec = std::fmax(ec, ulps(c, cosl((long double)x)));
        double u = un(g);
        if (i % 4 == 0) u = std::ldexp(u, -(i % 53));
        if (i % 7 == 0) u = 1.0 - std::ldexp(un(g), -40);
        if (u > 0) el = std::fmax(el, ulps(ss::log_unit_f64(u), logl((long double)u)));
    }
    // exact cases: quadrant boundaries, signed zero, log(1), the smallest uniform
    double s0, c0, s1, c1, s2, c2;
    ss::sincos_angle_f64(-0.0, &s0, &c0);
    ss::sincos_angle_f64(3.141592653589793, &s1, &c1);
    ss::sincos_angle_f64(-1.5707963267948966, &s2, &c2);
    std::printf("%.6f %.6f %.6f %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", es, ec, el,
                (int)std::signbit(s0), c0, s1, c1, s2, c2, ss::log_unit_f64(1.0),
                ss::log_unit_f64(std::ldexp(1.0, -53)));
    return 0;
}
"""


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    d = tmp_path_factory.mktemp("fp64math")
    src = d / "harness.hip"
    exe = d / "harness"
    src.write_text(HARNESS)
    subprocess.run([HIPCC, "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=fast", f"-I{CSRC}",
                    f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [float(v) for v in out]


def test_sincos_within_one_ulp(results):
    es, ec = results[0], results[1]
    assert es < 1.0, es
    assert ec < 1.0, ec


def test_log_within_one_ulp(results):
    assert results[2] < 1.0, results[2]


def test_special_values(results):
    _, _, _, sign_s0, c0, s_pi, c_pi, s_mhalf, c_mhalf, log1, log_min = results
    assert sign_s0 == 1 and c0 == 1.0               # sin(-0) = -0, cos 0 = 1
    assert abs(s_pi - 1.2246467991473532e-16) < 1e-31 and c_pi == -1.0
    assert s_mhalf == -1.0 and abs(c_mhalf - 6.123233995736766e-17) < 1e-32
    assert log1 == 0.0
    assert log_min == -53 * 0.6931471805599453094172321

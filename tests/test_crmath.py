"""The correctly rounded float functions of the cut algebra (sdl_crmath.h),
built for the host, against float(f(double(v))) of the host libm (glibc).

Every function must agree with the libm path bit for bit.  The CPU suite scans
a stride of all float bit patterns for the one-argument functions, every float
of the ranges the cut algebra uses densely, and seeded pairs for the two
argument ones; the full 2^32 scans run on the device (test_gpu_crmath.py).
The special cases are what the mathematics fixes (odd symmetry, exact zeros,
the pi/2 and pi limits of atan and atan2)."""
import math
import struct

import numpy as np
import pytest

from sdl_amd import build, crmath


@pytest.fixture(scope="module", autouse=True)
def _built():
    import os
    if not os.path.exists(crmath.LIB_PATH):
        build.build_check()


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


@pytest.mark.parametrize("name", ["asin", "sin", "cos", "atan"])
def test_one_arg_stride_of_all_floats(name):
    r = crmath.host_scan1(name, stride=127)
    assert r["checked"] > 3.3e7
    assert r["mismatches"] == 0, r


@pytest.mark.parametrize("name,lo,hi", [
    ("asin", 0.0, 1.0),      # |v| <= 0.95 in the cut algebra, both reduction branches
    ("sin", 0.0, 2.36),      # midpoint deltaPhi, beta angles, the strip-shift angles
    ("cos", 0.0, 2.36),
    ("atan", 0.0, 64.0),     # r/z and slopes of the strip shift
])
def test_one_arg_every_float_of_the_used_range(name, lo, hi):
    """Every float in [lo, hi) and its negation (stride 1 over the bit patterns
    above 2^-20, where the arguments of the cut algebra lie)."""
    a, b = max(_bits(2.0 ** -20), _bits(lo)), _bits(hi)
    for start in (a, a | 0x80000000):
        r = crmath.host_scan1(name, lo=start, hi=start + (b - a), stride=1)
        assert r["mismatches"] == 0, r
        # the rounding test falls back rarely (less than 1 call in 10^4)
        assert r["fallbacks"] < 1e-4 * r["checked"], r


@pytest.mark.parametrize("name", ["atan2", "divsin"])
@pytest.mark.parametrize("mode", [0, 1])
def test_two_arg_seeded_pairs(name, mode):
    r = crmath.host_scan2(name, mode, seed=3, n=1 << 24)
    assert r["mismatches"] == 0, r
    if mode == 1:
        assert r["fallbacks"] < 1e-4 * r["checked"], r


def test_special_values():
    f32 = np.float32
    for name in ("asin", "sin", "atan"):
        assert crmath.evaluate(name, 0.0)[1] == 0.0
        assert math.copysign(1.0, crmath.evaluate(name, -0.0)[1]) == -1.0  # odd functions keep -0
        for v in (1e-30, 0.25, 0.7, 0.95):
            assert crmath.evaluate(name, -v)[1] == -crmath.evaluate(name, v)[1]
    assert crmath.evaluate("cos", 0.0)[1] == 1.0
    assert crmath.evaluate("asin", 1.0)[1] == f32(math.pi / 2)
    assert math.isnan(crmath.evaluate("asin", 1.5)[1])
    assert crmath.evaluate("atan", float("inf"))[1] == f32(math.pi / 2)
    assert crmath.evaluate("atan", 1.0)[1] == f32(math.pi / 4)
    # atan2 quadrants and the signed zero of y
    assert crmath.evaluate("atan2", 0.0, -1.0)[1] == f32(math.pi)
    assert crmath.evaluate("atan2", -0.0, -1.0)[1] == -f32(math.pi)
    assert crmath.evaluate("atan2", 1.0, 1.0)[1] == f32(math.pi / 4)
    assert crmath.evaluate("atan2", -1.0, -1.0)[1] == f32(-3 * math.pi / 4)
    assert crmath.evaluate("atan2", 1.0, float("-inf"))[1] == f32(math.pi)
    # a tiny sine: sin(b) = b to double precision, so a / sin(b) = float(a / b)
    b = float(f32(1e-20))
    assert crmath.evaluate("divsin", 1.0, b)[1] == f32(1.0 / b)
    # outside the fast domain the libm decides
    ref, got, fast = crmath.evaluate("sin", 100.0)
    assert not fast and got == ref

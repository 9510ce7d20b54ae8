"""The correctly rounded float functions of the cut algebra on the device.

The kernels evaluate asin, sin, cos, atan, atan2 and a / sin(b) through
sdl_crmath.h; each result must equal float(f(double(v))) of the device libm,
which is what the chain computed before (and what the oracle's CR policy
restates).  The one-argument functions are checked on all 2^32 float bit
patterns, the two-argument ones on 2^30 seeded pairs of each kind."""
import pytest

from sdl_amd import crmath

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["asin", "sin", "cos", "atan"])
def test_every_float(name):
    print(f"case device scan {name} over 2^32 patterns", flush=True)
    r = crmath.device_scan1(name)
    assert r["checked"] == 1 << 32
    assert r["mismatches"] == 0, r


@pytest.mark.parametrize("name", ["atan2", "divsin"])
@pytest.mark.parametrize("mode", [0, 1])
def test_seeded_pairs(name, mode):
    print(f"case device scan {name} mode {mode}", flush=True)
    r = crmath.device_scan2(name, mode, seed=5, n=1 << 30)
    assert r["mismatches"] == 0, r
    if mode == 1:
        assert r["fallbacks"] < 1e-4 * r["checked"], r

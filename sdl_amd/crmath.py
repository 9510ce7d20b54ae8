"""ctypes binding of the checks of sdl_crmath.h (libsdl_crmath_check.so).

The cut algebra's float transcendentals (asin, sin, cos, atan, atan2 and
a / sin(b)) are evaluated by short double-precision polynomials with a
rounding test that falls back to the full libm call, and float(sqrt(double(v)))
by the float sqrt; the checks compare every result with float(f(double(v)))
of the libm, on the host or on the device.  Test
support only: the product library compiles the same header into its kernels."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsdl_crmath_check.so")

ONE_ARG = {"asin": 0, "sin": 1, "cos": 2, "atan": 3, "sqrt": 6}
TWO_ARG = {"atan2": 4, "divsin": 5}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with sdl_amd.build.build_check()")
        L = ctypes.CDLL(LIB_PATH)
        U, i, f = ctypes.c_ulonglong, ctypes.c_int, ctypes.c_float
        L.crm_host_scan1.argtypes = [i, U, U, U, i, ctypes.POINTER(U)]
        L.crm_host_scan2.argtypes = [i, i, U, U, i, ctypes.POINTER(U)]
        L.crm_device_scan1.argtypes = [i, U, U, ctypes.POINTER(U)]
        L.crm_device_scan2.argtypes = [i, i, U, U, ctypes.POINTER(U)]
        L.crm_pair.argtypes = [i, i, U, U, ctypes.POINTER(f), ctypes.POINTER(f)]
        L.crm_host_eval.argtypes = [i, f, f, ctypes.POINTER(f)]
        _lib = L
    return _lib


def _result(rc, out, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")
    return {"mismatches": int(out[0]), "fallbacks": int(out[1]), "first": int(out[2]), "checked": int(out[3])}


def host_scan1(name, lo=0, hi=1 << 32, stride=1, threads=8):
    """One-argument function over the float bit patterns lo, lo+stride, ... < hi."""
    out = (ctypes.c_ulonglong * 4)()
    return _result(lib().crm_host_scan1(ONE_ARG[name], lo, hi, stride, threads, out), out, "host scan " + name)


def host_scan2(name, mode, seed, n, threads=8):
    """atan2(a, b) or a / sin(b) on n seeded pairs (mode 0: uniform bit
    patterns, 1: the cut algebra's ranges)."""
    out = (ctypes.c_ulonglong * 4)()
    return _result(lib().crm_host_scan2(TWO_ARG[name], mode, seed, n, threads, out), out, "host scan " + name)


def device_scan1(name, lo=0, hi=1 << 32):
    out = (ctypes.c_ulonglong * 4)()
    return _result(lib().crm_device_scan1(ONE_ARG[name], lo, hi, out), out, "device scan " + name)


def device_scan2(name, mode, seed, n):
    out = (ctypes.c_ulonglong * 4)()
    return _result(lib().crm_device_scan2(TWO_ARG[name], mode, seed, n, out), out, "device scan " + name)


def evaluate(name, a, b=0.0):
    """(libm path, function, fast path decided) of one call on the host."""
    out = (ctypes.c_float * 3)()
    fn = ONE_ARG.get(name, TWO_ARG.get(name))
    lib().crm_host_eval(fn, a, b, out)
    return out[0], out[1], bool(out[2])

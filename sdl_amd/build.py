"""Builds the native libraries in-tree.

* ``sdl_amd/_lib/libsdl_amd.so`` -- the HIP kernels and the C-ABI (gfx950).
* ``sdl_amd/_lib/libsdl_gen.so`` -- the synthetic detector / event generator (host C++).
* ``sdl_amd/_lib/libsdl_crmath_check.so`` -- checks of the correctly rounded
  float functions against the host and device libm (test support).
* ``sdl_amd/_lib/libsdl_shim.so`` and ``sdl_amd/_lib/sdl_shim_demo`` -- the C++
  ``SDL::Event`` shim over the C-ABI and a small driver that uses it.
* ``sdl_amd/_lib/sdl_host_asan`` -- the host-only entry points under
  AddressSanitizer and UndefinedBehaviorSanitizer (test support).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_lib")
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # bit-exact parity with the reference's float/double arithmetic: no FMA contraction
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no hoisting of constant materialisations out of loops: the double-precision
    # libm polynomials of the cut kernels otherwise pin ~80 VGPRs of constants
    # across their loops and halve their occupancy
    "-mllvm", "-disable-machine-licm",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_gen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsdl_gen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", os.path.join(CSRC, "gen", "sdl_gen.cc"), "-o", out])
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsdl_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "sdl_amd.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_shim():
    os.makedirs(LIB, exist_ok=True)
    inc = [f"-I{os.path.join(HERE, 'include')}", f"-I{os.path.join(ROOT, 'include')}"]
    so = os.path.join(LIB, "libsdl_shim.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc, os.path.join(CSRC, "shim", "sdl_shim.cc"),
          f"-L{LIB}", "-lsdl_amd", "-Wl,-rpath,$ORIGIN", "-o", so])
    demo = os.path.join(LIB, "sdl_shim_demo")
    _run(["g++", "-O2", "-std=c++17", *inc, f"-I{os.path.join(CSRC, 'gen')}", os.path.join(CSRC, "shim", "shim_demo.cc"),
          os.path.join(CSRC, "gen", "sdl_gen.cc"), f"-L{LIB}", "-lsdl_shim", "-lsdl_amd", "-Wl,-rpath,$ORIGIN", "-o", demo])
    return so


def build_calib():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "copy_kernel")
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "calib", "copy_kernel.hip"), "-o", out])
    return out


def build_check():
    """Test support: the host and device scans of sdl_crmath.h."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libsdl_crmath_check.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "check", "crmath_check.hip"), "-o", out])
    return out


# host-code sanitizers for the test-support driver (the device code is built as usual)
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
CLANG = os.environ.get("SDL_CLANG", "/opt/rocm/llvm/bin/clang++")


def build_host_asan():
    """Test support: sdl_host_asan, the host-only C-ABI entry points (geometry
    parsing and cache, detId decoding, map builder) in a process whose host
    code carries AddressSanitizer and UndefinedBehaviorSanitizer."""
    os.makedirs(LIB, exist_ok=True)
    tmp = os.path.join(LIB, "_asan_obj")
    os.makedirs(tmp, exist_ok=True)
    lib_o = os.path.join(tmp, "sdl_amd.o")
    _run([HIPCC, *[f for f in HIP_FLAGS if f not in ("-O3", "-fPIC")], "-O1", "-g", *HOST_SAN, "-c",
          os.path.join(CSRC, "sdl_amd.hip"), "-o", lib_o])
    host = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=all"]
    objs = [lib_o]
    for src, name in ((os.path.join(CSRC, "check", "host_asan.cc"), "host_asan.o"),
                      (os.path.join(CSRC, "gen", "sdl_gen.cc"), "sdl_gen.o")):
        o = os.path.join(tmp, name)
        _run([CLANG, *host, "-c", src, "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "sdl_host_asan")
    _run([CLANG, "-fsanitize=address,undefined", *objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
          "-o", out])
    return out


def build_all(with_shim=True):
    build_gen()
    build_hip()
    build_calib()
    build_check()
    build_host_asan()
    if with_shim and os.path.exists(os.path.join(CSRC, "shim", "sdl_shim.cc")):
        build_shim()


if __name__ == "__main__":
    build_all()

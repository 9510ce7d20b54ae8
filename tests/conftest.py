import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def _ensure_host_libs():
    from sdl_amd import build
    if not os.path.exists(os.path.join(ROOT, "sdl_amd", "_lib", "libsdl_gen.so")):
        build.build_gen()
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)


_ensure_host_libs()


@pytest.fixture(scope="session")
def geom_paths():
    from sdl_amd import gen
    return {k: gen.geometry(k)[1] for k in (1, 2)}


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "ref_sets.npz")))

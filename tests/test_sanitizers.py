"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the oracle and the generator over every stage sequence the oracle restates
(oracle/asan_main.cc, built by `make -C oracle asan`), and the library's
host-only entry points -- geometry text parsing and the binary cache with
damaged files, detId decoding, the connection-map builder -- in a process whose
host code carries both sanitizers (sdl_amd/csrc/check/host_asan.cc, built by
sdl_amd/build.py).  A sanitizer report makes the process exit non-zero."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_ASAN = os.path.join(ROOT, "oracle", "_build", "oracle_asan")
HOST_ASAN = os.path.join(ROOT, "sdl_amd", "_lib", "sdl_host_asan")


@pytest.mark.skipif(not os.path.exists(ORACLE_ASAN), reason="oracle/_build/oracle_asan not built")
def test_oracle_under_sanitizers(tmp_path):
    r = subprocess.run([ORACLE_ASAN, str(tmp_path)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    last = r.stderr.strip().splitlines()[-1]
    assert last.startswith("ok ") and int(last.split()[1]) > 1_000_000, last


@pytest.mark.skipif(not os.path.exists(HOST_ASAN), reason="sdl_amd/_lib/sdl_host_asan not built")
def test_host_entry_points_under_sanitizers(tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")  # the HIP runtime's own start-up allocations
    r = subprocess.run([HOST_ASAN, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stderr.strip().splitlines()[-1] == "ok"

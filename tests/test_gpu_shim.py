"""The C++ SDL::Event shim (sdl_amd/include/SDL/Event.h) driven the way a user
of the reference drives SDL::Event (Event.h:110-300): hits added one by one
with addHitToModule, the six stages called in order, counters and objects read
back, with the MD, SG and TL objects' per-object values.  Its results must
equal the oracle's (CR libm policy, as for the C-ABI)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from sdl_amd import build, gen

pytestmark = pytest.mark.gpu

DEMO = os.path.join(build.LIB, "sdl_shim_demo")


def _checksum(rows):
    s = 0
    for r in np.asarray(rows, dtype=np.int64).reshape(len(rows), -1) if len(rows) else []:
        h = 1469598103934665603
        for x in r:
            h = ((h ^ (int(x) & 0xFFFFFFFF)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        s = (s + h) & 0xFFFFFFFFFFFFFFFF
    return s


def _checksum_with_values(rows, values):
    s = 0
    v = np.array(values, np.float32, copy=True)
    v[np.isnan(v)] = np.float32("nan")  # the demo hashes every NaN as the canonical one
    bits = v.view(np.uint32)
    for r, b in zip(np.asarray(rows, dtype=np.int64).reshape(len(rows), -1), bits):
        h = 1469598103934665603
        for x in list(r & 0xFFFFFFFF) + list(b):
            h = ((h ^ int(x)) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
        s = (s + h) & 0xFFFFFFFFFFFFFFFF
    return s


def _check(out, wl, e, stages):
    assert out.returncode == 0, out.stderr
    lines = dict(l.split(" ", 1) for l in out.stdout.strip().splitlines())
    counts = [int(v) for v in lines["counts"].split()]
    sums = [int(v) for v in lines["checksums"].split()]
    ref, cnt, vals = oracle.run(gen.geometry(gen.geom_kind(wl))[1], gen.event(wl, e), oracle.CR, stages, vars=True)
    assert counts == list(cnt.values())
    assert sums == [_checksum(ref[s]) for s in oracle.STAGES]
    assert [int(v) for v in lines["valuesums"].split()] == [_checksum_with_values(ref[k], vals[k])
                                                             for k in ("MD", "SG", "TL", "TCL")]
    assert lines["valuenames"].split(",") == [f"{t}.{n}" for t, k in (("MD", "MD"), ("SG", "SG"), ("TL", "TL"),
                                                                      ("TC", "TCL")) for n in oracle.VAR_NAMES[k]]
    return cnt


@pytest.mark.parametrize("wl,e", [(12, 0), (15, 1)])
def test_shim_matches_oracle(wl, e, tmp_path):
    print(f"case shim workload={wl} event={e}", flush=True)
    assert os.path.exists(DEMO), "shim demo not built"
    out = subprocess.run([DEMO, str(wl), str(e), str(tmp_path)], capture_output=True, text=True, timeout=120)
    _check(out, wl, e, 6)


def test_shim_layer_chain_matches_oracle(tmp_path):
    """createSegments, createTracklets and createTrackCandidates through the
    shim on a barrel-only event: counters and object sets equal the oracle's."""
    wl, e = 1, 2
    print(f"case shim layer chain workload={wl} event={e}", flush=True)
    out = subprocess.run([DEMO, str(wl), str(e), str(tmp_path), "layers"], capture_output=True, text=True,
                         timeout=120)
    cnt = _check(out, wl, e, oracle.LAYER_CHAIN)
    assert cnt["nTC"] > 0


def test_shim_pseudo_chain_matches_oracle(tmp_path):
    """createPseudoMiniDoubletsFromAnchorModule, then the module-map chain, through the shim."""
    wl, e = 12, 1
    print(f"case shim pseudo chain workload={wl} event={e}", flush=True)
    out = subprocess.run([DEMO, str(wl), str(e), str(tmp_path), "pseudo"], capture_output=True, text=True,
                         timeout=120)
    _check(out, wl, e, [10, 1, 2, 3, 4, 5])

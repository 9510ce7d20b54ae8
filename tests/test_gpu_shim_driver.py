"""The reference-style driver (oracle/ref_driver.cc, the harness that produced
the golden fixtures from the compiled reference) compiled unchanged against
this library's SDL headers and linked to the shim (oracle/_build/sdl_shim_driver).

It walks the objects the way a user of the reference does:
getModulePtrs() -> Module::getMiniDoubletPtrs() / getSegmentPtrs() /
getTripletPtrs() / getTrackletPtrs() -> MiniDoublet::lowerHitPtr()->idx(),
Segment::innerMiniDoubletPtr(), TrackletBase::innerSegmentPtr(), and
getLayerPtrs() -> Layer::getTrackletPtrs() / getTrackCandidatePtrs() with
dynamic_cast<SDL::Triplet*> on a TC's inner object; plus every counter getter
and, in dumpvars mode, every typed getter and reco variable of the MD, SG, TL
and TC objects.  Its dumps must equal the golden fixtures the same source
wrote when compiled against the reference (ref_*.npz), and the per-object
values the oracle's (CR policy, pinned to the compiled reference in
test_oracle.py::test_object_vars_match_reference)."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

import oracle
from sdl_amd import gen

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "oracle", "_build", "sdl_shim_driver")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
_mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mg)
COUNTERS = _mg.COUNTERS
# createTrackCandidatesFromTriplets after AllComb segments or triplets throws in
# the shim (covered through the C-ABI in test_gpu_allcomb.py)
TC_REFUSED = ("sg_allcomb", "tp_allcomb", "all_allcomb")


def _golden(name):
    with np.load(os.path.join(HERE, "golden", name)) as f:
        return {k: f[k] for k in f.files}


def _drive(tmp_path, mode, wl, e0, ne, *extra):
    assert os.path.exists(DRIVER), "oracle/_build/sdl_shim_driver not built (make -C oracle shimdriver)"
    out = os.path.join(tmp_path, f"{mode}_{wl}_{e0}.txt")
    print(f"case shim driver {mode} workload={wl} events={e0}..{e0 + ne - 1} {' '.join(extra)}", flush=True)
    r = subprocess.run([DRIVER, mode, str(wl), str(tmp_path), str(e0), str(ne), out, *extra], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return out


def _compare(parsed, golden, prefix, stages, layers):
    for e, (sets, counters) in parsed.items():
        key = prefix.format(e=e)
        for s in stages:
            assert np.array_equal(oracle.as_multiset(sets[s]), golden[f"{key}_{s}"]), f"{key} {s} differs"
        assert [counters[k] for k in COUNTERS] == list(golden[f"{key}_counters"]), key
        if layers:
            assert np.array_equal(counters["layers"], golden[f"{key}_layers"]), f"{key} per-layer counters differ"


@pytest.mark.parametrize("wl,e0,ne", _mg.CASES)
def test_module_map_chain(wl, e0, ne, tmp_path):
    out = _drive(tmp_path, "dump", wl, e0, ne)
    _compare(oracle.parse_ref_dump(out), _golden("ref_sets.npz"), f"w{wl}_e{{e}}", oracle.STAGES, False)


@pytest.mark.parametrize("wl,e0,ne", _mg.LAYER_CASES)
def test_layer_pair_chain(wl, e0, ne, tmp_path):
    out = _drive(tmp_path, "dumplayers", wl, e0, ne)
    _compare(oracle.parse_ref_dump(out), _golden("ref_layer_sets.npz"), f"w{wl}_e{{e}}", ("MD", "SG", "TL", "TCL"),
             True)


@pytest.mark.parametrize("wl,e0,ne", _mg.NAV_CASES)
def test_navigation_chain(wl, e0, ne, tmp_path):
    out = _drive(tmp_path, "dumpnav", wl, e0, ne)
    _compare(oracle.parse_ref_dump(out), _golden("ref_nav_sets.npz"), f"w{wl}_e{{e}}", ("MD", "SG", "TL"), False)


@pytest.mark.parametrize("wl,e0,ne", _mg.PSEUDO_CASES)
def test_pseudo_chain(wl, e0, ne, tmp_path):
    out = _drive(tmp_path, "dumppseudo", wl, e0, ne)
    _compare(oracle.parse_ref_dump(out), _golden("ref_pseudo_sets.npz"), f"w{wl}_e{{e}}", oracle.STAGES, True)


@pytest.mark.parametrize("wl,e0,ne", _mg.NM1_CASES)
def test_nm1_chain(wl, e0, ne, tmp_path):
    out = _drive(tmp_path, "dumpnm1", wl, e0, ne)
    _compare(oracle.parse_ref_dump(out), _golden("ref_nm1_sets.npz"), f"w{wl}_e{{e}}", oracle.STAGES, False)


@pytest.mark.parametrize("name", sorted(n for n in _mg.SEQS if n not in TC_REFUSED))
def test_stage_sequence(name, tmp_path):
    wl, e, spec, _ = _mg.SEQS[name]
    out = _drive(tmp_path, "dumpseq", wl, e, 1, spec)
    (parsed,) = oracle.parse_ref_dump(out).values()
    golden = _golden("ref_seq_sets.npz")
    _compare({e: parsed}, golden, name, oracle.STAGES, True)


def _rows(keys, values):
    v = np.array(values, np.float32, copy=True)
    v[np.isnan(v)] = np.float32("nan")  # every NaN as the canonical quiet NaN
    return _mg.sorted_rows(keys, v.view(np.uint32))


@pytest.mark.parametrize("wl,e0,ne", _mg.VAR_CASES)
def test_object_values(wl, e0, ne, tmp_path):
    """Every typed getter and reco variable of the MD, SG, TL and TC objects the
    driver reaches through the module and layer lists."""
    out = _drive(tmp_path, "dumpvars", wl, e0, ne)
    paths = gen.geometry(gen.geom_kind(wl))[1]
    for e, d in _mg.parse_var_dump(out).items():
        sets, _, values = oracle.run(paths, gen.event(wl, e), oracle.CR, 6, vars=True)
        for k, s in (("MV", "MD"), ("SV", "SG"), ("TV", "TL"), ("CV", "TCL")):
            keys, bits = d[k]
            got = _rows(keys, bits.view(np.float32))
            ref = _rows(sets[s], values[s])
            assert len(ref) > 0 or s == "TCL"
            assert got.shape == ref.shape, (wl, e, s, got.shape, ref.shape)
            bad = np.flatnonzero((got != ref).any(axis=1))
            assert len(bad) == 0, f"w{wl} e{e} {s}: {len(bad)} objects differ; first shim {got[bad[0]]} oracle {ref[bad[0]]}"

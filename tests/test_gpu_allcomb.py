"""Stage sequences with the AllComb algorithms (and DefaultNm1 tracklets in
every tracklet stage) on the device, through sdl_event_run's algo argument,
against the oracle (CR policy) and the compiled reference's sets and
per-layer counters (ref_seq_sets.npz, pinned to the oracle in
test_oracle.py::test_oracle_sequences_match_reference).

AllComb accepts every candidate (MiniDoublet.cc:223-226, Segment.cc:313-335,
Triplet.cc:45-48, Tracklet.cc:139-142, TrackCandidate.cc:138-141); the
objects keep their constructors' values.  TCs from triplets after AllComb
segments or triplets are refused (their bridges re-run the default algorithms
on objects the device does not store)."""
import importlib.util
import os

import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
_mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mg)
SEQS = _mg.SEQS
COUNTERS = runtime.COUNTER_NAMES
TC_REFUSED = ("sg_allcomb", "tp_allcomb", "all_allcomb")  # a TCP stage after AllComb segments or triplets


def _golden(name):
    with np.load(os.path.join(HERE, "golden", "ref_seq_sets.npz")) as f:
        return {k[len(name) + 1:]: f[k] for k in f.files if k.startswith(name + "_")}


def _geom(kind, cache={}):
    if kind not in cache:
        cache[kind] = runtime.Geometry(*gen.geometry(kind)[1])
    return cache[kind]


@pytest.mark.parametrize("name", sorted(SEQS))
def test_sequence_matches_oracle_and_reference(name):
    wl, e, _, codes = SEQS[name]
    print(f"case sequence {name} workload={wl} event={e} codes={codes}", flush=True)
    x = gen.event(wl, e)
    ev = runtime.Event(_geom(gen.geom_kind(wl)))
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    ran = []
    for code in codes:
        stage, algo = code % 100, code // 100
        if stage == runtime.STAGE_TCP and name in TC_REFUSED:
            with pytest.raises(runtime.SDLError):
                ev.run(stage, algo)
            continue
        ev.run(stage, algo)
        ran.append(code)
    ev.sync()
    paths = gen.geometry(gen.geom_kind(wl))[1]
    sets, counters = oracle.run(paths, x, oracle.CR, ran)
    golden = _golden(name)
    for i, s in enumerate(runtime.STAGES):
        try:
            got = oracle.as_multiset(ev.export(i))
        except runtime.SDLError:  # a stage that did not run (nor its layer-pair form)
            got = oracle.as_multiset(np.zeros((0, runtime.WIDTH[s]), np.int32))
        assert np.array_equal(got, oracle.as_multiset(sets[s])), f"{s} differs from the oracle"
        if not (s == "TCP" and name in TC_REFUSED):
            assert np.array_equal(got, golden[s]), f"{s} differs from the reference"
    assert ev.counts(0) == counters
    if name not in TC_REFUSED:
        assert [ev.counts(0)[k] for k in COUNTERS] == list(golden["counters"])
        assert np.array_equal(ev.layer_counts(0), golden["layers"])
    else:  # every counter except those of the TCs
        assert [ev.counts(0)[k] for k in COUNTERS[:4] + COUNTERS[5:9]] == \
            [int(v) for k, v in zip(COUNTERS, golden["counters"]) if k not in ("nTCCand", "nTC")]


def test_allcomb_object_values_are_the_constructors():
    """Per-object values of AllComb objects: typed fields 0, no reco variable
    (NaN); the oracle restates the same."""
    wl, e = 11, 0
    x = gen.event(wl, e)
    S = runtime
    ev = runtime.Event(_geom(gen.geom_kind(wl)))
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    for stage, algo in ((S.STAGE_MD, 1), (S.STAGE_SG, 1), (S.STAGE_TP, 1), (S.STAGE_TL, 2), (S.STAGE_TCL, 1)):
        ev.run(stage, algo)
    ev.sync()
    _, _, values = oracle.run(gen.geometry(1)[1], x, oracle.CR, [100, 101, 102, 203, 105], vars=True)
    for st, name in ((S.STAGE_MD, "MD"), (S.STAGE_SG, "SG"), (S.STAGE_TL, "TL"), (S.STAGE_TCL, "TCL")):
        got = ev.export_vars(st)
        assert got.shape == values[name].shape
        assert np.array_equal(np.isnan(got), np.isnan(values[name]))
        assert np.array_equal(np.nan_to_num(got), np.nan_to_num(values[name]))


def test_unknown_algorithm_values_are_rejected():
    x = gen.event(11, 0)
    ev = runtime.Event(_geom(1))
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.STAGE_MD, 2)
    ev.run(runtime.STAGE_MD, 0)
    ev.run(runtime.STAGE_SG, 0)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.STAGE_TP, 5)

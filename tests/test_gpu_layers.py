"""The layer-pair family on the GPU (createSegments, createTracklets,
createTrackCandidates; Event.cc:432-443, 599-609, 755-872) through the C-ABI,
against the oracle (CR libm policy) and the compiled reference's golden
per-layer counters.  Every object of an inner layer is paired with every
object of an outer layer, so the checks run on small events."""
import os

import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

LAYER_SEQ = [0, runtime.SG_LAYERS, runtime.TL_LAYERS, runtime.TC_LAYERS]


def _geom(kind, cache={}):
    if kind not in cache:
        cache[kind] = runtime.Geometry(*gen.geometry(kind)[1])
    return cache[kind]


def _run(wl, events, seq):
    ev = runtime.Event(_geom(gen.geom_kind(wl)))
    for i, e in enumerate(events):
        if i:
            ev.next_event()
        ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    for s in seq:
        ev.run(s)
    ev.sync()
    return ev


def _exported(ev, seq, e=0):
    have = {"MD"}
    for s in seq:
        have.add({1: "SG", 2: "TP", 3: "TL", 4: "TCP", 5: "TCL", 6: "SG", 7: "TL", 8: "TCL", 9: "TL"}.get(s, "MD"))
    return {s: ev.export(i, e) if s in have else np.zeros((0, oracle.WIDTH[s]), np.int32)
            for i, s in enumerate(oracle.STAGES)}


def _assert_same(got, ref, what):
    for s in oracle.STAGES:
        a, b = oracle.as_multiset(got[s]), oracle.as_multiset(ref[s])
        assert a.shape == b.shape, f"{what} {s}: gpu {a.shape} oracle {b.shape}"
        assert np.array_equal(a, b), f"{what} {s} differs"


@pytest.mark.parametrize("wl,e", [(1, 0), (1, 1), (1, 2), (12, 0), (15, 0)])
def test_layer_chain_matches_oracle(wl, e):
    print(f"case layer chain workload={wl} event={e}", flush=True)
    ev_in = gen.event(wl, e)
    ev = _run(wl, [ev_in], LAYER_SEQ)
    ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR, oracle.LAYER_CHAIN)
    _assert_same(_exported(ev, LAYER_SEQ), ref, f"layer chain w{wl} e{e}")
    assert ev.counts(0) == cnt
    assert cnt["nTL"] > 0


def test_layer_counters_match_reference_golden():
    """Per-layer counters against the compiled reference (ref_layer_sets.npz)."""
    print("case layer counters vs reference", flush=True)
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_layer_sets.npz")
    with np.load(path) as f:
        golden = {k: f[k] for k in f.files}
    for wl, e in ((1, 0), (1, 1), (12, 0)):
        ev = _run(wl, [gen.event(wl, e)], LAYER_SEQ)
        assert np.array_equal(ev.layer_counts(0), golden[f"w{wl}_e{e}_layers"]), (wl, e)
        got = ev.counts(0)
        assert [got[k] for k in runtime.COUNTER_NAMES] == list(golden[f"w{wl}_e{e}_counters"]), (wl, e)


@pytest.mark.parametrize("seq", [[0, 1, 7, 8], [0, 1, 2, 3, 8]])
def test_mixed_stage_orders(seq):
    """Module-map and layer-pair stages mixed as the reference allows: layer
    tracklets over module-map segments, layer track candidates over module-map
    tracklets."""
    print(f"case mixed stages {seq}", flush=True)
    wl, e = 12, 1
    ev_in = gen.event(wl, e)
    ev = _run(wl, [ev_in], seq)
    ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR, seq)
    _assert_same(_exported(ev, seq), ref, f"stages {seq}")
    assert ev.counts(0) == cnt


def test_layer_batch_equals_single_events():
    print("case layer chain batch of workload 1 events 0-1", flush=True)
    evs = [gen.event(1, 0), gen.event(1, 1)]
    batch = _run(1, evs, LAYER_SEQ)
    for i, e in enumerate(evs):
        single = _run(1, [e], LAYER_SEQ)
        _assert_same(_exported(batch, LAYER_SEQ, i), _exported(single, LAYER_SEQ), f"batch event {i}")
        assert batch.counts(i) == single.counts(0)
        assert np.array_equal(batch.layer_counts(i), single.layer_counts(0))


def test_layer_stage_order_rules():
    print("case layer stage order", flush=True)
    e = gen.event(1, 0)
    ev = runtime.Event(_geom(1))
    ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.TC_LAYERS)  # before any tracklets
    ev.run(0)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.TL_LAYERS)  # before any segments
    ev.run(runtime.SG_LAYERS)
    with pytest.raises(runtime.SDLError):
        ev.run(2)  # module-map triplets do not follow layer-pair segments
    ev.run(runtime.TL_LAYERS)
    with pytest.raises(runtime.SDLError):
        ev.run(4)  # TCs from triplets need module-map tracklets
    with pytest.raises(runtime.SDLError):
        ev.export(2)  # no triplets ran
    ev.run(runtime.TC_LAYERS)
    assert len(ev.export(5)) == ev.counts(0)["nTC"]


@pytest.mark.parametrize("seq", [[0, 1, 9], [0, 1, 2, 9, 5], [0, 6, 9, 8], [0, 1, 9, 8]])
def test_navigation_matches_oracle(seq):
    """createTrackletsViaNavigation over module-map or layer-pair segments, and
    the track candidates that may follow it, against the oracle."""
    print(f"case navigation stages {seq}", flush=True)
    for wl, e in ((12, 0), (1, 1)):
        ev_in = gen.event(wl, e)
        ev = _run(wl, [ev_in], seq)
        ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR, seq)
        have = {"MD", "SG", "TL"} | ({"TP"} if 2 in seq else set()) | ({"TCL"} if (5 in seq or 8 in seq) else set())
        got = {s: ev.export(i, 0) if s in have else np.zeros((0, oracle.WIDTH[s]), np.int32)
               for i, s in enumerate(oracle.STAGES)}
        _assert_same(got, ref, f"stages {seq} w{wl} e{e}")
        assert ev.counts(0) == cnt


def test_navigation_counters_match_reference_golden():
    print("case navigation counters vs reference", flush=True)
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_nav_sets.npz")
    with np.load(path) as f:
        golden = {k: f[k] for k in f.files}
    for wl, e in ((1, 0), (12, 1), (15, 0)):
        ev = _run(wl, [gen.event(wl, e)], [0, 1, runtime.TL_NAV])
        assert np.array_equal(ev.layer_counts(0), golden[f"w{wl}_e{e}_layers"]), (wl, e)


def test_navigation_order_rules():
    print("case navigation order", flush=True)
    e = gen.event(12, 0)
    ev = runtime.Event(_geom(2))
    ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    ev.run(0)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.TL_NAV)  # before any segments
    ev.run(runtime.SG_LAYERS)
    ev.run(runtime.TL_NAV)
    with pytest.raises(runtime.SDLError):
        ev.run(5)  # the TCL join needs module-map segments under navigation tracklets
    ev.run(runtime.TC_LAYERS)


@pytest.mark.parametrize("wl,e", [(1, 0), (1, 1), (12, 0), (15, 0)])
def test_pseudo_chain_matches_oracle_and_reference(wl, e):
    """createPseudoMiniDoubletsFromAnchorModule and the module-map chain on its
    MDs (zero cached values: the segment algorithms meet 0/0) against the
    oracle, and the per-layer counters against the compiled reference."""
    print(f"case pseudo chain workload={wl} event={e}", flush=True)
    seq = [runtime.MD_PSEUDO, 1, 2, 3, 4, 5]
    ev_in = gen.event(wl, e)
    ev = _run(wl, [ev_in], seq)
    ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR, seq)
    _assert_same(ev.export_all(0), ref, f"pseudo chain w{wl} e{e}")
    assert ev.counts(0) == cnt
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_pseudo_sets.npz")) as f:
        assert np.array_equal(ev.layer_counts(0), f[f"w{wl}_e{e}_layers"])


def test_pseudo_then_layer_pairs():
    print("case pseudo MDs then layer-pair chain", flush=True)
    seq = [runtime.MD_PSEUDO, runtime.SG_LAYERS, runtime.TL_LAYERS, runtime.TC_LAYERS]
    ev_in = gen.event(1, 2)
    ev = _run(1, [ev_in], seq)
    ref, cnt = oracle.run(gen.geometry(1)[1], ev_in, oracle.CR, seq)
    _assert_same(_exported(ev, seq), ref, "pseudo + layer pairs")
    assert ev.counts(0) == cnt


@pytest.mark.parametrize("seq", [[0, 1, runtime.TL_NAV, 5], [0, 1, 13], [runtime.MD_PSEUDO, 1, 9]])
def test_families_at_pu200_density(seq):
    """The navigation tracklets (default and DefaultNm1, the oracle's stage 13 =
    createTrackletsViaNavigation(DefaultNm1_TLAlgo)) and the pseudo mini-doublets
    on a PU200-density event (workload 2 event 0: 8000 tracks + 20 hard):
    object sets, counters and per-layer counters against the oracle."""
    print(f"case families at PU200 density, stages {seq}", flush=True)
    wl, e = 2, 0
    ev_in = gen.event(wl, e)
    ev = runtime.Event(_geom(2))
    ev.add_hits(ev_in.x, ev_in.y, ev_in.z, ev_in.idx, ev_in.detId)
    for s in seq:
        if s == 13:
            ev.run(runtime.TL_NAV, 1)
        else:
            ev.run(s)
    ev.sync()
    ref, cnt, lay = oracle.run(gen.geometry(2)[1], ev_in, oracle.CR, seq, layers=True)
    have = {"MD", "SG", "TL"} | ({"TCL"} if 5 in seq else set())
    got = {s: ev.export(i, 0) if s in have else np.zeros((0, oracle.WIDTH[s]), np.int32)
           for i, s in enumerate(oracle.STAGES)}
    _assert_same(got, ref, f"stages {seq} at PU200 density")
    assert ev.counts(0) == cnt
    assert np.array_equal(ev.layer_counts(0), lay)
    assert cnt["nTL"] > 100000


def test_layer_chain_at_pu200_matches_oracle_digest():
    """The layer-pair chain on a PU200-density event (workload 2 event 0; 2.5e9
    tracklet candidates, every segment of an inner layer against every segment of
    an outer layer) against the oracle's result precomputed on the CPU
    (oracle_pu200_digest.json: object counts, counters, per-layer counters,
    SHA-256 of each sorted object set; tests/golden/make_oracle_digests.py)."""
    import importlib.util
    import json
    print("case layer chain at PU200 density (workload 2 event 0)", flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(here, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    with open(os.path.join(here, "golden", "oracle_pu200_digest.json")) as f:
        ref = json.load(f)["layers_w2_e0"]
    ev = _run(2, [gen.event(2, 0)], LAYER_SEQ)
    got = _exported(ev, LAYER_SEQ)
    assert {s: len(got[s]) for s in oracle.STAGES} == ref["counts"]
    c = ev.counts(0)
    assert [c[k] for k in runtime.COUNTER_NAMES] == ref["counters"]
    assert ev.layer_counts(0).tolist() == ref["layers"]
    assert {s: mg.set_digest(got[s]) for s in oracle.STAGES} == ref["sha256"]

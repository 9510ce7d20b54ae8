"""DefaultNm1_TLAlgo on the device (sdl_event_run with algo 1 for the tracklet
stages; Tracklet.cc:127-131, the default cuts with dBeta waived) against the
oracle, whose restatement is pinned to the compiled reference in
test_oracle.py::test_oracle_nm1_tracklets_match_reference: object sets,
counters and the tracklets' per-object values."""
import os

import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

NM1 = 1


def _run(wl, e, seq):
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1])
    ev = runtime.Event(g)
    x = gen.event(wl, e)
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    for stage, algo in seq:
        ev.run(stage, algo)
    ev.sync()
    return ev, x


def _bits(v):
    v = np.array(v, np.float32, copy=True)
    v[np.isnan(v)] = np.float32("nan")
    return v.view(np.uint32).astype(np.int64)


@pytest.mark.parametrize("wl,e", [(1, 0), (15, 0), (2, 0)])
def test_nm1_module_map_chain_matches_oracle(wl, e):
    print(f"case nm1 module-map chain workload={wl} event={e}", flush=True)
    S = runtime
    # the reference's own sequence (ref_driver dumpnm1): TCs from triplets re-run
    # the default tracklet algorithm on their bridges, TCs from tracklets join
    # the stored (DefaultNm1) tracklets
    ev, x = _run(wl, e, [(S.STAGE_MD, 0), (S.STAGE_SG, 0), (S.STAGE_TP, 0), (S.STAGE_TL, NM1), (S.STAGE_TCP, 0),
                         (S.STAGE_TCL, 0)])
    paths = gen.geometry(gen.geom_kind(wl))[1]
    sets, counters, values = oracle.run(paths, x, oracle.CR, [0, 1, 2, 11, 4, 5], vars=True)
    got = {s: ev.export(i) for i, s in enumerate(runtime.STAGES)}
    for s in runtime.STAGES:
        assert np.array_equal(oracle.as_multiset(got[s]), oracle.as_multiset(sets[s])), s
    assert ev.counts(0) == counters
    if wl in (1, 15):  # the compiled reference's sets for this sequence
        with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_nm1_sets.npz")) as f:
            for s in runtime.STAGES:
                assert np.array_equal(oracle.as_multiset(got[s]), f[f"w{wl}_e{e}_{s}"]), s
    # TCP may also be skipped: TCs from tracklets directly after the tracklets
    ev2, _ = _run(wl, e, [(S.STAGE_MD, 0), (S.STAGE_SG, 0), (S.STAGE_TP, 0), (S.STAGE_TL, NM1), (S.STAGE_TCL, 0)])
    assert np.array_equal(oracle.as_multiset(ev2.export(S.STAGE_TCL)), oracle.as_multiset(sets["TCL"]))
    plain, _ = oracle.run(paths, x, oracle.CR, 4)
    assert len(sets["TL"]) > len(plain["TL"])  # the waived cut admits more tracklets
    if wl != 2:  # the values of every Nm1 tracklet, including those that fail dBeta
        keys, vals = ev.export(S.STAGE_TL), ev.export_vars(S.STAGE_TL)
        a = np.concatenate([keys.astype(np.int64), _bits(vals)], axis=1)
        b = np.concatenate([sets["TL"].astype(np.int64), _bits(values["TL"])], axis=1)
        a, b = a[np.lexsort(keys.T[::-1])], b[np.lexsort(sets["TL"].T[::-1])]
        assert np.array_equal(a, b)


def test_nm1_layer_and_navigation_tracklets_match_oracle():
    wl, e = 12, 0
    print(f"case nm1 layer-pair and navigation tracklets workload={wl} event={e}", flush=True)
    S = runtime
    paths = gen.geometry(gen.geom_kind(wl))[1]
    ev, x = _run(wl, e, [(S.STAGE_MD, 0), (S.SG_LAYERS, 0), (S.TL_LAYERS, NM1)])
    sets, _ = oracle.run(paths, x, oracle.CR, [0, 6, 12])
    assert len(sets["TL"]) > 0
    assert np.array_equal(oracle.as_multiset(ev.export(S.STAGE_TL)), oracle.as_multiset(sets["TL"]))
    ev, x = _run(wl, e, [(S.STAGE_MD, 0), (S.STAGE_SG, 0), (S.TL_NAV, NM1)])
    sets, _ = oracle.run(paths, x, oracle.CR, [0, 1, 13])
    assert len(sets["TL"]) > 0
    assert np.array_equal(oracle.as_multiset(ev.export(S.STAGE_TL)), oracle.as_multiset(sets["TL"]))


def test_chain_after_nm1_runs_default():
    """run_chain on a handle that ran DefaultNm1 uses the default cuts again."""
    wl, e = 15, 0
    S = runtime
    ev, x = _run(wl, e, [(S.STAGE_MD, 0), (S.STAGE_SG, 0), (S.STAGE_TP, 0), (S.STAGE_TL, NM1)])
    ev.reset()
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    ev.run_chain()
    ev.sync()
    sets, counters = oracle.run(gen.geometry(gen.geom_kind(wl))[1], x, oracle.CR)
    assert ev.counts(0) == counters

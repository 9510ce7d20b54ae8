"""Per-object values on the device (sdl_event_export_vars) against the CPU
oracle: for every stored MD, SG and TL, the reference's typed getters and reco
variables (the oracle's values are pinned to the compiled reference in
test_oracle.py::test_object_vars_match_reference).  The device evaluates the
float transcendentals correctly rounded, so it is compared bit for bit with the
oracle's CR policy, object by object (rows keyed by the hit-index tuple)."""
import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

STAGES = (("MD", runtime.STAGE_MD), ("SG", runtime.STAGE_SG), ("TL", runtime.STAGE_TL), ("TCL", runtime.STAGE_TCL))


def _bits(values):
    """Bit patterns, every NaN as the canonical quiet NaN: the device's and the
    host's libm may return NaNs of another sign (acos of a rounded cosine just
    above 1 in the tracklets' dBeta_midPoint)."""
    v = np.array(values, np.float32, copy=True)
    v[np.isnan(v)] = np.float32("nan")
    return v.view(np.uint32).astype(np.int64)


def _rows(keys, values):
    """Rows (tuple, value bits) ordered by the tuple (unique within a set)."""
    rows = np.concatenate([keys.astype(np.int64), _bits(values)], axis=1)
    return rows[np.lexsort(keys.T[::-1])] if len(rows) else rows


def _device(wl, e, seq=None):
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1])
    ev = runtime.Event(g)
    x = gen.event(wl, e)
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    if seq is None:
        ev.run_chain()
    else:
        for s in seq:
            ev.run(s)
    ev.sync()
    return ev, x


def _compare(ev, sets, values, what, stages=STAGES):
    for name, st in stages:
        got = _rows(ev.export(st), ev.export_vars(st))
        ref = _rows(sets[name], values[name])
        assert len(ref) > 0, f"{what} {name}: no objects"
        assert got.shape == ref.shape, (what, name, got.shape, ref.shape)
        bad = np.flatnonzero((got != ref).any(axis=1))
        assert len(bad) == 0, f"{what} {name}: {len(bad)} objects differ; first device {got[bad[0]]} oracle {ref[bad[0]]}"


def test_var_names_and_counts():
    assert runtime.var_names(runtime.STAGE_MD) == oracle.VAR_NAMES["MD"]
    assert runtime.var_names(runtime.STAGE_SG) == oracle.VAR_NAMES["SG"]
    assert runtime.var_names(runtime.STAGE_TL) == oracle.VAR_NAMES["TL"]
    assert runtime.var_names(runtime.STAGE_TCL) == oracle.VAR_NAMES["TCL"]
    assert runtime.var_names(runtime.STAGE_TP) == ()


@pytest.mark.parametrize("wl,e", [(1, 0), (12, 0), (15, 0), (15, 1)])
def test_object_vars_match_oracle(wl, e):
    print(f"case vars workload={wl} event={e}", flush=True)
    ev, x = _device(wl, e)
    sets, _, values = oracle.run(gen.geometry(gen.geom_kind(wl))[1], x, oracle.CR, vars=True)
    _compare(ev, sets, values, f"w{wl}e{e}")


def test_object_vars_pu200():
    print("case vars PU200 workload 2 event 0", flush=True)
    ev, x = _device(2, 0)
    sets, _, values = oracle.run(gen.geometry(2)[1], x, oracle.CR, vars=True)
    _compare(ev, sets, values, "pu200")


def test_object_vars_layer_chain_and_pseudo():
    """The layer-pair chain's SGs, TLs and TCs, and pseudo MDs (zero values, no
    miniCut) with the module-map chain on them."""
    print("case vars layer chain and pseudo MDs, workload 12 event 0", flush=True)
    wl, e = 12, 0
    ev, x = _device(wl, e, seq=[runtime.STAGE_MD, runtime.SG_LAYERS, runtime.TL_LAYERS, runtime.TC_LAYERS])
    sets, _, values = oracle.run(gen.geometry(gen.geom_kind(wl))[1], x, oracle.CR, stages=[0, 6, 7, 8], vars=True)
    _compare(ev, sets, values, "layers", STAGES if len(sets["TCL"]) else STAGES[:3])
    ev, x = _device(wl, e, seq=[runtime.MD_PSEUDO, runtime.STAGE_SG, runtime.STAGE_TP, runtime.STAGE_TL])
    sets, _, values = oracle.run(gen.geometry(gen.geom_kind(wl))[1], x, oracle.CR, stages=[10, 1, 2, 3], vars=True)
    _compare(ev, sets, values, "pseudo", STAGES[:3])
    md = ev.export_vars(runtime.STAGE_MD)
    assert (md[:, :6] == 0).all() and np.isnan(md[:, 6]).all()


def test_object_vars_of_batched_events():
    """Events 1 and 2 of a three-event batch (workload 12 events 0-2): the values
    of each event's objects equal those of the same event run alone, bit for bit
    and in the same object order (the per-event ranges of the export)."""
    print("case vars batch of 3 workload-12 events", flush=True)
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(12))[1])
    xs = [gen.event(12, e) for e in range(3)]
    batch = runtime.Event(g)
    for k, x in enumerate(xs):
        if k:
            batch.next_event()
        batch.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    batch.run_chain()
    batch.sync()
    for k in (1, 2):
        single, _ = _device(12, k)
        for name, st in STAGES:
            got, ref = batch.export_vars(st, k), single.export_vars(st)
            assert got.shape == ref.shape and len(ref) > 0, (k, name, got.shape, ref.shape)
            assert np.array_equal(_bits(got), _bits(ref)), f"batch event {k} {name} values differ"


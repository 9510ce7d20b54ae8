"""Parity of the HIP chain (through the C-ABI) with the CPU oracle.

The device evaluates the float transcendentals correctly rounded, so it is
compared bit for bit with the oracle's CR policy; test_oracle.py relates that
policy to the compiled reference.  Objects are compared as multisets of
hit-index tuples (MD 2, SG 4, TP/TL 8, TC 16 indices)."""
import os

import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

SMALL = [(1, 0), (1, 1), (1, 2), (12, 0), (12, 1), (15, 0), (15, 1)]


def _geom(kind, cache={}):
    if kind not in cache:
        cache[kind] = runtime.Geometry(*gen.geometry(kind)[1])
    return cache[kind]


def _run_gpu(wl, events):
    g = _geom(gen.geom_kind(wl))
    ev = runtime.Event(g)
    for i, e in enumerate(events):
        if i:
            ev.next_event()
        ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    ev.run_chain()
    ev.sync()
    return ev


def _assert_same(got_sets, ref_sets, what):
    for s in oracle.STAGES:
        a = oracle.as_multiset(got_sets[s])
        b = oracle.as_multiset(ref_sets[s])
        assert a.shape == b.shape, f"{what} {s}: gpu {a.shape} oracle {b.shape}"
        assert np.array_equal(a, b), f"{what} {s} differs"


@pytest.mark.parametrize("wl,e", SMALL)
def test_chain_matches_oracle(wl, e):
    print(f"case workload={wl} event={e}", flush=True)
    ev_in = gen.event(wl, e)
    ev = _run_gpu(wl, [ev_in])
    got = ev.export_all(0)
    ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR)
    _assert_same(got, ref, f"w{wl}e{e}")
    assert ev.counts(0) == cnt


@pytest.mark.parametrize("wl,e", [(1, 0), (12, 0), (15, 0)])
def test_chain_matches_reference_golden(wl, e, golden):
    """Where the two libm policies agree (checked on the CPU), the device must
    reproduce the compiled reference's sets directly."""
    print(f"case golden workload={wl} event={e}", flush=True)
    ev = _run_gpu(wl, [gen.event(wl, e)])
    got = ev.export_all(0)
    for s in oracle.STAGES:
        assert np.array_equal(oracle.as_multiset(got[s]), golden[f"w{wl}_e{e}_{s}"]), s


def test_batch_equals_single_events():
    print("case batch of 4 workload-12 events", flush=True)
    evs = [gen.event(12, e) for e in range(4)]
    batch = _run_gpu(12, evs)
    assert batch.num_events == 4
    for i, e in enumerate(evs):
        single = _run_gpu(12, [e])
        _assert_same(batch.export_all(i), single.export_all(0), f"batch event {i}")
        assert batch.counts(i) == single.counts(0)
        assert np.array_equal(batch.layer_counts(i), single.layer_counts(0))


def test_pu200_batch_equals_single_events():
    """Four PU200-density events in one batch: the scans over the batch's pair
    and candidate lists (tens of millions of items) take the multi-level path,
    the single events the one-level path; every object list and counter agrees."""
    print("case batch of 4 workload-2 events", flush=True)
    evs = [gen.event(2, e) for e in range(4)]
    batch = _run_gpu(2, evs)
    for i, e in enumerate(evs):
        single = _run_gpu(2, [e])
        got, ref = batch.export_all(i), single.export_all(0)
        for s in oracle.STAGES:
            assert np.array_equal(got[s], ref[s]), f"batch event {i} {s} differs"
        assert batch.counts(i) == single.counts(0)
        assert np.array_equal(batch.layer_counts(i), single.layer_counts(0))


def test_device_resident_input_and_rerun():
    from sdl_amd import hip
    print("case device-resident input, workload 15 event 1", flush=True)
    e = gen.event(15, 1)
    dev = [hip.DeviceArray(a) for a in (e.x, e.y, e.z, e.idx, e.detId)]
    hip.synchronize()
    g = _geom(2)
    ev = runtime.Event(g)
    ev.set_hits_device([0, e.nhits], *[d.address for d in dev])
    ref = _run_gpu(15, [e]).export_all(0)
    for rep in range(2):  # the handle reuses its device buffers
        ev.run_chain()
        ev.sync()
        _assert_same(ev.export_all(0), ref, f"device input rep {rep}")


def test_stage_order_and_errors():
    print("case error paths", flush=True)
    g = _geom(1)
    ev = runtime.Event(g)
    e = gen.event(1, 0)
    ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    with pytest.raises(runtime.SDLError):
        ev.run(runtime.STAGE_TL)  # stages must run in order
    ev.run(runtime.STAGE_MD)
    ev.run(runtime.STAGE_SG)
    with pytest.raises(runtime.SDLError):
        ev.export(runtime.STAGE_TL)  # not run yet
    bad = runtime.Event(g)
    bad.add_hits(e.x[:5], e.y[:5], e.z[:5], e.idx[:5], np.full(5, 7, np.uint32))
    with pytest.raises(runtime.SDLError, match="unknown"):
        bad.run_chain()


def test_empty_event():
    print("case empty event", flush=True)
    g = _geom(1)
    ev = runtime.Event(g)
    e = gen.event(1, 0)
    ev.add_hits(e.x[:0], e.y[:0], e.z[:0], e.idx[:0], e.detId[:0])
    ev.run_chain()
    ev.sync()
    assert all(len(v) == 0 for v in ev.export_all(0).values())
    assert all(v == 0 for v in ev.counts(0).values())


def test_duplicate_connection_rejected(tmp_path):
    print("case duplicate connection", flush=True)
    t, e, m = gen.geometry(1)[1]
    lines = open(m).read().splitlines()
    first = lines[0].split()
    first = [first[0], str(int(first[1]) + 1)] + first[2:] + [first[2]]
    bad = tmp_path / "map.txt"
    bad.write_text("\n".join([" ".join(first)] + lines[1:]) + "\n")
    with pytest.raises(runtime.SDLError, match="repeats"):
        runtime.Geometry(t, e, str(bad))


def test_pu200_event_full_parity():
    """One full PU200-density event (workload 2, ~1.1e5 hits): every object set
    and counter equals the oracle's."""
    print("case full PU200 event, workload 2 event 0", flush=True)
    e = gen.event(2, 0)
    ev = _run_gpu(2, [e])
    got = ev.export_all(0)
    ref, cnt = oracle.run(gen.geometry(2)[1], e, oracle.CR)
    _assert_same(got, ref, "pu200")
    assert ev.counts(0) == cnt


@pytest.mark.parametrize("stage", range(6))
def test_resume_after_short_prediction(stage):
    """Buffers predicted too small from the stage on: the kernels detect it, the
    chain resumes from that stage sized exactly, and the results are unchanged."""
    print(f"case resume from stage {stage}", flush=True)
    wl, e = 12, 1
    ev_in = gen.event(wl, e)
    ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR)
    g = _geom(gen.geom_kind(wl))
    ev = runtime.Event(g)
    for k in range(3):
        if k == 1:
            ev.set_headroom(stage, 0.5)
        ev.reset()
        ev.add_hits(ev_in.x, ev_in.y, ev_in.z, ev_in.idx, ev_in.detId)
        ev.run_chain()
        ev.sync()
        _assert_same(ev.export_all(0), ref, f"resume{stage} pass{k}")
        assert ev.counts(0) == cnt


def test_speculative_sizing_across_events():
    """One handle runs a small event, a PU200-size batch and a small event again:
    every call matches the oracle whichever sizing path it took."""
    print("case speculative sizing across events", flush=True)
    g = _geom(2)
    ev = runtime.Event(g)
    for wl, e in ((12, 0), (15, 0), (12, 1), (15, 1), (12, 0)):
        ev_in = gen.event(wl, e)
        ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR)
        ev.reset()
        ev.add_hits(ev_in.x, ev_in.y, ev_in.z, ev_in.idx, ev_in.detId)
        ev.run_chain()
        ev.sync()
        _assert_same(ev.export_all(0), ref, f"w{wl}e{e}")
        assert ev.counts(0) == cnt


def test_graph_replays_match_eager():
    """The speculative chain is captured into a HIP graph and replayed: every
    replay (same event several times, then events with fewer and more hits that
    fit the captured buffers) equals the uncaptured run and the oracle."""
    print("case graph replays", flush=True)
    g = _geom(2)
    eager = runtime.Event(g)
    eager.set_graphs(False)
    ev = runtime.Event(g)
    for k, (wl, e) in enumerate([(15, 0), (15, 0), (15, 0), (12, 1), (15, 0), (12, 0), (15, 1)]):
        ev_in = gen.event(wl, e)
        for h in (ev, eager):
            h.reset()
            h.add_hits(ev_in.x, ev_in.y, ev_in.z, ev_in.idx, ev_in.detId)
            h.run_chain()
            h.sync()
        ref, cnt = oracle.run(gen.geometry(gen.geom_kind(wl))[1], ev_in, oracle.CR)
        _assert_same(ev.export_all(0), ref, f"graph pass {k}")
        _assert_same(eager.export_all(0), ref, f"eager pass {k}")
        assert ev.counts(0) == cnt
        assert np.array_equal(ev.layer_counts(0), eager.layer_counts(0))


def test_graph_recaptured_after_short_prediction():
    """A captured graph whose sizes are too small for new input in the same
    device arrays (same pointers, so the same graph key): that call reruns
    sized exactly, the next call captures a new graph instead of replaying the
    old one into another rerun, and later calls replay it; every call equals
    the oracle."""
    from sdl_amd import hip
    print("case graph recaptured after a short prediction", flush=True)
    small, big = gen.event(12, 0), gen.event(5, 2)
    assert big.nhits > 5 * small.nhits
    fields = ("x", "y", "z", "idx", "detId")
    pad = lambda a: np.concatenate([a, np.zeros(big.nhits - len(a), a.dtype)])
    dev = [hip.DeviceArray(pad(getattr(small, f))) for f in fields]
    hip.synchronize()
    refs = {id(e): oracle.run(gen.geometry(2)[1], e, oracle.CR) for e in (small, big)}
    ev = runtime.Event(_geom(2))

    def run(e, what):
        for d, f in zip(dev, fields):
            d.upload(getattr(e, f))
        ev.set_hits_device([0, e.nhits], *[d.address for d in dev])
        ev.run_chain()
        ev.sync()
        ref, cnt = refs[id(e)]
        _assert_same(ev.export_all(0), ref, what)
        assert ev.counts(0) == cnt
        return ev.run_stats()

    run(big, "big exact")
    run(big, "big capture")  # every buffer now holds the big event
    ev.set_graphs(False)
    run(small, "small eager")
    ev.set_graphs(True)
    s0 = run(small, "small capture")  # a graph sized for the small event
    s0 = run(small, "small replay")
    assert s0["replays"] >= 1
    s1 = run(big, "big after small")  # replays the small graph, falls short, reruns
    assert s1["reruns"] == s0["reruns"] + 1
    s2 = run(big, "big recapture")
    assert s2["reruns"] == s1["reruns"] and s2["captures"] == s1["captures"] + 1
    s3 = run(big, "big replay")
    assert s3["reruns"] == s2["reruns"] and s3["replays"] == s2["replays"] + 1


def test_autotune_keeps_objects():
    """The autotuner only changes launch shapes: after tuning, and with every
    knob at another allowed value, a batch of uneven events gives the same
    objects; the choice is one of the allowed values; bad values are rejected."""
    from sdl_amd import hip
    print("case autotune, batch of workload 15 events 0-1 + workload 12 event 0", flush=True)
    evs = [gen.event(15, 0), gen.event(15, 1), gen.event(12, 0)]
    off = np.zeros(len(evs) + 1, np.int64)
    off[1:] = np.cumsum([e.nhits for e in evs])
    dev = [hip.DeviceArray(np.ascontiguousarray(np.concatenate([getattr(e, f) for e in evs])))
           for f in ("x", "y", "z", "idx", "detId")]
    hip.synchronize()
    ev = runtime.Event(_geom(2))
    ev.set_hits_device(off, *[d.address for d in dev])
    ev.run_chain()
    ev.sync()
    before = [ev.export_all(i) for i in range(len(evs))]
    ref, _ = oracle.run(gen.geometry(2)[1], evs[0], oracle.CR)
    _assert_same(before[0], ref, "before tuning")
    rep = ev.autotune(reps=1)
    chosen = ev.tuning()
    assert set(runtime.TUNING) <= set(rep)
    assert rep["best_ms"] <= rep["start_ms"]
    allowed = {"md_block": {64, 128, 256}, "sg_block": {64, 128, 256}, "tl1_block": {64, 128, 256},
               "tl2_block": {64, 128, 256}, "tl_expand_block": {64, 128, 256},
               "scored_waves": {2, 4, 8}}
    for k, v in chosen.items():
        assert v in allowed[k] and rep[k] == v, k
    for setting in (chosen, {k: min(v) for k, v in allowed.items()}, {k: max(v) for k, v in allowed.items()}):
        ev.set_tuning(setting)
        ev.run_chain()
        ev.sync()
        for i in range(len(evs)):
            _assert_same(ev.export_all(i), before[i], f"tuning {setting} event {i}")
    with pytest.raises(runtime.SDLError):
        ev.set_tuning(dict(chosen, md_block=100))
    for d in dev:
        d.free()


def test_binary_geometry_cache(tmp_path):
    """A geometry loaded from the binary cache gives the same objects as the
    text files and as the oracle (which reads the text files); a loaded
    geometry saves the same cache bytes as the converter."""
    print("case binary geometry cache, workload 15 event 0 and workload 2 event 1", flush=True)
    paths = gen.geometry(2)[1]
    conv = str(tmp_path / "conv.bin")
    runtime.convert_geometry(*paths, conv)
    gb = runtime.Geometry(binary=conv)
    saved = str(tmp_path / "saved.bin")
    gb.save(saved)
    assert open(conv, "rb").read() == open(saved, "rb").read()
    gt = _geom(2)
    assert gb.num_modules == gt.num_modules
    for wl, e in ((15, 0), (2, 1)):
        ev_in = gen.event(wl, e)
        out = []
        for g in (gt, gb):
            ev = runtime.Event(g)
            ev.add_hits(ev_in.x, ev_in.y, ev_in.z, ev_in.idx, ev_in.detId)
            ev.run_chain()
            ev.sync()
            out.append(ev.export_all(0))
            counts = ev.counts(0)
        ref, cnt = oracle.run(paths, ev_in, oracle.CR)
        _assert_same(out[1], ref, f"binary geometry vs oracle, workload {wl} event {e}")
        _assert_same(out[0], ref, f"text geometry vs oracle, workload {wl} event {e}")
        assert counts == cnt


@pytest.mark.gpu
def test_batch_offsets_follow_event_order():
    """Two events batched in the same device arrays, first as (a, b), then as
    (b, a): the batch has the same event count and hit total both times but
    other event offsets, which must reach the device (the library skips the
    upload only when the offsets are unchanged).  Each call, captured and
    replayed, equals the oracle event by event."""
    from sdl_amd import hip
    print("case batch offsets follow event order", flush=True)
    wl = 1
    a, b = gen.event(wl, 3), gen.event(wl, 4)
    assert a.nhits != b.nhits
    refs = {id(e): oracle.run(gen.geometry(gen.geom_kind(wl))[1], e, oracle.CR) for e in (a, b)}
    fields = ("x", "y", "z", "idx", "detId")
    n = a.nhits + b.nhits
    dev = [hip.DeviceArray(np.zeros(n, getattr(a, f).dtype)) for f in fields]
    hip.synchronize()
    ev = runtime.Event(_geom(gen.geom_kind(wl)))
    for order in ((a, b), (b, a), (b, a), (a, b), (a, b), (b, a)):
        for d, f in zip(dev, fields):
            d.upload(np.concatenate([getattr(e, f) for e in order]))
        ev.set_hits_device([0, order[0].nhits, n], *[d.address for d in dev])
        ev.run_chain()
        ev.sync()
        for i, e in enumerate(order):
            ref, cnt = refs[id(e)]
            _assert_same(ev.export_all(i), ref, f"event {i} of {[o.nhits for o in order]}")
            assert ev.counts(i) == cnt
    assert ev.run_stats()["replays"] >= 1

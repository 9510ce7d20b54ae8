"""The benchmark configurations beyond the single PU200 event, on the GPU.

* Uneven batch (workload 3: 2,000-10,000 tracks per event, full geometry):
  eight events of the benchmark's 64-event batch -- its smallest (2,303
  tracks) and largest (9,628) among them -- as one batch, replayed from the
  captured HIP graph; every event's six object sets, ten counters and
  per-layer counters equal the oracle's (CR policy, pinned to the compiled
  reference by the PU200 digest and the golden fixtures).
* The trained learned proposals on events they were not trained or validated
  on: the objects are a subset of the plain chain's, the track-candidate
  recall reaches the target the threshold was chosen for, and the
  second-stage tracklet candidates (the exact cuts' dominant work) fall by the
  stated factor: PU200 events (seeds 60-63) as one batch with
  proposal_w2.json, and a 32-event batch of 500-track barrel events (workload
  1, seeds 10000-10031) with proposal_w1.json, trained on 2,000 such events,
  at each of its 99, 99.5 and 99.9 % thresholds.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from sdl_amd import gen, hip, proposal, runtime

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UNEVEN = [0, 1, 2, 3, 4, 5, 9, 37]  # workload-3 events; 37: 2,303 tracks, 9: 9,628 tracks


class _Batch:
    def __init__(self, evs):
        self.off = np.zeros(len(evs) + 1, np.int64)
        self.off[1:] = np.cumsum([e.nhits for e in evs])
        cat = lambda f: np.ascontiguousarray(np.concatenate([getattr(e, f) for e in evs]))
        self.dev = [hip.DeviceArray(cat(f)) for f in ("x", "y", "z", "idx", "detId")]
        hip.synchronize()

    def run(self, ev):
        ev.set_hits_device(self.off, *[d.address for d in self.dev])
        ev.run_chain()
        ev.sync()

    def free(self):
        for d in self.dev:
            d.free()


def _events(wl, seeds):
    gen.geometry(gen.geom_kind(wl))
    with ThreadPoolExecutor(max_workers=8) as pool:
        return list(pool.map(lambda i: gen.event(wl, i), seeds))


def test_uneven_batch_matches_oracle():
    wl = 3
    evs = _events(wl, UNEVEN)
    tracks = [e.ntracks for e in evs]
    print(f"case uneven batch workload 3 events {UNEVEN} tracks {tracks}", flush=True)
    assert min(tracks) == 2303 and max(tracks) == 9628
    paths = gen.geometry(gen.geom_kind(wl))[1]
    with ThreadPoolExecutor(max_workers=8) as pool:  # the oracle releases the GIL inside its C++ chain
        ref = list(pool.map(lambda e: oracle.run(paths, e, oracle.CR, 6, layers=True), evs))
    b = _Batch(evs)
    ev = runtime.Event(runtime.Geometry(*paths))
    for rep in range(3):  # exact sizing, then the speculative chain captured into a graph, then a replay
        b.run(ev)
        for i, (sets, counters, layers) in enumerate(ref):
            for s in oracle.STAGES:
                got = oracle.as_multiset(ev.export(runtime.STAGES.index(s), i))
                assert np.array_equal(got, oracle.as_multiset(sets[s])), f"run {rep} event {UNEVEN[i]} {s}"
            assert ev.counts(i) == counters, f"run {rep} event {UNEVEN[i]} counters"
            assert np.array_equal(ev.layer_counts(i), layers), f"run {rep} event {UNEVEN[i]} per-layer counters"
    b.free()


def _plain_and_learned(wl, seeds, weights, recalls):
    """Plain chain and the learned chain at each target recall's threshold on one
    batch: ({"plain" | recall: (per-event sets, totals, tl_work)}, thresholds)."""
    evs = _events(wl, seeds)
    b = _Batch(evs)
    ev = runtime.Event(runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1]))
    m, meta = proposal.MLP.load(os.path.join(ROOT, "sdl_amd", "data", weights))
    assert not set(seeds) & set(meta["train_seeds"]) and not set(seeds) & set(meta["val_seeds"])
    out = {}
    for key in ["plain"] + list(recalls):
        if key == "plain":
            proposal.set_proposal(ev, None, 0.5)
        else:
            proposal.set_proposal(ev, m, meta["thresholds"][key])
        b.run(ev)
        sets = [{s: ev.export(k, i) for k, s in enumerate(runtime.STAGES)} for i in range(len(evs))]
        out[key] = (sets, ev.counts(-1), ev.tl_work())
    proposal.set_proposal(ev, None, 0.5)
    b.free()
    return out, meta["thresholds"]


def _rows(a):
    """Each row of an int32 object array as one opaque value (for set membership)."""
    a = np.ascontiguousarray(a)
    return a.view(np.dtype((np.void, a.dtype.itemsize * a.shape[1]))).ravel()


def _check_learned(res, key, target, min_cut):
    plain_sets, plain_tot, plain_work = res["plain"]
    sets, tot, work = res[key]
    for i, (p, l) in enumerate(zip(plain_sets, sets)):
        for s in runtime.STAGES:
            assert len(l[s]) <= len(p[s]), f"event {i} {s}"
            if len(l[s]):
                assert np.isin(_rows(l[s]), _rows(p[s])).all(), f"event {i} {s}: learned objects not in the plain chain"
    recall = tot["nTC"] / plain_tot["nTC"]
    cut = plain_work["stage2"] / max(work["stage2"], 1)
    print(f"recall target {key}: TC recall {recall:.5f}, TL recall {tot['nTL'] / plain_tot['nTL']:.5f}, "
          f"second-stage candidates {plain_work['stage2']} -> {work['stage2']} ({cut:.2f}x)", flush=True)
    assert recall >= target, (key, recall)
    assert cut >= min_cut, (key, cut)
    assert work["scored"] > 0
    return recall, cut


def test_trained_proposal_pu200_heldout():
    """proposal_w2.json (trained on PU200 seeds 100-111, threshold chosen on
    50-53) on PU200 seeds 60-63 as one batch: subset, TC recall >= 99.5 %,
    second-stage tracklet candidates cut at least 3x."""
    print("case learned proposal, workload 2 seeds 60-63", flush=True)
    res, _ = _plain_and_learned(2, [60, 61, 62, 63], "proposal_w2.json", ["0.995"])
    _check_learned(res, "0.995", 0.995, 3.0)


def test_trained_proposal_config1_batch():
    """proposal_w1.json (trained on 2,000 500-track barrel events) on a held-out
    32-event batch at the thresholds swept for 99, 99.5 and 99.9 % TC recall:
    subset, recall at least the target less 0.2 % (one batch is a small
    sample of the rare lost TCs), recall and work ordered by the thresholds."""
    print("case learned proposal, workload 1 seeds 10000-10031", flush=True)
    keys = ["0.99", "0.995", "0.999"]
    res, thr = _plain_and_learned(1, list(range(10000, 10032)), "proposal_w1.json", keys)
    assert thr["0.99"] >= thr["0.995"] >= thr["0.999"]
    got = [_check_learned(res, k, float(k) - 0.002, 1.0) for k in keys]
    recalls, cuts = zip(*got)
    assert recalls[0] <= recalls[1] <= recalls[2] and cuts[0] >= cuts[1] >= cuts[2]

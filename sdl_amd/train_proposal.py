"""Trains the learned proposal of the tracklet stage on the GPU and chooses its
score thresholds.

    python -m sdl_amd.train_proposal [--workload 2] [--train 12] [--val 4] [--epochs 4]
                                     [--events-per-batch 1] [--out sdl_amd/data/proposal_w2.json]

PU200-density events (workload 2) train one event at a time; small events (the
500-track barrel events of workload 1) run as batches of --events-per-batch
events, each batch one chain over device-resident hits.

1. The plain chain runs on seeded training events (seeds 100, 101, ...); every
   (inner segment, outer-inner MD) pair of its tracklet stage that passes the
   first-stage cuts (the pairs the proposal scores) becomes a row: 24 features
   and the label "the pair yields at least one tracklet" -- labels the exact
   cuts generate.
2. The network (24-64-64-64-1) trains on the GPU with this library's FP32
   forward/backward kernels and Adam (weighted cross entropy; positives are
   rare, so they are up-weighted).
3. On validation events (seeds 50, 51, ..., or the seeds after the training
   ones when there are more than 50; in batches like the training events), the chain runs with the proposal at a sweep of thresholds;
   track-candidate recall against the plain chain (the learned objects are a
   subset of the plain ones) picks, for each target recall (99, 99.5, 99.9 %),
   the largest threshold that reaches it on every validation batch.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from sdl_amd import gen, hip, proposal as P, runtime  # noqa: E402

TARGETS = (0.99, 0.995, 0.999)
# score thresholds swept (sigmoid outputs): 40 steps, geometric from 0.001 to 0.5
GRID = [float(t) for t in np.round(np.geomspace(0.001, 0.5, 40), 5)]


class Feed:
    """Events of one batch: host arrays for a single event, device-resident
    hits (one set of arrays, per-event offsets) for several."""

    def __init__(self, evs):
        self.evs = evs
        self.dev = None
        if len(evs) > 1:
            self.off = np.zeros(len(evs) + 1, np.int64)
            self.off[1:] = np.cumsum([e.nhits for e in evs])
            cat = lambda f: np.ascontiguousarray(np.concatenate([getattr(e, f) for e in evs]))
            self.dev = [hip.DeviceArray(cat(f)) for f in ("x", "y", "z", "idx", "detId")]

    def run(self, ev):
        if self.dev is None:
            e = self.evs[0]
            ev.reset()
            ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
        else:
            ev.set_hits_device(self.off, *[d.address for d in self.dev])
        ev.run_chain()
        ev.sync()

    def free(self):
        for d in self.dev or []:
            d.free()


def feeds(workload, first, n, per):
    """Batches of `per` events (seeds first .. first+n-1), generated on host threads."""
    gen.geometry(gen.geom_kind(workload))
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        evs = list(pool.map(lambda i: gen.event(workload, i), range(first, first + n)))
    return [Feed(evs[k:k + per]) for k in range(0, n, per)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--train", type=int, default=12)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--events-per-batch", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="default sdl_amd/data/proposal_w<workload>.json")
    a = ap.parse_args()
    out_path = a.out or os.path.join(ROOT, "sdl_amd", "data", f"proposal_w{a.workload}.json")
    t0 = time.time()
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(a.workload))[1])
    ev = runtime.Event(g)
    ev.set_graphs(False)
    per = max(1, a.events_per_batch)

    # 1. rows from the plain chain
    rows = []
    npos = nneg = 0
    train = feeds(a.workload, 100, a.train, per)
    for k, f in enumerate(train):
        f.run(ev)
        X, Y, n = P.pair_rows(ev)
        y = Y.to_host()[:n]
        scored = np.flatnonzero(y >= 0).astype(np.int32)  # pairs that pass the first stage
        pos = int((y == 1).sum())
        npos += pos
        nneg += len(scored) - pos
        rows.append((X, Y, scored))
        f.free()
        if k % max(1, len(train) // 10) == 0 or k == len(train) - 1:
            print(f"train batch {k} ({len(f.evs)} events): {n} pairs, {len(scored)} scored, {pos} positive",
                  flush=True)
    pos_weight = float(max(1.0, nneg / max(npos, 1) / 8.0))

    # 2. training
    m = P.MLP(P.init_params(a.seed))
    rng = np.random.default_rng(a.seed)
    history = []
    for epoch in range(a.epochs):
        lr = a.lr * (0.5 ** epoch)
        losses = []
        for k in rng.permutation(len(rows)):
            X, Y, scored = rows[k]
            steps = len(scored) // a.batch
            if steps == 0:
                continue
            perm = hip.DeviceArray(rng.permutation(scored))
            losses.append(m.train(X.address, Y.address, perm.address, 0, a.batch, steps, lr, pos_weight))
            perm.free()
        history.append(float(np.mean(losses)))
        print(f"epoch {epoch}: lr {lr:.2e} mean last-batch loss {history[-1]:.5f}", flush=True)
    for X, Y, _ in rows:
        X.free()
        Y.free()

    # 3. thresholds from track-candidate recall on validation batches
    val_first = 50 if a.val <= 50 else 100 + a.train  # never a training seed
    val = feeds(a.workload, val_first, a.val, per)
    plain_tc = []
    for f in val:
        P.set_proposal(ev, None, 0.5)
        f.run(ev)
        plain_tc.append(ev.counts(-1)["nTC"])
    sweep = []
    for t in GRID:
        P.set_proposal(ev, m, t)
        rec = []
        for f, base in zip(val, plain_tc):
            f.run(ev)
            rec.append(ev.counts(-1)["nTC"] / max(base, 1))
        sweep.append({"threshold": t, "tc_recall_min": float(min(rec)), "tc_recall_mean": float(np.mean(rec))})
        print(f"threshold {t}: TC recall min {min(rec):.5f} mean {np.mean(rec):.5f}", flush=True)
    P.set_proposal(ev, None, 0.5)
    for f in val:
        f.free()
    thresholds = {}
    for target in TARGETS:
        ok = [s["threshold"] for s in sweep if s["tc_recall_min"] >= target]
        thresholds[str(target)] = max(ok) if ok else None
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    m.save(out_path, workload=a.workload, train_seeds=list(range(100, 100 + a.train)),
           val_seeds=list(range(val_first, val_first + a.val)),
           events_per_batch=per, epochs=a.epochs, batch=a.batch, lr=a.lr, pos_weight=pos_weight, positives=npos,
           negatives=nneg, loss_history=history, sweep=sweep, thresholds=thresholds,
           seconds=round(time.time() - t0, 1))
    print(json.dumps({"thresholds": thresholds, "out": out_path}), flush=True)


if __name__ == "__main__":
    main()

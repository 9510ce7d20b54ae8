"""Trains the learned proposal of the tracklet stage on the GPU and chooses its
score thresholds.

    python -m sdl_amd.train_proposal [--workload 2] [--train 12] [--val 4] [--epochs 4]
                                     [--out sdl_amd/data/proposal_w2.json]

1. The plain chain runs on seeded training events (seeds 100, 101, ...); every
   (inner segment, outer-inner MD) pair of its tracklet stage that passes the
   first-stage cuts (the pairs the proposal scores) becomes a row: 24 features
   and the label "the pair yields at least one tracklet" -- labels the exact
   cuts generate.
2. The network (24-64-64-64-1) trains on the GPU with this library's FP32
   forward/backward kernels and Adam (weighted cross entropy; positives are
   rare, so they are up-weighted).
3. On validation events (seeds 50, 51, ...), the chain runs with the proposal at
   a sweep of thresholds; track-candidate recall against the plain chain (the
   learned objects are a subset of the plain ones) picks, for each target
   recall (99, 99.5, 99.9 %), the largest threshold that reaches it on every
   validation event.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from sdl_amd import gen, hip, proposal as P, runtime  # noqa: E402

TARGETS = (0.99, 0.995, 0.999)


def tc_count(ev):
    c = ev.counts(0)
    return c["nTC"]


def run(ev, e):
    ev.reset()
    ev.add_hits(e.x, e.y, e.z, e.idx, e.detId)
    ev.run_chain()
    ev.sync()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--train", type=int, default=12)
    ap.add_argument("--val", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(ROOT, "sdl_amd", "data", f"proposal_w2.json"))
    a = ap.parse_args()
    t0 = time.time()
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(a.workload))[1])
    ev = runtime.Event(g)
    ev.set_graphs(False)

    # 1. rows from the plain chain
    rows = []
    npos = nneg = 0
    for k in range(a.train):
        run(ev, gen.event(a.workload, 100 + k))
        X, Y, n = P.pair_rows(ev)
        y = Y.to_host()[:n]
        scored = np.flatnonzero(y >= 0).astype(np.int32)  # pairs that pass the first stage
        pos = int((y == 1).sum())
        npos += pos
        nneg += len(scored) - pos
        rows.append((X, Y, scored))
        print(f"train event {100 + k}: {n} pairs, {len(scored)} scored, {pos} positive", flush=True)
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

    # 3. thresholds from track-candidate recall on validation events
    val = [gen.event(a.workload, 50 + k) for k in range(a.val)]
    plain_tc = []
    for e in val:
        P.set_proposal(ev, None, 0.5)
        run(ev, e)
        plain_tc.append(tc_count(ev))
    grid = [0.001, 0.002, 0.005, 0.01, 0.02, 0.03, 0.05, 0.08, 0.12, 0.18, 0.25, 0.35, 0.5]
    sweep = []
    for t in grid:
        P.set_proposal(ev, m, t)
        rec = []
        for e, base in zip(val, plain_tc):
            run(ev, e)
            rec.append(tc_count(ev) / max(base, 1))
        sweep.append({"threshold": t, "tc_recall_min": float(min(rec)), "tc_recall_mean": float(np.mean(rec))})
        print(f"threshold {t}: TC recall min {min(rec):.5f} mean {np.mean(rec):.5f}", flush=True)
    P.set_proposal(ev, None, 0.5)
    thresholds = {}
    for target in TARGETS:
        ok = [s["threshold"] for s in sweep if s["tc_recall_min"] >= target]
        thresholds[str(target)] = max(ok) if ok else None
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    m.save(a.out, workload=a.workload, train_seeds=[100 + k for k in range(a.train)],
           val_seeds=[50 + k for k in range(a.val)], epochs=a.epochs, batch=a.batch, lr=a.lr,
           pos_weight=pos_weight, positives=npos, negatives=nneg, loss_history=history, sweep=sweep,
           thresholds=thresholds, seconds=round(time.time() - t0, 1))
    print(json.dumps({"thresholds": thresholds, "out": a.out}), flush=True)


if __name__ == "__main__":
    main()

"""Throughput benchmark of the SDL chain on one MI355X.

Workload: PU200-density events of the full tilted-barrel + endcap detector
(8000 tracks with p(pT) ~ pT exp(-pT/0.6 GeV) above 0.8 GeV + 20 hard tracks,
|eta| < 2.4, multiple scattering, 5% noise; ~1.1e5 hits per event), generated
synthetically and uploaded to HBM before the timed region.  A step runs the
whole module-map chain (MD, SG, TP, TL, TC from triplets, TC from tracklets) on
a batch of --events-per-step events through the C-ABI.

Prints one JSON line (see the keys below); exits when asked for more than one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3  # vector FP32 (packed); 78.6 without packed FP32
PEAK_BF16_TFLOPS = 2500.0  # dense MFMA
# frozen record sizes of the chain-level roofline (bytes): hit 16 + 16 derived,
# MD 8 + 32, SG 8 + 48, TP 8, TL 8 + 24, TC 8
REC = {"hit": 32, "md": 40, "sg": 56, "tp": 8, "tl": 32, "tc": 8}
F_MLP = 2 * (24 * 64 + 64 * 64 + 64 * 64 + 64 * 1)  # FLOP per scored pair


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--events-per-step", type=int, default=1)
    p.add_argument("--workload", type=int, default=2)
    p.add_argument("--cpu-baseline", type=int, default=1, help="0 skips the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (at most 16)")
    p.add_argument("--learned", default=os.path.join(ROOT, "sdl_amd", "data", "proposal_w2.json"),
                   help="trained proposal (skipped if the file is absent)")
    p.add_argument("--recall", default="0.995", help="target TC recall whose threshold the learned leg uses")
    p.add_argument("--batch", type=int, default=64,
                   help="events of the uneven-batch leg (workload --batch-workload), 0 skips it")
    p.add_argument("--batch-workload", type=int, default=3)
    p.add_argument("--batch-steps", type=int, default=3)
    p.add_argument("--autotune", type=int, default=1, help="0 skips the autotuner of the uneven-batch leg")
    p.add_argument("--hard-mix", type=int, default=256,
                   help="events of the hard-mix leg (workload 5, plain and learned, in sub-batches), 0 skips it")
    p.add_argument("--hard-mix-batch", type=int, default=64)
    p.add_argument("--learned-batch", type=int, default=32,
                   help="events per batch of the batched plain-vs-learned leg on the main workload, 0 skips it")
    p.add_argument("--config1-batch", type=int, default=32,
                   help="events per batch of the plain-vs-learned leg on 500-track barrel events (workload 1, "
                        "proposal_w1.json), 0 skips it")
    p.add_argument("--scaling", default="1,4,16,64",
                   help="events per step of the batch-scaling leg on the main workload (comma list; 0 skips it)")
    p.add_argument("--families", type=int, default=1,
                   help="0 skips the timing of the other stage families (Nm1, navigation, pseudo MDs, layer pairs)")
    return p.parse_args()


def cpu_baseline(workload, threads):
    """The CPU oracle (the C++ restatement of the reference, port) timed on a
    bounded sample: one event per thread, all threads busy at once; and one
    event on one thread (ms/event)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from sdl_amd import gen
    n = threads
    evs = [gen.event(workload, 1000 + i) for i in range(n)]
    paths = gen.geometry(gen.geom_kind(workload))[1]
    wall = oracle.time_events(paths, evs, oracle.GLIBC, threads)
    single = oracle.time_events(paths, evs[:1], oracle.GLIBC, 1)
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n / wall, "unit": "events/s", "cores": threads, "kind": "port",
            "single_thread_ms_per_event": round(1000.0 * single, 1), "cpu_model": cpu,
            "sample": f"{n} workload-{workload} events (seeds 1000..{1000 + n - 1}), one per thread, "
                      f"full chain, glibc libm, {wall:.1f} s wall; single thread: seed 1000, {single:.1f} s"}


def tl_roofline(workload, nevents, tl_ms):
    """Roofline of the dominant cost, the tracklet stage: the k_tl_* launches of
    one chain (pair lists, first-stage cuts, expansion, second-stage cuts,
    scatter, and their scans), timed together with HIP events.  Its work is
    floating-point cut algebra, so its ceiling is the FP32 peak: achieved =
    algorithmic FP operations of the tracklet stage on the step's events
    (bench_work.json, counted by the oracle's op-counting build along each
    candidate's path in the reference's loops) / measured stage duration.
    traffic = HBM bytes of the stage per chain from the FETCH_SIZE / WRITE_SIZE
    counters of the profiled run recorded in bench_work.json, scaled by the
    copy-kernel calibration."""
    path = os.path.join(ROOT, "bench_work.json")
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    rows = [table.get("events", {}).get(f"{workload}:{e}") for e in range(nevents)]
    flops = sum(r["flops"]["tl"] for r in rows) if all(rows) else None
    tl_s = tl_ms / 1000.0
    achieved = flops / tl_s / 1e12 if flops and tl_s > 0 else None
    traffic = table.get("traffic", {}).get(f"k_tl:{workload}:{nevents}")
    return {"bound": "valu", "achieved": None if achieved is None else round(achieved, 4),
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": None if achieved is None else round(achieved / PEAK_FP32_TFLOPS, 6),
            "traffic": None if traffic is None else traffic["bytes"],
            "kernel": "tracklet stage (k_tl_* launches of one chain)", "kernel_ms": round(tl_ms, 4),
            "peak_note": "FP32 vector peak 157.3 TFLOP/s (78.6 without packed FP32, which this build does "
                         "not use); the stage issues FP32/FP64 VALU and no MFMA, and is VALU-issue bound "
                         "(profiles/chain_pu200_counters.md)",
            "frac_nonpacked": None if achieved is None else round(achieved / 78.6, 6),
            "flops_per_launch": flops}


def chain_roofline(workload, nevents, ms_per_step, scored=0):
    """Bound of the whole chain: T_roof = sum over stages of max(F_s / P_FP32,
    B_s / P_HBM) (+ F_mlp x scored pairs / P_BF16 for the learned form), with
    F_s the algorithmic FP operations of bench_work.json and B_s each stage's
    input records read once plus its output records written once at the frozen
    record sizes (REC); frac = T_roof / measured time per step."""
    path = os.path.join(ROOT, "bench_work.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    rows = [table.get("events", {}).get(f"{workload}:{e}") for e in range(nevents)]
    if not all(rows):
        return None
    stages = {}
    for r in rows:
        F = r["flops"]
        nh = r["hits"]
        nmd, nsg, ntp, ntl, ntcp, ntcl = r["objects"]
        B = {"hits": 32 * nh,
             "md": REC["hit"] * nh + REC["md"] * nmd,
             "sg": REC["md"] * nmd + REC["sg"] * nsg,
             "tp": REC["sg"] * nsg + REC["tp"] * ntp,
             "tl": REC["sg"] * nsg + REC["tl"] * ntl,
             "tcp": REC["tp"] * ntp + REC["tc"] * ntcp,
             "tcl": REC["tl"] * ntl + REC["tc"] * ntcl}
        for k in B:
            t_f = F[k] / (PEAK_FP32_TFLOPS * 1e12)
            t_b = B[k] / (PEAK_HBM_GBS * 1e9)
            st = stages.setdefault(k, {"flops": 0, "bytes": 0, "t_roof_us": 0.0})
            st["flops"] += F[k]
            st["bytes"] += B[k]
            st["t_roof_us"] += 1e6 * max(t_f, t_b)
    t_roof = sum(v["t_roof_us"] for v in stages.values()) + 1e6 * F_MLP * scored / (PEAK_BF16_TFLOPS * 1e12)
    for v in stages.values():
        v["t_roof_us"] = round(v["t_roof_us"], 3)
        v["bound"] = "valu" if v["flops"] / (PEAK_FP32_TFLOPS * 1e12) >= v["bytes"] / (PEAK_HBM_GBS * 1e9) else "hbm"
    return {"t_roof_us": round(t_roof, 3), "t_measured_us": round(1000.0 * ms_per_step, 3),
            "frac": round(t_roof / (1000.0 * ms_per_step), 5), "scored_pairs": scored, "stages": stages,
            "formula": "sum_s max(F_s/P_FP32, B_s/P_HBM) + F_mlp*scored/P_BF16; P_FP32 157.3 TFLOP/s, "
                       "P_HBM 8 TB/s, P_BF16 2.5 PFLOP/s"}


def learned_leg(a, ev, step, plain_counts, plain_ms, plain_work):
    """The same steps with the learned proposal at the threshold chosen for the
    target TC recall; recall = TCs kept / TCs of the plain chain on the step's
    events (the learned objects are a subset of the plain ones)."""
    from sdl_amd import hip, proposal
    m, meta = proposal.MLP.load(a.learned)
    t = (meta.get("thresholds") or {}).get(a.recall)
    if t is None:
        return {"skipped": f"no threshold for recall {a.recall} in {os.path.basename(a.learned)}"}
    proposal.set_proposal(ev, m, t)
    for _ in range(2 + a.warmup):  # set-up of the learned form (buffers, graph), then the warmup
        step()
    ev.sync()
    hip.synchronize()
    s0 = ev.run_stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ev.sync()
    hip.synchronize()
    wall = time.perf_counter() - t0
    calls = stats_delta(s0, ev.run_stats())
    c = ev.counts(-1)
    work = ev.tl_work()
    proposal.set_proposal(ev, None, 0.5)
    ms = 1000.0 * wall / a.steps
    return {"value": round(a.events_per_step * a.steps / wall, 3), "ms_per_step": round(ms, 4),
            "speedup_vs_plain": round(plain_ms / ms, 4), "threshold": t, "target_tc_recall": float(a.recall),
            "tc_recall": round(c["nTC"] / max(plain_counts["nTC"], 1), 6),
            "tl_recall": round(c["nTL"] / max(plain_counts["nTL"], 1), 6),
            "tl_work_learned": work, "tl_work_plain": plain_work,
            "weights": os.path.relpath(a.learned, ROOT), "trained_on": meta.get("train_seeds"),
            "validated_on": meta.get("val_seeds"), "timed_calls": calls}


def generate(workload, first, n):
    """Events first .. first+n-1 of a workload, generated on host threads."""
    from concurrent.futures import ThreadPoolExecutor
    from sdl_amd import gen
    gen.geometry(gen.geom_kind(workload))
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as pool:
        return list(pool.map(lambda i: gen.event(workload, i), range(first, first + n)))


class Batch:
    """A batch of generated events resident in HBM, fed to an event handle."""

    def __init__(self, evs):
        from sdl_amd import hip
        self.n = len(evs)
        self.off = np.zeros(self.n + 1, np.int64)
        self.off[1:] = np.cumsum([e.nhits for e in evs])
        cat = lambda f: np.ascontiguousarray(np.concatenate([getattr(e, f) for e in evs]))
        self.dev = [hip.DeviceArray(cat(f)) for f in ("x", "y", "z", "idx", "detId")]
        self.tracks = [e.ntracks for e in evs]
        hip.synchronize()

    def run(self, ev):
        ev.set_hits_device(self.off, *[d.address for d in self.dev])
        ev.run_chain()

    def free(self):
        for d in self.dev:
            d.free()


def stats_delta(before, after):
    """Graph replays, captures and exact reruns between two run_stats() readings."""
    return {k: after[k] - before[k] for k in ("replays", "captures", "reruns")}


def timed(ev, fn, warmup, steps):
    """Seconds per step over `steps` calls after `warmup` calls, and the
    handle's graph replays, captures and exact reruns during the timed calls."""
    from sdl_amd import hip
    for _ in range(warmup):
        fn()
    ev.sync()
    hip.synchronize()
    s0 = ev.run_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ev.sync()
    hip.synchronize()
    return (time.perf_counter() - t0) / steps, stats_delta(s0, ev.run_stats())


def batch_leg(a, geom):
    """The plain chain on one batch of events of uneven size (workload
    --batch-workload: 2,000-10,000 tracks per event) as one set of launches,
    after the autotuner has chosen the launch shapes on this batch."""
    from sdl_amd import runtime
    evs = generate(a.batch_workload, 0, a.batch)
    b = Batch(evs)
    ev = runtime.Event(geom)
    b.run(ev)  # the batch the tuner times
    tune = None
    if a.autotune:
        tune = ev.autotune(reps=3)
    sec, st = timed(ev, lambda: b.run(ev), 2, a.batch_steps)
    counts = ev.counts(-1)
    out = {"workload": f"workload {a.batch_workload}: {a.batch} events, tracks per event uniform in "
                       f"[2000, 10000] + 20 hard, one batch per step",
           "events_per_step": a.batch, "value": round(a.batch / sec, 3), "unit": "events/s",
           "ms_per_step": round(1000 * sec, 3), "hits_per_step": int(b.off[-1]),
           "tracks_min_max": [min(b.tracks), max(b.tracks)], "objects_per_step": counts,
           "tuning": ev.tuning(), "autotune": tune, "timed_calls": st}
    b.free()
    return out


def plain_vs_learned(a, geom, workload, nevents, batch, steps, weights=None, recall=None, first_seed=0):
    """Events first_seed .. first_seed + nevents - 1 of a workload (none of them a
    training or validation seed of the weights) in batches of `batch` events:
    events/s of the plain chain and of the learned proposal (at the threshold
    chosen for the target TC recall) on the same batches, and the per-stage
    recall of the learned chain against the plain one (its objects are a subset
    of the plain chain's)."""
    from sdl_amd import proposal, runtime
    weights = weights or a.learned
    recall = recall or a.recall
    m, meta = proposal.MLP.load(weights)
    t = (meta.get("thresholds") or {}).get(recall)
    if t is None:
        return {"skipped": f"no threshold for recall {recall} in {os.path.basename(weights)}"}
    seeds = set(range(first_seed, first_seed + nevents))
    if meta.get("workload") == workload:
        used = set(meta.get("train_seeds") or []) | set(meta.get("val_seeds") or [])
        if seeds & used:
            raise SystemExit(f"bench: events {first_seed}..{first_seed + nevents - 1} overlap the seeds the "
                             f"proposal was trained or validated on")
    ev = runtime.Event(geom)
    keys = ("nMD", "nSG", "nTP", "nTL", "nTC")
    tot = {"plain": dict.fromkeys(keys, 0), "learned": dict.fromkeys(keys, 0)}
    sec = {"plain": 0.0, "learned": 0.0}
    work = {"plain": None, "learned": None}
    calls = {f: {"replays": 0, "captures": 0, "reruns": 0} for f in ("plain", "learned")}
    for first in range(first_seed, first_seed + nevents, batch):
        b = Batch(generate(workload, first, min(batch, first_seed + nevents - first)))
        for form in ("plain", "learned"):
            proposal.set_proposal(ev, m if form == "learned" else None, t if form == "learned" else 0.5)
            # three warmup calls: the first call of a new batch or form may rerun
            # sized exactly (its predictions come from the last batch or form),
            # the next one grows the buffers for the speculative form and
            # captures its graph; only calls after that are timed
            dt, st = timed(ev, lambda: b.run(ev), 3, steps)
            sec[form] += dt
            for k in st:
                calls[form][k] += st[k]
            c = ev.counts(-1)
            for k in keys:
                tot[form][k] += c[k]
            work[form] = ev.tl_work()
        proposal.set_proposal(ev, None, 0.5)
        b.free()
    rec = {k[1:]: round(tot["learned"][k] / max(tot["plain"][k], 1), 6) for k in keys}
    return {"events": nevents, "events_per_step": batch,
            "plain_events_per_s": round(nevents / sec["plain"], 3),
            "learned_events_per_s": round(nevents / sec["learned"], 3),
            "speedup_vs_plain": round(sec["plain"] / sec["learned"], 4), "threshold": t,
            "target_tc_recall": float(recall), "recall_per_stage": rec,
            "weights": os.path.relpath(weights, ROOT),
            "objects_plain": tot["plain"], "objects_learned": tot["learned"],
            "tl_work_last_batch": work, "timed_calls": calls}


def scaling_leg(a, geom):
    """The plain chain on batches of 1, 4, 16, 64 ... events of the main
    workload (seeds 0 .. n-1), one batch per step: events/s against the batch
    size."""
    from sdl_amd import runtime
    sizes = [int(x) for x in a.scaling.split(",") if x.strip()]
    evs = generate(a.workload, 0, max(sizes))
    ev = runtime.Event(geom)
    rows = []
    for n in sizes:
        b = Batch(evs[:n])
        sec, st = timed(ev, lambda: b.run(ev), 3, 5)
        rows.append({"events_per_step": n, "events_per_s": round(n / sec, 3), "ms_per_step": round(1000 * sec, 3),
                     "timed_calls": st})
        b.free()
    return {"workload": f"workload {a.workload}: the first n events as one batch per step", "rows": rows}


def config1_leg(a, geom):
    """500-track barrel events (workload 1, seeds from 10000: neither training
    nor validation seeds), --config1-batch events per batch, 4 batches: plain
    vs the learned proposal trained on 2,000 such events (proposal_w1.json) at
    each threshold of its 99 / 99.5 / 99.9 % TC-recall sweep."""
    w = os.path.join(ROOT, "sdl_amd", "data", "proposal_w1.json")
    if not os.path.exists(w):
        return {"skipped": "sdl_amd/data/proposal_w1.json absent"}
    out = {"workload": f"workload 1 (500-track barrel events): {4 * a.config1_batch} events (seeds 10000..."
                       f"{10000 + 4 * a.config1_batch - 1}) in batches of {a.config1_batch}", "by_target_recall": {}}
    for r in ("0.99", "0.995", "0.999"):
        res = plain_vs_learned(a, geom, 1, 4 * a.config1_batch, a.config1_batch, 5, weights=w, recall=r,
                               first_seed=10000)
        out["by_target_recall"][r] = res
    return out


FAMILIES = {  # name -> [(stage, algo), ...] through sdl_event_run
    "module_map_nm1": [(0, 0), (1, 0), (2, 0), (3, 1), (4, 0), (5, 0)],
    "navigation": [(0, 0), (1, 0), (9, 0), (5, 0)],
    "navigation_nm1": [(0, 0), (1, 0), (9, 1)],
    "pseudo_md_module_map": [(10, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0)],
    "layer_pairs": [(0, 0), (6, 0), (7, 0), (8, 0)],
}


def families_leg(a, geom):
    """The other stage families of the SDL::Event API on one PU200-density event
    (workload 2 event 0), stage by stage through sdl_event_run: ms per event
    (median of 3 after a warmup) and the objects it stores."""
    from sdl_amd import gen, hip, runtime
    b = Batch([gen.event(2, 0)])
    out = {}
    for name, seq in FAMILIES.items():
        ev = runtime.Event(geom)

        def once():
            ev.set_hits_device(b.off, *[d.address for d in b.dev])
            for st, al in seq:
                ev.run(st, al)
            ev.sync()

        once()
        ts = []
        for _ in range(3):
            hip.synchronize()
            t0 = time.perf_counter()
            once()
            ts.append(time.perf_counter() - t0)
        out[name] = {"stages": seq, "ms_per_event": round(1000.0 * sorted(ts)[1], 3), "objects": ev.counts(0)}
        del ev
    b.free()
    return out


def hard_mix_leg(a, geom):
    """Workload 5: 15 % low-pT and 5 % displaced tracks, 10 % noise, 2,000-10,000
    tracks per event; --hard-mix events in sub-batches, plain and learned."""
    out = plain_vs_learned(a, geom, 5, a.hard_mix, a.hard_mix_batch, 2)
    out["workload"] = f"workload 5 (hard mix): {a.hard_mix} events in batches of {a.hard_mix_batch}"
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the chain shards by event, "
              "see the design notes, but no multi-GPU run is part of this build.", file=sys.stderr)
        sys.exit(2)
    from sdl_amd import gen, hip, runtime

    B = a.events_per_step
    evs = [gen.event(a.workload, i) for i in range(B)]
    off = np.zeros(B + 1, np.int64)
    off[1:] = np.cumsum([e.nhits for e in evs])
    cat = lambda f: np.ascontiguousarray(np.concatenate([getattr(e, f) for e in evs]))
    dev = [hip.DeviceArray(cat(f)) for f in ("x", "y", "z", "idx", "detId")]
    hip.synchronize()

    geom = runtime.Geometry(*gen.geometry(gen.geom_kind(a.workload))[1])
    ev = runtime.Event(geom)
    ptrs = [d.address for d in dev]
    stream = ev.stream

    def step():
        ev.set_hits_device(off, *ptrs)
        ev.run_chain()

    # Handle set-up, before the W warmup steps: the first call sizes every buffer
    # exactly, the second grows them for the speculative form and captures its
    # graph; from then on a call is one graph replay and one host round trip.
    for _ in range(2):
        step()
    for _ in range(a.warmup):
        step()
    ev.sync()
    hip.synchronize()

    # Timed region: K steps back to back, no per-kernel timing.
    start, stop = hip.Event(), hip.Event()
    hip.synchronize()
    stats0 = ev.run_stats()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(a.steps):
        step()
    stop.record(stream)
    ev.sync()
    hip.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop)
    ms_per_step = 1000.0 * wall / a.steps
    timed_calls = stats_delta(stats0, ev.run_stats())
    counts = ev.counts(-1)

    # Per-kernel durations, measured in a second pass (HIP events around the
    # hit-preparation, MD, SG and TL kernels on the handle's stream).
    ev.set_timing(True)
    kms = {k: 0.0 for k in runtime.TIMERS}
    for _ in range(a.steps):
        step()
        for k, v in ev.kernel_ms().items():
            kms[k] += v
    ev.set_timing(False)
    for k in kms:
        kms[k] /= a.steps

    roofline = tl_roofline(a.workload, B, kms["tl"])
    chain = chain_roofline(a.workload, B, gpu_ms / a.steps)
    plain_work = ev.tl_work()
    learned = learned_leg(a, ev, step, counts, ms_per_step, plain_work) if os.path.exists(a.learned) else None
    out = {
        "metric": "events/s on 1 MI355X at PU200 density: plain vs learned, >=99.5% TC recall",
        "value": round(B * a.steps / wall, 3),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"workload {a.workload}: PU200-density event(s), full tilted barrel + endcap, "
                               f"~{int(off[-1] / B)} hits/event, plain module-map chain",
                   "events_per_step": B, "hits_per_step": int(off[-1])},
        "gpu_ms_per_step": round(gpu_ms / a.steps, 4),
        "timed_calls": timed_calls,
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "objects_per_step": counts,
        "roofline": roofline,
        "chain_roofline": chain,
    }
    if learned is not None and "ms_per_step" in learned:
        learned["chain_roofline"] = chain_roofline(a.workload, B, learned["ms_per_step"],
                                                   scored=learned["tl_work_learned"]["scored"])
    if learned is not None:
        out["learned"] = learned
    if a.learned_batch > 0 and os.path.exists(a.learned):
        lb = plain_vs_learned(a, geom, a.workload, a.learned_batch, a.learned_batch, 3)
        lb["workload"] = f"workload {a.workload}: {a.learned_batch} events as one batch per step"
        out["learned_batched"] = lb
    if a.scaling.strip() not in ("", "0"):
        out["batch_scaling"] = scaling_leg(a, geom)
    if a.batch > 0:
        out["batched"] = batch_leg(a, runtime.Geometry(*gen.geometry(gen.geom_kind(a.batch_workload))[1]))
    if a.config1_batch > 0:
        out["config1_learned"] = config1_leg(a, runtime.Geometry(*gen.geometry(gen.geom_kind(1))[1]))
    if a.families:
        out["families"] = families_leg(a, geom)
    if a.hard_mix > 0 and os.path.exists(a.learned):
        out["hard_mix"] = hard_mix_leg(a, runtime.Geometry(*gen.geometry(gen.geom_kind(5))[1]))
    if a.cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(a.workload, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

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
PEAK_FP32_TFLOPS = 157.3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--events-per-step", type=int, default=1)
    p.add_argument("--workload", type=int, default=2)
    p.add_argument("--cpu-baseline", type=int, default=1, help="0 skips the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (at most 16)")
    return p.parse_args()


def cpu_baseline(workload, threads):
    """The CPU oracle (the C++ restatement of the reference, port) timed on a
    bounded sample: one event per thread, all threads busy at once."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from sdl_amd import gen
    n = threads
    evs = [gen.event(workload, 1000 + i) for i in range(n)]
    paths = gen.geometry(gen.geom_kind(workload))[1]
    wall = oracle.time_events(paths, evs, oracle.GLIBC, threads)
    return {"value": n / wall, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"{n} workload-{workload} events (seeds 1000..{1000 + n - 1}), one per thread, "
                      f"full chain, glibc libm, {wall:.1f} s wall"}


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
    ev.set_timing(True)
    ptrs = [d.address for d in dev]
    stream = ev.stream

    def step():
        ev.set_hits_device(off, *ptrs)
        ev.run_chain()

    for _ in range(a.warmup):
        step()
    ev.sync()
    hip.synchronize()
    kms = {k: 0.0 for k in runtime.TIMERS}
    start, stop = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(a.steps):
        step()
        for k, v in ev.kernel_ms().items():
            kms[k] += v
    stop.record(stream)
    ev.sync()
    hip.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop)
    ms_per_step = 1000.0 * wall / a.steps
    counts = ev.counts(-1)
    for k in kms:
        kms[k] /= a.steps

    # Dominant kernel: the tracklet kernel.  Algorithmic bytes per launch: every
    # inner SG record (8 B ids + 48 B floats) read once and every TL written once
    # (8 B ids + 24 B floats), the record sizes of the design notes.
    tl_bytes = counts["nSG"] * 56 + counts["nTL"] * 32
    tl_s = kms["tl"] / 1000.0
    achieved = tl_bytes / tl_s / 1e9 if tl_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 6), "traffic": None,
                "kernel": "k_tl", "kernel_ms": round(kms["tl"], 4), "tl_candidates_per_launch": counts["nTLCand"]}

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
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "objects_per_step": counts,
        "roofline": roofline,
    }
    if a.cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(a.workload, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Writes digests of the oracle's results (CR libm policy, the device's) for
runs too slow for the oracle inside a GPU test: the layer-pair chain
(createMiniDoublets, createSegments, createTracklets, createTrackCandidates;
Event.cc:432-443, 599-609, 823-837) on a PU200-density event pairs every
segment of an inner layer with every segment of an outer layer (2.5e9 tracklet
candidates), minutes of CPU time.  The oracle is pinned to the compiled
reference on smaller events (test_oracle.py) and at PU200 by the module-map
digest; its layer-pair chain is pinned by the layer fixtures.

Each digest: object counts, the ten counters, the per-layer counters and a
SHA-256 of each sorted object set (make_golden.set_digest).

Usage:  python tests/golden/make_oracle_digests.py   ->  tests/golden/oracle_pu200_digest.json
"""
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402
from sdl_amd import gen  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
_mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mg)

# name -> (workload, event, oracle stage list or LAYER_CHAIN)
RUNS = {"layers_w2_e0": (2, 0, oracle.LAYER_CHAIN)}


def main():
    out = {}
    for name, (wl, e, stages) in RUNS.items():
        t0 = time.time()
        sets, counters, layers = oracle.run(gen.geometry(gen.geom_kind(wl))[1], gen.event(wl, e), oracle.CR, stages,
                                            layers=True)
        out[name] = {"workload": wl, "event": e, "stages": stages,
                     "counts": {s: int(len(sets[s])) for s in oracle.STAGES},
                     "sha256": {s: _mg.set_digest(sets[s]) for s in oracle.STAGES},
                     "counters": [int(counters[k]) for k in _mg.COUNTERS], "layers": layers.tolist(),
                     "oracle_seconds": round(time.time() - t0, 1)}
        print(name, out[name]["counts"], out[name]["oracle_seconds"], flush=True)
    with open(os.path.join(HERE, "oracle_pu200_digest.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

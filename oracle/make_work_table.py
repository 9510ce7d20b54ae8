"""Writes bench_work.json: the algorithmic work of each stage on the benchmark's
events, counted by the op-counting build of the oracle (oracle/opcount.h).

    python oracle/make_work_table.py [--workload 2] [--events 8] [--jobs 8]

Each event runs in its own process and geometry directory.  The table holds,
per (workload, event): floating-point operations per stage along the path each
candidate takes (the reference's loops, Event.cc:245-1062), candidate counts and
object counts.  bench.py divides the tracklet stage's figure by the measured
duration of the tracklet kernel."""
import argparse
import json
import os
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(HERE, "_build", "oracle_opcount")


def one(workload, event):
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([EXE, str(workload), d, str(event), "1"], capture_output=True, text=True, check=True)
    return json.loads(out.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--events", type=int, default=8)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    subprocess.run(["make", "-C", HERE, "opcount"], check=True, capture_output=True)
    with ThreadPoolExecutor(a.jobs) as ex:
        rows = list(ex.map(lambda e: one(a.workload, e), range(a.events)))
    path = os.path.join(ROOT, "bench_work.json")
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table["convention"] = ("floating-point operations along each candidate's path in the reference's loops: "
                           "add/sub/mul/div/compare/abs/copysign = 1, sqrt/sin/cos/tan/asin/atan/atan2 = 1 "
                           "(float or double), casts/negation/selects = 0; counted by oracle/opcount.cc")
    table["command"] = f"python oracle/make_work_table.py --workload {a.workload} --events {a.events}"
    ev = table.setdefault("events", {})
    for r in rows:
        ev[f"{r['workload']}:{r['event']}"] = {k: r[k] for k in ("hits", "flops", "candidates", "objects")}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"wrote {path} ({len(rows)} events)")


if __name__ == "__main__":
    main()

"""Lists every object on which the two libm policies of the oracle disagree.

The compiled reference evaluates float asin/atan/sin/cos/tan with glibc's float
functions; the device evaluates them in double and rounds once.  The oracle
implements both (GLIBC is pinned to the compiled reference, CR is what the
device computes bit for bit), so the objects that differ between the two
policies are exactly the objects on which the device and the reference may
disagree.  For each such object this script records the cut that decided it,
the cut variable and threshold under both policies, and the distance between
variable and threshold in float ulps.  Objects whose inputs already differ
(an SG built on a differing MD, a TL on a differing SG, and every TP / TC built
on those) are consequences and are counted, not listed.

    python tests/golden/list_libm_differences.py [--quick]

writes tests/golden/libm_differences.json."""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from sdl_amd import gen  # noqa: E402

FULL = [(1, e) for e in range(8)] + [(12, e) for e in range(8)] + [(15, e) for e in range(4)] + [(2, 0), (2, 1)]
QUICK = [(1, 0), (12, 0), (15, 0)]
CUT_NAMES = {0: "MD barrel dz", 1: "MD barrel dPhi", 2: "MD barrel dPhiChange", 10: "MD endcap dz",
             11: "MD endcap drt", 12: "MD endcap dPhi", 13: "MD endcap dPhiChange",
             20: "SG barrel z window", 21: "SG barrel dPhi", 22: "SG barrel dPhiChange",
             24: "SG dAlpha inner", 25: "SG dAlpha outer", 26: "SG dAlpha inner-outer",
             30: "SG endcap map", 31: "SG endcap z sign", 32: "SG endcap rt window", 33: "SG endcap dPhi",
             34: "SG endcap dPhiChange"}


def differences(wl, e, policy=oracle.CR):
    """Objects that differ between the GLIBC policy and `policy` (CR: the
    device's arithmetic; the tests also pass a planted error)."""
    paths = gen.geometry(gen.geom_kind(wl))[1]
    ev = gen.event(wl, e)
    a, _ = oracle.run(paths, ev, oracle.GLIBC)
    b, _ = oracle.run(paths, ev, policy)
    listed, consequences = [], {}
    common = {}
    for s, kind in (("MD", 0), ("SG", 1), ("TL", 2)):
        sa = {tuple(int(v) for v in r) for r in a[s]}
        sb = {tuple(int(v) for v in r) for r in b[s]}
        common[s] = sa & sb
        for t in sorted(sa ^ sb):
            if kind == 1 and not (t[:2] in common["MD"] and t[2:] in common["MD"]):
                consequences[s] = consequences.get(s, 0) + 1
                continue
            if kind == 2 and not (t[:4] in common["SG"] and t[4:] in common["SG"]):
                consequences[s] = consequences.get(s, 0) + 1
                continue
            info = oracle.explain(paths, ev, kind, t[:2] if kind == 0 else (t if kind == 1 else t[:4]),
                                  t[4:] if kind == 2 else None, policy)
            cut = int(info["cut_glibc"]) if info["cut_glibc"] >= 0 else int(info["cut_cr"])
            listed.append({"workload": wl, "event": e, "stage": s, "hits": list(t),
                           "in": "glibc" if t in sa else "cr", "cut": cut, "cut_name": CUT_NAMES.get(cut, f"TL cut {cut}"),
                           "var_glibc": info["var"], "var_cr": info["var_cr"],
                           "threshold_glibc": [info["lo"], info["hi"]], "threshold_cr": [info["lo_cr"], info["hi_cr"]],
                           "ulps": info["ulps"]})
    for s in ("TP", "TCP", "TCL"):
        n = len({tuple(r) for r in a[s]} ^ {tuple(r) for r in b[s]})
        if n:
            consequences[s] = n
    return listed, consequences


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args()
    cases = QUICK if args.quick else FULL
    out = {"cases": [f"{w}:{e}" for w, e in cases], "differences": [], "consequences": {}}
    for wl, e in cases:
        listed, cons = differences(wl, e)
        out["differences"] += listed
        if cons:
            out["consequences"][f"{wl}:{e}"] = cons
        print(f"workload {wl} event {e}: {len(listed)} listed, consequences {cons}", flush=True)
    out["max_ulps"] = max([d["ulps"] for d in out["differences"]], default=0.0)
    path = os.path.join(HERE, "libm_differences.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {len(out['differences'])} objects, max {out['max_ulps']} ulps")


if __name__ == "__main__":
    main()

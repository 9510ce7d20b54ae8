"""Writes the golden fixtures of the SDL chain from the compiled reference.

Runs in the build container only (it needs /root/reference): builds
oracle/_ref (the unmodified reference sources + this repository's driver,
oracle/ref_driver.cc) through oracle/Makefile, runs the reference's module-map
chain on seeded generated events and stores every object set as hit-index
tuples plus the stage counters in ``tests/golden/ref_sets.npz``; and the
layer-pair chain (createMiniDoublets, createSegments, createTracklets,
createTrackCandidates) with the per-layer counters in
``tests/golden/ref_layer_sets.npz``; and createMiniDoublets,
createSegmentsWithModuleMap, createTrackletsViaNavigation in
``tests/golden/ref_nav_sets.npz``; and createPseudoMiniDoubletsFromAnchorModule
followed by the module-map chain in ``tests/golden/ref_pseudo_sets.npz``; and
the per-object values of every MD, SG and TL of the module-map chain (typed
getters and reco variables, as float bit patterns) in
``tests/golden/ref_vars.npz``.

and stage sequences with other algorithms (AllComb, DefaultNm1; the
reference driver's dumpseq mode) in ``tests/golden/ref_seq_sets.npz``; and a
digest of a PU200-density event's module-map chain (object counts, counters,
per-layer counters, SHA-256 of each sorted object set) in
``tests/golden/ref_pu200_digest.json``.

Usage:  python tests/golden/make_golden.py [layers|nav|pseudo|vars|nm1|seq|digest]   (only those fixtures)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

# (workload, first event, number of events)
CASES = [(1, 0, 2), (12, 0, 2), (15, 0, 1)]
LAYER_CASES = [(1, 0, 2), (12, 0, 1)]
NAV_CASES = [(1, 0, 2), (12, 0, 2), (15, 0, 1)]
PSEUDO_CASES = [(1, 0, 2), (12, 0, 1), (15, 0, 1)]
VAR_CASES = [(1, 0, 2), (12, 0, 1), (15, 0, 1)]
VAR_WIDTH = {"MV": (2, 7), "SV": (4, 22), "TV": (8, 114), "CV": (16, 2)}  # tuple width, values per object
COUNTERS = ("nMDCand", "nSGCand", "nTPCand", "nTLCand", "nTCCand", "nMD", "nSG", "nTP", "nTL", "nTC")


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which == "all":
        module_map_fixtures()
    if which in ("all", "layers"):
        layer_fixtures()
    if which in ("all", "nav"):
        nav_fixtures()
    if which in ("all", "pseudo"):
        pseudo_fixtures()
    if which in ("all", "vars"):
        var_fixtures()
    if which in ("all", "nm1"):
        nm1_fixtures()
    if which in ("all", "seq"):
        seq_fixtures()
    if which in ("all", "digest"):
        digest_fixture()


def sorted_rows(keys, bits):
    """Rows (tuple, value bits) in lexicographic order: a canonical multiset."""
    rows = np.concatenate([keys.astype(np.int64), bits.astype(np.int64)], axis=1)
    if len(rows) == 0:
        return rows
    return rows[np.lexsort(rows.T[::-1])]


def parse_var_dump(path):
    """{event: {"MV"|"SV"|"TV": (keys int32 [n, w], bits uint32 [n, v])}} of a dumpvars file."""
    out, cur = {}, None
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "E":
            cur = out.setdefault(int(t[1]), {k: [] for k in VAR_WIDTH})
            continue
        w, nv = VAR_WIDTH[t[0]]
        cur[t[0]].append(([int(x) for x in t[1:1 + w]], [int(x, 16) for x in t[1 + w:1 + w + nv]]))
    res = {}
    for e, d in out.items():
        res[e] = {}
        for k, rows in d.items():
            w, nv = VAR_WIDTH[k]
            keys = np.array([r[0] for r in rows], np.int32).reshape(-1, w)
            bits = np.array([r[1] for r in rows], np.uint32).reshape(-1, nv)
            res[e][k] = (keys, bits)
    return res


def var_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in VAR_CASES:
            dump = os.path.join(tmp, f"v{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumpvars", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, d in parse_var_dump(dump).items():
                for k, (keys, bits) in d.items():
                    out[f"w{wl}_e{e}_{k}"] = sorted_rows(keys, bits).astype(np.uint32)  # indices >= 0, bits
    np.savez_compressed(os.path.join(HERE, "ref_vars.npz"), **out)


def module_map_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in CASES:
            dump = os.path.join(tmp, f"w{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dump", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, (sets, counters) in oracle.parse_ref_dump(dump).items():
                for s in oracle.STAGES:
                    out[f"w{wl}_e{e}_{s}"] = oracle.as_multiset(sets[s])
                out[f"w{wl}_e{e}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_sets.npz"), **out)


def layer_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in LAYER_CASES:
            dump = os.path.join(tmp, f"w{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumplayers", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, (sets, counters) in oracle.parse_ref_dump(dump).items():
                for s in ("MD", "SG", "TL", "TCL"):
                    out[f"w{wl}_e{e}_{s}"] = oracle.as_multiset(sets[s])
                out[f"w{wl}_e{e}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
                out[f"w{wl}_e{e}_layers"] = counters["layers"]
    np.savez_compressed(os.path.join(HERE, "ref_layer_sets.npz"), **out)


def nav_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in NAV_CASES:
            dump = os.path.join(tmp, f"w{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumpnav", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, (sets, counters) in oracle.parse_ref_dump(dump).items():
                for s in ("MD", "SG", "TL"):
                    out[f"w{wl}_e{e}_{s}"] = oracle.as_multiset(sets[s])
                out[f"w{wl}_e{e}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
                out[f"w{wl}_e{e}_layers"] = counters["layers"]
    np.savez_compressed(os.path.join(HERE, "ref_nav_sets.npz"), **out)


NM1_CASES = [(1, 0, 1), (15, 0, 1)]


def nm1_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in NM1_CASES:
            dump = os.path.join(tmp, f"w{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumpnm1", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, (sets, counters) in oracle.parse_ref_dump(dump).items():
                for s in oracle.STAGES:
                    out[f"w{wl}_e{e}_{s}"] = oracle.as_multiset(sets[s])
                out[f"w{wl}_e{e}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_nm1_sets.npz"), **out)


# Stage sequences with the Algo.h algorithms (sdl_ref_driver dumpseq): name ->
# (workload, event, driver sequence, oracle stage codes: stage + 100 * algo).
SEQS = {
    "md_allcomb": (11, 0, "md:1,sg:0,tp:0,tl:0,tcp:0,tcl:0", [100, 1, 2, 3, 4, 5]),
    "sg_allcomb": (11, 0, "md:0,sg:1,tp:0,tl:0,tcp:0,tcl:0", [0, 101, 2, 3, 4, 5]),
    "tp_allcomb": (11, 0, "md:0,sg:0,tp:1,tl:0,tcp:0,tcl:0", [0, 1, 102, 3, 4, 5]),
    "tl_allcomb": (11, 0, "md:0,sg:0,tp:0,tl:2,tcp:0,tcl:0", [0, 1, 2, 203, 4, 5]),
    "tc_allcomb": (11, 0, "md:0,sg:0,tp:0,tl:0,tcp:1,tcl:1", [0, 1, 2, 3, 104, 105]),
    "all_allcomb": (11, 0, "md:1,sg:1,tp:1,tl:2,tcp:0,tcl:1", [100, 101, 102, 203, 4, 105]),
    "layers_sg_allcomb": (11, 0, "md:0,sgl:1,tll:0,tcll:0", [0, 106, 7, 8]),
    "layers_tl_allcomb": (11, 0, "md:0,sgl:0,tll:2,tcll:0", [0, 6, 207, 8]),
    "layers_tc_allcomb": (11, 1, "md:0,sgl:0,tll:0,tcll:1", [0, 6, 7, 108]),
    "nav_allcomb": (11, 0, "md:0,sg:0,tln:2,tcl:1", [0, 1, 209, 105]),
    "nm1_chain": (1, 0, "md:0,sg:0,tp:0,tl:1,tcp:0,tcl:0", [0, 1, 2, 103, 4, 5]),
    "nm1_layers": (1, 0, "md:0,sgl:0,tll:1,tcll:0", [0, 6, 107, 8]),
    "nm1_nav": (1, 0, "md:0,sg:0,tln:1,tcl:0", [0, 1, 109, 5]),
}


def seq_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (wl, e, spec, _) in SEQS.items():
            dump = os.path.join(tmp, f"{name}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumpseq", str(wl), tmp, str(e), "1", dump, spec], check=True)
            (sets, counters), = oracle.parse_ref_dump(dump).values()
            for s in oracle.STAGES:
                out[f"{name}_{s}"] = oracle.as_multiset(sets[s])
            out[f"{name}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
            out[f"{name}_layers"] = counters["layers"]
    np.savez_compressed(os.path.join(HERE, "ref_seq_sets.npz"), **out)


# PU200-density events (workload 2: 8000 tracks + 20 hard, full geometry),
# too large to store as tuples: the module-map chain's object counts, counters,
# per-layer counters and a SHA-256 of each sorted object set.
DIGEST_CASES = [(2, 0)]
DIGEST_SEQ = "md:0,sg:0,tp:0,tl:0,tcp:0,tcl:0"


def set_digest(rows):
    """SHA-256 of a sorted object set: its rows in lexicographic order as
    little-endian int32, row after row."""
    import hashlib
    a = np.ascontiguousarray(oracle.as_multiset(np.asarray(rows, np.int32)), dtype="<i4")
    return hashlib.sha256(a.tobytes()).hexdigest()


def digest_fixture():
    import json
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e in DIGEST_CASES:
            dump = os.path.join(tmp, f"d{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumpseq", str(wl), tmp, str(e), "1", dump, DIGEST_SEQ], check=True)
            (sets, counters), = oracle.parse_ref_dump(dump).values()
            out[f"w{wl}_e{e}"] = {
                "counts": {s: int(len(sets[s])) for s in oracle.STAGES},
                "sha256": {s: set_digest(sets[s]) for s in oracle.STAGES},
                "counters": [int(counters[k]) for k in COUNTERS],
                "layers": counters["layers"].tolist(),
            }
    with open(os.path.join(HERE, "ref_pu200_digest.json"), "w") as f:
        json.dump(out, f, indent=1)


def pseudo_fixtures():
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for wl, e0, ne in PSEUDO_CASES:
            dump = os.path.join(tmp, f"w{wl}.txt")
            subprocess.run([oracle.REF_DRIVER, "dumppseudo", str(wl), tmp, str(e0), str(ne), dump], check=True)
            for e, (sets, counters) in oracle.parse_ref_dump(dump).items():
                for s in oracle.STAGES:
                    out[f"w{wl}_e{e}_{s}"] = oracle.as_multiset(sets[s])
                out[f"w{wl}_e{e}_counters"] = np.array([counters[k] for k in COUNTERS], np.int64)
                out[f"w{wl}_e{e}_layers"] = counters["layers"]
    np.savez_compressed(os.path.join(HERE, "ref_pseudo_sets.npz"), **out)


if __name__ == "__main__":
    main()

"""The CPU oracle against the compiled reference's golden object sets, and
the libm policy comparison (CPU only)."""
import os

import numpy as np
import pytest

import oracle
from sdl_amd import gen

CASES = [(1, 0), (1, 1), (12, 0), (12, 1), (15, 0)]
COUNTERS = ("nMDCand", "nSGCand", "nTPCand", "nTLCand", "nTCCand", "nMD", "nSG", "nTP", "nTL", "nTC")


@pytest.mark.parametrize("wl,e", CASES)
def test_oracle_matches_reference(wl, e, geom_paths, golden):
    ev = gen.event(wl, e)
    sets, counters = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC)
    for s in oracle.STAGES:
        ref = golden[f"w{wl}_e{e}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape, (s, got.shape, ref.shape)
        assert np.array_equal(got, ref), s
        assert len(ref) > 0, f"{s}: empty golden set makes the check vacuous"
    assert [counters[k] for k in COUNTERS] == list(golden[f"w{wl}_e{e}_counters"])


@pytest.mark.parametrize("wl,e", CASES)
def test_libm_policies(wl, e, geom_paths):
    """Correctly rounded float transcendentals (the device's) against glibc's
    (the reference's): any object that differs must have its deciding cut
    variable within 4 FP32 ulp of the threshold."""
    paths = geom_paths[gen.geom_kind(wl)]
    ev = gen.event(wl, e)
    a, _ = oracle.run(paths, ev, oracle.GLIBC)
    b, _ = oracle.run(paths, ev, oracle.CR)
    common = {}
    for s, kind in (("MD", 0), ("SG", 1), ("TL", 2)):
        sa = {tuple(r) for r in a[s]}
        sb = {tuple(r) for r in b[s]}
        common[s] = sa & sb
        for t in sa ^ sb:
            # an object whose inputs already differ is a consequence, not a new difference
            if kind == 1 and not (t[:2] in common["MD"] and t[2:] in common["MD"]):
                continue
            if kind == 2 and not (t[:4] in common["SG"] and t[4:] in common["SG"]):
                continue
            if kind == 0:
                info = oracle.explain(paths, ev, 0, t)
            elif kind == 1:
                info = oracle.explain(paths, ev, 1, t)
            else:
                info = oracle.explain(paths, ev, 2, t[:4], t[4:])
            assert info["ulps"] <= 4.0, (s, t, info)


def test_generator_is_deterministic():
    a = gen.event(12, 3)
    b = gen.event(12, 3)
    c = gen.event(12, 4)
    for f in ("x", "y", "z", "detId", "idx", "sim"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    assert not np.array_equal(a.x[:100], c.x[:100])
    assert np.array_equal(a.idx, np.arange(a.nhits))


def test_pow_square_is_exact():
    """pow(float, 2) in double equals the exact product, which the device uses."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.pow.restype = ctypes.c_double
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(2000).astype(np.float32) * 10 ** rng.uniform(-8, 4, 2000).astype(np.float32),
                         np.float32([0.0, -0.0, 1e-30, 3.4e18])])
    for x in xs:
        d = float(x)
        assert libm.pow(d, 2.0) == d * d


def test_libm_listing_detects_planted_error():
    """The difference listing finds and explains objects when the second policy
    carries a planted error (float transcendentals 0.4 % too large), and those
    objects would fail the 4-ulp rule: the listing and its rule are not vacuous."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import list_libm_differences as L
    listed, consequences = L.differences(12, 0, oracle.NUDGED)
    assert len(listed) > 0
    assert all(d["cut"] >= 0 for d in listed)
    assert max(d["ulps"] for d in listed) > 4.0


def test_libm_listing_committed():
    """The committed listing covers the PU200 events and lists no object: on
    those events the device's arithmetic (CR) and the reference's (glibc) give
    identical object sets."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "libm_differences.json")) as f:
        d = json.load(f)
    assert "2:0" in d["cases"] and "2:1" in d["cases"]
    assert all(x["ulps"] <= 4.0 for x in d["differences"])
    assert d["differences"] == []


LAYER_CASES = [(1, 0), (1, 1), (12, 0)]


@pytest.fixture(scope="module")
def layer_golden():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_layer_sets.npz")
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("wl,e", LAYER_CASES)
def test_oracle_layer_chain_matches_reference(wl, e, geom_paths, layer_golden):
    """The layer-pair chain (createMiniDoublets, createSegments, createTracklets,
    createTrackCandidates) against the compiled reference: object sets and the
    ten totals (the object totals are list sizes, the candidate totals sums of
    the per-layer counters, Event.cc:1080-1112)."""
    ev = gen.event(wl, e)
    sets, counters, lay = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, oracle.LAYER_CHAIN, layers=True)
    assert np.array_equal(lay, layer_golden[f"w{wl}_e{e}_layers"])
    for s in ("MD", "SG", "TL", "TCL"):
        ref = layer_golden[f"w{wl}_e{e}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape, (s, got.shape, ref.shape)
        assert np.array_equal(got, ref), s
        assert len(ref) > 0, f"{s}: empty golden set makes the check vacuous"
    assert len(sets["TP"]) == 0 and len(sets["TCP"]) == 0
    assert [counters[k] for k in COUNTERS] == list(layer_golden[f"w{wl}_e{e}_counters"])
    # the per-layer counters of the reference: candidates and objects of the
    # layer-pair tracklets and track candidates sit on the inner layer only
    layers = layer_golden[f"w{wl}_e{e}_layers"]
    assert layers[1].sum() == 0 and layers[6].sum() == 0  # createSegmentsFromTwoLayers counts nothing
    assert layers[3][3:].sum() == 0 and layers[4][1:].sum() == 0 and layers[9][1:].sum() == 0


def test_oracle_layer_chain_closed_forms(geom_paths):
    """Invariants of the layer-pair chain that fix it independently of the
    reference: the TL candidates are the products of the layer SG counts of the
    three layer pairs, the TC candidates |TL(barrel 1)| x |TL(barrel 3)|, and a
    TC exists for exactly the (TL, TL) pairs sharing the middle segment."""
    wl, e = 1, 0
    ev = gen.event(wl, e)
    sets, counters = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, oracle.LAYER_CHAIN)
    layer_of_hit = {}
    for i, d in zip(ev.idx, ev.detId):
        layer_of_hit[int(i)] = ((int(d) >> 20) & 7) - 1  # barrel-only workload: barrel layer bins
    sg_layer = np.array([layer_of_hit[int(r[0])] for r in sets["SG"]])
    n_sg = np.bincount(sg_layer, minlength=6)
    assert counters["nTLCand"] == int(n_sg[0] * n_sg[2] + n_sg[1] * n_sg[3] + n_sg[2] * n_sg[4])
    tl = sets["TL"]
    tl_layer = np.array([layer_of_hit[int(r[0])] for r in tl])
    assert counters["nTCCand"] == int((tl_layer == 0).sum() * (tl_layer == 2).sum())
    inner = [tuple(r) for r, L in zip(tl, tl_layer) if L == 0]
    outer = [tuple(r) for r, L in zip(tl, tl_layer) if L == 2]
    by_first = {}
    for o in outer:
        by_first.setdefault(o[:4], []).append(o)
    expect = sorted(i + o for i in inner for o in by_first.get(i[4:], []))
    got = sorted(tuple(r) for r in sets["TCL"])
    assert got == expect and len(got) > 0


NAV_CASES = [(1, 0), (1, 1), (12, 0), (12, 1), (15, 0)]


@pytest.mark.parametrize("wl,e", [(1, 0), (15, 0)])
def test_oracle_nm1_tracklets_match_reference(wl, e, geom_paths):
    """createTrackletsWithModuleMap(DefaultNm1_TLAlgo) in the module-map chain,
    against the compiled reference (ref_nm1_sets.npz): all six object sets and
    the counters; the waived dBeta cut makes the TL set a strict superset of
    the default one."""
    import os
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_nm1_sets.npz")) as f:
        golden = {k: f[k] for k in f.files}
    ev = gen.event(wl, e)
    paths = geom_paths[gen.geom_kind(wl)]
    sets, counters = oracle.run(paths, ev, oracle.GLIBC, [0, 1, 2, 11, 4, 5])
    for s in oracle.STAGES:
        ref = golden[f"w{wl}_e{e}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape and np.array_equal(got, ref), s
    assert [counters[k] for k in COUNTERS] == list(golden[f"w{wl}_e{e}_counters"])
    plain, _ = oracle.run(paths, ev, oracle.GLIBC, 6)
    nm1 = {tuple(r) for r in sets["TL"]}
    assert {tuple(r) for r in plain["TL"]} < nm1


@pytest.mark.parametrize("wl,e", NAV_CASES)
def test_oracle_navigation_matches_reference(wl, e, geom_paths):
    """createTrackletsViaNavigation after the module-map segments, against the
    compiled reference (ref_nav_sets.npz): object sets and counters."""
    import os
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_nav_sets.npz")) as f:
        golden = {k: f[k] for k in f.files}
    ev = gen.event(wl, e)
    sets, counters, lay = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, [0, 1, 9], layers=True)
    assert np.array_equal(lay, golden[f"w{wl}_e{e}_layers"])
    for s in ("MD", "SG", "TL"):
        ref = golden[f"w{wl}_e{e}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape and np.array_equal(got, ref), s
    assert len(golden[f"w{wl}_e{e}_TL"]) > 0
    assert [counters[k] for k in COUNTERS] == list(golden[f"w{wl}_e{e}_counters"])


def test_oracle_navigation_closed_form(geom_paths):
    """Navigation candidates by brute force over the SG set: (s1, s3) for every
    SG s2 that starts at s1's outer MD and ends at s3's inner MD; the tracklets
    are the candidates that pass, so the module-map TL set (connected modules)
    and the navigation set agree on every candidate both consider."""
    wl, e = 1, 0
    ev = gen.event(wl, e)
    paths = geom_paths[gen.geom_kind(wl)]
    nav, cn = oracle.run(paths, ev, oracle.GLIBC, [0, 1, 9])
    sg = [tuple(r) for r in nav["SG"]]
    starts = {}
    for r in sg:
        starts.setdefault(r[:2], []).append(r)
    cand = set()
    for s1 in sg:
        for s2 in starts.get(s1[2:], []):
            for s3 in starts.get(s2[2:], []):
                cand.add(s1 + s3)
    assert cn["nTLCand"] == sum(len(starts.get(s2[2:], [])) for s1 in sg for s2 in starts.get(s1[2:], []))
    tl = {tuple(r) for r in nav["TL"]}
    assert tl <= cand and len(tl) > 0
    mm, _ = oracle.run(paths, ev, oracle.GLIBC, [0, 1, 2, 3])
    both = {tuple(r) for r in mm["TL"]} & cand
    assert both <= tl


PSEUDO_CASES = [(1, 0), (1, 1), (12, 0), (15, 0)]


@pytest.mark.parametrize("wl,e", PSEUDO_CASES)
def test_oracle_pseudo_chain_matches_reference(wl, e, geom_paths):
    """createPseudoMiniDoubletsFromAnchorModule followed by the module-map chain
    (whose algorithms then meet zero cached values and the 0/0 they produce)
    against the compiled reference (ref_pseudo_sets.npz)."""
    import os
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_pseudo_sets.npz")) as f:
        golden = {k: f[k] for k in f.files}
    ev = gen.event(wl, e)
    sets, counters, lay = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, [10, 1, 2, 3, 4, 5],
                                     layers=True)
    assert np.array_equal(lay, golden[f"w{wl}_e{e}_layers"])
    for s in oracle.STAGES:
        ref = golden[f"w{wl}_e{e}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape and np.array_equal(got, ref), s
    assert [counters[k] for k in COUNTERS] == list(golden[f"w{wl}_e{e}_counters"])


def test_oracle_pseudo_md_closed_form(geom_paths):
    """Pseudo MDs by brute force, without the module decoding: per module pair,
    either every hit of one sensor pairs with the first hit (input order) of
    the other, or the reverse; nothing is rejected."""
    wl, e = 12, 0
    ev = gen.event(wl, e)
    sets, counters = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, [10])
    det_of = {int(i): int(d) for i, d in zip(ev.idx, ev.detId)}
    hits_of = {}
    for i, d in zip(ev.idx, ev.detId):
        hits_of.setdefault(int(d), []).append(int(i))
    groups = {}
    for lo, up in sets["MD"]:
        groups.setdefault((det_of[int(lo)], det_of[int(up)]), []).append((int(lo), int(up)))
    assert len(groups) > 0
    for (dlo, dup), mds in groups.items():
        los = sorted({m[0] for m in mds})
        ups = sorted({m[1] for m in mds})
        if len(ups) == 1 and los == sorted(hits_of[dlo]):
            assert ups[0] == hits_of[dup][0]  # every lower hit with the first upper hit
        else:
            assert len(los) == 1 and los[0] == hits_of[dlo][0] and ups == sorted(hits_of[dup])
        assert len(mds) == len(set(mds))
    assert counters["nMDCand"] == counters["nMD"] == len(sets["MD"])


VAR_CASES = [(1, 0), (1, 1), (12, 0), (15, 0)]
VAR_KEYS = (("MV", "MD"), ("SV", "SG"), ("TV", "TL"), ("CV", "TCL"))


def canonical_rows(keys, values):
    """(hit-index tuple, value bit patterns) rows in lexicographic order."""
    rows = np.concatenate([keys.astype(np.int64), np.ascontiguousarray(values, np.float32).view(np.uint32).astype(np.int64)],
                          axis=1)
    return rows[np.lexsort(rows.T[::-1])] if len(rows) else rows


@pytest.mark.parametrize("wl,e", VAR_CASES)
def test_object_vars_match_reference(wl, e, geom_paths):
    """Per-object values of every MD, SG and TL (the typed getters and reco
    variables of MiniDoublet.h, Segment.h and Tracklet.h, as the reference's
    default algorithms leave them on stored objects, and the two reco variables
    of a TC from tracklets): the oracle under the
    GLIBC policy reproduces the compiled reference's bit patterns, object by
    object (barrel-only, full-geometry and endcap-heavy hard-mix events: every
    SG and TL variant)."""
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vars.npz")) as f:
        golden = {k: f[k] for k in f.files}
    ev = gen.event(wl, e)
    sets, _, values = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, stages=6, vars=True)
    for k, s in VAR_KEYS:
        ref = golden[f"w{wl}_e{e}_{k}"]
        assert len(ref) > 0, f"{k}: empty golden set makes the check vacuous"
        mine = canonical_rows(sets[s], values[s])
        assert mine.shape == ref.shape, (k, mine.shape, ref.shape)
        bad = np.flatnonzero((mine != ref).any(axis=1))
        assert len(bad) == 0, f"{k}: {len(bad)} objects differ, first oracle {mine[bad[0]]} reference {ref[bad[0]]}"


def test_object_vars_layout_and_defaults(geom_paths):
    """What the values must be by construction: the typed fields a variant
    never sets stay at the constructor's 0 (barrel SGs: rt window and the
    deltaPhi min/max; endcap SGs: z window), reco variables that are never set
    on a path are -999 (endcap sdMuls, sdPVoff), BBEE tracklets set only
    deltaBeta, and the values that are plain hit coordinates equal them."""
    wl, e = 15, 0
    ev = gen.event(wl, e)
    sets, _, v = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, stages=4, vars=True)
    names = oracle.VAR_NAMES
    assert v["MD"].shape[1] == len(names["MD"]) and v["SG"].shape[1] == len(names["SG"])
    assert v["TL"].shape[1] == len(names["TL"])
    sg = v["SG"]
    col = {n: i for i, n in enumerate(names["SG"])}
    endcap = sg[:, col["sdMuls"]] == -999.0
    barrel = ~endcap
    assert endcap.any() and barrel.any()
    assert (sg[barrel][:, [col["rtLo"], col["rtHi"], col["deltaPhi"], col["deltaPhiMin"], col["deltaPhiMax"]]] == 0).all()
    assert (sg[endcap][:, [col["zIn"], col["zOut"], col["zLo"], col["zHi"]]] == 0).all()
    assert (sg[endcap][:, col["sdPVoff"]] == -999.0).all()
    assert (sg[endcap][:, col["sdCut"]] == sg[endcap][:, col["sdSlope"]]).all()  # Segment.cc:740
    # barrel: zOut lies in [zLo, zHi] and |deltaPhi| <= sdCut (the cuts the SG passed)
    b = sg[barrel]
    assert ((b[:, col["zOut"]] >= b[:, col["zLo"]]) & (b[:, col["zOut"]] <= b[:, col["zHi"]])).all()
    assert (np.abs(b[:, col["reco_deltaPhi"]]) <= b[:, col["sdCut"]]).all()
    md = v["MD"]
    assert np.isfinite(md[:, 6]).all() and (np.abs(md[:, 2]) < md[:, 6]).all()  # |dphi| < miniCut
    tl = v["TL"]
    bbee = (tl[:, 0] == 0) & (tl[:, 1] == 0) & (tl[:, 5] == 0)
    assert bbee.any() and (tl[bbee][:, 4] != 0).all()
    assert (np.abs(tl[~bbee][:, 4]) <= tl[~bbee][:, 5]).all()  # |dBeta| <= sqrt(dBetaCut2)
    # the tracklet map (Tracklet.cc:152-296): "algo" names the variant; BBEE
    # writes only the delta-beta iteration keys; the ten _v2 keys are absent
    # elsewhere; mode 3 of the iterations changes nothing
    t = {n: tl[:, i] for i, n in enumerate(names["TL"])}
    algo = t["reco_algo"]
    assert set(np.unique(algo)) == {1.0, 2.0, 3.0}
    assert (algo[bbee] == 2).all() and (algo[~bbee] != 2).all()
    iter_keys = {n for n in names["TL"] if n.startswith("reco_") and
                 (n[5:].startswith(("betaIn_", "betaOut_", "betaAv_", "betaPt_", "dBeta_")) or n == "reco_betacormode")}
    iter_keys.discard("reco_dBeta_midPoint")
    iter_keys.discard("reco_betaIn_cut")
    for n in names["TL"][6:]:
        if n not in iter_keys and n != "reco_algo":
            col = t[n][algo == 2]
            assert ((col == -999) | np.isnan(col)).all(), n
    for n in ("alpha1GeV_OutLo", "rt_InLo", "z_OutLo", "zpitch_InLo"):
        assert np.isnan(t["reco_" + n][algo != 1]).all() and np.isfinite(t["reco_" + n][algo == 1]).all()
    ends = algo != 2
    assert (t["reco_dBeta"][ends] == t["deltaBeta"][ends]).all()
    assert (t["reco_dBeta"][ends] == (t["reco_betaIn"][ends] - t["reco_betaOut"][ends])).all()
    assert (t["reco_betacormode"][ends] == 0).all()  # the algorithms' own, never updated
    assert set(np.unique(t["reco_betacormode"][algo == 2])) <= {1.0, 2.0, 3.0}
    m3 = (algo == 2) & (t["reco_betacormode"] == 3)
    for k in ("betaIn", "betaOut", "betaAv", "betaPt", "dBeta"):
        assert (t[f"reco_{k}_4th"][m3] == t[f"reco_{k}_0th"][m3]).all()
    # the anchors' coordinates and modules are the tuple's hits'
    x = dict(zip(ev.idx.tolist(), ev.x.tolist()))
    det = dict(zip(ev.idx.tolist(), ev.detId.tolist()))
    tup = sets["TL"][ends]
    for j, (hx, dk) in enumerate((("hit1_x", "innerSgInnerMdDetId"), ("hit2_x", "innerSgOuterMdDetId"),
                                  ("hit3_x", "outerSgInnerMdDetId"), ("hit4_x", "outerSgOuterMdDetId"))):
        lo, up = tup[:, 2 * j], tup[:, 2 * j + 1]
        got = t["reco_" + hx][ends]
        assert all(g in (np.float32(x[a]), np.float32(x[b])) for g, a, b in zip(got, lo, up))
        got = t["reco_" + dk][ends]
        assert all(g in (np.float32(det[a]), np.float32(det[b])) for g, a, b in zip(got, lo, up))
    assert (np.abs(tl[~bbee][:, 2]) < tl[~bbee][:, 3]).all()   # |betaOut| < betaOut cut


def _seq_table():
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.SEQS


SEQS = _seq_table()


@pytest.mark.parametrize("name", sorted(SEQS))
def test_oracle_sequences_match_reference(name, geom_paths):
    """Stage sequences with the AllComb algorithms (MiniDoublet.cc:223-226,
    Segment.cc:313-335, Triplet.cc:45-48, Tracklet.cc:139-142,
    TrackCandidate.cc:138-141) and with DefaultNm1 tracklets in each tracklet
    stage, against the compiled reference (ref_seq_sets.npz, the reference
    driver's dumpseq mode): all six object sets and the counters."""
    import os
    wl, e, _, codes = SEQS[name]
    with np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_seq_sets.npz")) as f:
        golden = {k: f[k] for k in f.files if k.startswith(name + "_")}
    ev = gen.event(wl, e)
    sets, counters, lay = oracle.run(geom_paths[gen.geom_kind(wl)], ev, oracle.GLIBC, codes, layers=True)
    assert np.array_equal(lay, golden[f"{name}_layers"]), "per-layer counters"
    for s in oracle.STAGES:
        ref = golden[f"{name}_{s}"]
        got = oracle.as_multiset(sets[s])
        assert got.shape == ref.shape and np.array_equal(got, ref), s
    assert [counters[k] for k in COUNTERS] == list(golden[f"{name}_counters"])


def test_allcomb_accepts_every_candidate(geom_paths):
    """What holds by construction: with every stage AllComb the object counts
    equal the candidate counts, and an AllComb stage's objects contain the
    default stage's."""
    wl, e = 11, 0
    ev = gen.event(wl, e)
    paths = geom_paths[gen.geom_kind(wl)]
    sets, c = oracle.run(paths, ev, oracle.GLIBC, [100, 101, 102, 203, 105])
    assert (c["nMD"], c["nSG"], c["nTP"], c["nTL"], c["nTC"]) == (
        c["nMDCand"], c["nSGCand"], c["nTPCand"], c["nTLCand"], c["nTCCand"])
    plain, _ = oracle.run(paths, ev, oracle.GLIBC, 6)
    for codes, stage in (([100], "MD"), ([0, 101], "SG"), ([0, 1, 102], "TP"), ([0, 1, 2, 203], "TL")):
        ac, _ = oracle.run(paths, ev, oracle.GLIBC, codes)
        assert {tuple(r) for r in plain[stage]} <= {tuple(r) for r in ac[stage]}, stage


DECODE_CHECK = os.path.join(os.path.dirname(oracle.REF_DRIVER), "decode_check")


@pytest.mark.skipif(not os.path.exists(DECODE_CHECK), reason="oracle/_ref not built (needs the reference sources)")
def test_detid_decoding_matches_reference():
    """Module detId decoding: the compiled reference's Module::parse* (all eleven
    fields) against the oracle's restatement and the library's
    sdl_module_decode over every combination of the 25 bits the decoding reads
    (twice, the unread bits clear and hashed), plus the partner involution and
    the lower/pixel closed forms on real sensors (oracle/decode_check.cc)."""
    import subprocess
    r = subprocess.run([DECODE_CHECK, "8"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stderr.strip() == f"ok {2 << 25}"


def test_module_decode_closed_forms():
    """Hand-decoded detIds through the library's C-ABI (host only): a barrel
    tilted PS sensor, its partner, and an endcap 2S sensor."""
    from sdl_amd import runtime
    # barrel: subdet 5, layer 2, side 1 (NegZ), rod 7, module 12, bit0 = 1
    d = (5 << 25) | (2 << 20) | (1 << 18) | (7 << 10) | (12 << 2) | 1
    m = runtime.decode_module(d)
    assert (m["subdet"], m["layer"], m["side"], m["rod"], m["ring"], m["module"]) == (5, 2, 1, 7, 0, 12)
    # layer <= 2, NegZ, even module: not inverted, so bit0 = 1 is the lower (pixel) sensor
    assert (m["isInverted"], m["isLower"], m["partnerDetId"], m["moduleType"], m["moduleLayerType"]) == (0, 1, d + 1, 0, 0)
    p = runtime.decode_module(d + 1)
    assert (p["isLower"], p["partnerDetId"], p["moduleLayerType"]) == (0, d, 1)
    # endcap: subdet 4, side 2 (PosZ), layer 3, ring 9, module 5 (odd: not inverted), bit0 = 0 -> upper
    e = (4 << 25) | (2 << 23) | (3 << 18) | (9 << 12) | (5 << 2) | 2
    m = runtime.decode_module(e)
    assert (m["subdet"], m["side"], m["layer"], m["ring"], m["module"], m["rod"]) == (4, 2, 3, 9, 5, 0)
    assert (m["isInverted"], m["isLower"], m["partnerDetId"], m["moduleType"], m["moduleLayerType"]) == (0, 0, e - 1, 1, 1)


def test_oracle_matches_reference_digest_at_pu200(geom_paths):
    """A PU200-density event (workload 2 event 0: 8000 tracks + 20 hard, 2.6 M
    track candidates), too large for stored tuples: the GLIBC-policy oracle's
    module-map chain against the compiled reference's digest
    (ref_pu200_digest.json: object counts, the ten counters, the per-layer
    counters and a SHA-256 of each sorted object set)."""
    import importlib.util
    import json
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(here, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    with open(os.path.join(here, "golden", "ref_pu200_digest.json")) as f:
        digest = json.load(f)
    for wl, e in mg.DIGEST_CASES:
        ref = digest[f"w{wl}_e{e}"]
        sets, counters, lay = oracle.run(geom_paths[gen.geom_kind(wl)], gen.event(wl, e), oracle.GLIBC, 6, layers=True)
        assert {s: len(sets[s]) for s in oracle.STAGES} == ref["counts"]
        assert [counters[k] for k in COUNTERS] == ref["counters"]
        assert lay.tolist() == ref["layers"]
        assert {s: mg.set_digest(sets[s]) for s in oracle.STAGES} == ref["sha256"]


KAT_CHECK = os.path.join(os.path.dirname(oracle.REF_DRIVER), "kat_check")


def _tilted_with_flat_strips(geom_paths, tmp_path):
    """The workload-2 geometry with the strip slopes of 40 tilted modules set to
    0 (no such module in the generated detector) and of 40 others to SDL_INF:
    the two special branches of MiniDoublet::shiftStripHits (MiniDoublet.cc:1145-1159)."""
    tilted, endcap, cmap = geom_paths[gen.geom_kind(2)]
    rows = [l.split() for l in open(tilted) if l.strip()]
    tilt = [r for r in rows if float(r[1]) != 0.0]
    zero = {r[0] for r in tilt[:80]}
    inf = {r[0] for r in tilt[80:160]}
    out = tmp_path / "tilted_special.txt"
    with open(out, "w") as f:
        for r in rows:
            s = "0" if r[0] in zero else ("123456789" if r[0] in inf else r[2])
            f.write(f"{r[0]} {r[1]} {s}\n")
    return str(out), endcap, cmap


@pytest.mark.skipif(not os.path.exists(KAT_CHECK), reason="oracle/_ref not built (needs the reference sources)")
def test_single_functions_match_reference(geom_paths, tmp_path):
    """Single functions of the compiled reference against the oracle's, value
    for value (oracle/kat_check.cc): Phi_mpi_pi on about 9 M floats of
    [-1000, 1000], ATan2 on a signed grid with zeros, getCenterFromThreePoints
    on random and degenerate triples, moduleGapSize and isTighterTiltedModules
    on every barrel and endcap sensor, and shiftStripHits on random hit pairs
    of every tilted and endcap lower module -- once with the generated
    detector, once with strip slopes set to 0 and SDL_INF."""
    import subprocess
    for files in (geom_paths[gen.geom_kind(2)], _tilted_with_flat_strips(geom_paths, tmp_path)):
        r = subprocess.run([KAT_CHECK, *files], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        last = r.stderr.strip().splitlines()[-1]
        assert last.startswith("ok ") and int(last.split()[1]) > 10_000_000, last


def test_phi_and_atan2_closed_forms():
    """Phi_mpi_pi (MathUtil.cc:3-19) and ATan2 (MathUtil.cc:21-27) against what
    the mathematics fixes: the result lies in [-pi, pi) and differs from the
    argument by a whole number of turns; the float nearest pi (above pi) maps
    below -pi's neighbour and the float nearest -pi (below -pi) maps to the top
    of the range; ATan2's axis cases and octant values."""
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.uniform(-1000, 1000, 20000), rng.uniform(-7, 7, 20000)]).astype(np.float32):
        r = L.oracle_kat_phi_mpi_pi(float(x))
        assert -np.pi <= r < np.pi, (x, r)
        turns = (float(x) - r) / (2 * np.pi)
        # each of the whole turns is subtracted in double and rounded to float
        bound = (abs(round(turns)) + 1) * float(np.spacing(np.float32(abs(x) + 7)))
        assert abs(turns - round(turns)) * 2 * np.pi <= bound, (x, r)
    pi32 = float(np.float32(np.pi))  # 3.14159274 > pi
    assert L.oracle_kat_phi_mpi_pi(pi32) == pytest.approx(pi32 - 2 * np.pi, abs=1e-6)
    assert L.oracle_kat_phi_mpi_pi(-pi32) == pytest.approx(-pi32 + 2 * np.pi, abs=1e-6)
    assert L.oracle_kat_phi_mpi_pi(0.0) == 0.0 and np.isnan(L.oracle_kat_phi_mpi_pi(float("nan")))
    half = float(np.float32(np.pi / 2))
    assert L.oracle_kat_atan2(0.0, 0.0) == 0.0
    assert L.oracle_kat_atan2(2.5, 0.0) == half and L.oracle_kat_atan2(-2.5, 0.0) == -half
    assert L.oracle_kat_atan2(0.0, -1.0) == pi32
    assert L.oracle_kat_atan2(1.0, 1.0) == float(np.float32(np.pi / 4))
    assert L.oracle_kat_atan2(-1.0, -1.0) == float(np.float32(-3 * np.pi / 4))


def test_circle_centre_closed_forms():
    """getCenterFromThreePoints (MathUtil.cc:44-106): three points of a circle
    give its centre; points on one line (equal slopes) give the origin, the
    reference's default Hit."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        cx, cy = rng.uniform(-50, 50, 2)
        R = rng.uniform(20, 200)
        a = np.sort(rng.uniform(0, 2 * np.pi, 3))
        if np.min(np.diff(np.concatenate([a, [a[0] + 2 * np.pi]]))) < 0.3:
            continue
        xs, ys = cx + R * np.cos(a), cy + R * np.sin(a)
        if np.min(np.abs(np.diff(np.concatenate([xs, xs[:1]])))) < 1e-2:
            continue  # a vertical chord: the reference sets its slope to 0
        c = oracle.kat_center(xs, ys)
        assert np.hypot(c[0] - cx, c[1] - cy) < 2e-3 * R, (cx, cy, R, c)
    assert oracle.kat_center([0, 1, 2], [0, 2, 4]).tolist() == [0.0, 0.0]
    assert oracle.kat_center([5, 6, 8], [1, 3, 7]).tolist() == [0.0, 0.0]


def test_strip_shift_closed_forms(geom_paths, tmp_path):
    """shiftStripHits (MiniDoublet.cc:989-1209): the shifted strip hit is the
    foot of the perpendicular from the strip hit onto the strip line through
    the anchored pixel position -- with slope SDL_INF it keeps the strip hit's y,
    with slope 0 its x -- and keeps the sign of the pixel hit's z."""
    files = _tilted_with_flat_strips(geom_paths, tmp_path)
    slope = {int(r.split()[0]): float(r.split()[2]) for r in open(files[0]) if r.strip()}
    L = oracle.lib()
    rng = np.random.default_rng(9)
    seen = {"inf": 0, "zero": 0, "other": 0}
    for d in list(slope)[:4000]:
        m = np.zeros(11, np.uint32)
        L.oracle_module_decode(d, m.ctypes.data)
        subdet, side, is_lower, partner, mtype, ltype = int(m[0]), int(m[1]), int(m[6]), int(m[8]), int(m[9]), int(m[10])
        if subdet != 5 or side == 3 or not is_lower:
            continue
        strip_is_lower = mtype == 1 or ltype == 1
        strip_id = d if strip_is_lower else partner
        s = slope.get(strip_id, 0.0)
        phi, r, z = rng.uniform(-np.pi, np.pi), rng.uniform(20, 60), rng.uniform(-100, 100)
        lo = np.array([r * np.cos(phi), r * np.sin(phi), z], np.float32)
        up = (lo + rng.uniform(-0.3, 0.3, 3)).astype(np.float32)
        out = oracle.kat_shift(files, d, lo, up)
        strip, pix = (lo, up) if strip_is_lower else (up, lo)
        if s == np.float32(123456789):
            assert out[1] == strip[1], (d, out, strip)
            seen["inf"] += 1
        elif s == 0:
            assert out[0] == strip[0], (d, out, strip)
            seen["zero"] += 1
        else:
            # (out - strip) is perpendicular to the strip direction (1, s)
            dx, dy = float(out[0] - strip[0]), float(out[1] - strip[1])
            assert abs(dx + s * dy) <= 1e-4 * (1 + abs(s)) * (abs(dx) + abs(dy) + 1), (d, s, out, strip)
            seen["other"] += 1
        assert np.sign(out[2]) == np.sign(pix[2])
    assert min(seen.values()) > 10, seen

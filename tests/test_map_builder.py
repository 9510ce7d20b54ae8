"""The connection-map builder (sdl_map_build): module connections from
simulated tracks, in the reference's map format (ModuleConnectionMap.cc:18-47),
merged into an existing map exactly as ModuleConnectionMap::load followed by
ModuleConnectionMap::add (ModuleConnectionMap.cc:49-85) would merge it -- checked
against the compiled reference (oracle/_ref/map_check) -- and usable by the
chain (a device run on a built map equals the oracle on the same files)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from sdl_amd import gen, runtime

MAP_CHECK = os.path.join(os.path.dirname(oracle.REF_DRIVER), "map_check")


def tracks_of(events):
    """Simulated tracks of generated events as [(station, detId), ...]: each
    track's hits ordered by distance from the origin, one station per run of
    hits on the same layer or disk (subdet, side, layer)."""
    out = []
    for e in events:
        d = e.detId.astype(np.int64)
        sub = (d >> 25) & 7
        layer = np.where(sub == 5, (d >> 20) & 7, (d >> 18) & 7)
        side = np.where(sub == 5, (d >> 18) & 3, (d >> 23) & 3)
        r3 = e.x.astype(np.float64) ** 2 + e.y.astype(np.float64) ** 2 + e.z.astype(np.float64) ** 2
        for t in np.unique(e.sim[e.sim >= 0]):
            h = np.flatnonzero(e.sim == t)
            h = h[np.argsort(r3[h], kind="stable")]
            key = sub[h] * 100 + side[h] * 10 + layer[h]
            station = np.concatenate([[0], np.cumsum(key[1:] != key[:-1])])
            out.append(list(zip(station.tolist(), e.detId[h].tolist())))
    return out


def read_map(path):
    with open(path) as f:
        return [list(map(int, l.split())) for l in f if l.strip()]


@pytest.fixture(scope="module")
def tracks():
    evs = [gen.event(1, 500 + i) for i in range(12)]
    return tracks_of(evs[:6]), tracks_of(evs[6:])


def test_built_map_format_and_connections(tracks, tmp_path):
    t1, _ = tracks
    path = str(tmp_path / "map.txt")
    runtime.build_connection_map(t1, path)
    rows = read_map(path)
    assert len(rows) > 100
    keys = [r[0] for r in rows]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    conn = {}
    for r in rows:
        assert r[1] == len(r) - 2 and r[1] > 0
        lst = r[2:]
        assert all(a < b for a, b in zip(lst, lst[1:])), "a list is not sorted or has duplicates"
        conn[r[0]] = set(lst)
    lower = lambda d: d if runtime.decode_module(d)["isLower"] else runtime.decode_module(d)["partnerDetId"]
    for d in list(conn)[:50] + [c for v in list(conn.values())[:20] for c in v]:
        assert runtime.decode_module(d)["isLower"] == 1
    # every pair of consecutive stations of every track is connected
    for trk in t1[:300]:
        st = {}
        for s, d in trk:
            st.setdefault(s, set()).add(lower(d))
        order = sorted(st)
        for a, b in zip(order, order[1:]):
            for p in st[a]:
                assert st[b] <= conn[p]


def test_build_then_merge_is_the_union(tracks, tmp_path):
    t1, t2 = tracks
    both, merged = str(tmp_path / "both.txt"), str(tmp_path / "merged.txt")
    runtime.build_connection_map(t1 + t2, both)
    runtime.build_connection_map(t1, merged)
    runtime.build_connection_map(t2, merged, merge=True)
    assert read_map(both) == read_map(merged)


@pytest.mark.skipif(not os.path.exists(MAP_CHECK), reason="oracle/_ref not built (needs the reference sources)")
def test_merge_matches_reference_add(tracks, tmp_path):
    """Merging into a hand-written map -- unsorted lists, duplicates, a repeated
    detId line, detIds the tracks do and do not connect -- equals the compiled
    reference's load(old) + add(new lists)."""
    _, t2 = tracks
    fresh = str(tmp_path / "fresh.txt")
    runtime.build_connection_map(t2, fresh)
    rows = read_map(fresh)
    touched = rows[0][0], rows[len(rows) // 2][0]
    c0, c1 = rows[0][2], rows[len(rows) // 2][2]
    old = str(tmp_path / "old.txt")
    with open(old, "w") as f:
        f.write(f"{touched[0]} 3 {c0 + 7} {c0} {c0 + 7}\n")  # unsorted, with a duplicate and a known connection
        f.write("123456 2 99 98\n")  # a detId the tracks do not touch: kept as read
        f.write(f"{touched[1]} 1 5\n")
        f.write(f"{touched[1]} 2 {c1} 4\n")  # repeated detId: the last line counts
        f.write("77 0\n")
    merged = str(tmp_path / "merged.txt")
    shutil.copy(old, merged)
    runtime.build_connection_map(t2, merged, merge=True)
    ref = str(tmp_path / "ref.txt")
    subprocess.run([MAP_CHECK, old, fresh, ref], check=True, timeout=120)
    assert read_map(merged) == read_map(ref)
    got = {r[0]: r[2:] for r in read_map(merged)}
    assert got[123456] == [99, 98] and got[77] == []
    assert got[touched[0]] == sorted(set(got[touched[0]])) and c0 + 7 in got[touched[0]]
    assert 5 not in got[touched[1]] and 4 in got[touched[1]]


@pytest.mark.gpu
def test_chain_on_a_built_map_matches_oracle(tracks, tmp_path):
    """The chain on the device with a map built from simulated tracks of other
    events (the generator's tilted and endcap tables) equals the oracle on the
    same three files."""
    t1, t2 = tracks
    path = str(tmp_path / "built_map.txt")
    runtime.build_connection_map(t1 + t2, path)
    t, e, _ = gen.geometry(1)[1]
    paths = (t, e, path)
    g = runtime.Geometry(*paths)
    for ie in (0, 1):
        print(f"case built map, workload 1 event {ie}", flush=True)
        x = gen.event(1, ie)
        ev = runtime.Event(g)
        ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
        ev.run_chain()
        ev.sync()
        ref, cnt = oracle.run(paths, x, oracle.CR)
        assert cnt["nTL"] > 0
        for k, s in enumerate(runtime.STAGES):
            assert np.array_equal(oracle.as_multiset(ev.export(k)), oracle.as_multiset(ref[s])), s
        assert ev.counts(0) == cnt

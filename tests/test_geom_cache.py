"""The binary geometry cache on the host: conversion from the reference's text
formats, the module and connection counts against the text files, and
rejection of damaged files.  (Loading a cache onto the GPU and running the
chain on it is covered by test_gpu_parity.py.)"""
import os

import pytest

from sdl_amd import gen, runtime


def _text_counts(paths):
    """Modules (detIds of the tilted, endcap and map files) and map connections."""
    tilted, endcap, cmap = paths
    ids, conns = set(), 0
    for p in (tilted, endcap):
        with open(p) as f:
            ids.update(int(line.split()[0]) for line in f if line.strip())
    with open(cmap) as f:
        for line in f:
            t = line.split()
            if len(t) < 2:
                continue
            ids.add(int(t[0]))
            ids.update(int(x) for x in t[2:2 + int(t[1])])
            conns += int(t[1])
    return len(ids), conns


@pytest.fixture(scope="module")
def cache(tmp_path_factory):
    paths = gen.geometry(1)[1]
    out = str(tmp_path_factory.mktemp("geom") / "geom_1.bin")
    runtime.convert_geometry(*paths, out)
    return paths, out


def test_convert_and_check(cache):
    paths, out = cache
    nm, nc = runtime.check_geometry_cache(out)
    assert (nm, nc) == _text_counts(paths)
    # converting again gives the same bytes (the cache is a pure function of the files)
    again = out + ".again"
    runtime.convert_geometry(*paths, again)
    assert open(out, "rb").read() == open(again, "rb").read()


@pytest.mark.parametrize("damage", ["magic", "flip", "truncate", "extend"])
def test_damaged_cache_rejected(cache, tmp_path, damage):
    _, out = cache
    b = bytearray(open(out, "rb").read())
    if damage == "magic":
        b[0] ^= 0xFF
    elif damage == "flip":
        b[len(b) // 2] ^= 0x01
    elif damage == "truncate":
        b = b[:-5]
    else:
        b += b"\0"
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(b))
    with pytest.raises(runtime.SDLError, match="not a valid geometry cache"):
        runtime.check_geometry_cache(bad)


def test_missing_file(tmp_path):
    with pytest.raises(runtime.SDLError):
        runtime.check_geometry_cache(str(tmp_path / "absent.bin"))

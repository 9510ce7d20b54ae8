"""sdl_event_export_noalloc, the export that allocates no device memory (its
staging is a static 64 MiB per-GPU device array), against sdl_event_export on
the same events: every stage's hit-index tuples identical, element for element,
on a PU200 event (8000 tracks; its stages too large for the array are refused
with SDL_ERR_CAPACITY) and then on a smaller barrel event, so the later calls
run over the array the earlier ones filled."""
import ctypes

import numpy as np
import pytest

from sdl_amd import gen, runtime

pytestmark = pytest.mark.gpu

STAGES = (runtime.STAGE_MD, runtime.STAGE_SG, runtime.STAGE_TP, runtime.STAGE_TL, runtime.STAGE_TCP,
          runtime.STAGE_TCL)
SCRATCH_INTS = 16 << 20


def _export(ev, name, stage):
    f = getattr(runtime.lib(), name)
    n = ctypes.c_longlong(0)
    runtime._check(f(ev.h, 0, stage, None, 0, ctypes.byref(n)))
    out = np.full((n.value, runtime.WIDTH[runtime.STAGES[stage]]), -7, np.int32)
    if n.value:
        runtime._check(f(ev.h, 0, stage, out.ctypes.data, out.size, ctypes.byref(n)))
    return out


def _event(wl, e):
    g = runtime.Geometry(*gen.geometry(gen.geom_kind(wl))[1])
    ev = runtime.Event(g)
    x = gen.event(wl, e)
    ev.add_hits(x.x, x.y, x.z, x.idx, x.detId)
    ev.run_chain()
    ev.sync()
    return g, ev


def _count(ev, name, stage):
    n = ctypes.c_longlong(0)
    runtime._check(getattr(runtime.lib(), name)(ev.h, 0, stage, None, 0, ctypes.byref(n)))
    return n.value


def test_noalloc_export_equals_export():
    """Every stage whose tuples fit the staging array equals sdl_event_export; a
    PU200 stage that does not fit is refused with SDL_ERR_CAPACITY."""
    largest, refused = [], []
    for wl, e in ((2, 0), (1, 0)):
        print(f"case export_noalloc workload={wl} event={e}", flush=True)
        g, ev = _event(wl, e)
        fitted = []
        for st in STAGES:
            ints = _count(ev, "sdl_event_export_noalloc", st) * runtime.WIDTH[runtime.STAGES[st]]
            assert ints == _count(ev, "sdl_event_export", st) * runtime.WIDTH[runtime.STAGES[st]]
            if ints > SCRATCH_INTS:
                with pytest.raises(runtime.SDLError, match="exceed its staging array"):
                    _export(ev, "sdl_event_export_noalloc", st)
                refused.append((wl, st))
                continue
            got = _export(ev, "sdl_event_export_noalloc", st)
            ref = _export(ev, "sdl_event_export", st)
            assert got.shape == ref.shape, (wl, st)
            assert np.array_equal(got, ref), f"workload {wl}: stage {runtime.STAGES[st]} differs"
            fitted.append(got.size)
        assert len(fitted) >= 4, (wl, refused)
        largest.append(max(fitted))
        del ev, g
    assert largest[0] > largest[1] > 0  # the second event's exports are smaller
    assert all(wl == 2 for wl, _ in refused)


def test_noalloc_export_refuses_a_short_output():
    g, ev = _event(1, 0)
    n = ctypes.c_longlong(0)
    L = runtime.lib()
    runtime._check(L.sdl_event_export_noalloc(ev.h, 0, runtime.STAGE_MD, None, 0, ctypes.byref(n)))
    assert n.value > 1
    out = np.zeros(2 * n.value - 2, np.int32)
    with pytest.raises(runtime.SDLError):
        runtime._check(L.sdl_event_export_noalloc(ev.h, 0, runtime.STAGE_MD, out.ctypes.data, out.size,
                                                  ctypes.byref(n)))

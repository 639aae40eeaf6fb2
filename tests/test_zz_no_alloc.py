"""The allocation-free snapshot save (rc_cache_snapshot_save_noalloc, records staged in the library's
static device array) through the C ABI against rc_cache_snapshot_save (a buffer allocated per call):
the same 24-byte records bit for bit, on two calls in a row, the second with a smaller cache, and a
capacity below the record count."""
import ctypes as C

import numpy as np
import pytest

import render_engine_amd as rea
from render_engine_amd._ffi import lib

pytestmark = pytest.mark.gpu


def _save(fn, cache, cap):
    out = np.zeros(max(cap, 1), rea.SNAP_REC)
    n = C.c_size_t(0)
    rc = fn(cache._h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n))
    assert rc == 0, lib.rc_last_error().decode() if isinstance(lib.rc_last_error(), bytes) else lib.rc_last_error()
    return out[:n.value]


def test_snapshot_save_noalloc_matches_allocating_save(torch_gpu):
    room = rea.Scene.box_room(seed=1, target_tris=2000).upload(trav=rea.TRAV_WIDE)
    street = rea.Scene.street(seed=3, nboxes=150, target_tris=12000, nlamps=6, sun=True, target_points=30000, table_log2=0,
                              W=96, H=64, spp=2, bounces=3).upload(trav=rea.TRAV_WIDE)
    assert street.info.Np < room.info.Np
    for scene, seed in ((room, 4), (street, 9)):          # the second call: a smaller input
        cache = rea.Cache(scene, rea.methods.flush_snapshots())
        cache.begin(0, seed)
        for st in [0] * 3 + [1] * 2:                        # the flush at the event leaves part of the entries empty
            cache.frame(st)
        Np = scene.info.Np
        a = _save(lib.rc_cache_snapshot_save_noalloc, cache, Np)
        b = _save(lib.rc_cache_snapshot_save, cache, Np)
        print(f"noalloc save: {len(a)} records of {Np} entries", flush=True)
        assert 0 < len(a) <= Np and len(a) == len(b)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.array_equal(a["index"], np.nonzero(cache.get("N") > 0)[0])
        c = _save(lib.rc_cache_snapshot_save_noalloc, cache, 100)   # capacity below the count: the first records
        assert len(c) == 100 and np.array_equal(c.view(np.uint8), b[:100].view(np.uint8))

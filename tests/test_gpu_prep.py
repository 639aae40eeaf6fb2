"""Scene preparation on the GPU: the binned-SAH BVH (cuda/prep_mesh.cuh:42-114) and the cache cells
sampled, reduced and Morton-ordered (:129-157) by the device path equal the host path's bit for
bit, so every scene array built from them does too (nodes, triangle order, wide nodes, entry
points, normals, albedos, materials, zones). The hash table (:356-366) is built on the GPU with
64-bit CAS inserts: the same keys, every key looking up to the same entry, empty slots -1, every
key within 64 probes of its home slot with no empty slot in between; slots may differ."""
import time

import numpy as np
import pytest

import render_engine_amd as rea

pytestmark = pytest.mark.gpu

INFO = ("Np", "npst", "nbox", "NR", "hsize", "nmat", "nst", "nnodes", "nwnodes", "bvh_depth", "wide_depth")
ARRAYS = ("ALB", "REGION", "MAT", "MATALB", "ST", "STN", "BN", "WN", "ORDER")
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(x):
    """the slot function (SplitMix64 finaliser), vectorised"""
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint64(30)); x = x * np.uint64(0xBF58476D1CE4E5B9)
        x = x ^ (x >> np.uint64(27)); x = x * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def _same_table(a, b):
    HKa, HVa, HKb, HVb = a.array("HK"), a.array("HV"), b.array("HK"), b.array("HV")
    assert HKa.shape == HKb.shape
    oa, ob = HKa != 0, HKb != 0
    ia, ib = np.argsort(HKa[oa]), np.argsort(HKb[ob])
    assert np.array_equal(HKa[oa][ia], HKb[ob][ib]), "the tables hold different keys"
    assert np.array_equal(HVa[oa][ia], HVb[ob][ib]), "a key looks up to another entry"
    assert np.all(HVb[~ob] == -1)
    size = len(HKb)
    mask = np.uint64(size - 1)
    slot = np.nonzero(ob)[0].astype(np.int64)
    home = (mix64(HKb[ob]) & mask).astype(np.int64)
    dist = (slot - home) % size
    assert dist.max() < 64, "a key lies 64 or more probes from its home slot"
    # start of the occupied run holding each slot (cyclic): the slot after the last empty one before it
    empty = np.nonzero(~ob)[0]
    assert len(empty) > 0
    last_empty = np.full(size, -1, np.int64)
    last_empty[empty] = empty
    last_empty = np.maximum.accumulate(last_empty)
    last_empty[last_empty < 0] = empty[-1] - size     # runs that wrap around the end
    run = (slot - (last_empty[slot] + 1)) % size
    assert np.all(dist <= run), "an empty slot lies between a key's home slot and its slot"


def _both(make):
    rea.prep_device(False)
    t0 = time.perf_counter()
    a = make()
    t1 = time.perf_counter()
    rea.prep_device(True)
    try:
        b = make()
    finally:
        rea.prep_device(False)
    t2 = time.perf_counter()
    return a, b, t1 - t0, t2 - t1


def _same(a, b):
    ia, ib = a.info, b.info
    for f in INFO:
        assert getattr(ia, f) == getattr(ib, f), f
    assert np.float32(ia.h).tobytes() == np.float32(ib.h).tobytes()
    for name in ARRAYS:
        x, y = a.array(name), b.array(name)
        assert x.dtype == y.dtype and x.shape == y.shape, name
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name
    for st in range(ia.NS):
        for name in ("P", "N"):
            x, y = a.array(name, st), b.array(name, st)
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (name, st)
    _same_table(a, b)


def test_prep_device_small_street(torch_gpu):
    a, b, _, _ = _both(lambda: rea.Scene.street(seed=3, nboxes=60, target_tris=5000, nlamps=4, sun=True, target_points=8000,
                                                table_log2=0, W=64, H=48, spp=1, bounces=2))
    _same(a, b)


def test_prep_device_box_room_fixed_cells(torch_gpu):
    a, b, _, _ = _both(lambda: rea.Scene.box_room(seed=5, target_tris=3000))
    _same(a, b)


def test_prep_device_street_full_size(torch_gpu):
    """the 960x540 street (199,978 triangles, ~530k cells, cell size adjusted toward the target)"""
    a, b, th, td = _both(lambda: rea.Scene.street(1, 2000, 200_000, 32, True, 530_000, 22, 960, 540, 4, 3))
    _same(a, b)
    print(f"scene preparation: host {th:.2f} s, device BVH + cells {td:.2f} s")


def test_prep_device_double_sided_mesh(torch_gpu, tmp_path):
    """double-sided triangles sample both sides into separate cells (normal buckets)"""
    room = rea.Scene.box_room(seed=2, target_tris=1500)
    V, alb = room.array("MESH_V"), room.array("MESH_ALB")
    nt = len(alb)
    ds = (np.random.default_rng(4).random(nt) < 0.5).astype(np.uint8)
    path = str(tmp_path / "room.msh")
    with open(path, "wb") as f:
        f.write(b"MSH1")
        f.write(np.int32(nt).tobytes())
        f.write(V.astype(np.float32).tobytes()); f.write(alb.astype(np.float32).tobytes()); f.write(ds.tobytes())
    a, b, _, _ = _both(lambda: rea.Scene.mesh(path, target_points=12000)[0])
    _same(a, b)


def test_prep_device_street_1m(torch_gpu):
    """the 1920x1080 street (~1M triangles, ~1M cells): chunked node bounds and bins at the top levels"""
    a, b, th, td = _both(lambda: rea.Scene.street(1, 2000, 1_000_000, 64, True, 1_000_000, 23, 1920, 1080, 4, 3))
    _same(a, b)
    print(f"scene preparation (1M triangles): host {th:.2f} s, device BVH + cells {td:.2f} s")

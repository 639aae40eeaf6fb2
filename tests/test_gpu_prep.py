"""Scene preparation on the GPU: the binned-SAH BVH (cuda/prep_mesh.cuh:42-114) and the cache cells
sampled, reduced and Morton-ordered (:129-157) by the device path equal the host path's bit for
bit, so every scene array built from them does too (nodes, triangle order, wide nodes, entry
points, normals, albedos, materials, zones, hash table)."""
import time

import numpy as np
import pytest

import render_engine_amd as rea

pytestmark = pytest.mark.gpu

INFO = ("Np", "npst", "nbox", "NR", "hsize", "nmat", "nst", "nnodes", "nwnodes", "bvh_depth", "wide_depth")
ARRAYS = ("ALB", "REGION", "HK", "HV", "MAT", "MATALB", "ST", "STN", "BN", "WN", "ORDER")


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

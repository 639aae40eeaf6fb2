"""Host side of the library without a GPU: the C ABI exports every declared symbol, scene
preparation (cell sampling, hash table, zones) matches the oracle's restatement bit for bit,
argument errors come back as rc_status codes, and GPU calls fail loudly without a device."""
import os
import re

import numpy as np
import pytest

import pyoracle as po
import render_engine_amd as rea
from render_engine_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "rc.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(rc_[a-z_0-9]+)\s*\(", hdr, re.M))
    declared |= set(re.findall(r"^(?:void|int|size_t|rc_status|const char\*)\s+(rc_[a-z_0-9]+)\(", hdr, re.M))
    assert declared, "no declarations parsed"
    import ctypes as C
    lib = C.CDLL(_ffi.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in rc.h but not exported"
    assert declared == set(_ffi.EXPORTS), declared ^ set(_ffi.EXPORTS)


def test_no_gpu_fails_loudly():
    if rea.device_count() > 0:
        pytest.skip("a GPU is visible")
    s = rea.Scene.box_room(1, 500)
    with pytest.raises(rea.RcError, match="no HIP device"):
        s.upload()


def test_argument_errors():
    with pytest.raises(rea.RcError, match="target_tris"):
        rea.Scene.box_room(1, 3)
    with pytest.raises(rea.RcError, match="out of range"):
        rea.Scene.street(1, 10, 1000, 100, True, 1000, 0, 8, 8, 1, 1)   # 101 lights
    with pytest.raises(rea.RcError, match="too small"):
        rea.Scene.street(1, 10, 1000, 2, True, 5000, 10, 8, 8, 1, 1)   # 2^10 table for ~5000 cells
    s = rea.Scene.box_room(1, 500)
    with pytest.raises(rea.RcError, match="not available"):
        s.array("EDIR", 0)   # device-produced array before upload
    with pytest.raises(rea.RcError):
        s.array("P", 7)      # state out of range
    for levels in (0, 65):
        with pytest.raises(rea.RcError, match="stack levels"):
            s.set_stack_levels(levels)


@pytest.mark.parametrize("kind", ["room", "street"])
def test_cells_and_table_match_oracle(kind):
    if kind == "room":
        s = rea.Scene.box_room(seed=2, target_tris=1500)
    else:
        s = rea.Scene.street(seed=5, nboxes=80, target_tris=6000, nlamps=3, sun=False, target_points=6000, table_log2=0, W=8,
                             H=8, spp=1, bounces=1)
    i = s.info
    V, alb, ds = s.array("MESH_V"), s.array("MESH_ALB"), s.array("MESH_DS")
    lo = np.array(list(i.lo), np.float32)
    n = po.lib.orc_sample_cells(po.ptr(V), po.ptr(alb), po.ptr(ds), len(alb), i.h, po.ptr(lo), None, None, None, None, 0)
    assert n == i.npst
    keys = np.zeros(n, np.uint64); P = np.zeros(3 * n, np.float32); N = np.zeros(3 * n, np.float32); A = np.zeros(n, np.float32)
    po.lib.orc_sample_cells(po.ptr(V), po.ptr(alb), po.ptr(ds), len(alb), i.h, po.ptr(lo), po.ptr(keys), po.ptr(P), po.ptr(N),
                            po.ptr(A), n)
    assert np.array_equal(P.view(np.uint32), s.array("P", 0)[:3 * n].view(np.uint32))
    assert np.array_equal(N.view(np.uint32), s.array("N", 0)[:3 * n].view(np.uint32))
    assert np.array_equal(A.view(np.uint32), s.array("ALB")[:n].view(np.uint32))
    HK = np.zeros(i.hsize, np.uint64); HV = np.zeros(i.hsize, np.int32)
    po.lib.orc_hash_build(po.ptr(keys), n, int(np.log2(i.hsize)), po.ptr(HK), po.ptr(HV))
    assert np.array_equal(HK, s.array("HK")) and np.array_equal(HV, s.array("HV"))
    # every cell key is found at its own entry, load factor <= 1/2
    for k in range(0, n, max(1, n // 500)):
        assert po.lib.orc_hash_find(po.ptr(HK), po.ptr(HV), i.hsize - 1, int(keys[k])) == k
    assert 2 * n <= i.hsize
    # zones: 4x4x4 cells of one normal bucket share a region index, numbered in first-seen order
    cx, cy, cz, b = (np.zeros(1, np.int32) for _ in range(4))
    import ctypes as C
    reg = s.array("REGION")[:n]
    zone = {}
    for k in range(n):
        x, y, z, bb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        po.lib.orc_decode_key(int(keys[k]), C.byref(x), C.byref(y), C.byref(z), C.byref(bb))
        zk = (x.value >> 2, y.value >> 2, z.value >> 2, bb.value)
        zone.setdefault(zk, len(zone))
        assert reg[k] == zone[zk]


def test_street_events_and_quads():
    s = rea.Scene.street(seed=1, nboxes=50, target_tris=4000, nlamps=3, sun=True, target_points=4000, table_log2=0, W=8, H=8,
                         spp=1, bounces=1)
    i = s.info
    assert i.NS == 4 and i.nl == 4 and i.Nq == 6
    ns = i.NS
    # occluder present only in state 3: events 0<->3 carry its voxels, 0->1 none
    assert len(s.array("OCC", 0 * ns + 3)) > 0 and len(s.array("OCC", 3 * ns + 0)) > 0
    assert len(s.array("OCC", 0 * ns + 1)) == 0
    # the sun (last light) turning back on (1 -> 0) touches every transit voxel; lamp 0 at 2 -> 0
    assert len(s.array("EM", (1 * ns + 0) * i.nl + 3)) == i.vx * i.vy * i.vz
    assert len(s.array("EM", (2 * ns + 0) * i.nl + 0)) == 1
    assert len(s.array("TQ", 3)) == 12 and len(s.array("TQ", 0)) == 0
    # quad points follow the box: centre of each face grid cell
    P3 = s.array("P", 3).reshape(-1, 3)[i.npst:]
    ob = s.array("OBJBOX", 3)
    assert np.all(P3 >= ob[:3] - 1e-4) and np.all(P3 <= ob[3:] + 1e-4)


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _check_round_trip(s, path):
    s.save(path)
    t = rea.Scene.load(path)
    i, j = s.info, t.info
    for f in ("Np", "npst", "nst", "NS", "nl", "NR", "Nq", "hsize"):
        assert getattr(i, f) == getattr(j, f), f
    for name in ("ST", "STN", "BN", "HK", "HV", "L", "ALB", "REGION", "DIMS", "OFF"):
        assert _same(s.array(name), t.array(name)), name
    for st in range(i.NS):
        for name in ("P", "N", "Q", "T", "TQ", "INTEN"):
            assert _same(s.array(name, st), t.array(name, st)), (name, st)
    for e in range(i.NS * i.NS):
        assert _same(s.array("OCC", e), t.array("OCC", e))
        for l in range(i.nl):
            assert _same(s.array("EM", e * i.nl + l), t.array("EM", e * i.nl + l))
    return t


def test_scene_file_round_trip(tmp_path):
    """the format-3 writer and the reader agree on every array (cuda/prep_mesh.cuh:242-297 layout)"""
    _check_round_trip(rea.Scene.box_room(seed=1, target_tris=600), str(tmp_path / "room.bin"))
    street = rea.Scene.street(seed=3, nboxes=20, target_tris=2000, nlamps=2, sun=True, target_points=3000, table_log2=0,
                              W=32, H=16, spp=1, bounces=1)
    with pytest.raises(rea.RcError, match="point lights only"):
        street.save(str(tmp_path / "street.bin"))


def test_mesh_scene(tmp_path):
    """an MSH1 mesh becomes the prep mode's bench scene: 3 point lamps and a cube inside the bounds,
    5 states with the prep mode's intensities and cube positions, quads of the cube per state"""
    room = rea.Scene.box_room(seed=2, target_tris=800)
    V, alb, ds = room.array("MESH_V"), room.array("MESH_ALB"), room.array("MESH_DS")
    nt = len(alb)
    path = str(tmp_path / "room.msh")
    with open(path, "wb") as f:
        f.write(b"MSH1")
        f.write(np.int32(nt).tobytes())
        f.write(V.astype(np.float32).tobytes()); f.write(alb.astype(np.float32).tobytes()); f.write(ds.astype(np.uint8).tobytes())
    s, cube = rea.Scene.mesh(path, target_points=20000)
    i = s.info
    assert i.nl == 3 and i.NS == 5 and i.Nq == 6 and cube > 0
    I = np.array([s.array("INTEN", k) for k in range(5)])
    assert np.array_equal(I, [[1, 1, 0], [0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 0]])
    lo, hi = V.reshape(-1, 3).min(0), V.reshape(-1, 3).max(0)
    L = s.array("L").reshape(3, 3)
    assert np.all(L >= lo) and np.all(L <= hi)
    q0, q2 = s.array("Q", 0).reshape(6, 12), s.array("Q", 2).reshape(6, 12)
    assert not np.array_equal(q0, q2) and np.array_equal(q0, s.array("Q", 1).reshape(6, 12))
    assert len(s.array("OCC", 0 * 5 + 2)) > 0 and len(s.array("OCC", 0 * 5 + 1)) == 0
    _check_round_trip(s, str(tmp_path / "mesh.bin"))
    with pytest.raises(rea.RcError):
        rea.Scene.mesh(str(tmp_path / "room.bin"))


def test_prep_device_flag():
    """the cell-preparation switch takes 0 or 1 only (the device path itself is a GPU test)"""
    assert rea.lib.rc_scene_prep_device(2) != 0
    assert rea.lib.rc_scene_prep_device(0) == 0


def test_stack_code_culling_equivalence():
    """The wide traversal stores an 8-bit code of each pushed child's entry distance (exponent and
    top 3 mantissa bits, t >= 2^-15; 0 below) and culls a popped child when its code exceeds the
    code of the current hit distance. That equals culling on the code's lower bound (the least
    distance with that code) against the distance itself, for every code and every distance:
    checked here over all 256 codes and distances at and around every code boundary."""
    def tcode(t):
        b = np.asarray(t, np.float32).view(np.uint32).astype(np.int64)
        return np.clip(((b >> 20) & 0x7FF) - 895, 0, 255)

    def tlb(c):
        c = np.asarray(c, np.int64)
        v = (((c - 1) + 896) << 20).astype(np.uint32).view(np.float32)
        return np.where(c == 0, np.float32(0), v)

    codes = np.arange(256)
    bounds = tlb(codes[1:]).astype(np.float32)
    t = np.concatenate([bounds, np.nextafter(bounds, np.float32(0)), np.nextafter(bounds, np.float32(np.inf)),
                        np.float32([0.0, 1e-4, 1e-30, 3.4e38, np.inf]),
                        np.random.default_rng(1).random(4000).astype(np.float32) * np.float32(1e4)]).astype(np.float32)
    assert np.all(tlb(tcode(t)) <= t)                      # a distance's code bound never exceeds it
    for c in codes:
        assert np.array_equal(tlb(c) <= t, c <= tcode(t)), c

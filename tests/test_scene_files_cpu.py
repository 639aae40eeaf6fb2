"""The prototype scene-file reader (formats 2 / 3, cuda/poc_gpu.cu:751-801 and cuda/prep_mesh.cuh:242-297) on
corrupt input (CPU only): truncated files, counts larger than the file, BVH nodes whose leaves leave the triangle
range or whose children point backwards (a cycle), hash values and entry regions out of range, and bad header
values all come back as an RcError, never a crash or an allocation driven by a count in the file; the library
stays usable afterwards."""
import numpy as np
import pytest

import render_engine_amd as rea

HDR = 32            # int[8] = Np, Nq, Nt, NS, NL, NR, has_ref, fmt
MESH = HDR + 28     # lo[3], ext[3], h, then int m[4] = nst, nnodes, table size, static entries


@pytest.fixture(scope="module")
def room_file(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("scene") / "room.bin")
    rea.Scene.box_room(1, 600).save(path)
    return path, np.fromfile(path, np.uint8)


def _layout(b):
    h = b[:HDR].view(np.int32)
    m = b[MESH:MESH + 16].view(np.int32)
    nst, nn, hs = int(m[0]), int(m[1]), int(m[2])
    st = MESH + 16
    stn = st + 36 * nst
    bn = stn + 16 * nst
    hk = bn + 32 * nn
    hv = hk + 8 * hs
    light = hv + 4 * hs
    Np, NL = int(h[0]), int(h[4])
    alb = light + 12 * NL
    region = alb + 4 * Np
    return dict(h=h, nst=nst, nn=nn, hs=hs, bn=bn, hk=hk, hv=hv, region=region, Np=Np)


def _load_fails(tmp_path, data, match=None):
    p = str(tmp_path / "bad.bin")
    np.asarray(data, np.uint8).tofile(p)
    with pytest.raises(rea.RcError) as e:
        rea.Scene.load(p)
    if match:
        assert match in str(e.value), str(e.value)


def test_round_trip(room_file):
    path, b = room_file
    s = rea.Scene.load(path)
    L = _layout(b)
    assert (s.info.Np, s.info.nst, s.info.nnodes, s.info.hsize) == (L["Np"], L["nst"], L["nn"], L["hs"])


def test_truncated_files(room_file, tmp_path):
    path, b = room_file
    L = _layout(b)
    cuts = [0, 7, 31, 40, MESH + 8, L["bn"] + 100, L["hk"] + 12, L["hv"] + 3, L["region"] + 5, len(b) // 2, len(b) - 1]
    for c in cuts:
        _load_fails(tmp_path, b[:c])
    rea.Scene.load(path)   # still usable


def _patched(b, off, dtype, value):
    x = b.copy()
    x[off:off + np.dtype(dtype).itemsize] = np.array([value], dtype).view(np.uint8)
    return x


def test_counts_larger_than_the_file(room_file, tmp_path):
    path, b = room_file
    for field, value in ((0, 1 << 30), (1, 1 << 28), (2, 1 << 29)):       # Np, Nq, Nt
        _load_fails(tmp_path, _patched(b, 4 * field, np.int32, value))
    for field, value in ((0, 1 << 30), (1, 1 << 29), (2, 1 << 30)):       # nst, nnodes, table size
        _load_fails(tmp_path, _patched(b, MESH + 4 * field, np.int32, value))


def test_bad_header_values(room_file, tmp_path):
    path, b = room_file
    for field, value in ((0, -5), (1, -1), (3, 0), (4, 0), (4, 1000), (5, 0), (5, 1 << 30), (7, 4)):
        _load_fails(tmp_path, _patched(b, 4 * field, np.int32, value))
    _load_fails(tmp_path, _patched(b, MESH + 8, np.int32, 1000), "power of two")        # table size
    _load_fails(tmp_path, _patched(b, MESH + 12, np.int32, 1 << 30), "mesh header")     # static entries > Np


def test_bvh_nodes_out_of_range_or_cyclic(room_file, tmp_path):
    path, b = room_file
    L = _layout(b)
    nodes = b[L["bn"]:L["bn"] + 32 * L["nn"]].view(np.int32).reshape(-1, 8)   # lo[3], lf, hi[3], cnt
    leaf = int(np.nonzero(nodes[:, 7] > 0)[0][0])
    inner = int(np.nonzero(nodes[:, 7] == 0)[0][1])
    off = lambda k, w: L["bn"] + 32 * k + 4 * w
    _load_fails(tmp_path, _patched(b, off(leaf, 7), np.int32, L["nst"] + 5), "BVH node")        # leaf past the triangles
    _load_fails(tmp_path, _patched(b, off(leaf, 3), np.int32, -3), "BVH node")                  # negative first triangle
    _load_fails(tmp_path, _patched(b, off(inner, 3), np.int32, 0), "BVH node")                  # child = the root: a cycle
    _load_fails(tmp_path, _patched(b, off(inner, 3), np.int32, inner), "BVH node")              # child = itself
    _load_fails(tmp_path, _patched(b, off(inner, 3), np.int32, L["nn"] - 1), "BVH node")        # second child past the end
    _load_fails(tmp_path, _patched(b, off(inner, 7), np.int32, -2), "BVH node")                 # negative count


def test_hash_values_and_regions_out_of_range(room_file, tmp_path):
    path, b = room_file
    L = _layout(b)
    hk = b[L["hk"]:L["hk"] + 8 * L["hs"]].view(np.uint64)
    slot = int(np.nonzero(hk)[0][0])
    _load_fails(tmp_path, _patched(b, L["hv"] + 4 * slot, np.int32, L["Np"] + 7), "hash value")
    _load_fails(tmp_path, _patched(b, L["hv"] + 4 * slot, np.int32, -1), "hash value")
    nr = int(L["h"][5])
    _load_fails(tmp_path, _patched(b, L["region"] + 4 * 3, np.int32, nr), "region")
    _load_fails(tmp_path, _patched(b, L["region"], np.int32, -1), "region")
    rea.Scene.load(path)


def test_bad_mesh_files(tmp_path):
    p = str(tmp_path / "m.msh")
    open(p, "wb").write(b"MSH1" + np.int32(1 << 26).tobytes() + b"\0" * 64)   # count larger than the file
    with pytest.raises(rea.RcError):
        rea.Scene.mesh(p, 1000)
    open(p, "wb").write(b"XXXX" + np.int32(1).tobytes())
    with pytest.raises(rea.RcError):
        rea.Scene.mesh(p, 1000)
    open(p, "wb").write(b"MS")
    with pytest.raises(rea.RcError):
        rea.Scene.mesh(p, 1000)


def _long_leaf_file(tmp_path, b):
    """the room file with its binary BVH's root turned into one leaf holding every static triangle"""
    L = _layout(b)
    c = b.copy()
    node = c[L["bn"]:L["bn"] + 32].view(np.int32)
    node[3], node[7] = 0, L["nst"]   # lf, cnt
    p = str(tmp_path / "longleaf.bin")
    c.tofile(p)
    return p, L["nst"]


def test_long_leaf_split_over_wide_nodes(room_file, tmp_path):
    """a binary leaf of more than 127 triangles becomes a wide node of its own (recursively) whose
    leaf children are chunks of at most 127 of its triangles, each triangle in exactly one chunk,
    every chunk with the leaf's box"""
    path, b = room_file
    p, nst = _long_leaf_file(tmp_path, b)
    assert nst > 4 * 127
    s = rea.Scene.load(p)
    wn = s.array("WN").reshape(-1, 32)
    wf = wn.view(np.float32)
    ch, ce = wn[:, 24:28].view(np.uint32), wn[:, 28:32].view(np.int32)
    root_lo = b[_layout(b)["bn"]:_layout(b)["bn"] + 12].view(np.float32)
    seen = np.zeros(nst, np.int32)
    reached = np.zeros(len(wn), np.int32)
    for k in range(len(wn)):
        for c in range(4):
            e = int(ch[k, c])
            if e >> 31:
                first, count = (e >> 7) & 0xFFFFFF, e & 127
                assert 1 <= count <= 127 and ce[k, c] == first + count
                assert np.array_equal(wf[k, [c, 4 + c, 8 + c]], root_lo)
                seen[first:first + count] += 1
            elif e != 0:
                reached[e] += 1
    assert np.all(seen == 1)
    assert reached[0] == 0 and np.all(reached[1:] == 1)
    assert s.info.wide_depth >= 3

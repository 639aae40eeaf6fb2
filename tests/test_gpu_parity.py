"""GPU parity: the HIP library through its C ABI against the CPU oracle, bit for bit on hit
ids, keys, masks and cache values; radiance within 1e-4 relative RMS of the FP64 twin."""
import ctypes as C

import numpy as np
import pytest

import render_engine_amd as rea
import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(torch_gpu):
    s = rea.Scene.box_room(seed=1, target_tris=2000).upload(trav=rea.TRAV_WIDE)
    return s, po.OracleScene(s)


@pytest.fixture(scope="module")
def street(torch_gpu):
    s = rea.Scene.street(seed=3, nboxes=150, target_tris=12000, nlamps=6, sun=True, target_points=30000, table_log2=0,
                         W=96, H=64, spp=2, bounces=3).upload(trav=rea.TRAV_WIDE)
    return s, po.OracleScene(s)


_CAMERAS = {}


@pytest.fixture(scope="module")
def room_longleaf(torch_gpu, tmp_path_factory):
    """the room loaded from a file whose binary BVH root is one leaf holding every static triangle: the
    wide collapse splits it into nodes of chunks of at most 127 triangles"""
    path = str(tmp_path_factory.mktemp("longleaf") / "room.bin")
    room = rea.Scene.box_room(seed=1, target_tris=2000)
    room.save(path)
    _CAMERAS["room_longleaf"] = room.camera()   # scene files hold no camera
    b = np.fromfile(path, np.uint8)
    m = b[60:76].view(np.int32)   # header int[8], lo[3], ext[3], h, then nst, nnodes, table size, entries
    nst = int(m[0])
    bn = 76 + 36 * nst + 16 * nst
    node = b[bn:bn + 32].view(np.int32)
    node[3], node[7] = 0, nst
    b.tofile(path)
    s = rea.Scene.load(path).upload(trav=rea.TRAV_WIDE)
    assert s.info.nst == nst > 127
    return s, po.OracleScene(s)


def _rays(scene, n, seed, st=0):
    rng = np.random.default_rng(seed)
    i = scene.info
    lo, ext = np.array(list(i.lo)), np.array(list(i.ext))
    O = (lo + ext * (0.05 + 0.9 * rng.random((n, 3)))).astype(np.float32)
    P = scene.array("P", st).reshape(-1, 3)
    pick = rng.integers(0, len(P), n // 2)
    O[: n // 2] = P[pick] + scene.array("N", st).reshape(-1, 3)[pick] * 1e-3
    D = rng.normal(size=(n, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    # a few axis-aligned and grazing directions
    D[:64] = np.eye(3)[rng.integers(0, 3, 64)] * rng.choice([-1, 1], (64, 1))
    return O, D.astype(np.float32)


def _gpu_trace(torch, scene, O, D, trav, st=0):
    Ot = torch.from_numpy(O).cuda()
    Dt = torch.from_numpy(D).cuda()
    t = torch.empty(len(O), dtype=torch.float32, device="cuda")
    q = torch.empty(len(O), dtype=torch.int32, device="cuda")
    rea.trace(scene, Ot, Dt, t, q, state=st, trav=trav)
    torch.cuda.synchronize()
    return t.cpu().numpy(), q.cpu().numpy()


@pytest.mark.parametrize("which", ["room", "street", "room_longleaf"])
def test_trace_bit_exact(torch_gpu, which, request):
    scene, osc = request.getfixturevalue(which)
    for st in range(scene.info.NS):
        O, D = _rays(scene, 20000, 10 + st, st)
        t0, q0 = osc.trace(O, D, st, trav=0)
        t1, q1 = osc.trace(O, D, st, trav=1)
        t3, q3 = osc.trace(O, D, st, trav=3)   # brute force: the quantized nodes' reference
        for trav, (te, qe) in ((0, (t0, q0)), (1, (t1, q1)), (2, (t1, q1)), (3, (t3, q3))):
            tg, qg = _gpu_trace(torch_gpu, scene, O, D, trav, st)
            assert np.array_equal(qg, qe), f"trav {trav} state {st}: {np.sum(qg != qe)} hit ids differ"
            assert np.array_equal(tg.view(np.uint32), te.view(np.uint32)), f"trav {trav} state {st}: t differs"


@pytest.mark.parametrize("which", ["room", "street"])
def test_build_bit_exact(which, request):
    scene, osc = request.getfixturevalue(which)
    for st in range(scene.info.NS):
        assert np.array_equal(scene.array("EDIR", st).view(np.uint32), osc.Edir[st].view(np.uint32))
        assert np.array_equal(scene.array("SV", st), osc.SV[st])
        assert np.array_equal(scene.array("LM", st), osc.LM[st])
        if scene.info.fine > 0:
            assert np.array_equal(scene.array("VOL", st), osc.VOL[st]), f"fine volumes differ in state {st}"


def test_hash_point_and_dda(torch_gpu, street):
    torch = torch_gpu
    scene, osc = street
    rng = np.random.default_rng(5)
    P = scene.array("P", 0).reshape(-1, 3)
    N = scene.array("N", 0).reshape(-1, 3)
    pick = rng.integers(0, len(P), 5000)
    H = (P[pick] + rng.normal(scale=scene.info.h, size=(5000, 3))).astype(np.float32)
    Nn = N[pick].astype(np.float32)
    out = torch.empty(5000, dtype=torch.int32, device="cuda")
    rea.hash_point(scene, torch.from_numpy(H).cuda(), torch.from_numpy(Nn).cuda(), out)
    exp = np.array([po.lib.orc_hash_point(osc.state(0), po.ptr(H[k]), po.ptr(Nn[k])) for k in range(5000)], np.int32)
    assert np.array_equal(out.cpu().numpy(), exp)
    Aseg = H
    Bseg = P[rng.integers(0, len(P), 5000)].astype(np.float32)
    dm = torch.empty(5000 * scene.info.nw, dtype=torch.int32, device="cuda")
    rea.dda_mark(scene, torch.from_numpy(Aseg).cuda(), torch.from_numpy(Bseg).cuda(), dm)
    em = np.zeros(5000 * scene.info.nw, np.uint32)
    po.lib.orc_dda_segments(osc.state(0), po.ptr(Aseg), po.ptr(Bseg), 5000, po.ptr(em))
    assert np.array_equal(dm.cpu().numpy().view(np.uint32), em)


def test_dda_volume_bit_exact(torch_gpu, street):
    """rc_dda_volume (Amanatides-Woo over a fine bit volume, bits gathered per word before the atomicOr)
    against the oracle's walk: three volumes chosen per segment, a non-cubic 40 x 24 x 17 grid, segments
    inside, crossing and leaving the bounds, axis-aligned and degenerate ones"""
    torch = torch_gpu
    scene, osc = street
    i = scene.info
    rng = np.random.default_rng(8)
    n = 30000
    lo, ext = np.array(list(i.lo)), np.array(list(i.ext))
    A = (lo + ext * rng.uniform(-0.1, 1.1, (n, 3))).astype(np.float32)
    B = (lo + ext * rng.uniform(-0.1, 1.1, (n, 3))).astype(np.float32)
    B[:500] = A[:500]                                     # zero length
    B[500:1500, 1:] = A[500:1500, 1:]                     # along x only
    B[1500:2500, :2] = A[1500:2500, :2]                   # along z only
    sv = rng.integers(0, 3, n).astype(np.int32)
    gx, gy, gz = 40, 24, 17
    wpv = (gx * gy * gz + 31) // 32
    vol = torch.zeros(3 * wpv, dtype=torch.int32, device="cuda")
    rea.dda_volume(scene, (gx, gy, gz), torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(sv).cuda(), vol, wpv)
    exp = np.zeros(3 * wpv, np.uint32)
    po.lib.orc_dda_volume(osc.state(0), gx, gy, gz, po.ptr(A), po.ptr(B), po.ptr(sv), n, po.ptr(exp), wpv)
    got = vol.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, exp), f"{np.count_nonzero(got != exp)} words differ"
    assert np.count_nonzero(exp) > 0.5 * len(exp)


def test_hash_insert_lookup(torch_gpu, street):
    torch = torch_gpu
    scene, _ = street
    HKh = scene.array("HK")
    keys = HKh[HKh != 0]
    HVh = scene.array("HV")[HKh != 0]
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(keys))
    # duplicates: every key inserted twice with two values, the smaller must win
    k2 = np.concatenate([keys[perm], keys])
    v2 = np.concatenate([HVh[perm] + 1000000, HVh]).astype(np.int32)
    size = len(HKh)
    HK = torch.zeros(size, dtype=torch.int64, device="cuda")
    HV = torch.full((size,), 2**31 - 1, dtype=torch.int32, device="cuda")
    fail = torch.zeros(1, dtype=torch.int32, device="cuda")
    rea.hash_insert(HK, HV, torch.from_numpy(k2.view(np.int64)).cuda(), torch.from_numpy(v2).cuda(), fail)
    absent = (keys[:2000] ^ np.uint64(1 << 60)).astype(np.uint64)
    qk = np.concatenate([keys, absent])
    out = torch.empty(len(qk), dtype=torch.int32, device="cuda")
    rea.hash_query(HK, HV, torch.from_numpy(qk.view(np.int64)).cuda(), out)
    o = out.cpu().numpy()
    assert int(fail.item()) == 0
    assert np.array_equal(o[: len(keys)], HVh)
    assert np.all(o[len(keys):] == -1)
    hk = HK.cpu().numpy().view(np.uint64)
    assert np.array_equal(np.sort(hk[hk != 0]), np.sort(keys))
    # linear-probing invariant: no empty slot between a key's home slot and its slot
    mask = size - 1
    occ = hk != 0
    for slot in np.nonzero(occ)[0][:5000]:
        home = int(po.lib.orc_mix64(int(hk[slot]))) & mask
        k = home
        while k != slot:
            assert occ[k]
            k = (k + 1) & mask


def test_hash_insert_duplicates_in_one_wave(torch_gpu):
    """duplicate keys inside one wave (1 to 5 copies of a key among 64 lanes, in shuffled lanes, with
    different values): the ballot groups them and the group's minimum value is stored; every key is
    inserted once; the empty-key lanes (key 0) insert nothing"""
    torch = torch_gpu
    rng = np.random.default_rng(11)
    distinct = rng.integers(1, 2**63, 6000, dtype=np.uint64)
    keys, vals = [], []
    for w in range(0, len(distinct), 20):              # 20 distinct keys per wave of 64 lanes
        ks = distinct[w:w + 20]
        reps = rng.integers(1, 6, len(ks))
        lane_k = np.repeat(ks, reps)[:64]
        lane_k = np.concatenate([lane_k, np.zeros(64 - len(lane_k), np.uint64)])
        rng.shuffle(lane_k)
        keys.append(lane_k)
        vals.append(rng.integers(0, 1 << 30, 64).astype(np.int32))
    keys, vals = np.concatenate(keys), np.concatenate(vals)
    live = keys != 0
    exp = {}
    for k, v in zip(keys[live], vals[live]):
        exp[int(k)] = min(exp.get(int(k), 1 << 31), int(v))
    size = 1 << 15
    HK = torch.zeros(size, dtype=torch.int64, device="cuda")
    HV = torch.full((size,), 2**31 - 1, dtype=torch.int32, device="cuda")
    fail = torch.zeros(1, dtype=torch.int32, device="cuda")
    rea.hash_insert(HK, HV, torch.from_numpy(keys.view(np.int64)).cuda(), torch.from_numpy(vals).cuda(), fail)
    q = np.array(sorted(exp), np.uint64)
    out = torch.empty(len(q), dtype=torch.int32, device="cuda")
    rea.hash_query(HK, HV, torch.from_numpy(q.view(np.int64)).cuda(), out)
    assert int(fail.item()) == 0
    assert np.array_equal(out.cpu().numpy(), np.array([exp[int(k)] for k in q], np.int32))
    hk = HK.cpu().numpy().view(np.uint64)
    assert np.count_nonzero(hk) == len(exp)


RUN_METHODS = [
    ("plain", lambda: rea.methods.plain()),
    ("targeted", lambda: rea.methods.targeted()),
    ("targeted_grad", lambda: rea.methods.targeted_grad(rays=3)),
    ("flush_snapshots", lambda: rea.methods.flush_snapshots()),
    ("flush_snapshots_compact", lambda: rea.methods.flush_snapshots_compact()),
] + [(n, (lambda m=m: m)) for n, m in rea.methods.timing_list()] + [
    (n, (lambda m=m: m)) for n, m in rea.methods.scenario_list(dyn=0b11) if "Oracle" not in n]

SEQ = {"room": [0] * 22 + [1] * 22 + [0] * 22, "street": [0] * 21 + [3] * 12 + [2] * 12 + [0] * 12 + [1] * 6}


def _oracle_grad_masks(scene, osc, rays):
    """the gradient-support masks a grad_masks cache computes at begin: adjoint forward of every
    state with unit emission and the material albedos, 3 bounces, seed 0x5eed0000 + state"""
    i = scene.info
    alb = scene.array("MATALB")
    return np.stack([po.OracleAdjoint(osc, st, rays, 3, 0x5EED0000 + st).forward32(np.ones(i.nl, np.float32), alb)[1]
                     for st in range(i.NS)])


@pytest.mark.parametrize("name,mk", RUN_METHODS, ids=[n for n, _ in RUN_METHODS])
@pytest.mark.parametrize("which", ["room", "street"])
def test_runner_bit_exact(torch_gpu, which, name, mk, request):
    scene, osc = request.getfixturevalue(which)
    m = mk()
    cache = rea.Cache(scene, m)
    orun = po.OracleRunner(osc, m)
    if m.grad_masks:
        orun.set_grad_masks(_oracle_grad_masks(scene, osc, m.grad_masks))
    seq = SEQ[which]
    cache.begin(seq[0], 77)
    orun.begin(seq[0], 77)
    for f, st in enumerate(seq):
        fs = cache.frame(st)
        _, inv = orun.frame(st)
        assert fs.invalidated == inv, f"frame {f}: invalidated {fs.invalidated} vs {inv}"
        for arr in ("LC", "M", "R", "N", "KK", "PENDING", "CUR", "LCUR", "SUMPR", "NPEND", "AGE", "STALE"):
            g, e = cache.get(arr), orun.get(arr)
            if g.dtype == np.float32:
                g, e = g.view(np.uint32), e.view(np.uint32)
            assert np.array_equal(g, e), f"{name} frame {f} state {st}: {arr} differs in {np.sum(g != e)} entries"


@pytest.mark.parametrize("which", ["street", "room_longleaf"])
def test_query_bit_exact(torch_gpu, which, request):
    """frames of the plain cache, then the camera pass with and without the cache, bit for bit (on the room
    whose BVH is one long leaf too: the entry pass's and camera pass's walks over the split wide nodes)"""
    torch = torch_gpu
    scene, osc = request.getfixturevalue(which)
    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 9)
    orun.begin(0, 9)
    for _ in range(3):
        cache.frame(0)
        orun.frame(0)
        for name in ("LC", "N", "KK"):
            assert np.array_equal(cache.get(name), orun.get(name)), f"cache array {name} differs"
    cam = _CAMERAS.get(which) or scene.camera()
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    for no_cache in (0, 1):
        rays = cache.render(img, cam, no_cache=bool(no_cache), seed=4, count_rays=True)
        R = orun.get("R")
        exp = np.zeros(cam.W * cam.H, np.float32)
        er = (C.c_longlong * 2)()
        po.lib.orc_k_query(osc.state(0), 1, None if no_cache else po.ptr(R), C.byref(cam), 4, 3, 0, po.ptr(exp), er, 0,
                           cam.W * cam.H)
        got = img.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"no_cache={no_cache}: {np.sum(got != exp)} pixels differ"
        assert tuple(rays) == (er[0], er[1])


def test_render_on_a_side_stream(torch_gpu, street):
    """the camera pass is ordered against torch's current stream: a fill queued before it on a side
    stream does not overwrite the image, a copy queued after it sees the whole image, and a clear
    queued after it is not overwritten by the pass"""
    torch = torch_gpu
    scene, _ = street
    cache = rea.Cache(scene, rea.methods.plain())
    cache.begin(0, 3)
    cache.frame(0)
    cam = scene.camera()
    ref = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    cache.render(ref, cam, seed=17)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    img = torch.empty_like(ref)
    with torch.cuda.stream(side):
        img.fill_(-1.0)
        cache.render(img, cam, seed=17)
        snap = img.clone()
        img.zero_()
    side.synchronize()
    assert torch.equal(snap, ref)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(img)) == 0


def test_query_stack_levels(torch_gpu, street):
    """The camera pass's traversal stacks keep their first levels in LDS and the deeper ones in a
    global buffer: the image and ray counts do not depend on the split (1 and 2 levels put nearly
    every push of every ray in the global buffer; 64 keeps every level in LDS)"""
    torch = torch_gpu
    scene, _ = street
    cache = rea.Cache(scene, rea.methods.plain())
    cache.begin(0, 9)
    for _ in range(2):
        cache.frame(0)
    cam = scene.camera()
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    out = {}
    for levels in (64, 1, 2, 5, 16):
        cache.set_stack_levels(levels)
        rays = cache.render(img, cam, seed=6, count_rays=True)
        out[levels] = (img.cpu().numpy().view(np.uint32).copy(), tuple(rays))
    for levels, (im, rays) in out.items():
        assert np.array_equal(im, out[64][0]), f"levels={levels}: {np.sum(im != out[64][0])} pixels differ"
        assert rays == out[64][1], levels
    with pytest.raises(rea.RcError):
        cache.set_stack_levels(0)


@pytest.mark.parametrize("mk", [rea.methods.targeted, rea.methods.flush_snapshots, lambda: dict(rea.methods.timing_list())["v4-A2C suivi 1/4"]])
def test_event_call_equals_the_frames_event(torch_gpu, street, mk):
    """rc_cache_event run between frames gives, frame by frame, the arrays and reset counts of the event
    block that rc_cache_frame runs itself on a state change"""
    scene, _ = street
    seq = [0] * 21 + [3] * 2 + [0] * 2 + [2] * 2 + [1] * 2
    a, b = rea.Cache(scene, mk()), rea.Cache(scene, mk())
    a.begin(seq[0], 13)
    b.begin(seq[0], 13)
    names = ("N", "KK", "R", "AGE", "LC", "M", "NPEND", "SUMPR")
    for f, st in enumerate(seq):
        inv_a = a.event(st) if f > 0 and seq[f - 1] != st else 0
        fa = a.frame(st)
        fb = b.frame(st)
        assert fa.event == 0
        assert inv_a == fb.invalidated, f"frame {f}: {inv_a} vs {fb.invalidated}"
        for arr in names + (("LCUR", "CUR") if a.method.deps else ()):
            ga, gb = a.get(arr), b.get(arr)
            if ga.dtype == np.float32:
                ga, gb = ga.view(np.uint32), gb.view(np.uint32)
            assert np.array_equal(ga, gb), f"frame {f} state {st}: {arr} differs in {np.sum(ga != gb)} entries"


def test_invalidate_and_resolve_calls_against_oracle(torch_gpu, street):
    """rc_cache_invalidate with an explicit light set and voxel list resets exactly the oracle's entries
    (orc_k_invalidate on the same recorded masks), and rc_cache_resolve gives the oracle's resolved radiance"""
    scene, osc = street
    m = rea.methods.targeted()
    cache = rea.Cache(scene, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 21)
    orun.begin(0, 21)
    for _ in range(4):
        cache.frame(0)
        orun.frame(0)
    i = scene.info
    for mask, vox in ((1 << 2, []), (0, [5, 77, 100]), ((1 << 0) | (1 << 63), [3])):
        n = cache.invalidate(mask, vox)
        c = po.lib.orc_runner_cache(orun.h)
        npend = C.cast(c.contents.npend, C.POINTER(C.c_int32))
        before = int(npend[1])
        v = np.asarray(vox, np.int32)
        po.lib.orc_k_invalidate(c, i.Np, i.nl, mask, po.ptr(v) if len(v) else None, len(v), i.nw)
        assert n == int(npend[1]) - before and n > 0, (mask, vox, n)
        for arr in ("LC", "N", "AGE"):
            g, e = cache.get(arr), orun.get(arr)
            if g.dtype == np.float32:
                g, e = g.view(np.uint32), e.view(np.uint32)
            assert np.array_equal(g, e), f"{arr} after invalidating {mask:#x} {vox}: {np.sum(g != e)} entries differ"
    cache.resolve()
    R = np.zeros(i.Np, np.float32)
    Lc = orun.get("LC")
    po.lib.orc_k_resolve(osc.state(0), po.ptr(Lc), po.ptr(R), 0)
    assert np.array_equal(cache.get("R").view(np.uint32), R.view(np.uint32))
    with pytest.raises(rea.RcError):
        cache.invalidate(0, [32 * i.nw])


@pytest.mark.parametrize("which", ["room", "street"])
def test_query_quantized_nodes(torch_gpu, which, request):
    """The camera pass over the quantized wide nodes (traversal mode 3) against the oracle's camera
    pass with brute-force closest hits (every triangle tested), bit for bit with the same ray counts, in
    every state, with and without the cache, with the lane stacks in LDS and nearly all in global memory
    (1 LDS level). The quantized boxes hold one step of slack on every side, so no box holding the
    closest hit is culled; the fp32 slab test can cull one at a grazing edge (one path of the room's
    524,288 in state 1: the fp32 modes then follow the oracle's fp32 traversal instead)."""
    torch = torch_gpu
    scene, osc = request.getfixturevalue(which)
    cache = rea.Cache(scene, rea.methods.plain())
    cache.begin(0, 5)
    cam = scene.camera()
    npix = cam.W * cam.H
    try:
        scene.set_trav(rea.TRAV_WIDE_Q)
        for st in range(scene.info.NS):
            cache.frame(st)
            R = cache.get("R")
            for spp, nb, cb in ((4, 3, 1), (1, 3, 5), (2, 1, 1)):
                cam.spp, cam.bounces, cam.cache_bounce = spp, nb, cb
                for no_cache in (False, True):
                    exp = np.zeros(npix, np.float32)
                    er = (C.c_longlong * 2)()
                    po.lib.orc_k_query(osc.state(st), 3, None if no_cache else po.ptr(R), C.byref(cam), 21 + st, st + 1, 0,
                                       po.ptr(exp), er, 0, npix)
                    for levels in (64, 1):
                        cache.set_stack_levels(levels)
                        img = torch.full((npix,), -1.0, dtype=torch.float32, device="cuda")
                        rays = cache.render(img, cam, no_cache=no_cache, seed=21 + st, count_rays=True)
                        got = img.cpu().numpy()
                        tag = f"state {st} spp {spp} bounces {nb} cache_bounce {cb} no_cache {no_cache} levels {levels}"
                        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), \
                            f"{tag}: {np.sum(got != exp)} pixels differ"
                        assert tuple(rays) == (er[0], er[1]), tag
    finally:
        scene.set_trav(rea.TRAV_WIDE)


@pytest.mark.parametrize("which", ["room", "street"])
def test_query_staged_equals_fused(torch_gpu, which, request):
    """The staged camera pass (per-segment compacted ray lists, one trace kernel per stage)
    gives the fused pass's image bit for bit and the same ray counts, in every state (dynamic
    triangles included), with and without the cache, over spp / bounce depths."""
    torch = torch_gpu
    scene, osc = request.getfixturevalue(which)
    cache = rea.Cache(scene, rea.methods.plain())
    cache.begin(0, 5)
    cam = scene.camera()
    cam.W, cam.H = min(cam.W, 128), min(cam.H, 96)
    for st in range(scene.info.NS):
        cache.frame(st)
        for spp, nb, cb in ((1, 0, 1), (2, 1, 1), (3, 3, 2), (4, 3, 5), (1, 2, 0)):
            cam.spp, cam.bounces, cam.cache_bounce = spp, nb, cb
            for no_cache in (False, True):
                imgs, rays = [], []
                for mode in (rea.RENDER_FUSED, rea.RENDER_STAGED):
                    cache.set_render_mode(mode)
                    img = torch.full((cam.W * cam.H,), -1.0, dtype=torch.float32, device="cuda")
                    rays.append(cache.render(img, cam, no_cache=no_cache, seed=11 + st, count_rays=True))
                    imgs.append(img.cpu().numpy())
                tag = f"state {st} spp {spp} bounces {nb} cache_bounce {cb} no_cache {no_cache}"
                assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32)), \
                    f"{tag}: {np.sum(imgs[0] != imgs[1])} pixels differ"
                assert rays[0] == rays[1], tag
    cache.set_render_mode(rea.RENDER_FUSED)


@pytest.mark.parametrize("which", ["room", "street"])
def test_radiance_vs_fp64(torch_gpu, which, request):
    """FP32 cache channels, resolved radiance and camera image after several frames within 1e-4
    relative RMS of the FP64 twin driven by the same rays and hits (the FP32 state bit-equal to the
    oracle's first)."""
    torch = torch_gpu
    scene, osc = request.getfixturevalue(which)
    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    cache.begin(0, 21)
    orun = po.OracleRunner(osc, m)
    orun.begin(0, 21)
    twin = po.FP64Twin(osc, orun, 0)
    for f in range(8):
        cache.frame(0)
        twin.frame()
    assert np.array_equal(cache.get("LC").view(np.uint32), orun.get("LC").view(np.uint32))
    rl = po.rel_rms(cache.get("LC"), twin.L64)
    rr = po.rel_rms(cache.get("R"), twin.R64())
    cam = scene.camera()
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    cache.render(img, cam, seed=31)
    ri = po.rel_rms(img.cpu().numpy(), twin.image64(cam, 31, 8, 0, cam.W * cam.H))
    print(f"fp64 twin {which}: LC {rl:.2e}, R {rr:.2e}, image {ri:.2e}", flush=True)
    assert rl < 1e-4 and rr < 1e-4 and ri < 1e-4, (rl, rr, ri)


def test_query_any_spp(torch_gpu, room):
    """spp that does not divide 64 takes the per-pixel loop kernel; same result as the oracle"""
    torch = torch_gpu
    scene, osc = room
    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 2)
    orun.begin(0, 2)
    cam = scene.camera()
    for spp, nb in ((3, 1), (5, 2), (1, 0)):
        cam.spp, cam.bounces = spp, nb
        img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
        cache.render(img, cam, seed=8)   # before any frame: resolved from the warm-start channels
        R = np.zeros(osc.Np, np.float32)
        po.lib.orc_k_resolve(osc.state(0), po.ptr(orun.get("LC")), po.ptr(R), 0)
        exp = np.zeros(cam.W * cam.H, np.float32)
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(R), C.byref(cam), 8, 0, 0, po.ptr(exp), None, 0, cam.W * cam.H)
        assert np.array_equal(img.cpu().numpy().view(np.uint32), exp.view(np.uint32)), f"spp {spp}"


def test_snapshot_records_round_trip(torch_gpu, street):
    """24-byte records of the entries with samples, in entry order, radiance = the channel sum in
    light order, keys = the hash-grid keys; loading them into another cache restores channel 0,
    history and age after a flush, and skips records whose key does not match."""
    scene, osc = street
    i = scene.info
    cache = rea.Cache(scene, rea.methods.plain())
    cache.begin(0, 4)
    for _ in range(5):
        cache.frame(0)
    n, lc, age = cache.get("N"), cache.get("LC").reshape(i.Np, i.nl), cache.get("AGE")
    rec = cache.snapshot_save()
    assert rea.SNAP_REC.itemsize == 24
    valid = np.nonzero(n > 0)[0]
    assert np.array_equal(rec["index"], valid)
    s = np.zeros(i.Np, np.float32)
    for l in range(i.nl):
        s = (s + lc[:, l]).astype(np.float32)
    assert np.array_equal(rec["radiance"].view(np.uint32), s[valid].view(np.uint32))
    assert np.array_equal(rec["n"], n[valid]) and np.array_equal(rec["age"], age[valid])
    ek = (np.uint64(1) << np.uint64(63)) | np.arange(i.Np, dtype=np.uint64)
    HK, HV = scene.array("HK"), scene.array("HV")
    occ = HK != 0
    ek[HV[occ]] = HK[occ]
    assert np.array_equal(rec["key"], ek[valid])
    other = rea.Cache(scene, rea.methods.plain())
    other.begin(1, 9)
    other.frame(1)
    bad = rec.copy()
    bad["key"][:10] ^= np.uint64(1)          # ten records that no longer match their entry
    other.snapshot_load(bad)
    n2, lc2, age2 = other.get("N"), other.get("LC").reshape(i.Np, i.nl), other.get("AGE")
    ok = valid[10:]
    assert np.array_equal(lc2[ok, 0].view(np.uint32), s[ok].view(np.uint32))
    assert np.all(lc2[ok, 1:] == 0) and np.array_equal(n2[ok], n[ok]) and np.array_equal(age2[ok], age[ok])
    assert np.all(n2[valid[:10]] == 0) and np.all(lc2[valid[:10]] == 0)


def test_query_tiles_odd_sizes(torch_gpu, room):
    """the per-path camera pass takes square pixel tiles per wave and 2 x 2 tiles per block; image
    sizes that no tile divides leave part of the edge tiles idle, and every spp that divides 64
    (tile shapes 8 x 8 down to 1 x 1 pixels) gives the oracle's image"""
    torch = torch_gpu
    scene, osc = room
    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 3)
    orun.begin(0, 3)
    R = np.zeros(osc.Np, np.float32)
    po.lib.orc_k_resolve(osc.state(0), po.ptr(orun.get("LC")), po.ptr(R), 0)
    cam = scene.camera()
    for W, H, spp in ((67, 45, 4), (33, 7, 1), (19, 23, 2), (9, 5, 16), (13, 3, 64)):
        cam.W, cam.H, cam.spp, cam.bounces = W, H, spp, 1
        img = torch.full((W * H,), -1.0, dtype=torch.float32, device="cuda")
        cache.render(img, cam, seed=11)   # before any frame: resolved from the warm-start channels
        exp = np.zeros(W * H, np.float32)
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(R), C.byref(cam), 11, 0, 0, po.ptr(exp), None, 0, W * H)
        got = img.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"{W}x{H} spp {spp}: {np.sum(got != exp)} pixels differ"

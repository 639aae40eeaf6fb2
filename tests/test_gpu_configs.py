"""GPU parity at the full shapes of the measured workloads, through the C ABI against the CPU oracle.

* street 960x540: 2,000 boxes / ~200k triangles, 32 lamps + sun (33 lights: the 9-elements-per-thread
  update), ~530k cache entries in a 2^22 table, 128^3 light volumes;
* street 1920x1080: ~1M triangles, 64 lamps + sun (65 lights: the widest update instantiation; lamp 63
  and the sun share light-mask bit 63), ~1M entries in a 2^23 table, 256^3 light volumes;
* the adjoint at 63 lamps + sun (64 emission scalars) and 4,096 material albedos;
* the learned cache at a 2^19 table, a 65,536-sample batch and a 1920x1080 query.

The state build (direct light, transit masks, light masks, fine volumes) is checked on blocks of entries
spread over the range (the oracle's per-entry shadow rays at 65 lights x 1M entries would take minutes);
the frame sequence (entry pass, update, event blocks of the plain / targeted / flush + snapshot methods
over occluder, revert, lamp-off and night events) runs on every entry and is bit-exact frame by frame;
the camera pass is compared on row bands spread over the image."""
import ctypes as C
import time

import numpy as np
import pytest

import render_engine_amd as rea
import pyoracle as po

pytestmark = pytest.mark.gpu

SHAPES = {
    "street960": dict(seed=1, nboxes=2000, target_tris=200_000, nlamps=32, sun=True, target_points=530_000, table_log2=22,
                      W=960, H=540, spp=4, bounces=3),
    "street1080": dict(seed=1, nboxes=2000, target_tris=1_000_000, nlamps=64, sun=True, target_points=1_000_000, table_log2=23,
                       W=1920, H=1080, spp=4, bounces=3),
}
_CACHE = {}


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def _shape(torch_gpu, name):
    """scene of one shape, uploaded (GPU state build), with an oracle view whose state arrays are
    filled from the GPU after the block check (test_state_build_blocks checks them first)"""
    if name not in _CACHE:
        t0 = time.time()
        s = rea.Scene.street(**SHAPES[name]).upload(trav=rea.TRAV_WIDE)
        osc = po.OracleScene(s, build=False)
        _CACHE[name] = (s, osc, {"filled": False})
        _say(f"{name}: {s.info.nst} triangles, {s.info.Np} entries, {s.info.nl} lights, table {s.info.hsize}, "
             f"fine {s.info.fine}, {time.time() - t0:.1f} s")
    return _CACHE[name]


def _blocks(Np, nblk=16, size=2048):
    return [(int(b * Np // nblk), int(min(Np, b * Np // nblk + size))) for b in range(nblk)]


def _fill(s, osc, flags):
    if flags["filled"]:
        return
    for st in range(s.info.NS):
        osc.Edir[st][:] = s.array("EDIR", st)
        osc.SV[st][:] = s.array("SV", st)
        osc.LM[st][:] = s.array("LM", st)
        if s.info.fine > 0:
            osc.VOL[st][:] = s.array("VOL", st)
    flags["filled"] = True


@pytest.mark.parametrize("name", list(SHAPES))
def test_state_build_blocks(torch_gpu, name):
    """k_build on 16 blocks of 2,048 entries per state: direct light and transit / light masks bit for
    bit; every fine-volume bit the oracle marks for those entries' lit shadow segments is set on the GPU"""
    s, osc, flags = _shape(torch_gpu, name)
    i = s.info
    wpv = (i.fine ** 3 + 31) // 32 if i.fine > 0 else 0
    for st in range(i.NS):
        E = np.zeros(i.Np * i.nl, np.float32)
        SV = np.zeros(i.Np * i.nw, np.uint32)
        LM = np.zeros(i.Np, np.uint64)
        VOL = np.zeros(max(1, i.nl * wpv), np.uint32)
        for a, b in _blocks(i.Np):
            po.lib.orc_k_build_rows(osc.state(st), 1, a, b, po.ptr(E), po.ptr(SV), po.ptr(LM), po.ptr(VOL) if wpv else None)
        gE, gSV, gLM = s.array("EDIR", st), s.array("SV", st), s.array("LM", st)
        for a, b in _blocks(i.Np):
            assert np.array_equal(gE[a * i.nl:b * i.nl].view(np.uint32), E[a * i.nl:b * i.nl].view(np.uint32)), \
                f"{name} state {st}: direct light differs in entries {a}..{b}"
            assert np.array_equal(gSV[a * i.nw:b * i.nw], SV[a * i.nw:b * i.nw]), f"{name} state {st}: transit masks differ"
            assert np.array_equal(gLM[a:b], LM[a:b]), f"{name} state {st}: light masks differ"
        if wpv:
            gV = s.array("VOL", st)
            assert VOL.any()
            assert np.array_equal(gV & VOL, VOL), f"{name} state {st}: {np.count_nonzero(VOL & ~gV)} volume words miss bits"
        _say(f"{name} state {st}: build blocks equal")
    if i.nl > 64:   # lamp 63 and the sun share light-mask bit 63
        lm = s.array("LM", 0)
        assert np.count_nonzero(lm >> np.uint64(63)) > 0


TIMELINES = {
    # (method, states per frame): occluder inserted, reverted, lamp 0 off, night
    "plain": [0] * 3 + [3] * 2 + [2] * 2 + [1] * 2,
    "targeted": [0] * 3 + [3] * 2 + [0] * 2 + [2] * 2 + [1] * 2,
    # 21 frames in state 0 make it a snapshot slot; the revisit restores it
    "flush_snapshots": [0] * 21 + [3] * 2 + [0] * 2 + [2] * 2,
}
METHODS = {"plain": rea.methods.plain, "targeted": rea.methods.targeted, "flush_snapshots": rea.methods.flush_snapshots}


def _eq(g, e):
    if g.dtype == np.float32:
        g, e = g.view(np.uint32), e.view(np.uint32)
    return np.array_equal(g, e), int(np.count_nonzero(g != e))


@pytest.mark.parametrize("method", list(TIMELINES))
@pytest.mark.parametrize("name", list(SHAPES))
def test_runner_full_shape(torch_gpu, name, method):
    """every cache array of every entry after every frame, bit for bit against the oracle runner"""
    s, osc, flags = _shape(torch_gpu, name)
    _fill(s, osc, flags)
    m = METHODS[method]()
    cache = rea.Cache(s, m)
    orun = po.OracleRunner(osc, m)
    seq = TIMELINES[method]
    cache.begin(seq[0], 31)
    orun.begin(seq[0], 31)
    t0 = time.time()
    for f, st in enumerate(seq):
        fs = cache.frame(st)
        _, inv = orun.frame(st)
        assert fs.invalidated == inv, f"{name} {method} frame {f}: invalidated {fs.invalidated} vs {inv}"
        event = f > 0 and seq[f - 1] != st
        names = ("N", "KK", "R", "AGE", "SUMPR", "NPEND") + (("LCUR",) if m.deps else ())
        if event or f == len(seq) - 1 or f == 0:
            names += ("LC", "M") + (("CUR",) if m.deps else ())
        for arr in names:
            ok, nd = _eq(cache.get(arr), orun.get(arr))
            assert ok, f"{name} {method} frame {f} state {st}: {arr} differs in {nd} entries"
    _say(f"{name} {method}: {len(seq)} frames bit-exact ({time.time() - t0:.1f} s)")


@pytest.mark.parametrize("name", list(SHAPES))
def test_query_full_shape_bands(torch_gpu, name):
    """the full-size camera pass (fused, and staged) on 8 row bands of 8 rows against the oracle's
    paths, after three frames of the plain cache; R equal first"""
    torch = torch_gpu
    s, osc, flags = _shape(torch_gpu, name)
    _fill(s, osc, flags)
    m = rea.methods.plain()
    cache = rea.Cache(s, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 12)
    orun.begin(0, 12)
    for _ in range(3):
        cache.frame(0)
        orun.frame(0)
    R = orun.get("R")
    assert _eq(cache.get("R"), R)[0]
    cam = s.camera()
    W, H = cam.W, cam.H
    exp = np.zeros(W * H, np.float32)
    bands = [(b * H // 8, min(H, b * H // 8 + 8)) for b in range(8)]
    for r0, r1 in bands:
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(R), C.byref(cam), 40, 3, 0, po.ptr(exp), None, r0 * W, r1 * W)
    whole = {}
    for mode in (rea.RENDER_FUSED, rea.RENDER_STAGED):
        cache.set_render_mode(mode)
        img = torch.full((W * H,), -1.0, dtype=torch.float32, device="cuda")
        cache.render(img, cam, seed=40)
        got = img.cpu().numpy()
        whole[mode] = got
        for r0, r1 in bands:
            a, b = got[r0 * W:r1 * W], exp[r0 * W:r1 * W]
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                f"{name} mode {mode} rows {r0}..{r1}: {np.count_nonzero(a != b)} pixels differ"
        assert np.all(np.isfinite(got)) and np.all(got >= 0)
    # the quantized wide nodes (conservative boxes: brute-force closest hits) give the fp32 nodes' image
    # but where the fp32 slab test culls a box holding the closest hit at a grazing edge (a path in
    # several hundred thousand on the small scenes)
    cache.set_render_mode(rea.RENDER_FUSED)
    s.set_trav(rea.TRAV_WIDE_Q)
    try:
        img = torch.full((W * H,), -1.0, dtype=torch.float32, device="cuda")
        cache.render(img, cam, seed=40)
        got = img.cpu().numpy()
    finally:
        s.set_trav(rea.TRAV_WIDE)
    ref = whole[rea.RENDER_FUSED]
    nd = int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
    assert nd <= W * H // 20000, f"{name} quantized nodes: {nd} pixels differ"
    assert np.all(np.isfinite(got)) and np.all(got >= 0)
    _say(f"{name}: camera pass bands bit-exact ({W}x{H} x {cam.spp} spp); quantized nodes: {nd} pixels differ from the fp32 nodes'")


@pytest.mark.parametrize("name", list(SHAPES))
def test_radiance_vs_fp64_full_shape(torch_gpu, name):
    """FP32 cache channels, resolved radiance and the camera image (row bands) after three frames of the
    plain cache within 1e-4 relative RMS of the FP64 twin replaying the same rays and hits"""
    torch = torch_gpu
    s, osc, flags = _shape(torch_gpu, name)
    _fill(s, osc, flags)
    m = rea.methods.plain()
    cache = rea.Cache(s, m)
    orun = po.OracleRunner(osc, m)
    cache.begin(0, 44)
    orun.begin(0, 44)
    twin = po.FP64Twin(osc, orun, 0)
    for _ in range(3):
        cache.frame(0)
        twin.frame()
    assert _eq(cache.get("LC"), orun.get("LC"))[0]
    rl = po.rel_rms(cache.get("LC"), twin.L64)
    rr = po.rel_rms(cache.get("R"), twin.R64())
    cam = s.camera()
    W, H = cam.W, cam.H
    img = torch.zeros(W * H, dtype=torch.float32, device="cuda")
    cache.render(img, cam, seed=45)
    got = img.cpu().numpy()
    a, b = [], []
    for k in range(4):
        r0 = k * H // 4
        b.append(twin.image64(cam, 45, 3, r0 * W, (r0 + 8) * W)[r0 * W:(r0 + 8) * W])
        a.append(got[r0 * W:(r0 + 8) * W])
    ri = po.rel_rms(np.concatenate(a), np.concatenate(b))
    _say(f"{name} fp64 twin: LC {rl:.2e}, R {rr:.2e}, image {ri:.2e}")
    assert rl < 1e-4 and rr < 1e-4 and ri < 1e-4, (rl, rr, ri)


def test_query_stack_offsets_past_32_bits(torch_gpu):
    """1920x1080 at 16 spp with one LDS stack level: the global levels would span more than 4 GiB, past
    the split stacks' 32-bit byte offsets, so the pass keeps every level in LDS; the image equals the
    all-LDS pass's and a split that fits (16 LDS levels)"""
    torch = torch_gpu
    s, osc, flags = _shape(torch_gpu, "street1080")
    cache = rea.Cache(s, rea.methods.plain())
    cache.begin(0, 2)
    cache.frame(0)
    cam = s.camera()
    cam.spp = 16
    out = {}
    for levels in (64, 1, 16):
        cache.set_stack_levels(levels)
        img = torch.full((cam.W * cam.H,), -1.0, dtype=torch.float32, device="cuda")
        cache.render(img, cam, seed=3)
        out[levels] = img.cpu().numpy().view(np.uint32).copy()
    assert np.array_equal(out[1], out[64]) and np.array_equal(out[16], out[64])
    _say("1080p x 16 spp: split stacks past 32-bit offsets fall back to LDS")


# ---------------------------------------------------------------- adjoint at 64 emission scalars
def test_adjoint_64_lights_4096_materials(torch_gpu):
    """63 lamps + sun = 64 emission scalars (the sun on bit 63 alone), 4,096 materials, 3 bounces x 4 rays:
    unit direct light, forward radiance, L^(B) and support masks bit-exact; gradients within 1e-4
    relative L2 of the FP64 reverse mode on the same hits"""
    import render_engine_amd.adjoint as ad

    torch = torch_gpu
    s = rea.Scene.street(seed=1, nboxes=2000, target_tris=200_000, nlamps=63, sun=True, target_points=530_000, table_log2=22,
                         W=960, H=540, spp=2, bounces=3).upload(trav=rea.TRAV_WIDE)
    i = s.info
    assert i.nl == 64 and i.nmat == 4096
    osc = po.OracleScene(s, build=False)
    K, B, seed = 4, 3, 17
    t0 = time.time()
    adj = ad.Adjoint(s, state=0, K=K, bounces=B, seed=seed)
    oadj = po.OracleAdjoint(osc, 0, K, B, seed)
    _say(f"adjoint 64 lights: oracle unit light + hits {time.time() - t0:.1f} s")
    assert np.array_equal(adj.get(ad.ADJ_G).view(np.uint32), oadj.G.view(np.uint32))
    assert np.array_equal(adj.get(ad.ADJ_SG), oadj.SG)
    rng = np.random.default_rng(4)
    I = (s.array("INTEN", 0) * (0.5 + rng.random(i.nl))).astype(np.float32)
    alb = s.array("MATALB").astype(np.float32)
    g = rng.random(i.Np).astype(np.float32)
    R = adj.forward(torch.from_numpy(I).cuda(), torch.from_numpy(alb).cuda())
    R32, SR, LB = oadj.forward32(I, alb)
    assert np.array_equal(R.cpu().numpy().view(np.uint32), R32.view(np.uint32))
    assert np.array_equal(adj.get(ad.ADJ_LB).view(np.uint32), LB.view(np.uint32))
    assert np.array_equal(adj.support(), SR)
    assert np.count_nonzero(SR >> np.uint64(63)) > 0      # entries depending on the sun
    gI, galb = adj.backward(torch.from_numpy(g).cuda())
    _, R64, dI64, dalb64 = oadj.grad64(I.astype(np.float64), alb.astype(np.float64), g.astype(np.float64))
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    assert rel(R.cpu().numpy().astype(np.float64), R64) < 1e-5
    assert rel(gI.cpu().numpy().astype(np.float64), dI64) < 1e-4
    assert rel(galb.cpu().numpy().astype(np.float64), dalb64) < 1e-4
    assert np.count_nonzero(dalb64) > 1000


# ---------------------------------------------------------------- learned cache at its full shape
def test_mlp_full_shape(torch_gpu):
    """2^19-entry tables, 16 levels 16..2048, 65,536-sample batches: inference and gradients within 2e-2
    relative L2 of the float32 CPU restatement, online training from the 960x540 street lowers the
    error judged against the grid cache, and the 1920x1080 learned-cache camera pass equals the grid
    cache's image bit for bit with the network and the resolved radiance set to one constant"""
    import render_engine_amd.mlp as M
    from test_gpu_mlp import backward_ref, encode_ref, forward_ref

    torch = torch_gpu
    s, osc, flags = _shape(torch_gpu, "street960")
    i = s.info
    cfg = M.config(list(i.lo), list(i.ext), log2_table=19, base_res=16, max_res=2048, max_batch=1 << 16, lr=1e-2, seed=5)
    mlp = M.MLPCache(cfg)
    rng = np.random.default_rng(2)
    mlp.set(M.MLP_TABLE, rng.uniform(-1, 1, M.LEVELS * 2 * (1 << 19)).astype(np.float32))
    n = 1 << 16
    P = np.zeros((n, 4), np.float32); Nn = np.zeros((n, 4), np.float32); D = np.zeros((n, 4), np.float32)
    pick = rng.integers(0, i.Np, n)
    P[:, :3] = s.array("P", 0).reshape(-1, 3)[pick]
    Nn[:, :3] = s.array("N", 0).reshape(-1, 3)[pick]
    v = rng.normal(size=(n, 3)); D[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    Pt, Nt, Dt = (torch.from_numpy(a).cuda() for a in (P, Nn, D))
    y = mlp.infer(Pt, Nt, Dt).cpu().numpy()
    X, corners = encode_ref(cfg, mlp.get(M.MLP_TABLE), P, Nn, D)
    W = mlp.get(M.MLP_W)
    yr, acts = forward_ref(W, X)
    rel = lambda a, b: float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))
    assert np.all(np.isfinite(y)) and rel(y, yr) < 2e-2, rel(y, yr)
    dy = rng.normal(size=n).astype(np.float32) / n
    mlp.zero_grad()
    mlp.backward(Pt, Nt, Dt, torch.from_numpy(dy).cuda())
    gWr, gTr = backward_ref(cfg, W, acts, corners, dy)
    assert rel(mlp.get(M.MLP_GW), gWr) < 2e-2
    assert rel(mlp.get(M.MLP_GTABLE), gTr) < 2e-2
    mlp.zero_grad()
    _say("mlp 2^19 / 65536: inference and gradients match")
    # online training from the scene at the full batch
    cache = rea.Cache(s, rea.methods.plain())
    cache.begin(0, 3)
    for _ in range(8):
        cache.frame(0)
    fresh = M.MLPCache(M.config(list(i.lo), list(i.ext), log2_table=19, base_res=16, max_res=2048, max_batch=1 << 16, lr=1e-2))
    judges = []
    for it in range(60):
        cache.frame(0)
        loss, j = fresh.train_scene(cache, n=n, seed=300 + it)
        assert np.isfinite(loss) and np.isfinite(j)
        judges.append(j)
    assert np.mean(judges[-5:]) < 0.6 * np.mean(judges[:3]), judges[::10]
    # 1080p pass ending on a constant network == the grid's pass ending on the same constant R
    c0 = 0.625
    Wc = np.zeros(M.NW, np.float32)
    for L in range(M.HIDDEN):
        Wc[L * 4096 + (38 if L == 0 else 0) * 64] = 1.0
    Wc[M.HIDDEN * 4096] = c0
    mlp.set(M.MLP_W, Wc)
    cache.set("R", np.full(i.Np, c0, np.float32))
    cam = s.camera()
    cam.W, cam.H, cam.spp = 1920, 1080, 1
    for mode in (rea.RENDER_FUSED, rea.RENDER_STAGED):
        cache.set_render_mode(mode)
        a = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
        b = torch.zeros_like(a)
        cache.render(a, cam, seed=8)
        mlp.render(cache, b, cam, seed=8)
        ga, gb = a.cpu().numpy(), b.cpu().numpy()
        assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), f"mode {mode}: {np.count_nonzero(ga != gb)} pixels differ"
    _say("mlp 1080p constant pass equals the grid's")

"""Pins of the oracle functions that no reference fixture covers (CPU only):

* the camera pass (orc_k_query), by a closed form on one lit plane (each path's primary hit and
  next-event term computed analytically in double from the same jitter and light draws), and by the
  expectation identity between path tracing without a cache and the render ending on the converged
  cache (the cache's resolved radiance A (E + L) is the radiance an unbiased path tracer estimates);
* the FP64 twin of the camera pass against the FP32 pass;
* the cache-cell sampling (orc_sample_cells), by invariants recomputed from a brute-force list of every
  triangle sample: each key's cell and normal bucket hold its point and normal, the point is the
  sample nearest the cell centre, the albedo is the area-weighted mean, keys are in Morton order.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import render_engine_amd as rea


def _plane_scene(lights, inten, A=0.6, half=50.0):
    """an orc_scene with one dynamic quad (the floor y = 0, normal +y, 2 triangles), 4 x 4 entries of
    albedo A, point lights at `lights` with intensities `inten`; nothing else"""
    o = np.array([-half, 0.0, -half], np.float32)
    u = np.array([2 * half, 0.0, 0.0], np.float32)
    v = np.array([0.0, 0.0, 2 * half], np.float32)
    n = np.array([0.0, 1.0, 0.0], np.float32)
    # winding u x v = -y would face down; the quad's normal is given explicitly (Q row 3)
    keep = {}
    keep["Q"] = np.concatenate([o, u, v, n]).astype(np.float32)
    keep["T"] = np.concatenate([o, u, v, o + u + v, -u, -v]).astype(np.float32)
    keep["TQ"] = np.zeros(2, np.int32)
    keep["dims"] = np.array([4, 4], np.int32)
    keep["off"] = np.zeros(1, np.int32)
    Np = 16
    keep["P"] = np.zeros(3 * Np, np.float32)
    keep["N"] = np.tile(n, Np).astype(np.float32)
    keep["Alb"] = np.full(Np, A, np.float32)
    keep["L"] = np.asarray(lights, np.float32).ravel()
    keep["ltype"] = np.zeros(len(inten), np.int32)
    keep["inten"] = np.asarray(inten, np.float32)
    keep["region"] = np.zeros(Np, np.int32)
    s = po.OrcScene()
    s.Q, s.T, s.TQ = po.ptr(keep["Q"]), po.ptr(keep["T"]), po.ptr(keep["TQ"])
    s.dims, s.off = po.ptr(keep["dims"]), po.ptr(keep["off"])
    s.P, s.N, s.Alb, s.region = po.ptr(keep["P"]), po.ptr(keep["N"]), po.ptr(keep["Alb"]), po.ptr(keep["region"])
    s.L, s.ltype, s.inten = po.ptr(keep["L"]), po.ptr(keep["ltype"]), po.ptr(keep["inten"])
    s.nt, s.nst, s.nl, s.Np, s.NR = 2, 0, len(inten), Np, 1
    s.sun_len = 1e4
    for a in range(3):
        s.lo[a], s.ext[a] = -half, 2 * half
    s.vx = s.vy = s.vz = 4
    s.nw = 2
    return s, keep


def _camera(W, H, spp, bounces):
    cam = rea.Camera()
    pos = np.array([0.4, 4.0, 1.5])
    fwd = np.array([0.0, -1.0, -0.35]); fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0]); right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    for a in range(3):
        cam.pos[a], cam.fwd[a], cam.right[a], cam.up[a] = pos[a], fwd[a], right[a], up[a]
    cam.tanx, cam.tany = 0.6, 0.45
    cam.W, cam.H, cam.spp, cam.bounces, cam.cache_bounce = W, H, spp, bounces, 1
    return cam


def test_query_closed_form_on_a_lit_plane():
    """bounces = 0, no cache: pixel = mean over samples of A * nl * I_l * cos / max(r, 0.05)^2 for the
    light l drawn by the path (uniform over nl lights, weight nl), at the analytic hit of the jittered
    camera ray with the plane; a light of intensity 0 contributes 0. Bounce rays of a lone plane leave
    the scene, so bounces = 2 gives the same image."""
    lights = [[0.3, 2.0, -0.7], [-1.2, 0.8, 0.9], [2.0, 3.0, 1.0]]
    inten = [50.0, 20.0, 0.0]
    A = 0.6
    s, keep = _plane_scene(lights, inten, A)
    W, H, spp, seed, frame = 24, 18, 3, 5, 2
    cam = _camera(W, H, spp, 0)
    img = np.zeros(W * H, np.float32)
    po.lib.orc_k_query(C.byref(s), 1, None, C.byref(cam), seed, frame, 0, po.ptr(img), None, 0, W * H)
    pos, fwd = np.array(cam.pos[:], np.float64), np.array(cam.fwd[:], np.float64)
    right, up = np.array(cam.right[:], np.float64), np.array(cam.up[:], np.float64)
    L = np.asarray(lights, np.float64)
    nl = len(inten)
    exp = np.zeros(W * H)
    for pix in range(W * H):
        px, py = pix % W, pix // W
        acc = 0.0
        for sp in range(spp):
            rs = C.c_uint32(po.lib.orc_seed(seed, frame, pix * spp + sp))
            ju, jv = po.lib.orc_rnd(C.byref(rs)), po.lib.orc_rnd(C.byref(rs))
            sx = (2 * (px + ju) / W - 1) * cam.tanx
            sy = (1 - 2 * (py + jv) / H) * cam.tany
            d = fwd + right * sx + up * sy
            d /= np.linalg.norm(d)
            assert d[1] < 0
            Hp = pos + d * (-pos[1] / d[1])
            u = po.lib.orc_rnd(C.byref(rs))
            l = min(int(np.float32(u) * np.float32(nl)), nl - 1)
            dv = L[l] - Hp
            r = np.linalg.norm(dv)
            c = dv[1] / r
            if inten[l] > 0 and c > 0:
                acc += A * nl * inten[l] * c / max(r, 0.05) ** 2
        exp[pix] = acc / spp
    assert np.count_nonzero(exp) > W * H // 2
    np.testing.assert_allclose(img, exp, rtol=2e-5, atol=1e-6)
    cam2 = _camera(W, H, spp, 2)
    img2 = np.zeros(W * H, np.float32)
    po.lib.orc_k_query(C.byref(s), 1, None, C.byref(cam2), seed, frame, 0, po.ptr(img2), None, 0, W * H)
    np.testing.assert_allclose(img2, exp, rtol=2e-5, atol=1e-6)


@pytest.fixture(scope="module")
def room():
    s = rea.Scene.box_room(seed=1, target_tris=2000)
    return s, po.OracleScene(s)


def test_query_without_cache_matches_converged_cache(room):
    """E[path tracing without a cache, 14 bounces] = E[render ending on the converged cache at bounce 1]:
    the cache's R = A (E + L) is the outgoing radiance (direct light summed over every light, indirect
    by cosine-sampled gathering) that the path tracer estimates by next-event estimation and cosine
    bounces. A dropped cosine, 1/r^2, albedo or light weight on either side breaks the equality. The
    cache is converged by Jacobi iterations of the oracle's k_ref; the two image means must agree
    within 3 % over the image and over each of 4 x 4 tiles within 6 %, beyond the noise of either."""
    s, osc = room
    i = s.info
    K, warm, T = 16, 12, 24
    Lin = np.zeros(i.Np * i.nl, np.float32)
    Lout, acc, cl = np.zeros_like(Lin), np.zeros_like(Lin), np.zeros(i.Np, np.float32)
    for it in range(warm + T):
        po.lib.orc_k_ref(osc.state(0), 1, po.ptr(Lin), po.ptr(Lout), po.ptr(acc), po.ptr(cl), (7919 * it + 5) & 0xFFFFFFFF, K,
                         int(it >= warm), 0)
        Lin, Lout = Lout, Lin
    Lc = (acc / T).astype(np.float32)
    R = np.zeros(i.Np, np.float32)
    po.lib.orc_k_resolve(osc.state(0), po.ptr(Lc), po.ptr(R), 0)
    cam = s.camera()
    cam.W, cam.H = 96, 96
    imgs = {}
    for name, Rp, spp, nb in (("cache", R, 8, 3), ("path", None, 32, 14)):
        cam.spp, cam.bounces = spp, nb
        img = np.zeros(cam.W * cam.H, np.float32)
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(Rp), C.byref(cam), 3, 0, 0, po.ptr(img), None, 0, cam.W * cam.H)
        imgs[name] = img.reshape(cam.H, cam.W).astype(np.float64)
    a, b = imgs["cache"], imgs["path"]
    rel = a.mean() / b.mean() - 1
    tiles = a.reshape(4, 24, 4, 24).mean((1, 3)) / b.reshape(4, 24, 4, 24).mean((1, 3)) - 1
    print(f"cache / path tracing: image {rel:+.4f}, tiles max {np.abs(tiles).max():.4f}")
    assert abs(rel) < 0.03, rel
    assert np.abs(tiles).max() < 0.06, tiles


def test_query_fp64_twin_matches_fp32(room):
    """the FP64 twin of the camera pass (same rays and hits, radiometry in double) equals the FP32
    pass to float rounding, with and without a cache"""
    s, osc = room
    i = s.info
    rng = np.random.default_rng(0)
    R = rng.uniform(0.1, 2.0, i.Np).astype(np.float32)
    cam = s.camera()
    cam.W, cam.H, cam.spp, cam.bounces = 64, 48, 2, 3
    n = cam.W * cam.H
    for Rp in (R, None):
        a = np.zeros(n, np.float32)
        b = np.zeros(n)
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(Rp), C.byref(cam), 9, 1, 0, po.ptr(a), None, 0, n)
        R64 = None if Rp is None else Rp.astype(np.float64)
        po.lib.orc_k_query_f64(osc.state(0), 1, po.ptr(R64), C.byref(cam), 9, 1, 0, po.ptr(b), 0, n)
        assert np.count_nonzero(b) > n // 2
        assert po.rel_rms(a, b) < 1e-6, po.rel_rms(a, b)


# ------------------------------------------------------------------ cache-cell sampling
def _mesh(seed, nt=40):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1.0, (nt, 1, 3))
    V = (c + rng.normal(scale=0.12, size=(nt, 3, 3))).astype(np.float32)
    V[0] = [[0.1, 0.1, 0.1], [0.1, 0.1, 0.1 + 1e-7], [0.1, 0.1, 0.1]]   # degenerate: skipped
    alb = rng.uniform(0.2, 0.8, nt).astype(np.float32)
    ds = (rng.random(nt) < 0.3).astype(np.uint8)
    return V, alb, ds


def _all_samples(V, alb, ds, h, lo):
    """every (cell, bucket) sample of the sampling rule, in triangle / sample / side order"""
    out = []
    f32 = np.float32
    for t in range(len(V)):
        v = V[t].ravel()
        e1 = (v[3:6] - v[0:3]).astype(f32)
        e2 = (v[6:9] - v[0:3]).astype(f32)
        cr = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]], f32)
        l = f32(np.sqrt(f32(f32(cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2])))
        if l < f32(1e-12):
            continue
        ar = f32(0.5) * l
        n = (cr / l).astype(f32)
        ns = int(min(1e6, max(1.0, np.ceil(float(f32(ar / f32(h * h))) * 4.0))))   # float quotient, then double
        st = C.c_uint64(po.lib.orc_hostrng_init(t))
        for k in range(ns):
            a = b = f32(1.0) / f32(3.0)
            if k > 0:
                a, b = f32(po.lib.orc_hostrng_u(C.byref(st))), f32(po.lib.orc_hostrng_u(C.byref(st)))
                if a + b > f32(1):
                    a, b = f32(1) - a, f32(1) - b
            P = (v[0:3] + e1 * a + e2 * b).astype(f32)
            cc = np.floor((P - lo) / h).astype(np.int64)
            ctr = (lo + (cc.astype(f32) + f32(0.5)) * h).astype(f32)
            dd = (P - ctr).astype(f32)
            d2 = f32(f32(f32(0) + dd[0] * dd[0]) + dd[1] * dd[1]) + dd[2] * dd[2]
            for sd in range(2 if ds[t] else 1):
                nn = -n if sd else n
                key = int(po.lib.orc_cell_key(int(cc[0]), int(cc[1]), int(cc[2]), po.lib.orc_normal_bucket(*map(float, nn))))
                out.append((key, f32(d2), P.copy(), nn.copy(), float(alb[t]) * float(ar) / ns, float(ar) / ns))
    return out


@pytest.mark.parametrize("seed,h", [(1, 0.05), (2, 0.11)])
def test_sample_cells_invariants(seed, h):
    V, alb, ds = _mesh(seed)
    h = np.float32(h)
    lo = np.array([-0.5, -0.5, -0.5], np.float32)
    nt = len(V)
    n = po.lib.orc_sample_cells(po.ptr(V), po.ptr(alb), po.ptr(ds), nt, float(h), po.ptr(lo), None, None, None, None, 0)
    keys = np.zeros(n, np.uint64)
    P, N, A = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32), np.zeros(n, np.float32)
    assert po.lib.orc_sample_cells(po.ptr(V), po.ptr(alb), po.ptr(ds), nt, float(h), po.ptr(lo), po.ptr(keys), po.ptr(P), po.ptr(N),
                                   po.ptr(A), n) == n
    P, N = P.reshape(-1, 3), N.reshape(-1, 3)
    samples = _all_samples(V, alb, ds, h, lo)
    by_key = {}
    for rec in samples:
        by_key.setdefault(rec[0], []).append(rec)
    assert set(by_key) == set(int(k) for k in keys) and len(keys) == len(by_key)
    morton = []
    for j, k in enumerate(keys):
        cx, cy, cz, b = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        po.lib.orc_decode_key(int(k), C.byref(cx), C.byref(cy), C.byref(cz), C.byref(b))
        cell = np.floor((P[j] - lo) / h).astype(np.int64) & 0xFFFFF
        assert tuple(cell) == (cx.value, cy.value, cz.value), "the point lies outside its key's cell"
        assert po.lib.orc_normal_bucket(*map(float, N[j])) == b.value
        recs = by_key[int(k)]
        best = min(range(len(recs)), key=lambda r: (recs[r][1], r))       # nearest to the centre, first on ties
        assert np.array_equal(P[j], recs[best][2]) and np.array_equal(N[j], recs[best][3])
        wa = sum(r[4] for r in recs) / sum(r[5] for r in recs)
        assert A[j] == np.float32(wa)
        sp = [int(po.lib.orc_spread3(x)) for x in (cx.value, cy.value, cz.value)]
        morton.append((sp[0] | sp[1] << 1 | sp[2] << 2) * 8 + b.value)
    assert morton == sorted(morton)

"""The CPU oracle pinned against published values, closed forms, brute force and invariants
(no GPU). A plausible slip in the oracle (a dropped term, a wrong index or sign, a swapped
operand) fails at least one of these."""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle as po
import render_engine_amd as rea

L = po.lib


# ---------------------------------------------------------------- RNG / hashes
def test_splitmix64_published_outputs():
    # SplitMix64 from state 0: published first outputs of the generator (Steele, Lea, Flood 2014;
    # the reference outputs used by java.util.SplittableRandom and xoshiro seeding)
    g = 0x9E3779B97F4A7C15
    outs = [L.orc_mix64((k * g) & (2**64 - 1)) for k in (1, 2, 3)]
    assert outs[0] == 0xE220A8397B1DCDAF
    assert outs[1] == 0x6E789E6AA1B965F4
    assert outs[2] == 0x06C45D188009454F


def test_hostrng_stream():
    s = C.c_uint64(L.orc_hostrng_init(0))
    # init state = mix64(0 * g + 1); each u() re-mixes state + g and keeps the top 24 bits
    # (cuda/prep_mesh.cuh:32-36)
    st = L.orc_mix64(1)
    for _ in range(5):
        u = L.orc_hostrng_u(C.byref(s))
        st = L.orc_mix64((st + 0x9E3779B97F4A7C15) & (2**64 - 1))
        assert u == (st >> 40) / 16777216.0


def test_hash32_is_a_bijection_and_rnd_uniform():
    # lowbias32 is invertible: x -> hash32(x) must not collide on a 2^20 sample
    xs = np.arange(1 << 20, dtype=np.uint64)
    hs = np.array([L.orc_hash32(int(x)) for x in xs[:4096]], np.uint64)
    assert len(np.unique(hs)) == 4096
    assert L.orc_hash32(0) == 0   # a xorshift-multiply chain fixes 0
    s = C.c_uint32(12345)
    u = np.array([L.orc_rnd(C.byref(s)) for _ in range(200000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    assert np.all((u * 16777216.0) == np.floor(u * 16777216.0))   # 24-bit grid


M32 = 0xFFFFFFFF


def _lowbias32_inverse(x):
    # the published inverse of lowbias32 (C. Wellons, "Prospecting for Hash Functions", 2018,
    # hash-prospector: lowbias32_r); it undoes hash32 only if every constant and shift matches
    x ^= x >> 16
    x = (x * 0x43021123) & M32
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & M32
    x ^= x >> 16
    return x


def test_hash32_against_published_inverse():
    rng = np.random.default_rng(7)
    xs = [0, 1, 2, 0x7FFFFFFF, 0x80000000, M32] + [int(v) for v in rng.integers(0, 1 << 32, 20000, dtype=np.uint64)]
    for x in xs:
        assert _lowbias32_inverse(L.orc_hash32(x)) == x


def _fmix32(h, c1):
    # MurmurHash3's 32-bit finaliser with its second multiplier 0xc2b2ae35 (Appleby, SMHasher)
    h ^= h >> 16; h = (h * c1) & M32; h ^= h >> 13; h = (h * 0xC2B2AE35) & M32; h ^= h >> 16
    return h


def test_seed_chain_constants():
    # per-entry seed hash32(seed * G ^ hash32(frame * C ^ hash32(i + 1))) (cuda/poc_gpu.cu:442):
    # peel it with the published inverse of hash32 to read G and C back out of the oracle
    inv = _lowbias32_inverse
    G = inv(L.orc_seed(1, 0, 0)) ^ L.orc_hash32(L.orc_hash32(1))
    C_ = inv(inv(L.orc_seed(0, 1, 0))) ^ L.orc_hash32(1)
    # G = floor(2^32 / golden ratio) (Knuth's multiplicative hashing constant)
    assert G == int(math.floor(2**32 * (math.sqrt(5.0) - 1.0) / 2.0)) == 0x9E3779B9
    # C is MurmurHash3's fmix32 multiplier: with it the published MurmurHash3_x86_32 vectors of
    # the empty key come out ("" seed 1 -> 0x514E28B7, seed 0xffffffff -> 0x81F16F39)
    assert _fmix32(1, C_) == 0x514E28B7 and _fmix32(M32, C_) == 0x81F16F39
    # and the chain is the composition for other arguments too
    for seed, frame, i in ((3, 17, 999), (0xDEADBEEF, 123456, 0), (7, 0, 1 << 20)):
        want = L.orc_hash32(((seed * G) & M32) ^ L.orc_hash32(((frame * C_) & M32) ^ L.orc_hash32((i + 1) & M32)))
        assert L.orc_seed(seed, frame, i) == want


def _pcg_rxs_m_xs_inverse(w):
    # inverse of PCG's 32-bit RXS-M-XS output permutation, with the multiplier's published
    # inverse 2897767785 (pcg_random.hpp, mcg_unmultiplier<uint32_t>) for 277803737
    w ^= w >> 22
    w = (w * 2897767785) & M32
    s = w
    sh = (s >> 28) + 4          # the top 4 bits are untouched by the shift of at least 4
    x = s
    for _ in range(8):          # undo x ^= x >> sh
        x = s ^ (x >> sh)
    return x


def test_rnd_against_pcg():
    # state step: PCG's default 32-bit LCG (multiplier 747796405, increment 2891336453,
    # pcg_random.hpp default_multiplier / default_increment<uint32_t>); full period by Hull-Dobell
    # (c odd, a - 1 divisible by 4, modulus a power of two)
    a, c = 747796405, 2891336453
    assert c % 2 == 1 and (a - 1) % 4 == 0
    rng = np.random.default_rng(11)
    for s0 in [0, 1, M32] + [int(v) for v in rng.integers(0, 1 << 32, 3000, dtype=np.uint64)]:
        s = C.c_uint32(s0)
        u = L.orc_rnd(C.byref(s))
        assert s.value == (s0 * a + c) & M32
        # output: 24 high bits of RXS-M-XS(state); one of the 256 completions of the low byte
        # must invert, through the published unmultiplier, to the state itself
        hi = int(u * 16777216.0)
        assert any(_pcg_rxs_m_xs_inverse((hi << 8) | b) == s.value for b in range(256))


def test_sincos_against_libm():
    xs = np.linspace(0.0, 2 * math.pi, 100001, dtype=np.float32)
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for x in xs[::7]:
        L.orc_sincos(float(x), C.byref(s), C.byref(c))
        worst = max(worst, abs(s.value - math.sin(float(x))), abs(c.value - math.cos(float(x))))
    assert worst < 4e-7


def test_cosine_dir_moments():
    rng = np.random.default_rng(0)
    for N in ([0.0, 0.0, 1.0], [0.95, 0.1, np.sqrt(1 - 0.95**2 - 0.01)], [-0.3, -0.8, 0.52]):
        N = (np.array(N, np.float64) / np.linalg.norm(N)).astype(np.float32)
        D = np.zeros(3, np.float32)
        ds = []
        for u1, u2 in rng.random((20000, 2)).astype(np.float32):
            L.orc_cosine_dir(po.ptr(N), float(u1), float(u2), po.ptr(D))
            ds.append(D.copy())
        ds = np.array(ds, np.float64)
        assert np.allclose(np.linalg.norm(ds, axis=1), 1.0, atol=2e-6)
        c = ds @ N
        assert c.min() >= -1e-6
        assert abs(c.mean() - 2 / 3) < 0.01          # E[cos] under p = cos/pi
        assert abs((c * c).mean() - 0.5) < 0.01      # E[cos^2]
        assert np.linalg.norm(ds.mean(0) - (2 / 3) * N) < 0.015


# ---------------------------------------------------------------- keys / table
def test_cell_key_round_trip_and_buckets():
    rng = np.random.default_rng(1)
    cx, cy, cz, b = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    for x, y, z in rng.integers(0, 1 << 20, (2000, 3)):
        for bb in range(6):
            k = L.orc_cell_key(int(x), int(y), int(z), bb)
            assert k != 0
            L.orc_decode_key(k, C.byref(cx), C.byref(cy), C.byref(cz), C.byref(b))
            assert (cx.value, cy.value, cz.value, b.value) == (x, y, z, bb)
    axes = {(1, 0, 0): 0, (-1, 0, 0): 1, (0, 1, 0): 2, (0, -1, 0): 3, (0, 0, 1): 4, (0, 0, -1): 5}
    for n, bb in axes.items():
        assert L.orc_normal_bucket(*map(float, n)) == bb
    assert L.orc_normal_bucket(0.6, 0.6, 0.5) == 0   # ties go to x, then y


def test_spread3_is_morton_interleave():
    rng = np.random.default_rng(2)
    for x in rng.integers(0, 1 << 21, 500):
        r = L.orc_spread3(int(x))
        assert all(((r >> (3 * i)) & 1) == ((int(x) >> i) & 1) for i in range(21))
        assert r & ~int(sum(1 << (3 * i) for i in range(21))) == 0


def test_hash_table_invariants():
    rng = np.random.default_rng(3)
    keys = np.unique(rng.integers(1, 2**62, 30000, dtype=np.int64)).astype(np.uint64)
    log2 = 16
    HK = np.zeros(1 << log2, np.uint64)
    HV = np.zeros(1 << log2, np.int32)
    L.orc_hash_build(po.ptr(keys), len(keys), log2, po.ptr(HK), po.ptr(HV))
    mask = (1 << log2) - 1
    for i, k in enumerate(keys):
        assert L.orc_hash_find(po.ptr(HK), po.ptr(HV), mask, int(k)) == i
    for k in rng.integers(1, 2**62, 2000, dtype=np.int64).astype(np.uint64):
        if k not in set(keys[:0]):
            v = L.orc_hash_find(po.ptr(HK), po.ptr(HV), mask, int(k))
            assert v == -1 or keys[v] == k
    assert np.count_nonzero(HK) == len(keys)


# ---------------------------------------------------------------- scene + traversal
@pytest.fixture(scope="module")
def room():
    s = rea.Scene.box_room(seed=1, target_tris=2000)
    return s, po.OracleScene(s)


@pytest.fixture(scope="module")
def street():
    s = rea.Scene.street(seed=3, nboxes=60, target_tris=5000, nlamps=4, sun=True, target_points=8000, table_log2=0, W=64, H=48,
                         spp=1, bounces=2)
    return s, po.OracleScene(s)


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    i = scene.info
    lo, ext = np.array(list(i.lo)), np.array(list(i.ext))
    O = (lo + ext * (0.05 + 0.9 * rng.random((n, 3)))).astype(np.float32)
    D = rng.normal(size=(n, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    return O, D.astype(np.float32)


@pytest.mark.parametrize("which", ["room", "street"])
def test_bvh_closest_equals_brute_force(which, request):
    scene, osc = request.getfixturevalue(which)
    O, D = _rays(scene, 4000, 7)
    t1, q1 = osc.trace(O, D, 0, trav=1)
    t3, q3 = osc.trace(O, D, 0, trav=3)
    assert np.array_equal(q1, q3) and np.array_equal(t1.view(np.uint32), t3.view(np.uint32))
    t0, q0 = osc.trace(O, D, 0, trav=0)
    # the prototype's traversal order differs only on exact ties of t
    diff = q0 != q1
    assert np.array_equal(t0[diff].view(np.uint32), t1[diff].view(np.uint32))
    assert diff.mean() < 0.01


def test_trace_closed_form_room(room):
    """rays from the room centre along the axes hit the walls at the room's half extents"""
    scene, osc = room
    O = np.tile(np.array([[4.0, 2.0, 2.9]], np.float32), (4, 1))   # above both cubes
    D = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], np.float32)
    t, q = osc.trace(O, D, 0, trav=1)
    assert np.allclose(t, [4.0, 4.0, 2.0, 2.0], atol=1e-5) and np.all(q <= -2)
    # along +z from the floor centre: ceiling at 3 m
    t, q = osc.trace(np.array([[4.0, 2.0, 0.01]], np.float32), np.array([[0, 0, 1]], np.float32), 0, trav=1)
    assert abs(t[0] - 2.99) < 1e-5


def test_bvh_structure(street):
    scene, _ = street
    bn = scene.array("BN").reshape(-1, 8)
    lo = bn[:, 0:3].view(np.float32)
    hi = bn[:, 4:7].view(np.float32)
    lf, cnt = bn[:, 3], bn[:, 7]
    nst = scene.info.nst
    seen = np.zeros(nst, np.int32)
    ST = scene.array("ST").reshape(-1, 9)
    for k in range(len(bn)):
        if cnt[k] > 0:
            seen[lf[k]:lf[k] + cnt[k]] += 1
            v0 = ST[lf[k]:lf[k] + cnt[k], 0:3]
            assert np.all(v0 >= lo[k] - 1e-5) and np.all(v0 <= hi[k] + 1e-5)
        else:
            for c in (lf[k], lf[k] + 1):
                assert np.all(lo[c] >= lo[k]) and np.all(hi[c] <= hi[k])
    assert np.all(seen == 1)
    # 4-wide nodes: every triangle in exactly one leaf range (leaf child = 1 << 31 | first << 7 |
    # count, count 1..127, end in the last row), inner children reached once, empty slots (child 0)
    # carry the never-entered box (lo = +inf, hi = -inf)
    wn = scene.array("WN").reshape(-1, 32)
    wf = wn.view(np.float32)
    ch, ce = wn[:, 24:28].view(np.uint32), wn[:, 28:32].view(np.int32)
    wlo, whi = wf[:, 0:12].reshape(-1, 3, 4), wf[:, 12:24].reshape(-1, 3, 4)
    wseen = np.zeros(nst, np.int32)
    reached = np.zeros(len(wn), np.int32)
    for k in range(len(wn)):
        for c in range(4):
            empty = np.all(np.isposinf(wlo[k, :, c])) and np.all(np.isneginf(whi[k, :, c]))
            e = int(ch[k, c])
            if e >> 31:
                first, count = (e >> 7) & 0xFFFFFF, e & 127
                assert 1 <= count <= 127 and ce[k, c] == first + count
                wseen[first:ce[k, c]] += 1
                v0 = ST[first:ce[k, c], 0:3]
                assert np.all(v0 >= wlo[k, :, c] - 1e-5) and np.all(v0 <= whi[k, :, c] + 1e-5)
            elif e == 0:
                assert empty and ce[k, c] == 0
            else:
                assert not empty and e < len(wn) and ce[k, c] == 0
                reached[e] += 1
    assert np.all(wseen == 1)
    assert reached[0] == 0 and np.all(reached[1:] == 1)



def _area(lo, hi):
    d = np.maximum(np.asarray(hi, np.float64) - np.asarray(lo, np.float64), 0.0)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


@pytest.mark.parametrize("which", ["room", "street"])
def test_wide_collapse_is_area_optimal(which, request):
    """the 4-wide collapse's total node area (sum over wide nodes of the area of the union of their
    child boxes) equals the least one over every way of cutting the binary tree into nodes of 2..4
    children, found here by enumerating each node's child sets (repeatedly replacing an inner
    member by its two children while fewer than 4), and is at most the largest-area-first rule's"""
    scene, _ = request.getfixturevalue(which)
    bn = scene.array("BN").reshape(-1, 8)
    lo, hi = bn[:, 0:3].view(np.float32).astype(np.float64), bn[:, 4:7].view(np.float32).astype(np.float64)
    lf, cnt = bn[:, 3].astype(np.int64), bn[:, 7].astype(np.int64)
    A = [_area(lo[k], hi[k]) for k in range(len(bn))]
    W = {}

    def sets(n):   # every child set of a wide node for inner node n
        out, todo = set(), [frozenset((int(lf[n]), int(lf[n]) + 1))]
        while todo:
            f = todo.pop()
            if f in out:
                continue
            out.add(f)
            if len(f) < 4:
                for m in f:
                    if cnt[m] == 0:
                        todo.append((f - {m}) | {int(lf[m]), int(lf[m]) + 1})
        return out

    for n in range(len(bn) - 1, -1, -1):   # children follow their parent
        if cnt[n] == 0:
            W[n] = A[n] + min(sum(W[m] for m in f if cnt[m] == 0) for f in sets(n))
    best = W[0]
    # largest-area-first: open the inner member with the largest area while fewer than 4
    def greedy(n):
        f = [int(lf[n]), int(lf[n]) + 1]
        while len(f) < 4:
            inner = [m for m in f if cnt[m] == 0]
            if not inner:
                break
            m = max(inner, key=lambda x: A[x])
            i = f.index(m)
            f[i:i + 1] = [int(lf[m]), int(lf[m]) + 1]
        return A[n] + sum(greedy(m) for m in f if cnt[m] == 0)
    g = greedy(0)
    wn = scene.array("WN").reshape(-1, 32)
    wf = wn.view(np.float32).astype(np.float64)
    ch = wn[:, 24:28]
    got = 0.0
    for k in range(len(wn)):
        m = ch[k] != 0
        klo = [wf[k, 4 * a:4 * a + 4][m].min() for a in range(3)]
        khi = [wf[k, 12 + 4 * a:16 + 4 * a][m].max() for a in range(3)]
        got += _area(klo, khi)
    print(f"[wide] {which}: {len(wn)} wide nodes, area cost {got:.6g}, optimum {best:.6g}, largest-area-first {g:.6g}")
    assert abs(got - best) <= 1e-5 * best
    assert got <= g * (1 + 1e-6)


# ---------------------------------------------------------------- DDA
def _exact_voxels(osc_state, info, a, b):
    """voxels whose closed box meets the segment (slab test in double precision)"""
    g = np.array([info.vx, info.vy, info.vz])
    lo, ext = np.array(list(info.lo), np.float64), np.array(list(info.ext), np.float64)
    A, B = (np.asarray(a, np.float64) - lo) / ext * g, (np.asarray(b, np.float64) - lo) / ext * g
    d = B - A
    out = set()
    for k in range(info.vz):
        for j in range(info.vy):
            for i in range(info.vx):
                c0, c1 = np.array([i, j, k], np.float64), np.array([i + 1, j + 1, k + 1], np.float64)
                t0, t1 = 0.0, 1.0
                ok = True
                for ax in range(3):
                    if abs(d[ax]) < 1e-12:
                        if A[ax] < c0[ax] or A[ax] > c1[ax]:
                            ok = False
                    else:
                        u, v = (c0[ax] - A[ax]) / d[ax], (c1[ax] - A[ax]) / d[ax]
                        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
                if ok and t0 <= t1:
                    out.add((k * info.vy + j) * info.vx + i)
    return out


def test_dda_matches_exact_walk(room):
    scene, osc = room
    info = scene.info
    rng = np.random.default_rng(11)
    lo, ext = np.array(list(info.lo)), np.array(list(info.ext))
    n = 300
    A = (lo + ext * (0.01 + 0.98 * rng.random((n, 3)))).astype(np.float32)
    B = (lo + ext * (0.01 + 0.98 * rng.random((n, 3)))).astype(np.float32)
    out = np.zeros(n * info.nw, np.uint32)
    L.orc_dda_segments(osc.state(0), po.ptr(A), po.ptr(B), n, po.ptr(out))
    bad = 0
    for s in range(n):
        got = {w * 32 + bit for w in range(info.nw) for bit in range(32) if (out[s * info.nw + w] >> bit) & 1}
        exact = _exact_voxels(osc, info, A[s], B[s])
        assert got <= exact, f"segment {s}: DDA visited voxels the segment does not cross"
        bad += got != exact
    # equality except when the segment passes (within rounding) through a voxel edge/corner
    assert bad <= 3


# ---------------------------------------------------------------- frame kernels: closed forms
def test_runner_ema_closed_form(room):
    """with every ray landing on an entry of constant value, the plain EMA follows
    L_t = m + (L_0 - m)(1 - alpha)^t with alpha = 1/min(max(n+1,1), nmax) = 1/nmax at n = nmax"""
    scene, osc = room
    m = rea.methods.plain(nmax=8.0)
    run = po.OracleRunner(osc, m)
    run.begin(0, 3)
    lc0 = run.get("LC").copy()
    run.frame(0)
    M = run.get("M").reshape(osc.nl, osc.Np).T.ravel()   # this frame's sample means (SoA -> AoS)
    lc1 = run.get("LC")
    assert np.allclose(lc1, lc0 + (M - lc0) / 8.0, rtol=1e-6, atol=1e-7)
    assert np.all(run.get("N") == 8.0)


def test_targeted_invalidation_semantics(room):
    """after a light-off event, exactly the entries whose recorded light deps contain the light
    (or that have no recorded deps) restart from zero; the others keep their history"""
    scene, osc = room
    m = rea.methods.targeted()
    run = po.OracleRunner(osc, m)
    run.begin(0, 5)
    for _ in range(4):
        run.frame(0)
    lcur, lprev = run.get("LCUR"), run.get("LPREV")
    cur, prev = run.get("CUR"), run.get("PREV")
    Np = osc.Np
    deps = lcur | lprev
    known = (deps != 0) | ((cur | prev).reshape(osc.nw, Np) != 0).any(0)
    expect = (~known) | ((deps >> np.uint64(1)) & np.uint64(1)).astype(bool)
    n_before = run.get("N")
    _, inv = run.frame(1)   # light 1 switches off
    assert inv == int(expect.sum())
    n_after = run.get("N")
    kk = run.get("KK")
    # invalidated entries: n restarted from 0 then grew by this frame's rays
    assert np.all(n_after[expect] == np.minimum(kk[expect], 32).astype(np.float32))
    assert np.all(n_after[~expect] == np.minimum(n_before[~expect] + kk[~expect], 32))


def test_rescale_linearity(room):
    """per-light rescale on an intensity change scales exactly that channel (light off -> 0)"""
    scene, osc = room
    m = rea.methods.make(mode=0, nmax=16.0, budget=1.0, rescale=True)
    run = po.OracleRunner(osc, m)
    run.begin(0, 5)
    run.frame(0)
    # one frame in state 1 (light 1 off): channel 1 is rescaled by 0 before this frame's update
    lc_before = run.get("LC").reshape(-1, 2)
    run.frame(1)
    M = run.get("M").reshape(2, -1).T
    lc = run.get("LC").reshape(-1, 2)
    # EMA from a zeroed channel 1: L = 0 + (m - 0) / 16
    assert np.allclose(lc[:, 1], M[:, 1] / 16.0, rtol=1e-6, atol=1e-9)
    assert np.allclose(lc[:, 0], lc_before[:, 0] + (M[:, 0] - lc_before[:, 0]) / 16.0, rtol=1e-6, atol=1e-7)


def test_scenario_parsing_and_region_metric():
    from render_engine_amd import scenarios as sc
    txt = "# comment\nR1 on-off-on;0:50,1:30,0:50\nI1 off;0:50,1:50\n"
    parsed = sc.parse_scenarios(txt)
    assert parsed[0] == ("R1 on-off-on", [(0, 50), (1, 30), (0, 50)])
    assert len(sc.expand(parsed[0][1])) == 130
    a = np.random.default_rng(0).random((40, 60))
    k = 9
    naive = np.zeros_like(a)
    p = np.pad(a, ((k // 2, k // 2), (k // 2, k // 2)), mode="edge")
    for y in range(40):
        for x in range(60):
            naive[y, x] = p[y:y + k, x:x + k].mean()
    assert np.allclose(sc.box_blur(a, k), naive)
    assert sc.blur_kernel(960) == 9 and sc.blur_kernel(1920) == 19
    on = np.ones((50, 50)); off = np.ones((50, 50)); off[:, :25] = 0.2
    reg = sc.light_off_region(on, off)
    assert reg[:, :18].all() and not reg[:, 32:].any()
    assert sc.frames_to_below([0.3, 0.1, 0.04, 0.02]) == 2
    # geometry event: |a - b| / max(a, b) > 0.25 after the blur, both signs (darkened and brightened)
    a = np.ones((40, 60)); b = np.ones((40, 60)); b[:, :20] = 0.5; b[:, 40:] = 2.0
    pan = sc.panel_region(a, b)
    assert pan[:, :15].all() and pan[:, 45:].all() and not pan[:, 25:35].any()
    k = sc.blur_kernel(60)
    ba, bb = sc.box_blur(a, k), sc.box_blur(b, k)
    naive = np.array([[abs(x - y) / max(x, y, 1e-6) > 0.25 for x, y in zip(ra, rb)] for ra, rb in zip(ba, bb)])
    assert np.array_equal(pan, naive)


def test_resolved_radiance(room):
    """R after an update = A * sum_l (Edir_l + Lc_l) of the new channels"""
    scene, osc = room
    run = po.OracleRunner(osc, rea.methods.plain())
    run.begin(0, 9)
    run.frame(0)
    lc = run.get("LC").reshape(-1, osc.nl).astype(np.float64)
    E = osc.Edir[0].reshape(-1, osc.nl).astype(np.float64)
    A = scene.array("ALB").astype(np.float64)
    R = run.get("R")
    assert np.allclose(R, A * (E + lc).sum(1), rtol=2e-6, atol=1e-7)


def test_fine_volumes_hold_lit_segments(street):
    """each lit shadow segment marks the voxel of its entry and of the light in the light's volume"""
    scene, osc = street
    i = scene.info
    g = i.fine
    assert g > 0
    wpv = (g ** 3 + 31) // 32
    lo, ext = np.array(list(i.lo), np.float64), np.array(list(i.ext), np.float64)
    vox = lambda p: tuple(np.clip(((p - lo) / ext * g).astype(int), 0, g - 1))
    P = scene.array("P", 0).reshape(-1, 3)
    N = scene.array("N", 0).reshape(-1, 3)
    L = scene.array("L").reshape(-1, 3)
    lt = scene.array("LTYPE")
    LM = osc.LM[0]
    rng = np.random.default_rng(4)
    checked = 0
    for e in rng.integers(0, i.npst, 400):
        for l in range(i.nl):
            if lt[l] != 0 or not ((int(LM[e]) >> l) & 1):
                continue
            V = osc.VOL[0][l * wpv:(l + 1) * wpv]
            for p in (P[e] + 1e-3 * N[e], L[l]):
                x, y, z = vox(p.astype(np.float64))
                v = (z * g + y) * g + x
                assert (V[v >> 5] >> (v & 31)) & 1
            checked += 1
    assert checked > 50


def test_occluder_event_light_set(street):
    """inserting the wall next to lamp 0 affects lamp 0 and the sun (their lit segments cross it);
    the targeted method invalidates every entry that recorded one of them"""
    scene, osc = street
    i = scene.info
    ns = i.NS
    occf = scene.array("OCCF", 0 * ns + 3)
    assert len(occf) > 0
    wpv = (i.fine ** 3 + 31) // 32
    hit = set()
    for l in range(i.nl):
        V = osc.VOL[0][l * wpv:(l + 1) * wpv] | osc.VOL[3][l * wpv:(l + 1) * wpv]
        if any((V[v >> 5] >> (v & 31)) & 1 for v in occf):
            hit.add(l)
    assert 0 in hit and (i.nl - 1) in hit
    run = po.OracleRunner(osc, rea.methods.targeted())
    run.begin(0, 5)
    for _ in range(3):
        run.frame(0)
    deps = run.get("LCUR") | run.get("LPREV")
    mask = sum(1 << min(l, 63) for l in hit)
    _, inv = run.frame(3)
    assert inv >= int(((deps & np.uint64(mask)) != 0).sum())

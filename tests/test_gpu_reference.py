"""GPU converged-reference solve (k_ref Jacobi iterations, two chains, cuda/prep_mesh.cuh:182-213,
454-520) and the per-frame error pass (k_error, cuda/poc_gpu.cu:683-696) against the CPU oracle.

rc_scene_reference's chains are replayed on the oracle with the same per-iteration seeds
(mix64(0x51 + 7919 state + 104729 chain + 31 it)) and the same host arithmetic for the per-light
means, totals and validity: the stored references, per-light references and validity flags must be
bit-identical. The oracle's k_ref is itself pinned by the stored converged references of the
reference's scene_main.bin (tests/test_reference_data.py). The error pass sums in FP64 in another
order than the oracle's sequential loop: relative difference < 1e-9."""
import numpy as np
import pytest

import render_engine_amd as rea
import pyoracle as po

pytestmark = pytest.mark.gpu

K, WARM, BLOCK, MAX_ITER = 8, 2, 4, 8     # target 0: exactly warm + max_iter iterations per chain


@pytest.fixture(scope="module", params=["room", "street"])
def solved(request, torch_gpu):
    if request.param == "room":
        s = rea.Scene.box_room(seed=1, target_tris=2000)
    else:
        s = rea.Scene.street(seed=3, nboxes=150, target_tris=12000, nlamps=6, sun=True, target_points=30000, table_log2=0,
                             W=96, H=64, spp=2, bounces=3)
    s.upload(trav=rea.TRAV_WIDE)
    osc = po.OracleScene(s)
    noise, iters = s.reference(K=K, warm=WARM, block=BLOCK, max_iter=MAX_ITER, target=0.0)
    print(f"reference case {request.param}: noise {noise}, iterations {iters}", flush=True)
    return request.param, s, osc, noise, iters


def _present_quads(s, st):
    """quads of the objects present in state st (an object = 6 quads; present when any of its quads
    has dynamic triangles in that state)"""
    tq = set(int(q) for q in s.array("TQ", st))
    objs = {q // 6 for q in tq}
    return {q for q in range(s.info.Nq) if q // 6 in objs}


def _oracle_reference(s, osc, st):
    i = s.info
    Np, nl = i.Np, i.nl
    acc = [np.zeros(Np * nl, np.float32) for _ in range(2)]
    chain = [np.zeros(Np * nl, np.float32) for _ in range(2)]
    closec = np.zeros(Np, np.float32)
    it_done = 0
    while True:
        n_it = WARM + BLOCK if it_done == 0 else BLOCK
        for c in range(2):
            a, b = chain[c], np.zeros(Np * nl, np.float32)
            for k in range(n_it):
                it = it_done + k
                seed = int(po.lib.orc_mix64(0x51 + 7919 * st + 104729 * c + 31 * it)) & 0xFFFFFFFF
                po.lib.orc_k_ref(osc.state(st), 1, po.ptr(a), po.ptr(b), po.ptr(acc[c]), po.ptr(closec), seed, K, int(it >= WARM),
                                 int(c == 0))
                a, b = b, a
            chain[c] = a
        it_done += n_it
        T = it_done - WARM
        if T >= MAX_ITER:
            break
    x0 = (acc[0] / np.float32(T)).astype(np.float32).reshape(Np, nl)
    x1 = (acc[1] / np.float32(T)).astype(np.float32).reshape(Np, nl)
    Pw = (np.float32(0.5) * (x0 + x1)).astype(np.float32)
    e0 = np.zeros(Np, np.float32)
    e1 = np.zeros(Np, np.float32)
    for l in range(nl):
        e0 = (e0 + x0[:, l]).astype(np.float32)
        e1 = (e1 + x1[:, l]).astype(np.float32)
    ref = (np.float32(0.5) * (e0 + e1)).astype(np.float32)
    valid = (closec.astype(np.float64) / (float(K) * it_done) < 0.3).astype(np.uint8)
    present = _present_quads(s, st)
    dims, off = s.array("DIMS").reshape(-1, 2), s.array("OFF")
    for q in range(i.Nq):
        if q not in present:
            valid[off[q]:off[q] + dims[q, 0] * dims[q, 1]] = 0
    return ref, Pw.ravel(), valid, T


def test_reference_solve_bit_exact(solved):
    name, s, osc, noise, iters = solved
    for st in range(s.info.NS):
        ref, Pw, valid, T = _oracle_reference(s, osc, st)
        assert iters[st] == T
        assert np.array_equal(s.array("REF", st).view(np.uint32), ref.view(np.uint32)), f"{name} state {st}: references differ"
        assert np.array_equal(s.array("PW", st).view(np.uint32), Pw.view(np.uint32)), f"{name} state {st}: per-light refs differ"
        assert np.array_equal(s.array("VALID", st), valid), f"{name} state {st}: validity differs"
        assert valid.sum() > 0.5 * s.info.Np


def test_error_pass_against_oracle(solved):
    """relative RMS per frame (plain and targeted methods, an event in the middle) against the oracle
    runner warm-started from the same per-light references; cache arrays bit for bit"""
    name, s, osc, _, _ = solved
    i = s.info
    Pw = [s.array("PW", k) for k in range(i.NS)]
    ref = [s.array("REF", k) for k in range(i.NS)]
    val = [s.array("VALID", k) for k in range(i.NS)]
    seq = [0] * 4 + [1] * 4
    for m in (rea.methods.plain(), rea.methods.targeted()):
        cache = rea.Cache(s, m)
        orun = po.OracleRunner(osc, m, Pw=Pw, ref=ref, valid=val)
        cache.begin(seq[0], 5)
        orun.begin(seq[0], 5)
        for f, st in enumerate(seq):
            fs = cache.frame(st, want_err=True)
            e, _ = orun.frame(st, True)
            assert np.array_equal(cache.get("LC").view(np.uint32), orun.get("LC").view(np.uint32)), f"{name} frame {f}"
            assert e > 0 and abs(fs.err - e) <= 1e-9 * e, (name, f, fs.err, e)

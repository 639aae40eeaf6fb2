"""The oracle's adjoint of the cached radiance pinned by mathematics: its FP64 reverse-mode
gradient against central finite differences of its FP64 forward (h = 1e-3), linearity in the
emission, the closed form at zero bounces, the support masks against per-light unit
forwards, and the FP32 forward against the FP64 one."""
import numpy as np
import pytest

import render_engine_amd as rea
import pyoracle as po


@pytest.fixture(scope="module")
def room_adj():
    s = rea.Scene.box_room(seed=1, target_tris=600)
    osc = po.OracleScene(s)
    return s, osc, po.OracleAdjoint(osc, 0, K=2, B=3, seed=5)


def _params(s, seed=0):
    rng = np.random.default_rng(seed)
    i = s.info
    I = np.array(s.array("INTEN", 0), np.float64) * (0.5 + rng.random(i.nl))
    alb = np.array(s.array("MATALB"), np.float64)
    g = rng.random(i.Np)
    return I, alb, g


def test_grad64_matches_central_differences(room_adj):
    s, osc, adj = room_adj
    I, alb, g = _params(s)
    J, _, dI, dalb = adj.grad64(I, alb, g)
    h = 1e-3
    for l in range(s.info.nl):
        e = np.zeros_like(I); e[l] = h
        fd = (adj.grad64(I + e, alb, g)[0] - adj.grad64(I - e, alb, g)[0]) / (2 * h)
        assert abs(fd - dI[l]) <= 1e-7 * abs(J) + 1e-9 * abs(fd), (l, fd, dI[l])
    used = np.unique(adj.mat)
    for m in used[:: max(1, len(used) // 6)]:
        e = np.zeros_like(alb); e[m] = h
        fd = (adj.grad64(I, alb + e, g)[0] - adj.grad64(I, alb - e, g)[0]) / (2 * h)
        # J is a polynomial of degree B + 1 in each albedo: the central difference is exact up to O(h^2)
        assert abs(fd - dalb[m]) <= 1e-6 * max(abs(dalb[m]), 1e-6 * abs(J)), (m, fd, dalb[m])


def test_forward_is_linear_in_emission(room_adj):
    s, osc, adj = room_adj
    I, alb, g = _params(s, 1)
    R = adj.grad64(I, alb, g)[1]
    acc = np.zeros_like(R)
    for l in range(s.info.nl):
        e = np.zeros_like(I); e[l] = 1.0
        acc += I[l] * adj.grad64(e, alb, g)[1]
    assert np.allclose(R, acc, rtol=1e-12, atol=1e-15 * np.abs(R).max())


def test_zero_bounces_closed_form(room_adj):
    s, osc, _ = room_adj
    adj0 = po.OracleAdjoint(osc, 0, K=2, B=0, seed=5)
    I, alb, g = _params(s, 2)
    _, R, dI, dalb = adj0.grad64(I, alb, g)
    G = adj0.G64.reshape(-1, s.info.nl)
    E = G @ I
    A = alb[adj0.mat]
    assert np.allclose(R, A * E, rtol=1e-13)
    assert np.allclose(dI, G.T @ (g * A), rtol=1e-12)
    assert np.allclose(dalb, np.bincount(adj0.mat, weights=g * E, minlength=s.info.nmat), rtol=1e-12, atol=1e-300)


def test_support_is_the_nonzero_pattern(room_adj):
    s, osc, adj = room_adj
    I, alb, g = _params(s, 3)
    _, SR, _ = adj.forward32(I, alb)
    for l in range(s.info.nl):
        e = np.zeros_like(I); e[l] = 1.0
        Rl = adj.grad64(e, alb, g)[1]
        bit = ((SR >> np.uint64(min(l, 63))) & np.uint64(1)).astype(bool)
        assert np.array_equal(bit, Rl > 0), l


def test_fp32_forward_close_to_fp64(room_adj):
    s, osc, adj = room_adj
    I, alb, g = _params(s, 4)
    R32, _, _ = adj.forward32(I.astype(np.float32), alb.astype(np.float32))
    R64 = adj.grad64(I.astype(np.float32).astype(np.float64), alb.astype(np.float32).astype(np.float64), g)[1]
    rel = np.sqrt(np.mean((R32 - R64) ** 2)) / np.sqrt(np.mean(R64 ** 2))
    assert rel < 1e-5, rel


def test_hits_rows_are_slices_of_the_full_trace(room_adj):
    """the row-range tracer (used to count the rays' work on a sample of entries) gives the same
    hits as the full trace of every bounce"""
    s, osc, adj = room_adj
    Np, K = s.info.Np, adj.K
    for b in range(adj.B):
        for i0, i1 in ((0, Np), (Np // 3, Np // 3 + 17), (Np - 5, Np)):
            h = np.empty((i1 - i0) * K, np.int32)
            po.lib.orc_adj_hits_rows(osc.state(0), 1, K, b, 5, i0, i1, po.ptr(h))
            full = adj.hits[b * Np * K:(b + 1) * Np * K]
            assert np.array_equal(h, full[i0 * K:i1 * K]), (b, i0, i1)

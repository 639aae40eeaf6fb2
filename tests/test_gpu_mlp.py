"""Hash-encoded MLP cache kernels (BF16 MFMA, FP32 accumulation) against a float32 restatement
of the same network on the CPU with the same BF16 roundings: forward outputs, weight and table
gradients, the Adam step, and training that fits a smooth function. Tolerances: the encoding is
exact (same operation order, no contraction); each MFMA layer differs from the CPU sum in
accumulation order only, which can flip a BF16 rounding, so outputs and gradients are compared
by relative L2 (2e-2)."""
import numpy as np
import pytest

import render_engine_amd.mlp as M

pytestmark = pytest.mark.gpu


def bf(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).float().numpy()


def encode_ref(cfg, table, P, N, D):
    res, dense = M.level_resolutions(cfg)
    T = 1 << cfg.log2_table
    tb = table.reshape(M.LEVELS, T, 2)
    lo = np.array(list(cfg.lo), np.float32)
    iext = (np.float32(1) / np.array(list(cfg.ext), np.float32)).astype(np.float32)
    x = np.clip((P[:, :3] - lo) * iext, np.float32(0), np.float32(1)).astype(np.float32)
    n = len(P)
    X = np.zeros((n, 64), np.float32)
    corners = []
    for lv in range(M.LEVELS):
        r = np.float32(res[lv])
        s = (x * r).astype(np.float32)
        f = np.minimum(np.floor(s), r - np.float32(1)).astype(np.float32)
        i = f.astype(np.uint32)
        fr = (s - f).astype(np.float32)
        f0 = np.zeros(n, np.float32); f1 = np.zeros(n, np.float32)
        cl = []
        for c in range(8):
            cx, cy, cz = i[:, 0] + np.uint32(c & 1), i[:, 1] + np.uint32((c >> 1) & 1), i[:, 2] + np.uint32(c >> 2)
            wx = fr[:, 0] if c & 1 else np.float32(1) - fr[:, 0]
            wy = fr[:, 1] if (c >> 1) & 1 else np.float32(1) - fr[:, 1]
            wz = fr[:, 2] if c >> 2 else np.float32(1) - fr[:, 2]
            w = ((wx * wy).astype(np.float32) * wz).astype(np.float32)
            if dense[lv]:
                r1 = np.uint32(res[lv] + 1)
                idx = cx + r1 * (cy + r1 * cz)
            else:
                idx = (cx ^ (cy * np.uint32(2654435761)) ^ (cz * np.uint32(805459861))) & np.uint32(T - 1)
            f0 = (f0 + (w * tb[lv, idx, 0]).astype(np.float32)).astype(np.float32)
            f1 = (f1 + (w * tb[lv, idx, 1]).astype(np.float32)).astype(np.float32)
            cl.append((idx, w))
        X[:, 2 * lv], X[:, 2 * lv + 1] = f0, f1
        corners.append(cl)
    X[:, 32:35] = np.float32(0.5) * (N[:, :3] + np.float32(1))
    X[:, 35:38] = np.float32(0.5) * (D[:, :3] + np.float32(1))
    X[:, 38] = 1
    return bf(X), corners


def split_w(W):
    Ws = [W[L * 4096:(L + 1) * 4096].reshape(64, 64) for L in range(M.HIDDEN)]
    W5 = W[M.HIDDEN * 4096:].reshape(64, 16)
    return Ws, W5


def forward_ref(W, X):
    Ws, W5 = split_w(bf(W))
    acts = [X]
    a = X
    for L in range(M.HIDDEN):
        a = bf(np.maximum(a @ Ws[L], 0))
        acts.append(a)
    return (a @ W5)[:, 0], acts


def backward_ref(cfg, W, acts, corners, dy, nt):
    Ws, W5 = split_w(bf(W))
    gW = np.zeros_like(W)
    a5 = acts[-1]
    gW[M.HIDDEN * 4096:].reshape(64, 16)[:, 0] = a5.T @ dy
    gz = bf(np.where(a5 > 0, dy[:, None] * W5[:, 0][None, :], 0))
    gX = None
    for L in range(M.HIDDEN - 1, -1, -1):
        gW[L * 4096:(L + 1) * 4096] = (acts[L].T.astype(np.float64) @ gz.astype(np.float64)).ravel()
        ga = gz @ Ws[L].T
        if L > 0:
            gz = bf(np.where(acts[L] > 0, ga, 0))
        else:
            gX = ga[:, :32]
    T = 1 << cfg.log2_table
    gt = np.zeros((M.LEVELS, T, 2))
    for lv in range(M.LEVELS):
        for idx, w in corners[lv]:
            np.add.at(gt[lv, :, 0], idx, w.astype(np.float64) * gX[:, 2 * lv])
            np.add.at(gt[lv, :, 1], idx, w.astype(np.float64) * gX[:, 2 * lv + 1])
    return gW, gt.ravel()


def _inputs(torch, n, seed, lo, ext):
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 4), np.float32)
    P[:, :3] = lo + ext * rng.random((n, 3))
    N = np.zeros((n, 4), np.float32)
    v = rng.normal(size=(n, 3)); N[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    D = np.zeros((n, 4), np.float32)
    v = rng.normal(size=(n, 3)); D[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return P, N, D, [torch.from_numpy(a).cuda() for a in (P, N, D)]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def small(torch_gpu):
    lo, ext = np.array([-2.0, 1.0, 0.0], np.float32), np.array([8.0, 4.0, 3.0], np.float32)
    cfg = M.config(lo, ext, log2_table=12, base_res=4, max_res=256, max_batch=4096, seed=3)
    mlp = M.MLPCache(cfg)
    # larger table values than the initialisation so the encoding matters
    rng = np.random.default_rng(1)
    mlp.set(M.MLP_TABLE, rng.uniform(-1, 1, M.LEVELS * 2 * (1 << 12)).astype(np.float32))
    return cfg, mlp, lo, ext


@pytest.mark.parametrize("n", [1, 100, 1000])
def test_mlp_forward_matches_reference(torch_gpu, small, n):
    torch = torch_gpu
    cfg, mlp, lo, ext = small
    P, N, D, (Pt, Nt, Dt) = _inputs(torch, n, 10 + n, lo, ext)
    y = mlp.infer(Pt, Nt, Dt).cpu().numpy()
    X, _ = encode_ref(cfg, mlp.get(M.MLP_TABLE), P, N, D)
    yr, _ = forward_ref(mlp.get(M.MLP_W), X)
    assert np.all(np.isfinite(y))
    assert _rel(y, yr) < 2e-2, _rel(y, yr)


def test_mlp_gradients_match_reference(torch_gpu, small):
    torch = torch_gpu
    cfg, mlp, lo, ext = small
    n = 700
    P, N, D, (Pt, Nt, Dt) = _inputs(torch, n, 5, lo, ext)
    dy = np.random.default_rng(6).normal(size=n).astype(np.float32) / n
    mlp.zero_grad()
    mlp.backward(Pt, Nt, Dt, torch.from_numpy(dy).cuda())
    gW, gT = mlp.get(M.MLP_GW), mlp.get(M.MLP_GTABLE)
    W = mlp.get(M.MLP_W)
    X, corners = encode_ref(cfg, mlp.get(M.MLP_TABLE), P, N, D)
    _, acts = forward_ref(W, X)
    gWr, gTr = backward_ref(cfg, W, acts, corners, dy, None)
    for L in range(M.HIDDEN):
        sl = slice(L * 4096, (L + 1) * 4096)
        assert _rel(gW[sl], gWr[sl]) < 2e-2, (L, _rel(gW[sl], gWr[sl]))
    assert _rel(gW[M.HIDDEN * 4096:], gWr[M.HIDDEN * 4096:]) < 2e-2
    assert _rel(gT, gTr) < 2e-2, _rel(gT, gTr)
    mlp.zero_grad()


def test_mlp_adam_step(torch_gpu, small):
    torch = torch_gpu
    cfg = M.config(small[0].lo, small[0].ext, log2_table=12, base_res=4, max_res=256, max_batch=4096, seed=9, lr=1e-2)
    mlp = M.MLPCache(cfg)
    lo, ext = small[2], small[3]
    n = 512
    P, N, D, (Pt, Nt, Dt) = _inputs(torch, n, 7, lo, ext)
    dy = torch.from_numpy(np.random.default_rng(8).normal(size=n).astype(np.float32)).cuda()
    W0, T0 = mlp.get(M.MLP_W).astype(np.float64), mlp.get(M.MLP_TABLE).astype(np.float64)
    mlp.backward(Pt, Nt, Dt, dy)
    gW, gT = mlp.get(M.MLP_GW).astype(np.float64), mlp.get(M.MLP_GTABLE).astype(np.float64)
    mlp.step()
    W1, T1 = mlp.get(M.MLP_W), mlp.get(M.MLP_TABLE)
    b1, b2, lr, eps = 0.9, 0.99, 1e-2, 1e-15
    for p0, g, p1 in ((W0, gW, W1), (T0, gT, T1)):
        m = (1 - b1) * g
        v = (1 - b2) * g * g
        exp = p0 - (lr / (1 - b1)) * m / (np.sqrt(v) / np.sqrt(1 - b2) + eps)
        assert np.allclose(p1, exp, rtol=1e-5, atol=1e-6 * lr)
    assert np.all(mlp.get(M.MLP_GW) == 0) and np.all(mlp.get(M.MLP_GTABLE) == 0)


def test_mlp_training_fits_a_smooth_function(torch_gpu):
    torch = torch_gpu
    lo, ext = np.zeros(3, np.float32), np.ones(3, np.float32)
    mlp = M.MLPCache(M.config(lo, ext, log2_table=14, base_res=4, max_res=128, max_batch=8192, seed=2, lr=1e-2))
    rng = np.random.default_rng(0)
    losses = []
    for it in range(300):
        P, N, D, (Pt, Nt, Dt) = _inputs(torch, 8192, 1000 + it, lo, ext)
        t = (1.5 + np.sin(6 * P[:, 0]) * np.cos(4 * P[:, 1]) + 0.5 * P[:, 2]).astype(np.float32)
        losses.append(mlp.train(Pt, Nt, Dt, torch.from_numpy(t).cuda()))
    assert np.isfinite(losses[-1])
    assert np.mean(losses[-20:]) < 0.02 * np.mean(losses[:5]), (losses[:5], losses[-5:])

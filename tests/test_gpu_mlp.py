"""Hash-encoded MLP cache kernels (BF16 MFMA, FP32 accumulation) against the oracle's restatement
of the same network (oracle/mlp_oracle.py, pinned on the CPU by tests/test_mlp_oracle_cpu.py) with
the same BF16 roundings: forward outputs, weight and table gradients, the Adam step, and training
that fits a smooth function. Tolerances: the encoding is exact (same operation order, no
contraction); each MFMA layer differs from the CPU sum in accumulation order only, which can flip
a BF16 rounding, so outputs and gradients are compared by relative L2 (2e-2)."""
import numpy as np
import pytest

import mlp_oracle as MO
import render_engine_amd.mlp as M

pytestmark = pytest.mark.gpu


def encode_ref(cfg, table, P, N, D):
    return MO.encode(table, P, N, D, np.array(list(cfg.lo), np.float32), np.array(list(cfg.ext), np.float32), cfg.log2_table,
                     cfg.base_res, cfg.max_res)


def forward_ref(W, X):
    return MO.forward(W, X)


def backward_ref(cfg, W, acts, corners, dy):
    return MO.backward(W, acts, corners, dy, cfg.log2_table)


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
    print(f"[mlp] forward n={n}: relative L2 {_rel(y, yr):.2e}", flush=True)
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
    gWr, gTr = backward_ref(cfg, W, acts, corners, dy)
    print(f"[mlp] gradients: weights {_rel(gW, gWr):.2e}, table {_rel(gT, gTr):.2e}", flush=True)
    for L in range(M.HIDDEN):
        sl = slice(L * 4096, (L + 1) * 4096)
        assert _rel(gW[sl], gWr[sl]) < 2e-2, (L, _rel(gW[sl], gWr[sl]))
    assert _rel(gW[M.HIDDEN * 4096:], gWr[M.HIDDEN * 4096:]) < 2e-2
    assert _rel(gT, gTr) < 2e-2, _rel(gT, gTr)
    mlp.zero_grad()


def test_mlp_table_gradient_rounds_and_accumulation(torch_gpu):
    """150,000 samples: the table-gradient scatter runs in three rounds of at most 65,536 samples,
    each level's single 4,096-slot slice is summed in several parts (atomics between the parts); the
    gradient matches the oracle's, and a second backward on the same samples doubles it"""
    torch = torch_gpu
    lo, ext = np.array([-2.0, 1.0, 0.0], np.float32), np.array([8.0, 4.0, 3.0], np.float32)
    cfg = M.config(lo, ext, log2_table=12, base_res=4, max_res=256, max_batch=150_000, seed=4)
    mlp = M.MLPCache(cfg)
    rng = np.random.default_rng(12)
    mlp.set(M.MLP_TABLE, rng.uniform(-1, 1, M.LEVELS * 2 * (1 << 12)).astype(np.float32))
    n = 150_000
    P, N, D, (Pt, Nt, Dt) = _inputs(torch, n, 13, lo, ext)
    order = np.argsort(P[:, 0], kind="stable")   # partly sorted positions: runs of shared corners
    P, N, D = P[order], N[order], D[order]
    Pt, Nt, Dt = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (P, N, D))
    dy = (rng.normal(size=n) / n).astype(np.float32)
    dyt = torch.from_numpy(dy).cuda()
    mlp.zero_grad()
    mlp.backward(Pt, Nt, Dt, dyt)
    g1 = mlp.get(M.MLP_GTABLE).astype(np.float64)
    W = mlp.get(M.MLP_W)
    X, corners = encode_ref(cfg, mlp.get(M.MLP_TABLE), P, N, D)
    _, acts = forward_ref(W, X)
    _, gTr = backward_ref(cfg, W, acts, corners, dy)
    print(f"[mlp] 150k-sample table gradient: relative L2 {_rel(g1, gTr):.2e}", flush=True)
    assert _rel(g1, gTr) < 2e-2, _rel(g1, gTr)
    mlp.backward(Pt, Nt, Dt, dyt)
    g2 = mlp.get(M.MLP_GTABLE).astype(np.float64)
    assert _rel(g2, 2 * g1) < 1e-5, _rel(g2, 2 * g1)
    mlp.zero_grad()
    assert np.all(mlp.get(M.MLP_GTABLE) == 0)


@pytest.mark.parametrize("log2T,n", [(10, 1), (10, 37), (11, 300)])
def test_mlp_table_gradient_small_tables(torch_gpu, log2T, n):
    """tables smaller than one 4,096-slot slice (one slice of 2^log2T slots per level) and batches of
    a few samples: the table gradient matches the oracle's"""
    torch = torch_gpu
    lo, ext = np.array([-1.0, 0.0, 2.0], np.float32), np.array([3.0, 5.0, 2.0], np.float32)
    cfg = M.config(lo, ext, log2_table=log2T, base_res=4, max_res=128, max_batch=512, seed=6)
    mlp = M.MLPCache(cfg)
    rng = np.random.default_rng(20 + n)
    mlp.set(M.MLP_TABLE, rng.uniform(-1, 1, M.LEVELS * 2 * (1 << log2T)).astype(np.float32))
    P, N, D, (Pt, Nt, Dt) = _inputs(torch, n, 30 + n, lo, ext)
    dy = rng.normal(size=n).astype(np.float32)
    mlp.zero_grad()
    mlp.backward(Pt, Nt, Dt, torch.from_numpy(dy).cuda())
    gT = mlp.get(M.MLP_GTABLE)
    W = mlp.get(M.MLP_W)
    X, corners = encode_ref(cfg, mlp.get(M.MLP_TABLE), P, N, D)
    _, acts = forward_ref(W, X)
    _, gTr = backward_ref(cfg, W, acts, corners, dy)
    assert np.count_nonzero(gTr) > 0
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

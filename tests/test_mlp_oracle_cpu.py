"""The learned cache's CPU restatement (oracle/mlp_oracle.py) pinned without the GPU: bfloat16
rounding against torch's, trilinear interpolation against linear functions and grid corners, the
spatial hash against the primes of the Instant NGP paper, the forward against torch's matrix
products, both gradients against torch.autograd and central differences, Adam against
torch.optim.Adam."""
import numpy as np
import pytest
import torch

import mlp_oracle as MO

LO, EXT = np.array([-2.0, 1.0, 0.0], np.float32), np.array([8.0, 4.0, 2.0], np.float32)


def _dirs(rng, n):
    v = rng.normal(size=(n, 3))
    out = np.zeros((n, 4), np.float32)
    out[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return out


def _points(rng, n):
    P = np.zeros((n, 4), np.float32)
    P[:, :3] = LO + EXT * rng.random((n, 3))
    return P


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 20000),
                        # exact ties between two bfloat16 values (low half 0x8000) round to even
                        (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
                        np.array([0.0, -0.0, np.inf, -np.inf, 3.4e38, 1e-40], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    got = MO.bf16(x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_level_resolutions_geometric():
    res, dense = MO.level_resolutions(19, 16, 2048)
    assert res[0] == 16 and res[-1] == 2048
    g = (2048 / 16) ** (1 / 15)
    for lv, r in enumerate(res):
        assert r == int(np.floor(16 * g ** lv + 1e-9))
        assert dense[lv] == ((r + 1) ** 3 <= 1 << 19)
    assert dense[0] and not dense[-1]


def _dense_slot(r, x, y, z):
    # row-major corner grid of a dense level, x fastest
    return x + (r + 1) * (y + (r + 1) * z)


def test_trilinear_reproduces_linear_functions():
    """on a dense level whose table holds a linear function of the corner coordinates, every
    feature equals that function at the continuous grid position (interpolation is exact on linear
    functions): pins the corner order, the weights and the dense index"""
    rng = np.random.default_rng(1)
    log2T, base, top = 16, 4, 64
    res, dense = MO.level_resolutions(log2T, base, top)
    T = 1 << log2T
    table = np.zeros((MO.LEVELS, T, 2), np.float32)
    coef = {}
    for lv in range(MO.LEVELS):
        if not dense[lv]:
            continue
        r = res[lv]
        a, b = rng.uniform(-1, 1, (2, 3)), rng.uniform(-1, 1, 2)
        coef[lv] = (a, b)
        z, y, x = np.meshgrid(np.arange(r + 1), np.arange(r + 1), np.arange(r + 1), indexing="ij")
        slot = _dense_slot(r, x.ravel(), y.ravel(), z.ravel())
        for f in range(2):
            table[lv, slot, f] = a[f, 0] * x.ravel() + a[f, 1] * y.ravel() + a[f, 2] * z.ravel() + b[f]
    assert len(coef) >= 3
    n = 4000
    P, N, D = _points(rng, n), _dirs(rng, n), _dirs(rng, n)
    X, _ = MO.encode(table.ravel(), P, N, D, LO, EXT, log2T, base, top, bf=False)
    xn = np.clip((P[:, :3].astype(np.float64) - LO) / EXT, 0, 1)
    for lv, (a, b) in coef.items():
        s = xn * res[lv]
        for f in range(2):
            want = a[f, 0] * s[:, 0] + a[f, 1] * s[:, 1] + a[f, 2] * s[:, 2] + b[f]
            assert np.allclose(X[:, 2 * lv + f], want, rtol=2e-5, atol=2e-4 * res[lv])
    # the rest of the inputs: normal and view direction in [0, 1], a constant 1, zeros
    assert np.allclose(X[:, 32:35], 0.5 * (N[:, :3] + 1)) and np.allclose(X[:, 35:38], 0.5 * (D[:, :3] + 1))
    assert np.all(X[:, 38] == 1) and np.all(X[:, 39:] == 0)


def test_corners_one_hot_and_spatial_hash():
    """positions on grid corners read exactly one table entry per level: on hashed levels the entry
    of the Instant NGP spatial hash (x * 1 xor y * 2654435761 xor z * 805459861) mod T"""
    log2T = 12
    T = 1 << log2T
    res, dense = MO.level_resolutions(log2T, 16, 2048)
    rng = np.random.default_rng(2)
    table = rng.uniform(-1, 1, (MO.LEVELS, T, 2)).astype(np.float32)
    lo, ext = np.zeros(3, np.float32), np.ones(3, np.float32)
    for lv in (0, 5, 9, 15):
        r = res[lv]
        k = rng.integers(0, r, (50, 3)).astype(np.int64)   # corners below the top face
        P = np.zeros((50, 4), np.float32)
        P[:, :3] = (k / r).astype(np.float32)
        # only keep points whose float32 position maps back to exactly the corner
        xs = (P[:, :3] * np.float32(r)).astype(np.float32)
        keep = np.all(xs == k, axis=1)
        P, k = P[keep], k[keep]
        assert len(k) > 20
        X, corners = MO.encode(table.ravel(), P, _dirs(rng, len(k)), _dirs(rng, len(k)), lo, ext, log2T, 16, 2048, bf=False)
        if dense[lv]:
            slot = _dense_slot(r, k[:, 0], k[:, 1], k[:, 2])
        else:
            h = (k[:, 0].astype(np.uint64) * 1) ^ (k[:, 1].astype(np.uint64) * 2654435761 & 0xFFFFFFFF) ^ \
                (k[:, 2].astype(np.uint64) * 805459861 & 0xFFFFFFFF)
            slot = (h & (T - 1)).astype(np.int64)
        assert np.array_equal(X[:, 2 * lv], table[lv, slot, 0]) and np.array_equal(X[:, 2 * lv + 1], table[lv, slot, 1])
        ws = np.array([w for _, w in corners[lv]])
        assert np.array_equal(ws[0], np.ones(len(k), np.float32)) and np.all(ws[1:] == 0)
    # weights are a partition of unity everywhere
    P = _points(rng, 3000)
    _, corners = MO.encode(table.ravel(), P, _dirs(rng, 3000), _dirs(rng, 3000), LO, EXT, log2T, 16, 2048, bf=False)
    for lv in range(MO.LEVELS):
        s = np.sum([w.astype(np.float64) for _, w in corners[lv]], axis=0)
        assert np.allclose(s, 1.0, atol=3e-7)


def _torch_net(W, X):
    Wt = torch.from_numpy(np.asarray(W, np.float64))
    a = torch.from_numpy(np.asarray(X, np.float64))
    for L in range(MO.HIDDEN):
        a = torch.relu(a @ Wt[L * 4096:(L + 1) * 4096].reshape(64, 64))
    return (a @ Wt[MO.HIDDEN * 4096:].reshape(64, 16))[:, 0]


def _net_weights(rng):
    W = np.zeros(MO.NW, np.float32)
    W[:MO.HIDDEN * 4096] = rng.uniform(-1, 1, MO.HIDDEN * 4096) * np.sqrt(6 / 64)
    W[MO.HIDDEN * 4096:] = rng.uniform(-1, 1, 64 * 16) * 0.3
    return W


def test_forward_against_torch():
    rng = np.random.default_rng(3)
    W = _net_weights(rng)
    X = rng.uniform(0, 1, (500, 64)).astype(np.float32)
    y, acts = MO.forward(W, X, bf=False)
    assert np.allclose(y, _torch_net(W, X).numpy(), rtol=1e-12, atol=1e-12)
    assert len(acts) == MO.HIDDEN + 1 and np.all(acts[-1] >= 0)
    # with BF16 operands: the same network on torch's bfloat16-rounded inputs, weights and activations
    yb, actsb = MO.forward(W, X, bf=True)
    Wt = torch.from_numpy(W).to(torch.bfloat16).double()
    a = torch.from_numpy(X).to(torch.bfloat16).double()
    for L in range(MO.HIDDEN):
        a = torch.relu(a @ Wt[L * 4096:(L + 1) * 4096].reshape(64, 64)).float().to(torch.bfloat16).double()
    want = (a @ Wt[MO.HIDDEN * 4096:].reshape(64, 16))[:, 0].numpy()
    assert np.allclose(yb, want, rtol=1e-12, atol=1e-12)


@pytest.fixture(scope="module")
def small_net():
    rng = np.random.default_rng(4)
    log2T = 10
    T = 1 << log2T
    table = rng.uniform(-1, 1, MO.LEVELS * T * 2).astype(np.float32)
    W = _net_weights(rng)
    n = 300
    P, N, D = _points(rng, n), _dirs(rng, n), _dirs(rng, n)
    dy = rng.normal(size=n)
    return log2T, table, W, P, N, D, dy


def test_gradients_against_autograd(small_net):
    """weight and table gradients of sum(dy * y) equal torch.autograd's through a torch graph that
    gathers the features from the table at the oracle's corners"""
    log2T, table, W, P, N, D, dy = small_net
    X, corners = MO.encode(table, P, N, D, LO, EXT, log2T, 4, 256, bf=False)
    y, acts = MO.forward(W, X, bf=False)
    gW, gT = MO.backward(W, acts, corners, dy, log2T, bf=False)

    T = 1 << log2T
    tt = torch.from_numpy(table.astype(np.float64)).reshape(MO.LEVELS, T, 2).requires_grad_(True)
    Wt = torch.from_numpy(W.astype(np.float64)).requires_grad_(True)
    feats = []
    for lv in range(MO.LEVELS):
        f = 0
        for idx, w in corners[lv]:
            f = f + torch.from_numpy(w.astype(np.float64))[:, None] * tt[lv, torch.from_numpy(idx)]
        feats.append(f)
    Xt = torch.cat(feats + [torch.from_numpy(X[:, 32:].astype(np.float64))], dim=1)
    a = Xt
    for L in range(MO.HIDDEN):
        a = torch.relu(a @ Wt[L * 4096:(L + 1) * 4096].reshape(64, 64))
    yt = (a @ Wt[MO.HIDDEN * 4096:].reshape(64, 16))[:, 0]
    assert np.allclose(yt.detach().numpy(), y, rtol=1e-5, atol=1e-6)   # float32 encode vs float64 gather
    (yt * torch.from_numpy(dy)).sum().backward()
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(gW, Wt.grad.numpy()) < 1e-5
    assert rel(gT, tt.grad.numpy().ravel()) < 1e-5
    assert np.count_nonzero(gT) > 1000


def test_table_gradient_central_differences(small_net):
    log2T, table, W, P, N, D, dy = small_net
    X, corners = MO.encode(table, P, N, D, LO, EXT, log2T, 4, 256, bf=False)
    _, acts = MO.forward(W, X, bf=False)
    _, gT = MO.backward(W, acts, corners, dy, log2T, bf=False)
    h = 2e-3
    for j in np.argsort(-np.abs(gT))[:12]:
        tp, tm = table.copy(), table.copy()
        tp[j] += h
        tm[j] -= h
        fp = float(dy @ MO.forward(W, MO.encode(tp, P, N, D, LO, EXT, log2T, 4, 256, bf=False)[0], bf=False)[0])
        fm = float(dy @ MO.forward(W, MO.encode(tm, P, N, D, LO, EXT, log2T, 4, 256, bf=False)[0], bf=False)[0])
        fd = (fp - fm) / (tp[j] - tm[j])
        assert abs(fd - gT[j]) < 5e-3 * max(abs(gT[j]), 1e-3), (j, fd, gT[j])


def test_adam_against_torch():
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=1000)
    grads = [rng.normal(size=1000) * s for s in (1.0, 1e-3, 10.0)]
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-15
    pt = torch.nn.Parameter(torch.from_numpy(p0.copy()))
    opt = torch.optim.Adam([pt], lr=lr, betas=(b1, b2), eps=eps)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step, g in enumerate(grads, start=1):
        pt.grad = torch.from_numpy(g.copy())
        opt.step()
        p, m, v = MO.adam(p, g, m, v, step, lr, b1, b2, eps)
        assert np.allclose(p, pt.detach().numpy(), rtol=1e-12, atol=1e-14)

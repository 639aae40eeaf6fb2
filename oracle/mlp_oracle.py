"""CPU restatement of the hash-encoded MLP radiance cache (numpy) — TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the checker of the library's learned-cache kernels; never by the product
package. The reference has no code for a learned cache (its one is NVIDIA's closed NRC library,
`README.md:15`, `LINUX.md:9`), so this restates the network the library defines:

* encoding: a 16-level multiresolution grid over the position normalised to [0, 1]^3 by the scene
  bounds (Mueller et al. 2022, "Instant Neural Graphics Primitives", section 3): level resolutions
  floor(base * g^l) with g = (max / base)^(1/15); 2 features per level, 2^log2T entries per level;
  a level whose (res + 1)^3 corner grid fits the table is indexed densely (x + (r+1)(y + (r+1)z)),
  finer ones by the paper's spatial hash (x * 1 xor y * 2654435761 xor z * 805459861) mod 2^log2T;
  trilinear weights ((wx wy) wz) and features summed over the 8 corners in corner order
  c = x + 2y + 4z; then the normal and the view direction mapped to [0, 1], a constant 1, zeros:
  64 inputs;
* 5 hidden layers of 64 ReLU units and a scalar output, no biases, weights stored layer by layer
  as [k][j] (input k, unit j), the output column at j = 0 of a 64 x 16 block;
* with `bf16=True` every layer's inputs and weights are rounded to bfloat16 (round to nearest
  even), as the library's MFMA operands are, and products accumulate in float64 here;
* backward: weight gradients A^T G, the input-feature gradients spread over each level's corners
  with the trilinear weights (the table gradient);
* Adam in torch.optim.Adam's form.

Pinned by tests/test_mlp_oracle_cpu.py: trilinear interpolation reproduces linear functions of
the corner coordinates and is one-hot at corners, the hash against the paper's primes, the forward
against torch's matrix products, both gradients against torch.autograd and central differences,
the Adam step against torch.optim.Adam.
"""
import numpy as np

LEVELS, HIDDEN, WIDTH = 16, 5, 64
NW = HIDDEN * WIDTH * WIDTH + WIDTH * 16
PRIMES = (1, 2654435761, 805459861)


def bf16(x):
    """round float32 values to bfloat16 (nearest, ties to even), returned as float32"""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).reshape(np.shape(x))


def level_resolutions(log2_table, base_res, max_res):
    """grid resolution and dense flag of each level"""
    g = np.exp((np.log(float(max_res)) - np.log(float(base_res))) / (LEVELS - 1))
    res = [int(np.floor(base_res * g ** lv + 1e-9)) for lv in range(LEVELS)]
    dense = [(r + 1) ** 3 <= (1 << log2_table) for r in res]
    return res, dense


def corner_index(res, dense, log2_table, cx, cy, cz):
    """table slot of integer corner coordinates (uint32 arrays) at one level"""
    cx, cy, cz = (np.asarray(v, np.uint64) for v in (cx, cy, cz))
    if dense:
        r1 = np.uint64(res + 1)
        return (cx + r1 * (cy + r1 * cz)).astype(np.int64)
    m = np.uint64((1 << 32) - 1)
    h = (cx * np.uint64(PRIMES[0])) & m
    h ^= (cy * np.uint64(PRIMES[1])) & m
    h ^= (cz * np.uint64(PRIMES[2])) & m
    return (h & np.uint64((1 << log2_table) - 1)).astype(np.int64)


def normalise(lo, ext, P):
    lo = np.asarray(lo, np.float32)
    iext = (np.float32(1) / np.asarray(ext, np.float32)).astype(np.float32)
    return np.clip(((np.asarray(P, np.float32)[:, :3] - lo).astype(np.float32) * iext).astype(np.float32),
                   np.float32(0), np.float32(1))


def cell(x, r):
    """base corner and fractions of normalised positions x at resolution r (float32 steps)"""
    r = np.float32(r)
    s = (x * r).astype(np.float32)
    f = np.minimum(np.floor(s), r - np.float32(1)).astype(np.float32)
    return f.astype(np.uint32), (s - f).astype(np.float32)


def corner_weight(fr, c):
    wx = fr[:, 0] if c & 1 else np.float32(1) - fr[:, 0]
    wy = fr[:, 1] if (c >> 1) & 1 else np.float32(1) - fr[:, 1]
    wz = fr[:, 2] if c >> 2 else np.float32(1) - fr[:, 2]
    return ((wx * wy).astype(np.float32) * wz).astype(np.float32)


def encode(table, P, N, D, lo, ext, log2_table, base_res, max_res, bf=True):
    """64 network inputs per sample, and per level the 8 (slot, weight) corner pairs"""
    res, dense = level_resolutions(log2_table, base_res, max_res)
    T = 1 << log2_table
    tb = np.asarray(table, np.float32).reshape(LEVELS, T, 2)
    x = normalise(lo, ext, P)
    n = len(x)
    X = np.zeros((n, WIDTH), np.float32)
    corners = []
    for lv in range(LEVELS):
        i, fr = cell(x, res[lv])
        f0 = np.zeros(n, np.float32)
        f1 = np.zeros(n, np.float32)
        cl = []
        for c in range(8):
            idx = corner_index(res[lv], dense[lv], log2_table, i[:, 0] + np.uint32(c & 1), i[:, 1] + np.uint32((c >> 1) & 1),
                               i[:, 2] + np.uint32(c >> 2))
            w = corner_weight(fr, c)
            f0 = (f0 + (w * tb[lv, idx, 0]).astype(np.float32)).astype(np.float32)
            f1 = (f1 + (w * tb[lv, idx, 1]).astype(np.float32)).astype(np.float32)
            cl.append((idx, w))
        X[:, 2 * lv], X[:, 2 * lv + 1] = f0, f1
        corners.append(cl)
    X[:, 32:35] = np.float32(0.5) * (np.asarray(N, np.float32)[:, :3] + np.float32(1))
    X[:, 35:38] = np.float32(0.5) * (np.asarray(D, np.float32)[:, :3] + np.float32(1))
    X[:, 38] = 1
    return (bf16(X) if bf else X), corners


def split_weights(W, bf=True):
    W = bf16(W) if bf else np.asarray(W, np.float32)
    Ws = [W[L * WIDTH * WIDTH:(L + 1) * WIDTH * WIDTH].reshape(WIDTH, WIDTH).astype(np.float64) for L in range(HIDDEN)]
    return Ws, W[HIDDEN * WIDTH * WIDTH:].reshape(WIDTH, 16).astype(np.float64)


def forward(W, X, bf=True):
    """network output per sample and every layer's input activations"""
    Ws, W5 = split_weights(W, bf)
    a = (bf16(X) if bf else np.asarray(X, np.float32)).astype(np.float64)
    acts = [a]
    for L in range(HIDDEN):
        z = np.maximum(a @ Ws[L], 0.0)
        a = bf16(z.astype(np.float32)).astype(np.float64) if bf else z
        acts.append(a)
    return (a @ W5)[:, 0], acts


def backward(W, acts, corners, dy, log2_table, bf=True):
    """gradients of sum(dy * y) w.r.t. the weights (NW floats) and the table (LEVELS * 2^log2T * 2)"""
    Ws, W5 = split_weights(W, bf)
    rnd = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if bf else (lambda v: v)
    dy = np.asarray(dy, np.float64)
    gW = np.zeros(NW, np.float64)
    a5 = acts[-1]
    gW[HIDDEN * WIDTH * WIDTH:].reshape(WIDTH, 16)[:, 0] = a5.T @ dy
    gz = rnd(np.where(a5 > 0, dy[:, None] * W5[:, 0][None, :], 0.0))
    gX = None
    for L in range(HIDDEN - 1, -1, -1):
        gW[L * WIDTH * WIDTH:(L + 1) * WIDTH * WIDTH] = (acts[L].T @ gz).ravel()
        ga = gz @ Ws[L].T
        if L > 0:
            gz = rnd(np.where(acts[L] > 0, ga, 0.0))
        else:
            gX = ga[:, :2 * LEVELS]
    T = 1 << log2_table
    gt = np.zeros((LEVELS, T, 2), np.float64)
    for lv in range(LEVELS):
        for idx, w in corners[lv]:
            np.add.at(gt[lv, :, 0], idx, w.astype(np.float64) * gX[:, 2 * lv])
            np.add.at(gt[lv, :, 1], idx, w.astype(np.float64) * gX[:, 2 * lv + 1])
    return gW, gt.ravel()


def adam(p, g, m, v, step, lr, b1, b2, eps):
    """one Adam step (bias-corrected moments, torch.optim.Adam's update); returns (p, m, v)"""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    c1, c2 = 1.0 - b1 ** step, 1.0 - b2 ** step
    return p - (lr / c1) * m / (np.sqrt(v) / np.sqrt(c2) + eps), m, v

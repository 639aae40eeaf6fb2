"""The learned cache inside the frame: its camera pass against the grid cache's (bit for bit
when the network is set to a constant and the grid's resolved radiance to the same constant),
and online training from the scene lowering the error judged against the grid cache."""
import numpy as np
import pytest

import render_engine_amd as rea
import render_engine_amd.mlp as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def street(torch_gpu):
    return rea.Scene.street(seed=3, nboxes=150, target_tris=12000, nlamps=6, sun=True, target_points=30000, table_log2=0,
                            W=96, H=64, spp=2, bounces=3).upload(trav=rea.TRAV_WIDE)


def _mlp(scene, **kw):
    i = scene.info
    return M.MLPCache(M.config(list(i.lo), list(i.ext), **kw))


def test_mlp_render_constant_matches_grid(torch_gpu, street):
    torch = torch_gpu
    c0 = 0.625   # exact in bfloat16
    mlp = _mlp(street, log2_table=14, max_res=512, max_batch=4096)
    W = np.zeros(M.NW, np.float32)
    for L in range(M.HIDDEN):   # the constant input (feature 38) through unit 0 of every layer
        W[L * 4096 + (38 if L == 0 else 0) * 64 + 0] = 1.0
    W[M.HIDDEN * 4096 + 0 * 16 + 0] = c0
    mlp.set(M.MLP_W, W)
    cache = rea.Cache(street, rea.methods.plain())
    cache.begin(0, 3)
    cache.frame(0)
    cache.set("R", np.full(street.info.Np, c0, np.float32))
    cam = street.camera()
    for st, spp in ((0, 2), (3, 3)):
        if st != 0:
            cache.frame(st)
            cache.set("R", np.full(street.info.Np, c0, np.float32))
        cam.spp = spp
        for mode in (rea.RENDER_STAGED, rea.RENDER_FUSED):   # learned-cache pass: staged, then fused
            cache.set_render_mode(mode)
            a = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
            b = torch.zeros_like(a)
            cache.render(a, cam, seed=21)
            ms = mlp.render(cache, b, cam, seed=21, timed=True)
            ga, gb = a.cpu().numpy(), b.cpu().numpy()
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), f"state {st} mode {mode}: {np.sum(ga != gb)} pixels differ"
            assert 0.0 < mlp.last_infer_ms() < ms   # the batched inference, timed inside the timed pass


def test_mlp_render_grid_fallback_per_query(torch_gpu, street):
    """with the grid as per-query judge, a query whose prediction differs from the grid's resolved radiance
    at its entry by more than the threshold ends on the grid's value: a network that predicts c1 everywhere
    against a grid holding c1 on half of the entries and c0 far from it elsewhere gives the grid's own image
    (both schedules); a threshold above the gap gives the network's image; 0 turns the judge off"""
    torch = torch_gpu
    c1, c0 = 0.625, 0.25
    mlp = _mlp(street, log2_table=14, max_res=512, max_batch=4096)
    W = np.zeros(M.NW, np.float32)
    for L in range(M.HIDDEN):
        W[L * 4096 + (38 if L == 0 else 0) * 64 + 0] = 1.0
    W[M.HIDDEN * 4096] = c1
    mlp.set(M.MLP_W, W)
    cache = rea.Cache(street, rea.methods.plain())
    cache.begin(0, 3)
    cache.frame(0)
    Np = street.info.Np
    R = np.where(np.arange(Np) % 2 == 0, c1, c0).astype(np.float32)
    cam = street.camera()
    a = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    for mode in (rea.RENDER_FUSED, rea.RENDER_STAGED):
        cache.set_render_mode(mode)
        for thr, grid_R in ((0.5, R), (2.0, np.full(Np, c1, np.float32)), (0.0, np.full(Np, c1, np.float32))):
            mlp.set_grid_fallback(thr)
            cache.set("R", R)
            mlp.render(cache, b, cam, seed=5)
            cache.set("R", grid_R)
            cache.render(a, cam, seed=5)
            ga, gb = a.cpu().numpy(), b.cpu().numpy()
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32)), f"mode {mode} thr {thr}: {np.sum(ga != gb)} pixels differ"
    mlp.set_grid_fallback(0.0)


def test_mlp_online_training_against_the_grid(torch_gpu, street):
    torch = torch_gpu
    cache = rea.Cache(street, rea.methods.plain())
    cache.begin(0, 5)
    for _ in range(24):   # grid cache near its converged state
        cache.frame(0)
    mlp = _mlp(street, log2_table=16, max_res=1024, max_batch=8192, lr=1e-2)
    judges, losses = [], []
    for it in range(250):
        loss, judge = mlp.train_scene(cache, n=8192, seed=100 + it)
        losses.append(loss); judges.append(judge)
    assert np.all(np.isfinite(losses)) and np.all(np.isfinite(judges))
    first, last = np.mean(judges[:5]), np.mean(judges[-10:])
    print(f"judge: first {first:.3f} last {last:.3f}; loss {np.mean(losses[:5]):.4f} -> {np.mean(losses[-10:]):.4f}")
    assert last < 0.5 * first and last < 0.35, (first, last)
    img = torch.zeros(street.info.cam.W * street.info.cam.H, dtype=torch.float32, device="cuda")
    mlp.render(cache, img, seed=3)
    ref = torch.zeros_like(img)
    cache.render(ref, seed=3)
    a, b = img.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.isfinite(a))
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.35

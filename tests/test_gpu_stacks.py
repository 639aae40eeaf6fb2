"""Traversal stacks split between LDS and global memory: the scene's tracing kernels (direct
light, entry pass, adjoint) give the same results whatever the number of LDS levels (1 level
puts nearly every push of every ray in the global levels; 64 keeps them all in LDS)."""
import numpy as np
import pytest

import render_engine_amd as rea
import render_engine_amd.adjoint as ad

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(torch_gpu):
    out = {}
    for levels in (64, 1):   # set before the upload, whose direct-light kernel then uses it too
        s = rea.Scene.street(seed=3, nboxes=150, target_tris=12000, nlamps=6, sun=True, target_points=30000, table_log2=0,
                             W=96, H=64, spp=2, bounces=3)
        out[levels] = s.set_stack_levels(levels).upload(trav=rea.TRAV_WIDE)
    return out


def test_scene_kernels_same_results(torch_gpu, pair):
    a, b = pair[64], pair[1]
    for st in range(a.info.NS):
        for name in ("EDIR", "SV", "LM"):
            assert np.array_equal(a.array(name, st).view(np.uint8), b.array(name, st).view(np.uint8)), (name, st)
    res = {}
    for levels, s in pair.items():
        c = rea.Cache(s, rea.methods.plain())
        c.begin(0, 5)
        for f in range(4):
            c.frame(0 if f < 2 else 2)
        res[levels] = {k: c.get(k) for k in ("LC", "R", "M", "KK")}
    for k in res[64]:
        assert np.array_equal(res[64][k].view(np.uint8), res[1][k].view(np.uint8)), k


def test_adjoint_same_results(torch_gpu, pair):
    torch = torch_gpu
    out = {}
    for levels, s in pair.items():
        adj = ad.Adjoint(s, state=0, K=2, bounces=3, seed=4)
        I = torch.tensor(s.array("INTEN", 0), dtype=torch.float32, device="cuda")
        alb = torch.tensor(s.array("MATALB"), dtype=torch.float32, device="cuda")
        R = adj.forward(I, alb).cpu().numpy()
        SR = adj.support()
        gI, ga = adj.backward(torch.ones(s.info.Np, dtype=torch.float32, device="cuda"))
        out[levels] = (R, SR, gI.cpu().numpy(), ga.cpu().numpy())
    assert np.array_equal(out[64][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[64][1], out[1][1])
    for k in (2, 3):   # the reverse pass sums with float atomics, whose order may differ
        x, y = out[64][k], out[1][k]
        assert np.linalg.norm(x - y) <= 1e-5 * np.linalg.norm(x), k

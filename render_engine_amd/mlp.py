"""Hash-encoded MLP radiance cache over the library's C ABI (rc_mlp_*).

Encoding: 16-level multiresolution hash grid of the position (2 features per level), the normal
and the view direction; 5 hidden layers of 64 ReLU units; scalar output. BF16 MFMA kernels with
FP32 accumulation; Adam on FP32 master weights and table. Inputs are float4 device tensors
(xyz used) of positions, normals and view directions.
"""
import ctypes as C

import numpy as np
import torch

from ._ffi import Camera, check, lib  # noqa: F401

MLP_W, MLP_TABLE, MLP_GW, MLP_GTABLE = range(4)
LEVELS, HIDDEN, WIDTH = 16, 5, 64
NW = HIDDEN * WIDTH * WIDTH + WIDTH * 16


class Config(C.Structure):
    _fields_ = [("log2_table", C.c_int), ("base_res", C.c_int), ("max_res", C.c_int), ("max_batch", C.c_int),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint32),
                ("lo", C.c_float * 3), ("ext", C.c_float * 3)]


def config(lo, ext, log2_table=19, base_res=16, max_res=2048, max_batch=1 << 16, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15,
           seed=1):
    return Config(log2_table, base_res, max_res, max_batch, lr, beta1, beta2, eps, seed, (C.c_float * 3)(*lo), (C.c_float * 3)(*ext))


def level_resolutions(cfg):
    """grid resolution and dense flag of every level (the library's rule)"""
    g = np.exp((np.log(cfg.max_res) - np.log(cfg.base_res)) / (LEVELS - 1))
    res = [int(np.floor(cfg.base_res * g ** l + 1e-9)) for l in range(LEVELS)]
    dense = [(r + 1) ** 3 <= (1 << cfg.log2_table) for r in res]
    return res, dense


def _dev4(t, n, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and tuple(t.shape) == (n, 4)):
        raise ValueError(f"{name}: expected a contiguous float32 device tensor of shape ({n}, 4)")
    return C.c_void_p(t.data_ptr())


def _dev1(t, n, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.numel() == n):
        raise ValueError(f"{name}: expected a contiguous float32 device tensor of {n} elements")
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class MLPCache:
    def __init__(self, cfg):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib.rc_mlp_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def __del__(self, _destroy=lib.rc_mlp_destroy):
        h = getattr(self, "_h", None)
        if h:
            _destroy(h)
            self._h = None

    def infer(self, P, N, D, y=None):
        n = P.shape[0]
        y = torch.empty(n, dtype=torch.float32, device=P.device) if y is None else y
        check(lib.rc_mlp_infer(self._h, _dev4(P, n, "P"), _dev4(N, n, "N"), _dev4(D, n, "D"), n, _dev1(y, n, "y"), _stream()))
        return y

    def backward(self, P, N, D, dy):
        n = P.shape[0]
        check(lib.rc_mlp_backward(self._h, _dev4(P, n, "P"), _dev4(N, n, "N"), _dev4(D, n, "D"), _dev1(dy, n, "dy"), n, _stream()))

    def train(self, P, N, D, t, want_loss=True):
        n = P.shape[0]
        loss = C.c_float(0.0)
        check(lib.rc_mlp_train(self._h, _dev4(P, n, "P"), _dev4(N, n, "N"), _dev4(D, n, "D"), _dev1(t, n, "t"), n,
                               C.byref(loss) if want_loss else None, _stream()))
        return loss.value if want_loss else None

    def step(self):
        check(lib.rc_mlp_step(self._h, _stream()))

    def zero_grad(self):
        check(lib.rc_mlp_zero_grad(self._h, _stream()))

    def train_scene(self, cache, n=1 << 16, seed=1):
        """one online training step from the cache's scene and state; returns (loss, judge) where
        judge is the relative RMS of the predictions against the grid cache's resolved radiance"""
        loss, judge = C.c_float(0.0), C.c_float(0.0)
        check(lib.rc_mlp_train_scene(self._h, cache._h, n, seed, C.byref(loss), C.byref(judge)))
        return loss.value, judge.value

    def set_grid_fallback(self, thr):
        """queries whose prediction differs from the grid cache's resolved radiance at their entry by more
        than thr times it end on the grid's value (0: never)"""
        check(lib.rc_mlp_set_grid_fallback(self._h, float(thr)))

    def render(self, cache, img, cam=None, seed=7, timed=False):
        """camera pass of the cache's scene ending on the MLP at cam.cache_bounce (the cache's render
        mode: fused paths recording their queries, or the staged pass; the same image); ms when timed"""
        cam = cam or cache.scene.camera()
        ms = C.c_float(0.0)
        check(lib.rc_mlp_render(self._h, cache._h, C.byref(cam), seed, C.c_void_p(img.data_ptr()), C.byref(ms) if timed else None,
                                C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return ms.value if timed else None

    def last_infer_ms(self):
        """time of the batched inference inside the last render(..., timed=True) (HIP events on its stream)"""
        ms = C.c_float(0.0)
        check(lib.rc_mlp_last_infer_ms(self._h, C.byref(ms)))
        return ms.value

    def get(self, which):
        nb = lib.rc_mlp_array_bytes(self._h, which)
        out = np.empty(nb // 4, np.float32)
        check(lib.rc_mlp_get(self._h, which, out.ctypes.data_as(C.c_void_p), nb))
        return out

    def set(self, which, arr):
        a = np.ascontiguousarray(arr, np.float32)
        check(lib.rc_mlp_set(self._h, which, a.ctypes.data_as(C.c_void_p), a.nbytes))

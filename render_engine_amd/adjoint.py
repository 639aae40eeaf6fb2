"""Adjoint of the cached radiance with respect to per-light emission and per-material albedo,
as a torch.autograd.Function over the library's C ABI (rc_adjoint_*).

Model (the cache's transport with fixed rays): G = unit-intensity direct light of the state's
geometry, E_i = sum_l I_l G_il, A_i = alb[mat_i], L^(0) = 0,
L^(b+1)_i = (1/K) sum over entry i's K rays of bounce b that reach a valid entry h of
A_h (E_h + L^(b)_h), and the cached radiance R_i = A_i (E_i + L^(B)_i). The backward pass
re-traces every bounce's rays from their seeds (no stored paths) and scatters the adjoints to
the hit entries; the nonzero support of dR_i/dI_l is returned by `support()` as 64-bit masks
(bit min(l, 63)), the per-entry light dependencies that gradient-driven invalidation uses.
"""
import ctypes as C

import numpy as np
import torch

from ._ffi import check, lib

ADJ_G, ADJ_SG, ADJ_SR, ADJ_LB, ADJ_DA, ADJ_DE = range(6)
_ADT = {ADJ_G: np.float32, ADJ_SG: np.uint64, ADJ_SR: np.uint64, ADJ_LB: np.float32, ADJ_DA: np.float32, ADJ_DE: np.float32}


def _dev(t, n, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.numel() == n):
        raise ValueError(f"{name}: expected a contiguous float32 device tensor of {n} elements")
    return C.c_void_p(t.data_ptr())


class Adjoint:
    """rc_adjoint handle on one state of an uploaded scene (K rays per entry and bounce, B bounces)."""

    def __init__(self, scene, state=0, K=4, bounces=3, seed=1):
        if not scene.uploaded:
            raise ValueError("scene not uploaded")
        self.scene, self.state, self.K, self.B, self.seed = scene, state, K, bounces, seed
        i = scene.info
        self.Np, self.nl, self.nmat = i.Np, i.nl, i.nmat
        h = C.c_void_p()
        check(lib.rc_adjoint_create(scene._h, state, K, bounces, seed, C.byref(h)))
        self._h = h
        self.generation = 0   # forwards run so far (the backward needs its own forward's state)

    def __del__(self, _destroy=lib.rc_adjoint_destroy):
        h = getattr(self, "_h", None)
        if h:
            _destroy(h)
            self._h = None

    def forward(self, I, alb, R=None):
        """R[Np] (float32 device) from emission I[nl] and material albedos alb[nmat]"""
        if R is None:
            R = torch.empty(self.Np, dtype=torch.float32, device=I.device)
        check(lib.rc_adjoint_forward(self._h, _dev(I, self.nl, "I"), _dev(alb, self.nmat, "alb"), _dev(R, self.Np, "R"),
                                     C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        self.generation += 1
        return R

    def backward(self, gR):
        """gradient of sum_i gR_i R_i at the last forward: (dI[nl], dalb[nmat])"""
        gI = torch.empty(self.nl, dtype=torch.float32, device=gR.device)
        galb = torch.empty(self.nmat, dtype=torch.float32, device=gR.device)
        check(lib.rc_adjoint_backward(self._h, _dev(gR, self.Np, "gR"), _dev(gI, self.nl, "gI"), _dev(galb, self.nmat, "galb"),
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return gI, galb

    def get(self, which):
        nb = lib.rc_adjoint_array_bytes(self._h, which)
        if nb == 0:
            raise ValueError("array not available (not computed yet)")
        out = np.empty(nb // np.dtype(_ADT[which]).itemsize, _ADT[which])
        check(lib.rc_adjoint_get(self._h, which, out.ctypes.data_as(C.c_void_p), nb))
        return out

    def set_timed(self, on=True):
        """bracket every traced bounce of the following forward / backward calls with HIP events
        (each call then waits for its last bounce)"""
        check(lib.rc_adjoint_set_timed(self._h, 1 if on else 0))

    def trace_ms(self):
        """(forward, backward): GPU time of the traced bounces of the last timed calls, summed over the B launches"""
        ms = (C.c_float * 2)()
        check(lib.rc_adjoint_trace_ms(self._h, ms))
        return float(ms[0]), float(ms[1])

    def support(self):
        """per-entry 64-bit masks of the lights l with dR_i/dI_l != 0 at the last forward"""
        return self.get(ADJ_SR)


class CachedRadiance(torch.autograd.Function):
    """R = CachedRadiance.apply(adj, I, alb): cached radiance of every entry; backward gives
    dJ/dI and dJ/dalb from the adjoint kernels."""

    @staticmethod
    def forward(ctx, adj, I, alb):
        I, alb = I.detach().contiguous(), alb.detach().contiguous()
        R = adj.forward(I, alb)
        ctx.adj, ctx.generation = adj, adj.generation
        ctx.save_for_backward(I, alb)
        return R

    @staticmethod
    def backward(ctx, gR):
        adj = ctx.adj
        if adj.generation != ctx.generation:   # another forward ran since: replay this one's bounces
            I, alb = ctx.saved_tensors
            adj.forward(I, alb)
        gI, galb = adj.backward(gR.contiguous())
        return None, gI, galb

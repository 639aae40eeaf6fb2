"""render_engine_amd — dependency-aware radiance cache on AMD Instinct MI355X.

Python host mirror of the prototype's scene / event / method interface (cuda/poc_gpu.cu:
Scene + load :737-801, Method :803-805, Runner::run :826-1016, scenario files :1378-1394)
over the C ABI of librc.so (include/rc.h). All compute runs in the HIP library; this module
only marshals arguments, and raises RcError when the library or a GPU is missing.
"""
import ctypes as C

import numpy as np

from ._ffi import A, CA, Camera, FrameStats, Method, RcError, SceneInfo, check, lib
from . import methods, scenarios  # noqa: F401

__all__ = ["Scene", "Cache", "Method", "Camera", "RcError", "methods", "scenarios", "device_count", "trace", "hash_insert",
           "hash_query", "hash_point", "dda_mark", "dda_volume", "TRAV_REF", "TRAV_BINARY", "TRAV_WIDE", "TRAV_WIDE_Q"]

TRAV_REF, TRAV_BINARY, TRAV_WIDE, TRAV_WIDE_Q = 0, 1, 2, 3   # 3: camera pass and trace() on quantized wide nodes
RENDER_FUSED, RENDER_STAGED = 0, 1
# compact snapshot record (rc_snap_rec, 24 bytes)
SNAP_REC = np.dtype([("index", np.uint32), ("n", np.float32), ("key", np.uint64), ("radiance", np.float32), ("age", np.float32)])

_DT = {"ST": np.float32, "STN": np.float32, "HK": np.uint64, "HV": np.int32, "P": np.float32, "N": np.float32,
       "ALB": np.float32, "REGION": np.int32, "L": np.float32, "LTYPE": np.int32, "INTEN": np.float32, "Q": np.float32,
       "DIMS": np.int32, "OFF": np.int32, "T": np.float32, "TQ": np.int32, "TALB": np.float32, "EDIR": np.float32,
       "SV": np.uint32, "LM": np.uint64, "OCC": np.int32, "EM": np.int32, "REF": np.float32, "VALID": np.uint8,
       "PW": np.float32, "ORDER": np.int32, "OBJBOX": np.float32, "MESH_V": np.float32, "MESH_ALB": np.float32,
       "MESH_DS": np.uint8, "VOL": np.uint32, "OCCF": np.int32,
       "MAT": np.int32, "MATALB": np.float32}
_CDT = {"LC": np.float32, "M": np.float32, "N": np.float32, "KK": np.int32, "PENDING": np.uint8, "CUR": np.uint32,
        "PREV": np.uint32, "LCUR": np.uint64, "LPREV": np.uint64, "SUMPR": np.int64, "NPEND": np.int32, "AGE": np.float32,
        "STALE": np.uint64, "C0": np.float32, "RSTAT": np.int64, "LNEW": np.uint64, "DSTAT": np.int64,
        "R": np.float32}


def device_count():
    return int(lib.rc_device_count())


def prep_device(on):
    """Build the BVH and sample, reduce and Morton-order the cache cells of the scenes created
    from now on on the GPU (True; the same tree and cells bit for bit) or on the host (False,
    the default)."""
    check(lib.rc_scene_prep_device(1 if on else 0))


def _ptr(t):
    """device pointer of a torch tensor (or None -> NULL)"""
    if t is None:
        return None
    if not t.is_cuda:
        raise RcError("expected a device tensor")
    if not t.is_contiguous():
        raise RcError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    return C.c_void_p(stream.cuda_stream) if stream is not None else None


class Scene:
    """A prepared scene: static triangles + BVHs, cache entries in a 64-bit-key hash grid,
    lights, per-state intensities and dynamic quads, per state-pair event voxels."""

    def __init__(self, handle):
        self._h = handle
        self.uploaded = False

    @classmethod
    def box_room(cls, seed=1, target_tris=2000):
        h = C.c_void_p()
        check(lib.rc_scene_box_room(seed, target_tris, C.byref(h)))
        return cls(h)

    @classmethod
    def street(cls, seed=1, nboxes=2000, target_tris=200_000, nlamps=32, sun=True, target_points=530_000, table_log2=22,
               W=960, H=540, spp=4, bounces=3):
        h = C.c_void_p()
        check(lib.rc_scene_street(seed, nboxes, target_tris, nlamps, 1 if sun else 0, target_points, table_log2, W, H, spp,
                                  bounces, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        """prototype scene file (format 2 / 3); .gz files are decompressed to a temporary file"""
        import gzip
        import os
        import shutil
        import tempfile
        tmp = None
        if path.endswith(".gz"):
            fd, tmp = tempfile.mkstemp(suffix=".bin")
            with gzip.open(path, "rb") as src, os.fdopen(fd, "wb") as dst:
                shutil.copyfileobj(src, dst)
            path = tmp
        try:
            h = C.c_void_p()
            check(lib.rc_scene_load(path.encode(), C.byref(h)))
        finally:
            if tmp:
                os.unlink(tmp)
        return cls(h)

    @classmethod
    def mesh(cls, path, target_points=300000):
        """bench scene around an MSH1 mesh (3 lamps, a cube, 5 states); returns (scene, cube side in m)"""
        h, b = C.c_void_p(), C.c_float(0.0)
        check(lib.rc_scene_mesh(path.encode(), target_points, C.byref(h), C.byref(b)))
        return cls(h), b.value

    def save(self, path):
        """format-3 scene file (with the converged references when computed); point lights only"""
        check(lib.rc_scene_save(self._h, path.encode()))

    def __del__(self, _destroy=lib.rc_scene_destroy):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _destroy(h)
            self._h = None

    @property
    def info(self):
        i = SceneInfo()
        check(lib.rc_scene_info_get(self._h, C.byref(i)))
        return i

    def array(self, name, state=0):
        which = A[name]
        nb = lib.rc_scene_array_bytes(self._h, which, state)
        if name == "BN":
            out = np.empty(nb // 4, np.int32)
        elif name == "WN":
            out = np.empty(nb // 4, np.int32)
        else:
            out = np.empty(nb // np.dtype(_DT[name]).itemsize, _DT[name])
        check(lib.rc_scene_array(self._h, which, state, out.ctypes.data_as(C.c_void_p), nb))
        return out

    def upload(self, trav=TRAV_WIDE):
        check(lib.rc_scene_upload(self._h, trav))
        self.uploaded = True
        return self

    def set_fine_grid(self, g):
        check(lib.rc_scene_set_fine_grid(self._h, g))
        return self

    def set_trav(self, trav):
        check(lib.rc_scene_set_trav(self._h, trav))

    def set_stack_levels(self, levels):
        """keep exactly `levels` levels of the traversal stacks of the scene's tracing kernels (and of
        adjoint handles created afterwards) in LDS, the deeper ones in global memory (default: all
        in LDS, or 16 where that lets more waves reside); results do not depend on it"""
        check(lib.rc_scene_set_stack_levels(self._h, int(levels)))
        return self

    def reference(self, K=32, warm=16, block=64, max_iter=2048, target=0.03):
        ns = self.info.NS
        noise = (C.c_float * ns)()
        iters = (C.c_int * ns)()
        check(lib.rc_scene_reference(self._h, K, warm, block, max_iter, target, noise, iters))
        return list(noise), list(iters)

    def camera(self):
        return self.info.cam


class Cache:
    """Cache buffers + the per-frame runner of one method on a scene (Runner, :807-1016)."""

    def __init__(self, scene, method):
        if not scene.uploaded:
            raise RcError("scene not uploaded")
        self.scene = scene
        self.method = method
        h = C.c_void_p()
        check(lib.rc_cache_create(scene._h, C.byref(method), C.byref(h)))
        self._h = h

    def __del__(self, _destroy=lib.rc_cache_destroy):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _destroy(h)
            self._h = None

    def begin(self, state0, seed):
        check(lib.rc_cache_begin(self._h, state0, seed))

    def frame(self, state, want_err=False, timed=False):
        fs = FrameStats()
        check(lib.rc_cache_frame(self._h, state, 1 if want_err else 0, 1 if timed else 0, C.byref(fs)))
        return fs

    def run(self, seq, seed, want_err=True):
        """Runner::run over a state sequence: per-frame relative RMS errors (-1 without a reference)."""
        self.begin(seq[0], seed)
        return [self.frame(s, want_err).err for s in seq]

    def render(self, img, cam=None, no_cache=False, seed=7, count_rays=False, timed=False):
        """query pass into img, ordered after the work queued on torch's current stream and before the
        work queued on it afterwards; returns (ext_rays, shadow_rays) when count_rays, ms when timed"""
        import torch
        cam = cam or self.scene.camera()
        rays = (C.c_longlong * 2)() if count_rays else None
        ms = C.c_float(0.0) if timed else None
        check(lib.rc_cache_render(self._h, C.byref(cam), 1 if no_cache else 0, seed, _ptr(img), rays,
                                  C.byref(ms) if timed else None, _stream(torch.cuda.current_stream())))
        out = []
        if count_rays:
            out.append((rays[0], rays[1]))
        if timed:
            out.append(ms.value)
        return out[0] if len(out) == 1 else (tuple(out) if out else None)

    def snapshot_save(self):
        """compact snapshot of the entries with samples: structured array of 24-byte records
        (index, n, key, radiance, age) in entry order"""
        cap = self.scene.info.Np
        out = np.zeros(cap, SNAP_REC)
        n = C.c_size_t(0)
        check(lib.rc_cache_snapshot_save(self._h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return out[: n.value]

    def snapshot_load(self, records):
        """flush, then restore the records whose key matches its entry"""
        r = np.ascontiguousarray(records, SNAP_REC)
        check(lib.rc_cache_snapshot_load(self._h, r.ctypes.data_as(C.c_void_p), len(r)))

    def set_render_mode(self, mode):
        """RENDER_FUSED (default): one thread per path; RENDER_STAGED (4-wide BVH): stage kernels over
        per-segment compacted ray lists. Both give the same image bit for bit."""
        check(lib.rc_cache_set_render_mode(self._h, int(mode)))

    def set_stack_levels(self, levels):
        """keep exactly `levels` levels of the fused camera pass's traversal stacks in LDS, the
        deeper ones in a global buffer (default: all in LDS, or 16 where that lets more waves
        reside); the image does not depend on it"""
        check(lib.rc_cache_set_stack_levels(self._h, int(levels)))

    def get(self, name):
        which = CA[name]
        nb = lib.rc_cache_array_bytes(self._h, which)
        out = np.empty(nb // np.dtype(_CDT[name]).itemsize, _CDT[name])
        check(lib.rc_cache_get(self._h, which, out.ctypes.data_as(C.c_void_p), nb))
        return out

    def set(self, name, arr):
        which = CA[name]
        a = np.ascontiguousarray(arr, dtype=_CDT[name])
        check(lib.rc_cache_set(self._h, which, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def sync(self):
        check(lib.rc_cache_sync(self._h))

    def event(self, to_state):
        """the event block of the transition to to_state now (the next frame in to_state runs none);
        returns the entries the targeted invalidation reset"""
        n = C.c_int(0)
        check(lib.rc_cache_event(self._h, int(to_state), C.byref(n)))
        return n.value

    def invalidate(self, light_mask=0, voxels=()):
        """targeted invalidation on the recorded masks: entries depending on a light of light_mask (bit
        min(l, 63)) or a transit voxel of voxels, or on nothing recorded, restart; returns their count"""
        v = np.ascontiguousarray(voxels, np.int32)
        n = C.c_int(0)
        check(lib.rc_cache_invalidate(self._h, int(light_mask), v.ctypes.data_as(C.POINTER(C.c_int)), len(v), C.byref(n)))
        return n.value

    def resolve(self):
        """resolved radiance R = A sum_l (Edir_l + Lc_l) of every entry from the current channels"""
        check(lib.rc_cache_resolve(self._h))


# ---------------------------------------------------------------- kernel level (torch device tensors)
def trace(scene, O, D, t, q, state=0, trav=TRAV_WIDE, stream=None):
    check(lib.rc_trace(scene._h, state, _ptr(O), _ptr(D), int(O.shape[0]), _ptr(t), _ptr(q), trav, _stream(stream)))


def hash_insert(HK, HV, keys, vals, fail, stream=None):
    check(lib.rc_hash_insert(_ptr(HK), _ptr(HV), HK.shape[0] - 1, _ptr(keys), _ptr(vals), int(keys.shape[0]), _ptr(fail),
                             _stream(stream)))


def hash_query(HK, HV, keys, out, stream=None):
    check(lib.rc_hash_query(_ptr(HK), _ptr(HV), HK.shape[0] - 1, _ptr(keys), int(keys.shape[0]), _ptr(out), _stream(stream)))


def hash_point(scene, H, N, out, stream=None):
    check(lib.rc_hash_point(scene._h, _ptr(H), _ptr(N), int(H.shape[0]), _ptr(out), _stream(stream)))


def dda_mark(scene, Aseg, Bseg, out, stream=None):
    check(lib.rc_dda_mark(scene._h, _ptr(Aseg), _ptr(Bseg), int(Aseg.shape[0]), _ptr(out), _stream(stream)))


def dda_volume(scene, g, Aseg, Bseg, seg_vol, vol, words_per_vol, stream=None):
    check(lib.rc_dda_volume(scene._h, g[0], g[1], g[2], _ptr(Aseg), _ptr(Bseg), _ptr(seg_vol), int(Aseg.shape[0]), _ptr(vol),
                            words_per_vol, _stream(stream)))

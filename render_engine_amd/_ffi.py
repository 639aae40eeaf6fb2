"""ctypes view of the C ABI in include/rc.h (librc.so, built in-tree next to this file).

The library is the only compute path: importing this module raises if librc.so is missing,
and every call that returns a non-zero rc_status raises RcError with rc_last_error().
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librc.so")


class RcError(RuntimeError):
    pass


class Method(C.Structure):
    """rc_method (cuda/poc_gpu.cu:803-805 Method + SDK bench flush/targeted switches)."""
    _fields_ = [
        ("mode", C.c_int), ("nmax", C.c_float), ("budget", C.c_float), ("deps", C.c_int), ("verify", C.c_int),
        ("oracle", C.c_int), ("track_every", C.c_int), ("rescale", C.c_int), ("resp_mask", C.c_uint64),
        ("nmax_resp", C.c_float), ("ddgi", C.c_int), ("ddgi_h", C.c_float), ("ddgi_thr", C.c_float),
        ("ddgi_clamp", C.c_int), ("memory", C.c_int), ("ramp", C.c_int), ("wake_verify", C.c_int),
        ("snaps", C.c_int), ("snap_verify", C.c_int), ("flush", C.c_int), ("targeted", C.c_int), ("compact", C.c_int), ("grad_masks", C.c_int),
        ("eventboost", C.c_int),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
        ("tanx", C.c_float), ("tany", C.c_float), ("W", C.c_int32), ("H", C.c_int32), ("spp", C.c_int32),
        ("bounces", C.c_int32), ("cache_bounce", C.c_int32), ("pad0", C.c_int32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "Np", "npst", "nbox", "nst", "nnodes", "nwnodes", "NS", "nl", "NR", "Nq", "hsize", "vx", "vy", "vz", "nw",
        "bvh_depth", "wide_depth", "has_ref", "fine")] + [
        ("h", C.c_float), ("lo", C.c_float * 3), ("ext", C.c_float * 3), ("diag", C.c_float), ("sun_len", C.c_float),
        ("cam", Camera), ("nmat", C.c_int), ("pad1", C.c_int),
    ]


class FrameStats(C.Structure):
    _fields_ = [
        ("ms_swap", C.c_float), ("ms_trace", C.c_float), ("ms_decide", C.c_float), ("ms_update", C.c_float),
        ("ms_event", C.c_float), ("err", C.c_double), ("event", C.c_int), ("invalidated", C.c_int), ("rays", C.c_longlong),
    ]


# scene arrays (enum order of include/rc.h)
A_NAMES = ["ST", "STN", "BN", "WN", "HK", "HV", "P", "N", "ALB", "REGION", "L", "LTYPE", "INTEN", "Q", "DIMS", "OFF", "T",
           "TQ", "TALB", "EDIR", "SV", "LM", "OCC", "EM", "REF", "VALID", "PW", "ORDER", "OBJBOX", "MESH_V", "MESH_ALB",
           "MESH_DS", "VOL", "OCCF", "MAT", "MATALB"]
A = {n: i for i, n in enumerate(A_NAMES)}
C_NAMES = ["LC", "M", "N", "KK", "PENDING", "CUR", "PREV", "LCUR", "LPREV", "SUMPR", "NPEND", "AGE", "STALE", "C0", "RSTAT",
           "LNEW", "DSTAT", "R"]
CA = {n: i for i, n in enumerate(C_NAMES)}

# exported symbols and their signatures: (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "rc_last_error": (C.c_char_p, []),
    "rc_version": (C.c_int, []),
    "rc_device_count": (C.c_int, []),
    "rc_scene_box_room": (C.c_int, [C.c_uint64, C.c_int, C.POINTER(_P)]),
    "rc_scene_street": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.POINTER(_P)]),
    "rc_scene_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "rc_scene_mesh": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P), C.POINTER(C.c_float)]),
    "rc_scene_save": (C.c_int, [_P, C.c_char_p]),
    "rc_scene_destroy": (None, [_P]),
    "rc_scene_info_get": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "rc_scene_array_bytes": (C.c_size_t, [_P, C.c_int, C.c_int]),
    "rc_scene_array": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_size_t]),
    "rc_scene_upload": (C.c_int, [_P, C.c_int]),
    "rc_scene_set_trav": (C.c_int, [_P, C.c_int]),
    "rc_scene_set_fine_grid": (C.c_int, [_P, C.c_int]),
    "rc_scene_prep_device": (C.c_int, [C.c_int]),
    "rc_scene_reference": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float),
                                     C.POINTER(C.c_int)]),
    "rc_cache_create": (C.c_int, [_P, C.POINTER(Method), C.POINTER(_P)]),
    "rc_cache_destroy": (None, [_P]),
    "rc_cache_begin": (C.c_int, [_P, C.c_int, C.c_uint32]),
    "rc_cache_frame": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(FrameStats)]),
    "rc_cache_render": (C.c_int, [_P, C.POINTER(Camera), C.c_int, C.c_uint32, _P, C.POINTER(C.c_longlong),
                                 C.POINTER(C.c_float), _P]),
    "rc_cache_array_bytes": (C.c_size_t, [_P, C.c_int]),
    "rc_cache_get": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rc_cache_set": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rc_cache_sync": (C.c_int, [_P]),
    "rc_cache_event": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "rc_cache_invalidate": (C.c_int, [_P, C.c_ulonglong, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    "rc_cache_resolve": (C.c_int, [_P]),
    "rc_cache_set_render_mode": (C.c_int, [_P, C.c_int]),
    "rc_cache_set_stack_levels": (C.c_int, [_P, C.c_int]),
    "rc_scene_set_stack_levels": (C.c_int, [_P, C.c_int]),
    "rc_cache_snapshot_save": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rc_cache_snapshot_save_noalloc": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "rc_cache_snapshot_load": (C.c_int, [_P, _P, C.c_size_t]),
    "rc_adjoint_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.POINTER(_P)]),
    "rc_adjoint_destroy": (None, [_P]),
    "rc_adjoint_forward": (C.c_int, [_P, _P, _P, _P, _P]),
    "rc_adjoint_backward": (C.c_int, [_P, _P, _P, _P, _P]),
    "rc_adjoint_array_bytes": (C.c_size_t, [_P, C.c_int]),
    "rc_adjoint_get": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rc_adjoint_set_timed": (C.c_int, [_P, C.c_int]),
    "rc_adjoint_trace_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rc_mlp_create": (C.c_int, [_P, C.POINTER(_P)]),
    "rc_mlp_destroy": (None, [_P]),
    "rc_mlp_infer": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, _P]),
    "rc_mlp_backward": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, _P]),
    "rc_mlp_train": (C.c_int, [_P, _P, _P, _P, _P, C.c_int, C.POINTER(C.c_float), _P]),
    "rc_mlp_step": (C.c_int, [_P, _P]),
    "rc_mlp_zero_grad": (C.c_int, [_P, _P]),
    "rc_mlp_array_bytes": (C.c_size_t, [_P, C.c_int]),
    "rc_mlp_get": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rc_mlp_set": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "rc_mlp_train_scene": (C.c_int, [_P, _P, C.c_int, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rc_mlp_render": (C.c_int, [_P, _P, C.POINTER(Camera), C.c_uint32, _P, C.POINTER(C.c_float), _P]),
    "rc_mlp_set_grid_fallback": (C.c_int, [_P, C.c_float]),
    "rc_mlp_last_infer_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "rc_trace": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, _P, _P, C.c_int, _P]),
    "rc_hash_insert": (C.c_int, [_P, _P, C.c_uint32, _P, _P, C.c_int, _P, _P]),
    "rc_hash_query": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int, _P, _P]),
    "rc_hash_point": (C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    "rc_dda_mark": (C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    "rc_dda_volume": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_int, _P, C.c_longlong, _P]),
}
EXPORTS = sorted(_SIGS)


def load(path=LIB_PATH):
    # torch bundles its own HIP runtime (torch/lib/libamdhip64.so, soname libamdhip64.so.7) and
    # loads it by file name. Importing torch first makes librc's libamdhip64.so.7 dependency bind
    # to that same runtime, so device pointers and streams are shared by one HIP instance.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise RcError(f"librc.so not found at {path}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def check(status):
    if status != 0:
        raise RcError(f"rc_status {status}: {lib.rc_last_error().decode(errors='replace')}")

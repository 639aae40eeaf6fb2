"""ctypes wrapper of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, as the
checker / baseline of the HIP library; never by the product package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_P = C.c_void_p


class OrcScene(C.Structure):
    _fields_ = [
        ("ST", _P), ("STN", _P), ("BN", _P), ("WN", _P), ("HK", _P), ("HV", _P), ("hmask", C.c_uint32), ("hcell", C.c_float),
        ("L", _P), ("ltype", _P), ("Alb", _P), ("region", _P), ("lo", C.c_float * 3), ("ext", C.c_float * 3),
        ("sun_len", C.c_float), ("nst", C.c_int32), ("nl", C.c_int32), ("Np", C.c_int32), ("NR", C.c_int32),
        ("vx", C.c_int32), ("vy", C.c_int32), ("vz", C.c_int32), ("nw", C.c_int32),
        ("P", _P), ("N", _P), ("Q", _P), ("dims", _P), ("off", _P), ("T", _P), ("TQ", _P), ("nt", C.c_int32),
        ("Edir", _P), ("SV", _P), ("LM", _P), ("inten", _P), ("VOL", _P), ("fg", C.c_int32),
    ]


class OrcCache(C.Structure):
    _fields_ = [(n, _P) for n in ("Lc", "Lc2", "n", "c0", "m", "R", "kk", "pending", "cur", "prev", "lcur", "lprev", "rstat",
                                  "dstat", "stale", "age", "sumpr", "npend", "err")]


class OrcParams(C.Structure):
    _fields_ = [
        ("Np", C.c_int32), ("NR", C.c_int32), ("mode", C.c_int32), ("budget", C.c_float), ("nmax", C.c_float),
        ("seed", C.c_uint32), ("frame", C.c_uint32), ("pend_left", C.c_int32), ("deps", C.c_int32),
        ("resp_mask", C.c_uint64), ("nmax_resp", C.c_float), ("ddgi", C.c_int32), ("ddgi_h", C.c_float),
        ("ddgi_thr", C.c_float), ("ddgi_clamp", C.c_int32), ("dormant", C.c_uint64), ("ramp", C.c_int32),
        ("rays", C.c_float), ("nl", C.c_int32), ("pad0", C.c_int32),
    ]


class OrcMethod(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("nmax", C.c_float), ("budget", C.c_float), ("deps", C.c_int), ("verify", C.c_int),
        ("oracle", C.c_int), ("track_every", C.c_int), ("rescale", C.c_int), ("resp_mask", C.c_uint64),
        ("nmax_resp", C.c_float), ("ddgi", C.c_int), ("ddgi_h", C.c_float), ("ddgi_thr", C.c_float),
        ("ddgi_clamp", C.c_int), ("memory", C.c_int), ("ramp", C.c_int), ("wake_verify", C.c_int), ("snaps", C.c_int),
        ("snap_verify", C.c_int), ("flush", C.c_int), ("targeted", C.c_int), ("compact", C.c_int), ("grad_masks", C.c_int),
        ("eventboost", C.c_int),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    lib.orc_hash32.restype = C.c_uint32; lib.orc_hash32.argtypes = [C.c_uint32]
    lib.orc_rnd.restype = C.c_float; lib.orc_rnd.argtypes = [C.POINTER(C.c_uint32)]
    lib.orc_seed.restype = C.c_uint32; lib.orc_seed.argtypes = [C.c_uint32] * 3
    lib.orc_sincos.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_cosine_dir.argtypes = [_P, C.c_float, C.c_float, _P]
    lib.orc_mix64.restype = C.c_uint64; lib.orc_mix64.argtypes = [C.c_uint64]
    lib.orc_cell_key.restype = C.c_uint64; lib.orc_cell_key.argtypes = [C.c_int] * 4
    lib.orc_decode_key.argtypes = [C.c_uint64] + [C.POINTER(C.c_int)] * 4
    lib.orc_normal_bucket.restype = C.c_int; lib.orc_normal_bucket.argtypes = [C.c_float] * 3
    lib.orc_spread3.restype = C.c_uint64; lib.orc_spread3.argtypes = [C.c_uint64]
    lib.orc_hostrng_init.restype = C.c_uint64; lib.orc_hostrng_init.argtypes = [C.c_uint64]
    lib.orc_hostrng_u.restype = C.c_float; lib.orc_hostrng_u.argtypes = [C.POINTER(C.c_uint64)]
    lib.orc_hash_find.restype = C.c_int; lib.orc_hash_find.argtypes = [_P, _P, C.c_uint32, C.c_uint64]
    lib.orc_hash_point.restype = C.c_int; lib.orc_hash_point.argtypes = [C.POINTER(OrcScene), _P, _P]
    lib.orc_hash_build.argtypes = [_P, C.c_int, C.c_int, _P, _P]
    lib.orc_sample_cells.restype = C.c_int
    lib.orc_sample_cells.argtypes = [_P, _P, _P, C.c_int, C.c_float, _P, _P, _P, _P, _P, C.c_int]
    lib.orc_trace_rays.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, C.c_int, _P, _P]
    lib.orc_vox_of.restype = C.c_int; lib.orc_vox_of.argtypes = [C.POINTER(OrcScene), _P]
    lib.orc_dda_reg.argtypes = [C.POINTER(OrcScene), _P, _P, _P]
    lib.orc_dda_segments.argtypes = [C.POINTER(OrcScene), _P, _P, C.c_int, _P]
    lib.orc_dda_volume.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_int, _P, C.c_longlong]
    lib.orc_k_build.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, _P]
    lib.orc_k_build_rows.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]
    lib.orc_k_trace2.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.POINTER(OrcCache), C.POINTER(OrcParams), _P, _P, C.c_int]
    lib.orc_k_update.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcCache), C.POINTER(OrcParams), _P]
    lib.orc_k_ref.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, _P, _P, C.c_uint32, C.c_int, C.c_int, C.c_int]
    lib.orc_k_resolve.argtypes = [C.POINTER(OrcScene), _P, _P, C.c_uint64]
    lib.orc_k_invalidate.argtypes = [C.POINTER(OrcCache), C.c_int, C.c_int, C.c_uint64, _P, C.c_int, C.c_int]
    lib.orc_count_enable.argtypes = [C.c_int]
    lib.orc_count_get.argtypes = [_P]
    lib.orc_k_query.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P, _P, C.c_int, C.c_int]
    lib.orc_k_build_f64.argtypes = [C.POINTER(OrcScene), C.c_int, _P]
    lib.orc_k_trace2_f64.argtypes = [C.POINTER(OrcScene), _P, _P, _P, _P, C.c_int, C.c_uint64, _P]
    lib.orc_k_update_f64.argtypes = [C.POINTER(OrcScene), _P, _P, C.POINTER(OrcParams), _P, _P, _P]
    lib.orc_runner_create.restype = _P
    lib.orc_runner_create.argtypes = [_P, C.c_int, C.POINTER(OrcMethod), C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_runner_destroy.argtypes = [_P]
    lib.orc_runner_set_grad_masks.argtypes = [_P, _P]
    lib.orc_runner_begin.argtypes = [_P, C.c_int, C.c_uint32]
    lib.orc_runner_frame.restype = C.c_double
    lib.orc_runner_frame.argtypes = [_P, C.c_int, C.c_int, C.POINTER(C.c_int)]
    lib.orc_runner_cache.restype = C.POINTER(OrcCache); lib.orc_runner_cache.argtypes = [_P]
    lib.orc_runner_params.restype = C.POINTER(OrcParams); lib.orc_runner_params.argtypes = [_P]
    lib.orc_k_resolve_f64.argtypes = [C.POINTER(OrcScene), _P, _P, C.c_uint64, _P]
    lib.orc_k_query_f64.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P, C.c_int, C.c_int]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_adj_seed.restype = C.c_uint32; lib.orc_adj_seed.argtypes = [C.c_uint32, C.c_int]
    lib.orc_adj_unit.argtypes = [C.POINTER(OrcScene), C.c_int, _P, _P, _P]
    lib.orc_adj_hits.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_int, C.c_uint32, _P]
    lib.orc_adj_hits_rows.argtypes = [C.POINTER(OrcScene), C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, _P]
    lib.orc_adj_forward.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_adj_grad64.restype = C.c_double
    lib.orc_adj_grad64.argtypes = [C.c_int] * 5 + [_P] * 9
    return lib


lib = _load()


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class OracleScene:
    """Host copies of a render_engine_amd.Scene's arrays, one orc_scene per state. The state
    build arrays (Edir, SV, LM) come from the oracle's own k_build unless given."""

    def __init__(self, scene, build=True, trav=1):
        self.scene = scene
        self.info = i = scene.info
        self.keep = {}
        g = lambda name, st=0: self.keep.setdefault((name, st), scene.array(name, st))
        self.NS, self.Np, self.nl, self.nw = i.NS, i.Np, i.nl, i.nw
        self.states = (OrcScene * i.NS)()
        for st in range(i.NS):
            s = self.states[st]
            s.ST, s.STN, s.BN, s.WN = ptr(g("ST")), ptr(g("STN")), ptr(g("BN")), ptr(g("WN"))
            s.HK, s.HV = ptr(g("HK")), ptr(g("HV"))
            s.hmask, s.hcell = i.hsize - 1, i.h
            s.L, s.ltype, s.Alb, s.region = ptr(g("L")), ptr(g("LTYPE")), ptr(g("ALB")), ptr(g("REGION"))
            for a in range(3):
                s.lo[a], s.ext[a] = i.lo[a], i.ext[a]
            s.sun_len = i.sun_len
            s.nst, s.nl, s.Np, s.NR, s.vx, s.vy, s.vz, s.nw = i.nst, i.nl, i.Np, i.NR, i.vx, i.vy, i.vz, i.nw
            s.P, s.N, s.Q = ptr(g("P", st)), ptr(g("N", st)), ptr(g("Q", st))
            s.dims, s.off = ptr(g("DIMS")), ptr(g("OFF"))
            s.T, s.TQ = ptr(g("T", st)), ptr(g("TQ", st))
            s.nt = len(g("TQ", st))
            s.inten = ptr(g("INTEN", st))
        self.Edir = [np.zeros(i.Np * i.nl, np.float32) for _ in range(i.NS)]
        self.SV = [np.zeros(i.Np * i.nw, np.uint32) for _ in range(i.NS)]
        self.LM = [np.zeros(i.Np, np.uint64) for _ in range(i.NS)]
        self.fg = i.fine
        wpv = (i.fine ** 3 + 31) // 32 if i.fine > 0 else 0
        self.VOL = [np.zeros(i.nl * wpv, np.uint32) for _ in range(i.NS)]
        for st in range(i.NS):
            s = self.states[st]
            s.Edir, s.SV, s.LM = ptr(self.Edir[st]), ptr(self.SV[st]), ptr(self.LM[st])
            s.VOL, s.fg = (ptr(self.VOL[st]) if wpv else None), i.fine
            if build:
                lib.orc_k_build(C.byref(s), trav, s.Edir, s.SV, s.LM)
        occ = [g("OCC", k) for k in range(i.NS * i.NS)]
        em = [g("EM", k) for k in range(i.NS * i.NS * i.nl)]
        self.occ_flat = np.concatenate(occ + [np.zeros(1, np.int32)]).astype(np.int32)
        self.occ_off = np.concatenate([[0], np.cumsum([len(x) for x in occ])]).astype(np.int64)
        self.em_flat = np.concatenate(em + [np.zeros(1, np.int32)]).astype(np.int32)
        self.em_off = np.concatenate([[0], np.cumsum([len(x) for x in em])]).astype(np.int64)
        occf = [g("OCCF", k) for k in range(i.NS * i.NS)]
        self.occf_flat = np.concatenate(occf + [np.zeros(1, np.int32)]).astype(np.int32)
        self.occf_off = np.concatenate([[0], np.cumsum([len(x) for x in occf])]).astype(np.int64)

    def state(self, st):
        return C.byref(self.states[st])

    def trace(self, O, D, st=0, trav=1):
        n = O.shape[0]
        t = np.empty(n, np.float32)
        q = np.empty(n, np.int32)
        lib.orc_trace_rays(self.state(st), trav, ptr(np.ascontiguousarray(O, np.float32)), ptr(np.ascontiguousarray(D, np.float32)), n,
                           ptr(t), ptr(q))
        return t, q


def method_from(m):
    """oracle method from a render_engine_amd Method"""
    o = OrcMethod()
    for name, _ in OrcMethod._fields_:
        setattr(o, name, getattr(m, name))
    return o


class OracleRunner:
    def __init__(self, osc, method, trav=1, Pw=None, ref=None, valid=None):
        self.osc = osc
        self.keep = [Pw, ref, valid]
        arr = lambda lst, T: (T * osc.NS)(*[x.ctypes.data_as(T) for x in lst]) if lst is not None else None
        self.pw = arr(Pw, C.POINTER(C.c_float)) if Pw is not None else None
        self.rf = arr(ref, C.POINTER(C.c_float)) if ref is not None else None
        self.vd = arr(valid, C.POINTER(C.c_uint8)) if valid is not None else None
        self.m = method_from(method)
        self.h = lib.orc_runner_create(C.cast(osc.states, _P), osc.NS, C.byref(self.m), trav, ptr(osc.occ_flat), ptr(osc.occ_off),
                                       ptr(osc.em_flat), ptr(osc.em_off), C.cast(self.pw, _P) if self.pw else None,
                                       C.cast(self.rf, _P) if self.rf else None, C.cast(self.vd, _P) if self.vd else None,
                                       ptr(osc.occf_flat), ptr(osc.occf_off))

    def __del__(self, _destroy=lib.orc_runner_destroy):
        if getattr(self, "h", None):
            _destroy(self.h)
            self.h = None

    def set_grad_masks(self, masks):
        """per-state gradient-support masks [NS][Np] for the grad_masks methods"""
        self.gmask = np.ascontiguousarray(masks, np.uint64)
        lib.orc_runner_set_grad_masks(self.h, ptr(self.gmask))

    def begin(self, st, seed):
        lib.orc_runner_begin(self.h, st, seed)

    def frame(self, st, want_err=False):
        inv = C.c_int(0)
        e = lib.orc_runner_frame(self.h, st, 1 if want_err else 0, C.byref(inv))
        return e, inv.value

    def get(self, name):
        c = lib.orc_runner_cache(self.h).contents
        Np, nl, nw = self.osc.Np, self.osc.nl, self.osc.nw
        spec = {"LC": (np.float32, nl * Np), "M": (np.float32, nl * Np), "N": (np.float32, Np), "KK": (np.int32, Np),
                "PENDING": (np.uint8, Np), "CUR": (np.uint32, nw * Np), "PREV": (np.uint32, nw * Np), "LCUR": (np.uint64, Np),
                "LPREV": (np.uint64, Np), "SUMPR": (np.int64, 2), "NPEND": (np.int32, 2), "AGE": (np.float32, Np),
                "STALE": (np.uint64, Np), "C0": (np.float32, Np), "R": (np.float32, Np)}
        dt, n = spec[name]
        field = {"LC": "Lc", "M": "m", "N": "n", "KK": "kk", "PENDING": "pending", "CUR": "cur", "PREV": "prev",
                 "LCUR": "lcur", "LPREV": "lprev", "SUMPR": "sumpr", "NPEND": "npend", "AGE": "age", "STALE": "stale",
                 "C0": "c0", "R": "R"}[name]
        addr = getattr(c, field)
        buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(addr)
        return np.frombuffer(buf, dt).copy()


class FP64Twin:
    """FP64 twin of an OracleRunner's cache in one state (no events): each frame's entry-pass rays and
    hits are recorded from the FP32 pass (orc_k_trace2 on a copy of the runner's cache), the sample
    means and the EMA update are replayed in double with the FP32 run's ray counts and history, and
    the direct light is the FP32 build's visibility with the radiometry in double. Create it right
    after the runner's begin and call frame() instead of the runner's frame()."""

    def __init__(self, osc, orun, st, kmax=4):
        self.osc, self.orun, self.st, self.kmax, self.f = osc, orun, st, kmax, 0
        i = osc.info
        self.E64 = np.zeros(i.Np * i.nl)
        lib.orc_k_build_f64(osc.state(st), 1, ptr(self.E64))
        self.L64 = orun.get("LC").astype(np.float64)

    def frame(self):
        i, orun = self.osc.info, self.orun
        s = self.osc.state(self.st)
        n_before = orun.get("N")
        c = lib.orc_runner_cache(orun.h).contents
        p = OrcParams()
        C.memmove(C.byref(p), lib.orc_runner_params(orun.h), C.sizeof(OrcParams))
        p.frame, p.pend_left = self.f, 0
        hits = np.full(i.Np * self.kmax, -2, np.int32)
        mcopy = np.zeros(i.Np * i.nl, np.float32)
        kk = np.zeros(i.Np, np.int32)
        tmp = OrcCache()
        for fld, _ in OrcCache._fields_:
            setattr(tmp, fld, getattr(c, fld))
        tmp.m, tmp.kk = mcopy.ctypes.data, kk.ctypes.data
        lib.orc_k_trace2(s, 1, 0, C.byref(tmp), C.byref(p), None, ptr(hits), self.kmax)
        assert kk.max(initial=0) <= self.kmax, "more rays per entry than recorded"
        m64 = np.zeros(i.Np * i.nl)
        lib.orc_k_trace2_f64(s, ptr(self.E64), ptr(self.L64), ptr(kk), ptr(hits), self.kmax, p.dormant, ptr(m64))
        out = np.zeros_like(self.L64)
        lib.orc_k_update_f64(s, ptr(n_before), ptr(kk), C.byref(p), ptr(self.L64), ptr(m64), ptr(out))
        self.L64 = out
        orun.frame(self.st)
        self.f += 1

    def R64(self, dormant=0):
        R = np.zeros(self.osc.info.Np)
        lib.orc_k_resolve_f64(self.osc.state(self.st), ptr(self.E64), ptr(self.L64), dormant, ptr(R))
        return R

    def image64(self, cam, seed, frame, pix0, pix1, no_cache=False):
        img = np.zeros(cam.W * cam.H)
        R = None if no_cache else self.R64()
        lib.orc_k_query_f64(self.osc.state(self.st), 1, ptr(R), C.byref(cam), seed, frame, 0, ptr(img), pix0, pix1)
        return img


def rel_rms(a, b):
    """relative RMS of a against the reference b: sqrt(mean((a - b)^2)) / mean(|b|)"""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.mean(np.abs(b)), 1e-300))


class OracleAdjoint:
    """CPU adjoint of the cached radiance on state `st` of an OracleScene: unit direct light
    (FP32 and FP64), the hit entry of every ray of the B fixed-seed bounces, then the FP32 forward
    (the kernels' operation order) and the FP64 forward + reverse-mode gradient."""

    def __init__(self, osc, st, K, B, seed, trav=1):
        i = osc.info
        self.Np, self.nl, self.K, self.B = i.Np, i.nl, K, B
        self.nmat = i.nmat
        self.mat = np.ascontiguousarray(osc.scene.array("MAT"), np.int32)
        self.G = np.zeros(i.Np * i.nl, np.float32)
        self.SG = np.zeros(i.Np, np.uint64)
        self.G64 = np.zeros(i.Np * i.nl, np.float64)
        lib.orc_adj_unit(osc.state(st), trav, ptr(self.G), ptr(self.SG), ptr(self.G64))
        self.hits = np.zeros(max(B, 1) * i.Np * K, np.int32)
        if B > 0:
            lib.orc_adj_hits(osc.state(st), trav, K, B, seed, ptr(self.hits))

    def forward32(self, I, alb):
        I = np.ascontiguousarray(I, np.float32); alb = np.ascontiguousarray(alb, np.float32)
        R = np.zeros(self.Np, np.float32); SR = np.zeros(self.Np, np.uint64); LB = np.zeros(self.Np, np.float32)
        lib.orc_adj_forward(self.Np, self.nl, self.K, self.B, ptr(self.hits), ptr(self.mat), ptr(alb), ptr(I), ptr(self.G),
                            ptr(self.SG), ptr(R), ptr(SR), ptr(LB))
        return R, SR, LB

    def grad64(self, I, alb, g):
        I = np.ascontiguousarray(I, np.float64); alb = np.ascontiguousarray(alb, np.float64)
        g = np.ascontiguousarray(g, np.float64)
        R = np.zeros(self.Np); dI = np.zeros(self.nl); dalb = np.zeros(self.nmat)
        J = lib.orc_adj_grad64(self.Np, self.nl, self.nmat, self.K, self.B, ptr(self.hits), ptr(self.mat), ptr(alb), ptr(I),
                               ptr(self.G64), ptr(g), ptr(R), ptr(dI), ptr(dalb))
        return J, R, dI, dalb

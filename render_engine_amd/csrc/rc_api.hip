// C ABI of librc (include/rc.h): scene handles, device upload, converged references, the cache
// runner (per-frame kernel sequence and event block of Runner::run, cuda/poc_gpu.cu:826-1016)
// and the kernel-level entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rc.h"
#include "rc_kernels.h"
#include "rc_scene.h"

static thread_local std::string g_err;

static rc_status fail(rc_status st, const char* fmt, ...) {
    char buf[512];
    va_list a;
    va_start(a, fmt);
    vsnprintf(buf, sizeof buf, fmt, a);
    va_end(a);
    g_err = buf;
    return st;
}
#define HIPCK(x)                                                                                        \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) return fail(RC_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

extern "C" const char* rc_last_error(void) { return g_err.c_str(); }
extern "C" int rc_version(void) { return 1; }
extern "C" int rc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return n;
}

static inline size_t fine_words(int g) { return ((size_t)g * g * g + 31) / 32; }

template <class T> static hipError_t dput(T** d, const std::vector<T>& v, std::vector<void*>& owned) {
    size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    hipError_t e = hipMalloc((void**)d, bytes);
    if (e != hipSuccess) return e;
    owned.push_back(*d);
    if (!v.empty()) e = hipMemcpy(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}
template <class T> static hipError_t dalloc(T** d, size_t n, std::vector<void*>& owned) {
    hipError_t e = hipMalloc((void**)d, std::max<size_t>(n * sizeof(T), 16));
    if (e != hipSuccess) return e;
    owned.push_back(*d);
    return hipMemset(*d, 0, std::max<size_t>(n * sizeof(T), 16));
}

// traversal modes 2 and 3 share every kernel but the camera pass and rc_trace, which walk the
// quantized nodes in mode 3; the others take mode 2 (fp32 wide nodes) there
static inline int trav_fp(int trav) { return trav == 3 ? 2 : trav; }

struct rc_scene {
    rc::Scene H;
    bool uploaded = false;
    int trav = 2;
    hipStream_t stream = nullptr;
    rc_dscene ds{};
    std::vector<rc_dstate> dst;
    std::vector<int*> dOcc, dEm, dOccF;
    unsigned long long* d_lights = nullptr;   // event light set (k_vol_lights)
    uint64_t* d_ekeys = nullptr;              // per entry: its hash-grid key (1 << 63 | index for quad points)
    std::vector<void*> owned;
    // global levels of the lane stacks of the wide-traversal kernels on the scene stream (levels
    // beyond RC_LCAP, one column per thread of a launch; grown on demand, the stream orders its users)
    uint32_t* ovf = nullptr;
    size_t ovf_words = 0;
    int lcap = RC_LCAP;           // LDS levels of those split lane stacks (and of adjoint handles created later)
    bool lforce = false;          // split always (set by rc_scene_set_stack_levels) or where it raises occupancy
    // converged references (host + device)
    bool has_ref = false;
    std::vector<std::vector<float>> ref, Pw;
    std::vector<std::vector<uint8_t>> valid;
    ~rc_scene() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (ovf) (void)hipFree(ovf);
        for (void* p : owned) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// ------------------------------------------------------------------------------------ lane stacks
// State `base` for a wide-traversal launch of nthreads threads whose lane stacks may keep cap levels
// in LDS and the deeper ones in *buf (grown to what the launch needs; `sync`, the only stream that
// uses *buf, is drained before the old buffer is freed): always (force) or where that lets more
// blocks reside than all levels in LDS. cap 0 or a binary-BVH scene: all in LDS.
static hipError_t split_dst(const rc_dstate& base, const rc_dscene& ds, int trav, int cap, bool force, uint32_t** buf, size_t* words,
                            long long nthreads, hipStream_t sync, rc_dstate* out) {
    rc_dstate d = base;
    d.ocap = trav >= 2 ? cap : 0;
    d.oforce = force ? 1 : 0;
    size_t need = rc_ovf_words(&d, &ds, nthreads);
    if (need > ((size_t)1 << 30)) {   // beyond the 32-bit byte offsets of the global levels: every level in LDS
        d.ocap = 0; d.oforce = 0;
        need = 0;
    }
    if (need > *words) {
        hipError_t e = hipStreamSynchronize(sync);
        if (e != hipSuccess) return e;
        if (*buf) { (void)hipFree(*buf); *buf = nullptr; *words = 0; }
        e = hipMalloc(buf, need * sizeof(uint32_t));
        if (e != hipSuccess) { *buf = nullptr; return e; }
        *words = need;
    }
    d.ovf = *buf;
    d.ovf_words = *words;
    *out = d;
    return hipSuccess;
}
static long long grid_threads(long long n) { return rc_grid_threads(n); }
// state si for a launch of nthreads threads on the scene stream
static hipError_t scene_dst(rc_scene* s, int si, long long nthreads, rc_dstate* out) {
    return split_dst(s->dst[si], s->ds, s->trav, s->lcap, s->lforce, &s->ovf, &s->ovf_words, nthreads, s->stream, out);
}

// ------------------------------------------------------------------------------------ scenes
static rc_status make_scene(const rc::SceneDesc& d, rc_scene** out) {
    if (!out) return fail(RC_ERR_ARG, "out is NULL");
    rc_scene* s = new rc_scene();
    std::string e = rc::prepare(d, s->H);
    if (!e.empty()) { delete s; return fail(RC_ERR_SCENE, "scene preparation: %s", e.c_str()); }
    *out = s;
    return RC_OK;
}

extern "C" rc_status rc_scene_box_room(uint64_t seed, int target_tris, rc_scene** out) {
    if (target_tris < 12) return fail(RC_ERR_ARG, "target_tris must be >= 12");
    return make_scene(rc::gen_box_room(seed, target_tris), out);
}

extern "C" rc_status rc_scene_street(uint64_t seed, int nboxes, int target_tris, int nlamps, int sun, int target_points, int table_log2,
                                     int W, int H, int spp, int bounces, rc_scene** out) {
    if (nboxes < 1 || target_tris < 12 || nlamps < 1 || nlamps + (sun ? 1 : 0) > RC_MAX_LIGHTS || target_points < 16 || W < 1 ||
        H < 1 || spp < 1 || bounces < 0 || table_log2 < 0 || table_log2 > 30)
        return fail(RC_ERR_ARG, "rc_scene_street: argument out of range");
    return make_scene(rc::gen_street(seed, nboxes, target_tris, nlamps, sun != 0, target_points, table_log2, W, H, spp, bounces), out);
}

extern "C" rc_status rc_scene_load(const char* path, rc_scene** out) {
    if (!path || !out) return fail(RC_ERR_ARG, "NULL argument");
    rc_scene* s = new rc_scene();
    rc::FileRefs R;
    std::string e = rc::load_scene_file(path, s->H, R);
    if (!e.empty()) { delete s; return fail(RC_ERR_SCENE, "scene file %s: %s", path, e.c_str()); }
    if (R.has) { s->has_ref = true; s->ref = R.ref; s->Pw = R.Pw; s->valid = R.valid; }
    *out = s;
    return RC_OK;
}

extern "C" rc_status rc_scene_mesh(const char* msh_path, int target_points, rc_scene** out, float* cube_m) {
    if (!msh_path || !out) return fail(RC_ERR_ARG, "NULL argument");
    if (target_points < 16) return fail(RC_ERR_ARG, "target_points must be >= 16");
    rc::Mesh M;
    std::string e = rc::read_msh(msh_path, M);
    if (!e.empty()) return fail(RC_ERR_SCENE, "mesh %s: %s", msh_path, e.c_str());
    return make_scene(rc::desc_from_mesh(M, target_points, cube_m), out);
}

extern "C" rc_status rc_scene_save(const rc_scene* s, const char* path) {
    if (!s || !path) return fail(RC_ERR_ARG, "NULL argument");
    rc::FileRefs R;
    if (s->has_ref) { R.has = true; R.ref = s->ref; R.Pw = s->Pw; R.valid = s->valid; }
    std::string e = rc::save_scene_file(s->H, R, path);
    if (!e.empty()) return fail(RC_ERR_SCENE, "scene file %s: %s", path, e.c_str());
    return RC_OK;
}

extern "C" void rc_scene_destroy(rc_scene* s) { delete s; }

extern "C" rc_status rc_scene_info_get(const rc_scene* s, rc_scene_info* o) {
    if (!s || !o) return fail(RC_ERR_ARG, "NULL argument");
    const rc::Scene& H = s->H;
    memset(o, 0, sizeof *o);
    o->Np = H.Np; o->npst = H.npst; o->nbox = H.nbox; o->nst = H.nst; o->nnodes = (int)H.BN.size(); o->nwnodes = (int)H.WN.size();
    o->NS = H.NS; o->nl = H.nl; o->NR = H.NR; o->Nq = H.Nq; o->hsize = (int)H.HK.size();
    o->vx = H.vx; o->vy = H.vy; o->vz = H.vz; o->nw = H.nw; o->bvh_depth = H.bvh_depth; o->wide_depth = H.wide_depth;
    o->has_ref = s->has_ref ? 1 : 0;
    o->h = H.h; o->diag = H.diag; o->sun_len = H.sun_len; o->fine = H.fine;
    for (int a = 0; a < 3; ++a) { o->lo[a] = H.lo[a]; o->ext[a] = H.ext[a]; }
    memcpy(&o->cam, &H.cam, sizeof(rc_camera));
    o->nmat = H.nmat;
    return RC_OK;
}

// host view of one array: pointer + bytes (device-produced ones are fetched separately)
static bool host_array(const rc_scene* s, int which, int st, const void** p, size_t* bytes) {
    const rc::Scene& H = s->H;
    auto ok_state = [&](int n) { return st >= 0 && st < n; };
#define RET(vec) do { *p = (vec).data(); *bytes = (vec).size() * sizeof((vec)[0]); return true; } while (0)
    switch (which) {
        case RC_A_ST: RET(H.ST);
        case RC_A_STN: RET(H.STN);
        case RC_A_BN: RET(H.BN);
        case RC_A_WN: RET(H.WN);
        case RC_A_HK: RET(H.HK);
        case RC_A_HV: RET(H.HV);
        case RC_A_ALB: RET(H.A);
        case RC_A_REGION: RET(H.region);
        case RC_A_L: RET(H.L);
        case RC_A_LTYPE: RET(H.ltype);
        case RC_A_DIMS: RET(H.dims);
        case RC_A_OFF: RET(H.off);
        case RC_A_TALB: RET(H.Talb);
        case RC_A_ORDER: RET(H.order);
        case RC_A_MESH_V: RET(H.mesh.V);
        case RC_A_MESH_ALB: RET(H.mesh.alb);
        case RC_A_MESH_DS: RET(H.mesh.ds);
        case RC_A_MAT: RET(H.mat);
        case RC_A_MATALB: RET(H.matalb);
        case RC_A_P: if (!ok_state(H.NS)) return false; RET(H.P[st]);
        case RC_A_N: if (!ok_state(H.NS)) return false; RET(H.N[st]);
        case RC_A_INTEN: if (!ok_state(H.NS)) return false; RET(H.inten[st]);
        case RC_A_Q: if (!ok_state(H.NS)) return false; RET(H.Q[st]);
        case RC_A_T: if (!ok_state(H.NS)) return false; RET(H.T[st]);
        case RC_A_TQ: if (!ok_state(H.NS)) return false; RET(H.TQ[st]);
        case RC_A_OBJBOX: if (!ok_state(H.NS)) return false; RET(H.objbox[st]);
        case RC_A_OCC: if (!ok_state(H.NS * H.NS)) return false; RET(H.occ[st]);
        case RC_A_OCCF: if (!ok_state(H.NS * H.NS)) return false; RET(H.occ_fine[st]);
        case RC_A_EM: if (!ok_state(H.NS * H.NS * H.nl)) return false; RET(H.em[st]);
        case RC_A_REF: if (!s->has_ref || !ok_state(H.NS)) return false; RET(s->ref[st]);
        case RC_A_VALID: if (!s->has_ref || !ok_state(H.NS)) return false; RET(s->valid[st]);
        case RC_A_PW: if (!s->has_ref || !ok_state(H.NS)) return false; RET(s->Pw[st]);
        default: return false;
    }
#undef RET
}

static bool device_array(const rc_scene* s, int which, int st, const void** dp, size_t* bytes) {
    if (!s->uploaded || st < 0 || st >= s->H.NS) return false;
    const rc_dstate& d = s->dst[st];
    size_t Np = (size_t)s->H.Np;
    switch (which) {
        case RC_A_EDIR: *dp = d.Edir; *bytes = Np * s->H.nl * sizeof(float); return true;
        case RC_A_SV: *dp = d.SV; *bytes = Np * s->H.nw * sizeof(uint32_t); return true;
        case RC_A_LM: *dp = d.LM; *bytes = Np * sizeof(uint64_t); return true;
        case RC_A_VOL: if (!d.VOL) return false; *dp = d.VOL; *bytes = (size_t)s->H.nl * fine_words(s->H.fine) * 4; return true;
        default: return false;
    }
}

extern "C" size_t rc_scene_array_bytes(const rc_scene* s, int which, int state) {
    if (!s) return 0;
    const void* p; size_t b = 0;
    if (host_array(s, which, state, &p, &b)) return b;
    if (device_array(s, which, state, &p, &b)) return b;
    return 0;
}

extern "C" rc_status rc_scene_array(const rc_scene* s, int which, int state, void* dst, size_t bytes) {
    if (!s || (!dst && bytes)) return fail(RC_ERR_ARG, "NULL argument");
    const void* p; size_t b = 0;
    if (host_array(s, which, state, &p, &b)) {
        if (bytes != b) return fail(RC_ERR_ARG, "array %d: %zu bytes requested, %zu available", which, bytes, b);
        if (b) memcpy(dst, p, b);
        return RC_OK;
    }
    if (device_array(s, which, state, &p, &b)) {
        if (bytes != b) return fail(RC_ERR_ARG, "array %d: %zu bytes requested, %zu available", which, bytes, b);
        HIPCK(hipStreamSynchronize(s->stream));
        HIPCK(hipMemcpy(dst, p, b, hipMemcpyDeviceToHost));
        return RC_OK;
    }
    return fail(RC_ERR_ARG, "array %d (state %d) not available%s", which, state, s->uploaded ? "" : " (scene not uploaded)");
}

static rc_status need_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) { (void)hipGetLastError(); return fail(RC_ERR_NODEV, "no HIP device visible"); }
    return RC_OK;
}

extern "C" rc_status rc_scene_prep_device(int on) {
    if (on != 0 && on != 1) return fail(RC_ERR_ARG, "rc_scene_prep_device: on must be 0 or 1");
    rc::g_prep_device = on != 0;
    return RC_OK;
}

extern "C" rc_status rc_scene_set_fine_grid(rc_scene* s, int g) {
    if (!s || g < 0 || g > 1024) return fail(RC_ERR_ARG, "fine grid must be 0..1024");
    if (s->uploaded) return fail(RC_ERR_STATE, "fine grid must be set before upload");
    rc::fine_events(s->H, g);
    return RC_OK;
}

extern "C" rc_status rc_scene_set_trav(rc_scene* s, int trav) {
    if (!s || trav < 0 || trav > 3) return fail(RC_ERR_ARG, "trav must be 0, 1, 2 or 3");
    s->trav = trav;
    return RC_OK;
}

extern "C" rc_status rc_scene_upload(rc_scene* s, int trav) {
    if (!s) return fail(RC_ERR_ARG, "NULL scene");
    if (trav < 0 || trav > 3) return fail(RC_ERR_ARG, "trav must be 0, 1, 2 or 3");
    if (s->uploaded) return fail(RC_ERR_STATE, "scene already uploaded");
    rc_status st = need_device();
    if (st != RC_OK) return st;
    s->trav = trav;
    rc::Scene& H = s->H;
    HIPCK(hipStreamCreate(&s->stream));
    rc_dscene& d = s->ds;
    float *dST, *dST4, *dSTN, *dL, *dA; rc_bvh_node* dBN; rc_wnode* dWN; rc_qnode* dQN; uint64_t* dHK; int *dHV, *dLT, *dR;
    std::vector<float> st4((size_t)12 * H.nst, 0.f);
    for (int j = 0; j < H.nst; ++j)
        for (int r = 0; r < 3; ++r)
            for (int x = 0; x < 3; ++x) st4[12 * (size_t)j + 4 * r + x] = H.ST[9 * (size_t)j + 3 * r + x];
    HIPCK(dput(&dST, H.ST, s->owned)); HIPCK(dput(&dST4, st4, s->owned)); HIPCK(dput(&dSTN, H.STN, s->owned)); HIPCK(dput(&dBN, H.BN, s->owned));
    {   // quantized wide nodes
        std::vector<rc_qnode> qn;
        rc::build_qnodes(H.WN, rc::quant_step_floor(H.BN, H.lo, H.ext), qn);
        HIPCK(dput(&dQN, qn, s->owned));
    }
    HIPCK(dput(&dWN, H.WN, s->owned)); HIPCK(dput(&dHK, H.HK, s->owned)); HIPCK(dput(&dHV, H.HV, s->owned));
    HIPCK(dput(&dL, H.L, s->owned)); HIPCK(dput(&dLT, H.ltype, s->owned)); HIPCK(dput(&dA, H.A, s->owned));
    HIPCK(dput(&dR, H.region, s->owned));
    d.ST = dST; d.ST4 = dST4; d.STN = dSTN; d.BN = dBN; d.WN = dWN; d.QN = dQN; d.HK = dHK; d.HV = dHV; d.L = dL; d.ltype = dLT; d.Alb = dA; d.region = dR;
    d.hmask = H.HK.empty() ? 0u : (uint32_t)(H.HK.size() - 1); d.hcell = H.h;
    for (int a = 0; a < 3; ++a) { d.lo[a] = H.lo[a]; d.ext[a] = H.ext[a]; }
    d.sun_len = H.sun_len; d.nst = H.nst; d.nl = H.nl; d.Np = H.Np; d.NR = H.NR;
    d.vx = H.vx; d.vy = H.vy; d.vz = H.vz; d.nw = H.nw;
    d.wstack = 3 * H.wide_depth + 1;
    d.fg = H.fine;
    int *dDims, *dOff;
    HIPCK(dput(&dDims, H.dims, s->owned)); HIPCK(dput(&dOff, H.off, s->owned));
    s->dst.resize(H.NS);
    for (int k = 0; k < H.NS; ++k) {
        rc_dstate& x = s->dst[k];
        memset(&x, 0, sizeof x);
        float *P, *N, *Q, *T, *I, *E; int* TQ; uint32_t* SV; uint64_t* LM;
        HIPCK(dput(&P, H.P[k], s->owned)); HIPCK(dput(&N, H.N[k], s->owned)); HIPCK(dput(&Q, H.Q[k], s->owned));
        HIPCK(dput(&T, H.T[k], s->owned)); HIPCK(dput(&TQ, H.TQ[k], s->owned)); HIPCK(dput(&I, H.inten[k], s->owned));
        HIPCK(dalloc(&E, (size_t)H.nl * H.Np, s->owned)); HIPCK(dalloc(&SV, (size_t)H.nw * H.Np, s->owned));
        HIPCK(dalloc(&LM, (size_t)H.Np, s->owned));
        if (H.fine > 0) { uint32_t* V; HIPCK(dalloc(&V, (size_t)H.nl * fine_words(H.fine), s->owned)); x.VOL = V; }
        x.P = P; x.N = N; x.Q = Q; x.dims = dDims; x.off = dOff; x.T = T; x.TQ = TQ; x.Edir = E; x.SV = SV; x.LM = LM; x.inten = I;
        x.nt = H.nt[k];
    }
    for (auto& v : H.occ) { int* p; std::vector<int> w = v.empty() ? std::vector<int>{0} : v; HIPCK(dput(&p, w, s->owned)); s->dOcc.push_back(p); }
    for (auto& v : H.em) { int* p; std::vector<int> w = v.empty() ? std::vector<int>{0} : v; HIPCK(dput(&p, w, s->owned)); s->dEm.push_back(p); }
    for (auto& v : H.occ_fine) { int* p; std::vector<int> w = v.empty() ? std::vector<int>{0} : v; HIPCK(dput(&p, w, s->owned)); s->dOccF.push_back(p); }
    {   // entry keys for the compact snapshot records
        std::vector<uint64_t> ek(H.Np);
        for (int i = 0; i < H.Np; ++i) ek[i] = (1ull << 63) | (uint64_t)i;
        for (size_t k = 0; k < H.HK.size(); ++k)
            if (H.HK[k] != 0ull && H.HV[k] >= 0 && H.HV[k] < H.Np) ek[H.HV[k]] = H.HK[k];
        HIPCK(dput(&s->d_ekeys, ek, s->owned));
    }
    for (int k = 0; k < H.NS; ++k) {
        rc_dstate dk;
        HIPCK(scene_dst(s, k, grid_threads(H.Np), &dk));
        HIPCK(rc_launch_build(trav_fp(trav), &s->ds, &dk, s->stream));
    }
    HIPCK(hipStreamSynchronize(s->stream));
    if (s->has_ref) {   // references read from a scene file
        for (int si = 0; si < H.NS; ++si) {
            float *r, *pw; uint8_t* v;
            HIPCK(dput(&r, s->ref[si], s->owned)); HIPCK(dput(&pw, s->Pw[si], s->owned)); HIPCK(dput(&v, s->valid[si], s->owned));
            s->dst[si].ref = r; s->dst[si].Pw = pw; s->dst[si].valid = v;
        }
    }
    s->uploaded = true;
    return RC_OK;
}

// Converged reference per state (cuda/prep_mesh.cuh:454-520): warm iterations, then blocks
// of iterations averaged per chain until the two chains agree within `target` (relative RMS
// over valid points) or max_iter. Points whose rays hit within 2 cm more than 30 % of the
// time are not valid (inside geometry). Pw = per-light mean of both chains, ref = sum.
extern "C" rc_status rc_scene_reference(rc_scene* s, int K, int warm, int block, int max_iter, float target, float* noise_out, int* iters_out) {
    if (!s || K < 1 || warm < 0 || block < 1 || max_iter < block) return fail(RC_ERR_ARG, "rc_scene_reference: bad arguments");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    rc::Scene& H = s->H;
    const size_t Np = H.Np, nl = H.nl;
    std::vector<void*> tmp;
    float *Lout, *closec, *chain[2], *acc[2];
    auto cleanup = [&]() { for (void* p : tmp) (void)hipFree(p); };
#define TCK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return fail(RC_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
    TCK(dalloc(&Lout, nl * Np, tmp)); TCK(dalloc(&closec, Np, tmp));
    for (int c = 0; c < 2; ++c) { TCK(dalloc(&chain[c], nl * Np, tmp)); TCK(dalloc(&acc[c], nl * Np, tmp)); }
    s->ref.assign(H.NS, {}); s->Pw.assign(H.NS, {}); s->valid.assign(H.NS, {});
    std::vector<float> a0(nl * Np), a1(nl * Np), cc(Np);
    for (int si = 0; si < H.NS; ++si) {
        TCK(hipMemsetAsync(closec, 0, Np * 4, s->stream));
        for (int c = 0; c < 2; ++c) { TCK(hipMemsetAsync(chain[c], 0, nl * Np * 4, s->stream)); TCK(hipMemsetAsync(acc[c], 0, nl * Np * 4, s->stream)); }
        int it_done = 0, T = 0; double nz = 1.0;
        std::vector<float> ref(Np), Pw(nl * Np); std::vector<uint8_t> val(Np);
        while (true) {
            int n_it = it_done == 0 ? warm + block : block;
            for (int c = 0; c < 2; ++c) {
                float* a = chain[c];
                float* b = Lout;
                for (int k = 0; k < n_it; ++k) {
                    int it = it_done + k;
                    uint32_t seed = (uint32_t)rc::mix64(0x51ull + 7919ull * si + 104729ull * c + 31ull * it);
                    rc_dstate dk;
                    TCK(scene_dst(s, si, grid_threads(Np), &dk));
                    TCK(rc_launch_ref(trav_fp(s->trav), &s->ds, &dk, a, b, acc[c], closec, seed, K, it >= warm, c == 0, s->stream));
                    std::swap(a, b);
                }
                if (a != chain[c]) TCK(hipMemcpyAsync(chain[c], a, nl * Np * 4, hipMemcpyDeviceToDevice, s->stream));
            }
            it_done += n_it;
            T = it_done - warm;
            TCK(hipStreamSynchronize(s->stream));
            TCK(hipMemcpy(a0.data(), acc[0], a0.size() * 4, hipMemcpyDeviceToHost));
            TCK(hipMemcpy(a1.data(), acc[1], a1.size() * 4, hipMemcpyDeviceToHost));
            TCK(hipMemcpy(cc.data(), closec, Np * 4, hipMemcpyDeviceToHost));
            double d2 = 0, rs = 0; int nv = 0;
            for (size_t i = 0; i < Np; ++i) {
                float e0 = 0, e1 = 0;
                for (size_t l = 0; l < nl; ++l) {
                    float x0 = a0[nl * i + l] / T, x1 = a1[nl * i + l] / T;
                    Pw[nl * i + l] = 0.5f * (x0 + x1); e0 += x0; e1 += x1;
                }
                ref[i] = 0.5f * (e0 + e1);
                val[i] = cc[i] / ((double)K * it_done) < 0.3;
                if (val[i]) { d2 += 0.25 * (e0 - e1) * (e0 - e1); rs += ref[i]; nv++; }
            }
            nz = nv && rs > 0 ? std::sqrt(d2 / nv) / (rs / nv) : 0.0;
            if (nz <= target || T >= max_iter) break;
        }
        // points of absent objects are never valid
        for (size_t i = (size_t)H.npst; i < Np; ++i) {
            int q = -1;
            for (int qq = 0; qq < H.Nq; ++qq) if ((int)i >= H.off[qq] && (int)i < H.off[qq] + H.dims[2 * qq] * H.dims[2 * qq + 1]) q = qq;
            if (q >= 0 && !H.present[si][q / 6]) val[i] = 0;
        }
        s->ref[si] = ref; s->Pw[si] = Pw; s->valid[si] = val;
        if (noise_out) noise_out[si] = (float)nz;
        if (iters_out) iters_out[si] = T;
    }
    cleanup();
#undef TCK
    // device copies for the error kernel and the warm start
    for (int si = 0; si < H.NS; ++si) {
        float *r, *pw; uint8_t* v;
        HIPCK(dput(&r, s->ref[si], s->owned)); HIPCK(dput(&pw, s->Pw[si], s->owned)); HIPCK(dput(&v, s->valid[si], s->owned));
        s->dst[si].ref = r; s->dst[si].Pw = pw; s->dst[si].valid = v;
    }
    s->has_ref = true;
    return RC_OK;
}

// ------------------------------------------------------------------------------------ cache
// row pitch (entries) of the SoA sample means: a multiple of 64, so every update tile's rows are aligned
static size_t mpitch_of(size_t Np) { return (Np + 63) & ~(size_t)63; }

struct rc_cache {
    rc_scene* S = nullptr;
    rc_method M{};
    rc_dcache c{};
    uint64_t* lnew = nullptr;
    float* d_query_img = nullptr;
    unsigned long long* d_counters = nullptr;
    std::vector<void*> owned;
    hipEvent_t ev[8]{};
    // runner state
    bool begun = false;
    bool r_valid = false;         // R matches Lc of the last frame
    int prev_state = -1;
    int f = 0;
    uint32_t seed = 0;
    float nmax = 0.f;
    int boost = 0;                // frames left of the post-event history-cap ramp
    int pend_left = 0;
    uint64_t dormant = 0;
    std::vector<float> ilearn;
    int nslots = 0, slot_next = 0, last_event_f = 0;
    int slot_state[RC_KSNAP];
    int* d_slotcnt = nullptr;     // [RC_KSNAP + 1] records of each compact snapshot slot (device), then a spare count
    int* d_blk = nullptr;         // per-block counts / offsets of the snapshot packing
    int* h_inv = nullptr;         // pinned: entries reset by the last targeted invalidation
    hipEvent_t ev_inv = nullptr;  // recorded after its copy to h_inv
    bool inv_pending = false;
    rc_params p{};
    int render_mode = RC_RENDER_FUSED;
    // The camera pass runs on its own stream (rs) after the frame's update (ev_upd): the next
    // frame's entry pass, which touches nothing the camera pass reads, can then overlap it on
    // the scene stream. Whatever writes what the camera pass reads (R, the staged workspace,
    // event handling, a new run) first makes the scene stream wait for the last pass (ev_q).
    hipStream_t rs = nullptr;
    hipEvent_t ev_upd = nullptr, ev_q = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;   // ordering against a caller's stream (render calls)
    bool q_pending = false;
    // R is double-buffered: a frame's update writes the buffer the camera pass two frames back
    // read (waiting for that pass's ev_r only), so it need not wait for the last camera pass.
    float* Rb[2] = {nullptr, nullptr};
    int rb = 0;                   // the buffer d.R points to
    hipEvent_t ev_r[2] = {nullptr, nullptr};
    bool r_pend[2] = {false, false};
    std::vector<uint64_t*> gmask;   // per state: gradient-support masks (grad_masks methods)
    uint32_t* qovf = nullptr;     // lane-stack overflow of the fused camera pass (qovf_words words)
    size_t qovf_words = 0;
    int lcap = RC_LCAP;           // LDS levels of the camera pass's split lane stacks
    bool lforce = false;          // split always (set by rc_cache_set_stack_levels) or where it raises occupancy
    rc_qwork qw{};                // staged camera-pass workspace, sized for qw_paths paths
    long long qw_paths = 0;
    std::vector<void*> qw_owned;
    void free_qwork() {
        for (void* q : qw_owned) (void)hipFree(q);
        qw_owned.clear(); qw = rc_qwork{}; qw_paths = 0;
    }
    ~rc_cache() {
        if (rs) { (void)hipStreamSynchronize(rs); (void)hipStreamDestroy(rs); }
        if (S) (void)hipStreamSynchronize(S->stream);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (ev_upd) (void)hipEventDestroy(ev_upd);
        if (ev_q) (void)hipEventDestroy(ev_q);
        if (ev_inv) (void)hipEventDestroy(ev_inv);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (h_inv) (void)hipHostFree(h_inv);
        for (hipEvent_t e : ev_r) if (e) (void)hipEventDestroy(e);
        for (void* q : owned) (void)hipFree(q);
        free_qwork();
        if (qovf) (void)hipFree(qovf);
    }
};

// The scene stream waits for the last camera pass (before anything writes what it reads).
static hipError_t wait_query(rc_cache* c) {
    if (!c->q_pending) return hipSuccess;
    c->q_pending = false;
    return hipStreamWaitEvent(c->S->stream, c->ev_q, 0);
}

extern "C" rc_status rc_cache_create(rc_scene* s, const rc_method* m, rc_cache** out) {
    if (!s || !m || !out) return fail(RC_ERR_ARG, "NULL argument");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    if (m->mode < 0 || m->mode > 1 || !(m->nmax > 0.f) || m->budget < 0.f || m->track_every < 1)
        return fail(RC_ERR_ARG, "method: mode must be 0/1, nmax > 0, budget >= 0, track_every >= 1");
    rc_cache* c = new rc_cache();
    c->S = s;
    c->M = *m;
    const rc::Scene& H = s->H;
    const size_t Np = H.Np, nl = H.nl, nw = H.nw;
    rc_dcache& d = c->c;
#define CK2(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { delete c; return fail(RC_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
    CK2(dalloc(&d.Lc, nl * Np, c->owned)); CK2(dalloc(&d.Lc2, nl * Np, c->owned)); CK2(dalloc(&d.m, nl * mpitch_of(Np), c->owned));
    CK2(dalloc(&d.n, Np, c->owned)); CK2(dalloc(&d.c0, Np, c->owned)); CK2(dalloc(&d.kk, Np, c->owned));
    CK2(dalloc(&c->Rb[0], Np, c->owned)); CK2(dalloc(&c->Rb[1], Np, c->owned));
    d.R = c->Rb[0];
    CK2(dalloc(&d.pending, Np, c->owned));
    CK2(dalloc(&d.cur, nw * Np, c->owned)); CK2(dalloc(&d.prev, nw * Np, c->owned));
    CK2(dalloc(&d.lcur, Np, c->owned)); CK2(dalloc(&d.lprev, Np, c->owned)); CK2(dalloc(&c->lnew, Np, c->owned));
    CK2(dalloc(&d.rstat, 4 * (size_t)H.NR, c->owned)); CK2(dalloc(&d.dstat, 3 * (size_t)H.NR, c->owned));
    CK2(dalloc(&d.stale, Np, c->owned)); CK2(dalloc(&d.age, Np, c->owned));
    if (m->snaps && !m->compact) CK2(dalloc(&d.snap, (size_t)RC_KSNAP * nl * Np, c->owned));
    if (m->snaps && m->compact) { rc_snaprec* cs = nullptr; CK2(dalloc(&cs, (size_t)RC_KSNAP * Np, c->owned)); d.csnap = cs; }
    CK2(dalloc(&c->d_blk, (size_t)(Np + 255) / 256, c->owned));
    CK2(dalloc(&c->d_slotcnt, (size_t)RC_KSNAP + 1, c->owned));
    CK2(hipHostMalloc((void**)&c->h_inv, sizeof(int), hipHostMallocDefault));
    CK2(hipEventCreateWithFlags(&c->ev_inv, hipEventDisableTiming));
    CK2(dalloc(&d.sumpr, 2, c->owned)); CK2(dalloc(&d.prpart, (size_t)(Np + 63) / 64, c->owned)); CK2(dalloc(&d.npend, 2, c->owned)); CK2(dalloc(&d.err, 3, c->owned));
    CK2(dalloc(&d.errpart, 3 * (size_t)rc_grid_threads(Np) / 256, c->owned));   // 256-thread blocks of the error pass
    CK2(dalloc(&d.blkcnt, (size_t)rc_grid_threads(Np) / 256 + 1, c->owned));
    CK2(dalloc(&d.rsum, Np, c->owned));
    CK2(dalloc(&c->d_counters, 2, c->owned));
    for (auto& e : c->ev) CK2(hipEventCreate(&e));
    CK2(hipStreamCreate(&c->rs));
    CK2(hipEventCreateWithFlags(&c->ev_upd, hipEventDisableTiming)); CK2(hipEventCreateWithFlags(&c->ev_q, hipEventDisableTiming));
    CK2(hipEventCreateWithFlags(&c->ev_in, hipEventDisableTiming)); CK2(hipEventCreateWithFlags(&c->ev_out, hipEventDisableTiming));
    for (auto& e : c->ev_r) CK2(hipEventCreateWithFlags(&e, hipEventDisableTiming));
#undef CK2
    *out = c;
    return RC_OK;
}

extern "C" void rc_cache_destroy(rc_cache* c) { delete c; }

static rc_status cache_grad_masks(rc_cache* c);

// pack the entries with samples into compact records at dst (entry order), their number into *d_count
// (device): per-block counts, a device scan of them, then the ballot-ranked writes (no host round trip)
static hipError_t snap_pack(rc_cache* c, rc_snaprec* dst, int* d_count) {
    rc_scene* S = c->S;
    const int Np = S->H.Np, nb = (Np + 255) / 256;
    hipStream_t st = S->stream;
    hipError_t e = rc_launch_snap_count(&c->c, Np, c->d_blk, st);
    if (e == hipSuccess) e = rc_launch_scan_blocks(c->d_blk, nb, d_count, st);
    if (e == hipSuccess) e = rc_launch_snap_pack(&c->c, Np, S->H.nl, S->d_ekeys, c->d_blk, dst, st);
    return e;
}

extern "C" rc_status rc_cache_begin(rc_cache* c, int state0, uint32_t seed) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    rc_scene* S = c->S;
    const rc::Scene& H = S->H;
    if (state0 < 0 || state0 >= H.NS) return fail(RC_ERR_ARG, "state %d out of range", state0);
    hipStream_t st = S->stream;
    const size_t Np = H.Np, nl = H.nl, nw = H.nw;
    rc_dcache& d = c->c;
    HIPCK(wait_query(c));
    if (S->has_ref) HIPCK(hipMemcpyAsync(d.Lc, S->dst[state0].Pw, nl * Np * 4, hipMemcpyDeviceToDevice, st));
    else HIPCK(hipMemsetAsync(d.Lc, 0, nl * Np * 4, st));
    c->nmax = c->M.nmax;
    std::vector<float> nm(Np, c->nmax), ag(Np, 1e4f);
    HIPCK(hipMemcpyAsync(d.n, nm.data(), Np * 4, hipMemcpyHostToDevice, st));
    HIPCK(hipMemcpyAsync(d.age, ag.data(), Np * 4, hipMemcpyHostToDevice, st));
    HIPCK(hipMemsetAsync(d.pending, 0, Np, st));
    HIPCK(hipMemsetAsync(d.cur, 0, nw * Np * 4, st)); HIPCK(hipMemsetAsync(d.prev, 0, nw * Np * 4, st));
    HIPCK(hipMemsetAsync(d.lcur, 0, Np * 8, st)); HIPCK(hipMemsetAsync(d.lprev, 0, Np * 8, st));
    HIPCK(hipMemsetAsync(c->lnew, 0, Np * 8, st));
    HIPCK(hipMemsetAsync(d.stale, 0, Np * 8, st));
    HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
    HIPCK(hipMemsetAsync(d.sumpr, 0, 16, st));
    HIPCK(rc_launch_sumpr(&d, (int)Np, st));   // sum of 1/(n+1) at n = nmax
    HIPCK(hipStreamSynchronize(st));
    if (c->M.targeted && c->M.grad_masks > 0 && c->gmask.empty()) {
        rc_status r = cache_grad_masks(c);
        if (r != RC_OK) return r;
    }
    c->begun = true;
    c->r_valid = false;
    c->prev_state = state0;
    c->f = 0;
    c->seed = seed;
    c->pend_left = 0;
    c->boost = 0;
    c->dormant = 0;
    c->ilearn.assign(nl, 0.f);
    c->nslots = 0; c->slot_next = 0; c->last_event_f = 0;
    for (int k = 0; k < RC_KSNAP; ++k) c->slot_state[k] = -1;
    HIPCK(hipMemsetAsync(c->d_slotcnt, 0, (RC_KSNAP + 1) * sizeof(int), st));
    rc_params& p = c->p;
    memset(&p, 0, sizeof p);
    p.Np = (int)Np; p.NR = H.NR; p.mode = c->M.mode; p.budget = c->M.budget; p.nmax = c->nmax; p.seed = seed;
    p.deps = c->M.deps; p.resp_mask = c->M.resp_mask; p.nmax_resp = c->M.nmax_resp;
    p.ddgi = c->M.ddgi; p.ddgi_h = c->M.ddgi_h; p.ddgi_thr = c->M.ddgi_thr; p.ddgi_clamp = c->M.ddgi_clamp;
    p.ramp = c->M.ramp; p.rays = 1.f; p.nl = (int)nl; p.mpitch = (int)mpitch_of(Np);
    return RC_OK;
}

static float ms_between(hipEvent_t a, hipEvent_t b) { float ms = 0.f; (void)hipEventElapsedTime(&ms, a, b); return ms; }

// event block of Runner::run for the transition e -> si (cuda/poc_gpu.cu:856-943)
// event voxels of the transit grid (a voxel list) as the per-word mask the event kernels test
static rc_vmask vox_mask(const std::vector<int>& V) {
    rc_vmask m;
    memset(&m, 0, sizeof m);
    for (int v : V) if (v >= 0 && (v >> 5) < RC_MAX_NW) m.w[v >> 5] |= 1u << (v & 31);
    return m;
}

// The event block queues its work on the scene stream without waiting for it: the number of entries
// the targeted invalidation reset is copied to pinned host memory and read at the end of the frame.
static rc_status run_event(rc_cache* c, int e, int si) {
    rc_scene* S = c->S;
    const rc::Scene& H = S->H;
    const rc_method& M = c->M;
    hipStream_t st = S->stream;
    rc_dcache& d = c->c;
    const int Np = H.Np, nl = H.nl;
    const int pe = e * H.NS + si;
    const int f = c->f;
    if (M.snaps && f - c->last_event_f >= 20) {   // snapshot of the state being left
        int k = -1;
        for (int a = 0; a < c->nslots; ++a) if (c->slot_state[a] == e) k = a;
        if (k < 0) { if (c->nslots < RC_KSNAP) k = c->nslots++; else { k = c->slot_next; c->slot_next = (c->slot_next + 1) % RC_KSNAP; } }
        c->slot_state[k] = e;
        if (M.compact) {
            HIPCK(snap_pack(c, (rc_snaprec*)d.csnap + (size_t)k * Np, c->d_slotcnt + k));
        } else {
            HIPCK(hipMemcpyAsync(d.snap + (size_t)k * Np * nl, d.Lc, (size_t)Np * nl * 4, hipMemcpyDeviceToDevice, st));
        }
    }
    if (M.oracle && S->has_ref) HIPCK(rc_launch_oracle_method(&d, Np, S->dst[e].ref, S->dst[si].ref, st));
    if (M.rescale && !M.deps) {
        for (int l = 0; l < nl; ++l) {
            float a = H.inten[e][l], b = H.inten[si][l];
            if (a == b) continue;
            bool dm = l < 64 && ((c->dormant >> l) & 1ull);
            if (M.memory && l < 64 && a > 0 && b == 0) { c->dormant |= 1ull << l; c->ilearn[l] = a; }
            else if (M.memory && a == 0 && b > 0 && dm) { HIPCK(rc_launch_rescale(&d, Np, nl, l, b / c->ilearn[l], st)); c->dormant &= ~(1ull << l); }
            else if (a > 0) HIPCK(rc_launch_rescale(&d, Np, nl, l, b / a, st));
        }
    }
    if (M.deps && !M.targeted) {
        uint64_t woke = 0;
        for (int l = 0; l < nl; ++l) {
            float a = H.inten[e][l], b = H.inten[si][l];
            if (a == b) continue;
            bool dm = l < 64 && ((c->dormant >> l) & 1ull);
            if (M.memory && l < 64 && a > 0 && b == 0) { c->dormant |= 1ull << l; c->ilearn[l] = a; }
            else if (M.memory && a == 0 && b > 0 && dm) {
                HIPCK(rc_launch_rescale(&d, Np, nl, l, b / c->ilearn[l], st));
                c->dormant &= ~(1ull << l);
                woke |= 1ull << l;
                if (M.wake_verify) {
                    HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
                    HIPCK(hipMemsetAsync(d.rstat, 0, 4 * (size_t)H.NR * 8, st));
                    HIPCK(rc_launch_wake(&d, Np, nl, l, 1, c->dormant, st));
                    HIPCK(hipMemcpyAsync(d.npend, d.npend + 1, 4, hipMemcpyDeviceToDevice, st));
                    c->pend_left = RC_VERIF_FRAMES;
                } else HIPCK(rc_launch_wake(&d, Np, nl, l, 0, c->dormant, st));
            } else if (a > 0) HIPCK(rc_launch_rescale(&d, Np, nl, l, b / a, st));
        }
        for (int l = 0; l < nl; ++l) {   // new light without memory: touched entries restart
            int ix = pe * nl + l;
            if (!H.em[ix].empty() && !(l < 64 && ((woke >> l) & 1ull)))
                HIPCK(rc_launch_touched(&d, Np, nl, vox_mask(H.em[ix]), H.nw, 0, c->dormant, st));
        }
        if (!H.occ[pe].empty()) {
            if (M.verify) {
                HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
                HIPCK(hipMemsetAsync(d.rstat, 0, 4 * (size_t)H.NR * 8, st));
                HIPCK(rc_launch_touched(&d, Np, nl, vox_mask(H.occ[pe]), H.nw, 1, c->dormant, st));
                HIPCK(hipMemcpyAsync(d.npend, d.npend + 1, 4, hipMemcpyDeviceToDevice, st));
                c->pend_left = RC_VERIF_FRAMES;
            } else if (c->dormant) HIPCK(rc_launch_touched(&d, Np, nl, vox_mask(H.occ[pe]), H.nw, 2, c->dormant, st));
        }
    }
    int hit = -1;
    if (M.snaps)
        for (int a = 0; a < c->nslots; ++a) if (c->slot_state[a] == si) hit = a;
    if (M.targeted) {   // resolve-time dependency invalidation (rtxgi_bench.patch:368-386, :1028-1050)
        uint64_t inv = 0;
        for (int l = 0; l < nl; ++l) if (H.inten[e][l] != H.inten[si][l]) inv |= 1ull << (l < 63 ? l : 63);
        const int nv = (int)H.occ[pe].size();
        const unsigned long long* extra = nullptr;
        if (H.fine > 0 && !H.occ_fine[pe].empty()) {   // lights whose lit segments cross the object (before or after)
            if (!S->d_lights) { HIPCK(hipMalloc((void**)&S->d_lights, 8)); S->owned.push_back(S->d_lights); }
            HIPCK(hipMemsetAsync(S->d_lights, 0, 8, st));
            HIPCK(rc_launch_vol_lights(nl, S->dst[e].VOL, S->dst[si].VOL, (long long)fine_words(H.fine), S->dOccF[pe],
                                       (int)H.occ_fine[pe].size(), S->d_lights, st));
            extra = S->d_lights;   // read by the invalidation kernel: no host round trip
        }
        if (inv || nv || extra) {
            const bool gm = !c->gmask.empty();
            int* cnt = d.blkcnt + (Np + 255) / 256;
            HIPCK(rc_launch_invalidate(&d, Np, nl, inv, extra, vox_mask(H.occ[pe]), H.nw, gm ? c->gmask[e] : nullptr,
                                       gm ? c->gmask[si] : nullptr, cnt, st));
            HIPCK(hipMemcpyAsync(c->h_inv, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCK(hipEventRecord(c->ev_inv, st));
            c->inv_pending = true;
        }
    }
    if (M.flush && hit < 0) HIPCK(rc_launch_flush(&d, Np, nl, st));   // new state: clear (rtxgi_bench.patch:363-366)
    if (M.snaps && M.compact && hit >= 0) {   // revisited state: flush, then its records (rtxgi_bench.patch:1337-1345)
        HIPCK(rc_launch_flush(&d, Np, nl, st));
        HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
        if (M.snap_verify) HIPCK(hipMemsetAsync(d.rstat, 0, 4 * (size_t)H.NR * 8, st));
        HIPCK(rc_launch_snap_restore(&d, Np, nl, S->d_ekeys, (const rc_snaprec*)d.csnap + (size_t)hit * Np, c->d_slotcnt + hit, Np,
                                     M.snap_verify ? 1 : 0, c->dormant, st));
        if (M.snap_verify) {
            HIPCK(hipMemcpyAsync(d.npend, d.npend + 1, 4, hipMemcpyDeviceToDevice, st));
            c->pend_left = RC_VERIF_FRAMES;
        } else {
            c->pend_left = 0;
        }
    } else if (M.snaps && hit >= 0) {   // revisited state: restore its snapshot
        HIPCK(hipMemcpyAsync(d.Lc, d.snap + (size_t)hit * Np * nl, (size_t)Np * nl * 4, hipMemcpyDeviceToDevice, st));
        if (M.snap_verify) {
            HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
            HIPCK(hipMemsetAsync(d.rstat, 0, 4 * (size_t)H.NR * 8, st));
            HIPCK(rc_launch_restored(&d, Np, nl, c->nmax, c->dormant, 1, st));
            HIPCK(hipMemcpyAsync(d.npend, d.npend + 1, 4, hipMemcpyDeviceToDevice, st));
            c->pend_left = RC_VERIF_FRAMES;
        } else {
            HIPCK(rc_launch_restored(&d, Np, nl, c->nmax, c->dormant, 0, st));
            HIPCK(hipMemsetAsync(d.npend, 0, 8, st));
            c->pend_left = 0;
        }
    }
    c->last_event_f = f;
    if (M.eventboost > 0 && !(M.snaps && hit >= 0)) c->boost = M.eventboost;   // rtxgi_bench.patch:355-356
    if (M.mode == 1) {   // ray allocation must see this frame's invalidation
        HIPCK(rc_launch_sumpr(&d, Np, st));
    }
    return RC_OK;
}

extern "C" rc_status rc_cache_frame(rc_cache* c, int si, int want_err, int timed, rc_frame_stats* out) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    const rc::Scene& H = S->H;
    if (si < 0 || si >= H.NS) return fail(RC_ERR_ARG, "state %d out of range", si);
    const rc_method& M = c->M;
    hipStream_t st = S->stream;
    rc_dcache& d = c->c;
    const int Np = H.Np, nl = H.nl, nw = H.nw;
    rc_params& p = c->p;
    rc_frame_stats fs;
    memset(&fs, 0, sizeof fs);
    fs.err = -1.0;
    p.frame = (uint32_t)c->f;
    if (c->f > 0 && c->prev_state != si) {
        HIPCK(wait_query(c));   // event handling rewrites what the last camera pass reads
        if (timed) HIPCK(hipEventRecord(c->ev[4], st));
        rc_status r = run_event(c, c->prev_state, si);
        if (r != RC_OK) return r;
        fs.event = 1;
        if (timed) { HIPCK(hipEventRecord(c->ev[5], st)); HIPCK(hipEventSynchronize(c->ev[5])); fs.ms_event = ms_between(c->ev[4], c->ev[5]); }
    }
    p.pend_left = c->pend_left;
    p.dormant = c->dormant;
    if (c->boost > 0) {   // post-event history cap: min(nmax, 2 + k/2) frames, k frames after the event (:409-421)
        const int k = M.eventboost - c->boost;
        p.nmax = fminf(c->nmax, (2.f + (float)(k / 2)) * p.rays);
        c->boost--;
    } else {
        p.nmax = c->nmax;
    }
    if (M.ddgi) HIPCK(hipMemsetAsync(d.dstat, 0, 3 * (size_t)H.NR * 8, st));
    if (timed) HIPCK(hipEventRecord(c->ev[0], st));
    if (M.deps && c->f % 32 == 0) {   // dependency banks: swap every 32 frames, clear the new current bank
        std::swap(d.cur, d.prev);
        std::swap(d.lcur, d.lprev);
        HIPCK(hipMemsetAsync(d.cur, 0, (size_t)nw * Np * 4, st));
        HIPCK(hipMemsetAsync(d.lcur, 0, (size_t)Np * 8, st));
    }
    if (timed) HIPCK(hipEventRecord(c->ev[1], st));
    bool track = M.deps && (c->f % M.track_every == 0);
    rc_dstate dk;
    HIPCK(scene_dst(S, si, grid_threads(p.Np), &dk));
    HIPCK(rc_launch_trace2(trav_fp(S->trav), track ? 1 : 0, &S->ds, &dk, &d, &p, c->lnew, st));
    if (timed) HIPCK(hipEventRecord(c->ev[2], st));
    if (c->pend_left > 0) {
        HIPCK(hipMemsetAsync(d.npend + 1, 0, 4, st));
        HIPCK(rc_launch_decide(&S->ds, &d, &p, st));
        HIPCK(hipMemcpyAsync(d.npend, d.npend + 1, 4, hipMemcpyDeviceToDevice, st));
        c->pend_left--;
    }
    {   // the update writes the other R buffer: wait only for the camera pass that last read it
        const int k = c->rb ^ 1;
        if (c->r_pend[k]) { HIPCK(hipStreamWaitEvent(st, c->ev_r[k], 0)); c->r_pend[k] = false; }
        c->rb = k;
        d.R = c->Rb[k];
    }
    if (timed) HIPCK(hipEventRecord(c->ev[3], st));
    HIPCK(rc_launch_update(&S->ds, &S->dst[si], &d, &p, track ? c->lnew : nullptr, st));   // also sets sumpr[0] = sumpr[1]
    std::swap(d.Lc, d.Lc2);
    c->r_valid = true;
    if (timed) {
        HIPCK(hipEventRecord(c->ev[6], st));
        HIPCK(hipEventSynchronize(c->ev[6]));
        fs.ms_swap = ms_between(c->ev[0], c->ev[1]);
        fs.ms_trace = ms_between(c->ev[1], c->ev[2]);
        fs.ms_decide = ms_between(c->ev[2], c->ev[3]);
        fs.ms_update = ms_between(c->ev[3], c->ev[6]);
    }
    if (want_err && S->has_ref) {
        HIPCK(rc_launch_error(&d, Np, nl, S->dst[si].ref, S->dst[si].valid, c->dormant, st));
        double h3[3];
        HIPCK(hipMemcpyAsync(h3, d.err, 24, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        fs.err = h3[2] > 0 ? std::sqrt(h3[0] / h3[2]) / (h3[1] / h3[2]) : 0.0;
    }
    if (c->inv_pending) {   // the event's reset count: its copy was queued before this frame's trace
        HIPCK(hipEventSynchronize(c->ev_inv));
        fs.invalidated = *c->h_inv;
        c->inv_pending = false;
    }
    c->prev_state = si;
    c->f++;
    if (out) *out = fs;
    return RC_OK;
}

// ---- the event block, targeted invalidation and the resolve as calls of their own

extern "C" rc_status rc_cache_event(rc_cache* c, int to_state, int* invalidated_out) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    if (to_state < 0 || to_state >= c->S->H.NS) return fail(RC_ERR_ARG, "state %d out of range", to_state);
    if (invalidated_out) *invalidated_out = 0;
    if (to_state == c->prev_state) return RC_OK;
    if (c->f > 0) {   // as rc_cache_frame: before the first frame a state change runs no event block
        HIPCK(wait_query(c));
        rc_status r = run_event(c, c->prev_state, to_state);
        if (r != RC_OK) return r;
        if (c->inv_pending) {
            HIPCK(hipEventSynchronize(c->ev_inv));
            if (invalidated_out) *invalidated_out = *c->h_inv;
            c->inv_pending = false;
        }
    }
    c->prev_state = to_state;
    return RC_OK;
}

extern "C" rc_status rc_cache_invalidate(rc_cache* c, unsigned long long light_mask, const int* voxels, int nvox, int* invalidated_out) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    const rc::Scene& H = S->H;
    if (nvox < 0 || (nvox > 0 && !voxels)) return fail(RC_ERR_ARG, "rc_cache_invalidate: bad voxel list");
    std::vector<int> V(voxels, voxels + nvox);
    for (int v : V)
        if (v < 0 || v >= 32 * H.nw) return fail(RC_ERR_ARG, "rc_cache_invalidate: voxel %d outside the transit grid (%d voxels)", v, 32 * H.nw);
    const int Np = H.Np;
    hipStream_t st = S->stream;
    int* cnt = c->c.blkcnt + (Np + 255) / 256;
    HIPCK(rc_launch_invalidate(&c->c, Np, H.nl, (uint64_t)light_mask, nullptr, vox_mask(V), H.nw, nullptr, nullptr, cnt, st));
    HIPCK(hipMemcpyAsync(c->h_inv, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (invalidated_out) *invalidated_out = *c->h_inv;
    return RC_OK;
}

extern "C" rc_status rc_cache_resolve(rc_cache* c) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    HIPCK(wait_query(c));   // the last camera pass may still read R
    HIPCK(rc_launch_resolve(&S->ds, &S->dst[c->prev_state], c->c.Lc, c->c.R, c->dormant, S->stream));
    c->r_valid = true;
    return RC_OK;
}

// Staged camera-pass workspace for npath paths (grown on demand, kept between calls).
static hipError_t qwork_reserve(rc_cache* c, long long npath) {
    if (npath <= c->qw_paths) return hipSuccess;
    hipError_t es = hipStreamSynchronize(c->rs);   // a staged pass may still use the old workspace
    if (es == hipSuccess) es = hipStreamSynchronize(c->S->stream);
    if (es != hipSuccess) return es;
    c->free_qwork();
    rc_qwork& w = c->qw;
    const size_t n = (size_t)npath;
    hipError_t e;
#define QA(ptr, bytes) do { void* p_ = nullptr; e = hipMalloc(&p_, (bytes)); if (e != hipSuccess) { c->free_qwork(); return e; } \
        c->qw_owned.push_back(p_); ptr = decltype(ptr)(p_); } while (0)
    QA(w.ro, 16 * n); QA(w.rd, 16 * n); QA(w.so, 16 * n); QA(w.sd, 16 * n);
    QA(w.sc, 4 * n); QA(w.ht, 4 * n); QA(w.hq, 4 * n);
    QA(w.rs, 4 * n); QA(w.thr, 4 * n); QA(w.rad, 4 * n);
    const size_t nseg = (n + 1023) / 1024;
    QA(w.idx, 4 * nseg * 1024); QA(w.sidx, 4 * nseg * 1024); QA(w.seg, 8 * nseg);
    QA(w.mq0, 16 * n); QA(w.mq1, 16 * n); QA(w.mq2, 16 * n); QA(w.mpred, 4 * n);
#undef QA
    c->qw_paths = npath;
    return hipSuccess;
}

extern "C" rc_status rc_scene_set_stack_levels(rc_scene* s, int levels) {
    if (!s) return fail(RC_ERR_ARG, "NULL argument");
    if (levels < 1 || levels > 64) return fail(RC_ERR_ARG, "stack levels must be in [1, 64]");
    s->lcap = levels;
    s->lforce = true;
    return RC_OK;
}

extern "C" rc_status rc_cache_set_stack_levels(rc_cache* c, int levels) {
    if (!c) return fail(RC_ERR_ARG, "NULL argument");
    if (levels < 1 || levels > 64) return fail(RC_ERR_ARG, "stack levels must be in [1, 64]");
    c->lcap = levels;
    c->lforce = true;
    return RC_OK;
}

extern "C" rc_status rc_cache_set_render_mode(rc_cache* c, int mode) {
    if (!c) return fail(RC_ERR_ARG, "NULL argument");
    if (mode != RC_RENDER_FUSED && mode != RC_RENDER_STAGED) return fail(RC_ERR_ARG, "render mode must be 0 (fused) or 1 (staged)");
    c->render_mode = mode;
    return RC_OK;
}

extern "C" rc_status rc_cache_render(rc_cache* c, const rc_camera* cam, int no_cache, uint32_t seed, float* d_img, long long* rays_out,
                                     float* ms_out, void* stream) {
    if (!c || !cam || !d_img) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    if (cam->W < 1 || cam->H < 1 || cam->spp < 1 || cam->bounces < 0 || cam->cache_bounce < 0 || (long long)cam->W * cam->H > (1ll << 30))
        return fail(RC_ERR_ARG, "camera: bad size / spp / bounces");
    if ((long long)cam->W * cam->H * cam->spp > (1ll << 30)) return fail(RC_ERR_ARG, "camera: W * H * spp above 2^30 paths");
    rc_scene* S = c->S;
    hipStream_t st = S->stream;
    int si = c->prev_state;
    if (!no_cache && !c->r_valid) {   // no frame yet: resolve the warm-start channels
        HIPCK(wait_query(c));
        HIPCK(rc_launch_resolve(&S->ds, &S->dst[si], c->c.Lc, c->c.R, c->dormant, st));
        c->r_valid = true;
    }
    const bool staged = c->render_mode == RC_RENDER_STAGED && S->trav == 2;
    if (staged) HIPCK(qwork_reserve(c, (long long)cam->W * cam->H * cam->spp));
    rc_dstate dq = S->dst[si];   // the fused pass: lane stacks split at c->lcap levels, the rest in c->qovf
    if (!staged)
        HIPCK(split_dst(S->dst[si], S->ds, S->trav, c->lcap, c->lforce, &c->qovf, &c->qovf_words, rc_query_threads(cam), c->rs, &dq));
    // the pass runs on the cache's render stream, after everything queued on the scene stream
    hipStream_t qs = c->rs;
    HIPCK(hipEventRecord(c->ev_upd, st));
    HIPCK(hipStreamWaitEvent(qs, c->ev_upd, 0));
    if (stream) {   // and after the caller's work on its stream (what wrote or reads d_img)
        HIPCK(hipEventRecord(c->ev_in, (hipStream_t)stream));
        HIPCK(hipStreamWaitEvent(qs, c->ev_in, 0));
    }
    unsigned long long* ctr = nullptr;
    if (rays_out) { HIPCK(hipMemsetAsync(c->d_counters, 0, 16, qs)); ctr = c->d_counters; }
    if (ms_out) HIPCK(hipEventRecord(c->ev[0], qs));
    if (staged)
        HIPCK(rc_launch_query_staged(&S->ds, &S->dst[si], no_cache ? nullptr : c->c.R, cam, seed, (uint32_t)c->f, c->dormant, d_img, ctr,
                                     &c->qw, nullptr, qs));
    else
        HIPCK(rc_launch_query(S->trav, &S->ds, &dq, no_cache ? nullptr : c->c.R, cam, seed, (uint32_t)c->f, c->dormant, d_img, ctr, qs));
    HIPCK(hipEventRecord(c->ev_q, qs));
    c->q_pending = true;
    if (stream) HIPCK(hipStreamWaitEvent((hipStream_t)stream, c->ev_q, 0));   // the caller's later work sees d_img
    HIPCK(hipEventRecord(c->ev_r[c->rb], qs));   // the pass read R buffer rb
    c->r_pend[c->rb] = true;
    if (ms_out) {
        HIPCK(hipEventRecord(c->ev[1], qs));
        HIPCK(hipEventSynchronize(c->ev[1]));
        *ms_out = ms_between(c->ev[0], c->ev[1]);
    }
    if (rays_out) {
        unsigned long long h2[2];
        HIPCK(hipMemcpyAsync(h2, c->d_counters, 16, hipMemcpyDeviceToHost, qs));
        HIPCK(hipStreamSynchronize(qs));
        rays_out[0] = (long long)h2[0]; rays_out[1] = (long long)h2[1];
    }
    return RC_OK;
}

static bool cache_array(const rc_cache* c, int which, void** dp, size_t* bytes) {
    const rc::Scene& H = c->S->H;
    const size_t Np = H.Np, nl = H.nl, nw = H.nw, NR = H.NR;
    const rc_dcache& d = c->c;
    switch (which) {
        case RC_C_LC: *dp = d.Lc; *bytes = nl * Np * 4; return true;
        case RC_C_M: *dp = d.m; *bytes = nl * Np * 4; return true;
        case RC_C_N: *dp = d.n; *bytes = Np * 4; return true;
        case RC_C_KK: *dp = d.kk; *bytes = Np * 4; return true;
        case RC_C_PENDING: *dp = d.pending; *bytes = Np; return true;
        case RC_C_CUR: *dp = d.cur; *bytes = nw * Np * 4; return true;
        case RC_C_PREV: *dp = d.prev; *bytes = nw * Np * 4; return true;
        case RC_C_LCUR: *dp = d.lcur; *bytes = Np * 8; return true;
        case RC_C_LPREV: *dp = d.lprev; *bytes = Np * 8; return true;
        case RC_C_SUMPR: *dp = d.sumpr; *bytes = 16; return true;
        case RC_C_NPEND: *dp = d.npend; *bytes = 8; return true;
        case RC_C_AGE: *dp = d.age; *bytes = Np * 4; return true;
        case RC_C_STALE: *dp = d.stale; *bytes = Np * 8; return true;
        case RC_C_C0: *dp = d.c0; *bytes = Np * 4; return true;
        case RC_C_RSTAT: *dp = d.rstat; *bytes = 4 * NR * 8; return true;
        case RC_C_LNEW: *dp = c->lnew; *bytes = Np * 8; return true;
        case RC_C_DSTAT: *dp = d.dstat; *bytes = 3 * NR * 8; return true;
        case RC_C_R: *dp = d.R; *bytes = Np * 4; return true;
        default: return false;
    }
}

extern "C" size_t rc_cache_array_bytes(const rc_cache* c, int which) {
    void* p; size_t b = 0;
    if (!c || !cache_array(c, which, &p, &b)) return 0;
    return b;
}
extern "C" rc_status rc_cache_get(const rc_cache* c, int which, void* host_dst, size_t bytes) {
    void* p; size_t b = 0;
    if (!c || !cache_array(c, which, &p, &b)) return fail(RC_ERR_ARG, "unknown cache array %d", which);
    if (bytes != b) return fail(RC_ERR_ARG, "cache array %d: %zu bytes requested, %zu available", which, bytes, b);
    HIPCK(hipStreamSynchronize(c->rs));
    HIPCK(hipStreamSynchronize(c->S->stream));
    const size_t Np = c->S->H.Np;
    if (which == RC_C_M) HIPCK(hipMemcpy2D(host_dst, Np * 4, p, mpitch_of(Np) * 4, Np * 4, c->S->H.nl, hipMemcpyDeviceToHost));
    else HIPCK(hipMemcpy(host_dst, p, b, hipMemcpyDeviceToHost));
    return RC_OK;
}
extern "C" rc_status rc_cache_set(rc_cache* c, int which, const void* host_src, size_t bytes) {
    void* p; size_t b = 0;
    if (!c || !cache_array(c, which, &p, &b)) return fail(RC_ERR_ARG, "unknown cache array %d", which);
    if (bytes != b) return fail(RC_ERR_ARG, "cache array %d: %zu bytes given, %zu expected", which, bytes, b);
    HIPCK(hipStreamSynchronize(c->rs));
    HIPCK(hipStreamSynchronize(c->S->stream));
    const size_t Np = c->S->H.Np;
    if (which == RC_C_M) HIPCK(hipMemcpy2D(p, mpitch_of(Np) * 4, host_src, Np * 4, Np * 4, c->S->H.nl, hipMemcpyHostToDevice));
    else HIPCK(hipMemcpy(p, host_src, b, hipMemcpyHostToDevice));
    return RC_OK;
}
extern "C" rc_status rc_cache_sync(rc_cache* c) {
    if (!c) return fail(RC_ERR_ARG, "NULL cache");
    HIPCK(hipStreamSynchronize(c->rs));
    HIPCK(hipStreamSynchronize(c->S->stream));
    return RC_OK;
}

// ------------------------------------------------------------------------------------ kernel level
static hipStream_t pick(rc_scene* s, void* stream) { return stream ? (hipStream_t)stream : (s ? s->stream : nullptr); }

extern "C" rc_status rc_trace(rc_scene* s, int state, const float* d_O, const float* d_D, int n, float* d_t, int* d_q, int trav, void* stream) {
    if (!s || n < 0 || (n > 0 && (!d_O || !d_D || !d_t || !d_q))) return fail(RC_ERR_ARG, "rc_trace: bad arguments");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    if (state < 0 || state >= s->H.NS || trav < 0 || trav > 3) return fail(RC_ERR_ARG, "rc_trace: state / trav out of range");
    HIPCK(rc_launch_trace_rays(trav, &s->ds, &s->dst[state], d_O, d_D, n, d_t, d_q, pick(s, stream)));
    return RC_OK;
}

extern "C" rc_status rc_hash_insert(uint64_t* d_HK, int32_t* d_HV, uint32_t hmask, const uint64_t* d_keys, const int32_t* d_vals, int n,
                                    int32_t* d_fail, void* stream) {
    if (n < 0 || (n > 0 && (!d_HK || !d_HV || !d_keys || !d_vals || !d_fail))) return fail(RC_ERR_ARG, "rc_hash_insert: bad arguments");
    if ((hmask & (hmask + 1)) != 0) return fail(RC_ERR_ARG, "rc_hash_insert: table size must be a power of two");
    HIPCK(rc_launch_hash_insert((unsigned long long*)d_HK, d_HV, hmask, (const unsigned long long*)d_keys, d_vals, n, d_fail, (hipStream_t)stream));
    return RC_OK;
}
extern "C" rc_status rc_hash_query(const uint64_t* d_HK, const int32_t* d_HV, uint32_t hmask, const uint64_t* d_keys, int n, int32_t* d_out,
                                   void* stream) {
    if (n < 0 || (n > 0 && (!d_HK || !d_HV || !d_keys || !d_out))) return fail(RC_ERR_ARG, "rc_hash_query: bad arguments");
    if ((hmask & (hmask + 1)) != 0) return fail(RC_ERR_ARG, "rc_hash_query: table size must be a power of two");
    HIPCK(rc_launch_hash_query((const unsigned long long*)d_HK, d_HV, hmask, (const unsigned long long*)d_keys, n, d_out, (hipStream_t)stream));
    return RC_OK;
}
extern "C" rc_status rc_hash_point(rc_scene* s, const float* d_H, const float* d_N, int n, int32_t* d_out, void* stream) {
    if (!s || n < 0 || (n > 0 && (!d_H || !d_N || !d_out))) return fail(RC_ERR_ARG, "rc_hash_point: bad arguments");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    HIPCK(rc_launch_hash_point(&s->ds, d_H, d_N, n, d_out, pick(s, stream)));
    return RC_OK;
}
extern "C" rc_status rc_dda_mark(rc_scene* s, const float* d_A, const float* d_B, int n, uint32_t* d_out, void* stream) {
    if (!s || n < 0 || (n > 0 && (!d_A || !d_B || !d_out))) return fail(RC_ERR_ARG, "rc_dda_mark: bad arguments");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    HIPCK(rc_launch_dda_segments(&s->ds, d_A, d_B, n, d_out, pick(s, stream)));
    return RC_OK;
}
extern "C" rc_status rc_dda_volume(rc_scene* s, int gx, int gy, int gz, const float* d_A, const float* d_B, const int32_t* d_seg_vol, int n,
                                   uint32_t* d_vol, long long words_per_vol, void* stream) {
    if (!s || n < 0 || gx < 1 || gy < 1 || gz < 1 || (n > 0 && (!d_A || !d_B || !d_vol))) return fail(RC_ERR_ARG, "rc_dda_volume: bad arguments");
    if (words_per_vol < ((long long)gx * gy * gz + 31) / 32) return fail(RC_ERR_ARG, "rc_dda_volume: volume smaller than the grid");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    HIPCK(rc_launch_dda_volume(&s->ds, gx, gy, gz, d_A, d_B, d_seg_vol, n, d_vol, words_per_vol, pick(s, stream)));
    return RC_OK;
}

// ------------------------------------------------------------------------------------ adjoint
struct rc_adjoint {
    rc_scene* S = nullptr;
    int state = 0, K = 1, B = 0;
    uint32_t seed = 0;
    rc_adjbuf w{};
    bool forward_done = false, backward_done = false;
    std::vector<void*> owned;
    bool timed = false;
    hipEvent_t tev[2 * 16] = {};   // around each traced bounce (B <= 16)
    float trace_ms[2] = {0.f, 0.f};
    rc_dstate dst{};               // the state with this handle's lane-stack split (global levels in ovf)
    uint32_t* ovf = nullptr;
    size_t ovf_words = 0;
    ~rc_adjoint() {
        for (hipEvent_t e : tev) if (e) (void)hipEventDestroy(e);
        if (ovf) (void)hipFree(ovf);
        for (void* p : owned) (void)hipFree(p);
    }
};

// summed GPU time of the B traced bounces bracketed by a->tev
static rc_status adj_trace_time(rc_adjoint* a, float* out) {
    *out = 0.f;
    if (a->B < 1) return RC_OK;
    for (int k = 0; k < a->B; ++k) {
        float ms = 0.f;
        // wait for both events: a completed later event does not make an earlier one report ready
        HIPCK(hipEventSynchronize(a->tev[2 * k]));
        HIPCK(hipEventSynchronize(a->tev[2 * k + 1]));
        HIPCK(hipEventElapsedTime(&ms, a->tev[2 * k], a->tev[2 * k + 1]));
        *out += ms;
    }
    return RC_OK;
}

extern "C" rc_status rc_adjoint_create(rc_scene* s, int state, int K, int bounces, uint32_t seed, rc_adjoint** out) {
    if (!s || !out) return fail(RC_ERR_ARG, "NULL argument");
    if (!s->uploaded) return fail(RC_ERR_STATE, "scene not uploaded");
    const rc::Scene& H = s->H;
    if (state < 0 || state >= H.NS || K < 1 || K > 1024 || bounces < 0 || bounces > 16)
        return fail(RC_ERR_ARG, "rc_adjoint_create: state in [0, NS), 1 <= K <= 1024, 0 <= bounces <= 16");
    if (H.nmat < 1 || (int)H.mat.size() != H.Np) return fail(RC_ERR_SCENE, "scene has no material table");
    for (int m : H.mat) if (m < 0 || m >= H.nmat) return fail(RC_ERR_SCENE, "entry material out of range");
    rc_adjoint* a = new rc_adjoint();
    a->S = s; a->state = state; a->K = K; a->B = bounces; a->seed = seed;
    rc_adjbuf& w = a->w;
    const size_t Np = H.Np, nl = H.nl, nblk = (Np + 255) / 256;
#define AK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { delete a; return fail(RC_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
    AK(dalloc(&w.G, Np * nl, a->owned)); AK(dalloc(&w.SG, Np, a->owned)); AK(dalloc(&w.mat, Np, a->owned));
    AK(dalloc(&w.A, Np, a->owned)); AK(dalloc(&w.E, Np, a->owned)); AK(dalloc(&w.L, Np * (size_t)std::max(bounces, 1), a->owned));
    AK(dalloc(&w.S[0], Np, a->owned)); AK(dalloc(&w.S[1], Np, a->owned)); AK(dalloc(&w.SR, Np, a->owned));
    AK(dalloc(&w.dA, Np, a->owned)); AK(dalloc(&w.dE, Np, a->owned));
    AK(dalloc(&w.lam[0], Np, a->owned)); AK(dalloc(&w.lam[1], Np, a->owned));
    AK(dalloc(&w.part, nblk * nl, a->owned));
    AK(dalloc(&w.perm, Np, a->owned)); AK(dalloc(&w.moff, (size_t)H.nmat + 1, a->owned));
    AK(dalloc(&w.pmat, Np, a->owned)); AK(dalloc(&w.tail, Np, a->owned));
    w.nmat = H.nmat;
    // entries grouped by material (counting sort, entry order kept inside a material)
    std::vector<int> moff(H.nmat + 1, 0), perm(Np);
    for (int m : H.mat) ++moff[m + 1];
    for (int m = 0; m < H.nmat; ++m) moff[m + 1] += moff[m];
    {
        std::vector<int> pos(moff.begin(), moff.end() - 1);
        for (size_t i = 0; i < Np; ++i) perm[pos[H.mat[i]]++] = (int)i;
    }
    AK(hipMemcpyAsync(w.mat, H.mat.data(), Np * 4, hipMemcpyHostToDevice, s->stream));
    AK(hipMemcpyAsync(w.perm, perm.data(), Np * 4, hipMemcpyHostToDevice, s->stream));
    AK(hipMemcpyAsync(w.moff, moff.data(), moff.size() * 4, hipMemcpyHostToDevice, s->stream));
    std::vector<int> pmat(Np);
    for (int m = 0; m < H.nmat; ++m) for (int j = moff[m]; j < moff[m + 1]; ++j) pmat[j] = m;
    AK(hipMemcpyAsync(w.pmat, pmat.data(), Np * 4, hipMemcpyHostToDevice, s->stream));
    // the handle's kernels run on the caller's stream: a global stack buffer of their own
    AK(split_dst(s->dst[state], s->ds, s->trav, s->lcap, s->lforce, &a->ovf, &a->ovf_words, grid_threads(H.Np), s->stream, &a->dst));
    AK(rc_launch_adj_unit(trav_fp(s->trav), &s->ds, &a->dst, w.G, w.SG, s->stream));
    AK(hipStreamSynchronize(s->stream));
#undef AK
    *out = a;
    return RC_OK;
}

extern "C" void rc_adjoint_destroy(rc_adjoint* a) { delete a; }

extern "C" rc_status rc_adjoint_forward(rc_adjoint* a, const float* d_I, const float* d_alb, float* d_R, void* stream) {
    if (!a || !d_I || !d_alb) return fail(RC_ERR_ARG, "NULL argument");
    rc_scene* s = a->S;
    HIPCK(rc_launch_adj_forward(trav_fp(s->trav), &s->ds, &a->dst, &a->w, d_I, d_alb, a->K, a->B, a->seed, d_R, (hipStream_t)stream,
                                a->timed ? a->tev : nullptr));
    a->forward_done = true;
    return a->timed ? adj_trace_time(a, &a->trace_ms[0]) : RC_OK;
}

extern "C" rc_status rc_adjoint_backward(rc_adjoint* a, const float* d_gR, float* d_gI, float* d_galb, void* stream) {
    if (!a || !d_gR || !d_gI || !d_galb) return fail(RC_ERR_ARG, "NULL argument");
    if (!a->forward_done) return fail(RC_ERR_STATE, "rc_adjoint_backward before rc_adjoint_forward");
    rc_scene* s = a->S;
    HIPCK(rc_launch_adj_backward(trav_fp(s->trav), &s->ds, &a->dst, &a->w, d_gR, a->K, a->B, a->seed, d_gI, d_galb, (hipStream_t)stream,
                                 a->timed ? a->tev : nullptr));
    a->backward_done = true;
    return a->timed ? adj_trace_time(a, &a->trace_ms[1]) : RC_OK;
}

extern "C" rc_status rc_adjoint_set_timed(rc_adjoint* a, int timed) {
    if (!a) return fail(RC_ERR_ARG, "NULL argument");
    if (timed && !a->tev[0])
        for (int k = 0; k < 2 * a->B; ++k) HIPCK(hipEventCreate(&a->tev[k]));
    a->timed = timed != 0;
    a->trace_ms[0] = a->trace_ms[1] = 0.f;
    return RC_OK;
}

extern "C" rc_status rc_adjoint_trace_ms(const rc_adjoint* a, float ms_out[2]) {
    if (!a || !ms_out) return fail(RC_ERR_ARG, "NULL argument");
    ms_out[0] = a->trace_ms[0];
    ms_out[1] = a->trace_ms[1];
    return RC_OK;
}

static bool adj_array(const rc_adjoint* a, int which, const void** dp, size_t* bytes) {
    const size_t Np = a->S->H.Np, nl = a->S->H.nl;
    switch (which) {
        case RC_ADJ_G: *dp = a->w.G; *bytes = Np * nl * 4; return true;
        case RC_ADJ_SG: *dp = a->w.SG; *bytes = Np * 8; return true;
        case RC_ADJ_SR: if (!a->forward_done) return false; *dp = a->w.SR; *bytes = Np * 8; return true;
        case RC_ADJ_LB: if (!a->forward_done || a->B < 1) return false; *dp = a->w.L + (size_t)(a->B - 1) * Np; *bytes = Np * 4; return true;
        case RC_ADJ_DA: if (!a->backward_done) return false; *dp = a->w.dA; *bytes = Np * 4; return true;
        case RC_ADJ_DE: if (!a->backward_done) return false; *dp = a->w.dE; *bytes = Np * 4; return true;
        default: return false;
    }
}

extern "C" size_t rc_adjoint_array_bytes(const rc_adjoint* a, int which) {
    const void* p; size_t b = 0;
    if (!a || !adj_array(a, which, &p, &b)) return 0;
    return b;
}

extern "C" rc_status rc_adjoint_get(const rc_adjoint* a, int which, void* host_dst, size_t bytes) {
    const void* p; size_t b = 0;
    if (!a || !host_dst) return fail(RC_ERR_ARG, "NULL argument");
    if (!adj_array(a, which, &p, &b)) return fail(RC_ERR_ARG, "rc_adjoint_get: unknown array or not computed yet");
    if (bytes != b) return fail(RC_ERR_ARG, "rc_adjoint_get: %zu bytes given, %zu expected", bytes, b);
    HIPCK(hipDeviceSynchronize());
    HIPCK(hipMemcpy(host_dst, p, b, hipMemcpyDeviceToHost));
    return RC_OK;
}

// Gradient-support masks of every state for the grad_masks methods: one adjoint forward per
// state (unit emission, the material albedos; the support does not depend on their values
// while they are positive), grad_masks rays per entry and bounce, 3 bounces.
static rc_status cache_grad_masks(rc_cache* c) {
    rc_scene* S = c->S;
    const rc::Scene& H = S->H;
    const size_t Np = H.Np;
    float *dI = nullptr, *dalb = nullptr;
    HIPCK(dalloc(&dI, (size_t)H.nl, c->owned));
    HIPCK(dalloc(&dalb, (size_t)H.nmat, c->owned));
    std::vector<float> ones(H.nl, 1.f);
    HIPCK(hipMemcpyAsync(dI, ones.data(), ones.size() * 4, hipMemcpyHostToDevice, S->stream));
    HIPCK(hipMemcpyAsync(dalb, H.matalb.data(), H.matalb.size() * 4, hipMemcpyHostToDevice, S->stream));
    HIPCK(hipStreamSynchronize(S->stream));
    for (int st = 0; st < H.NS; ++st) {
        rc_adjoint* a = nullptr;
        rc_status r = rc_adjoint_create(S, st, c->M.grad_masks, 3, 0x5eed0000u + (uint32_t)st, &a);
        if (r != RC_OK) return r;
        uint64_t* g = nullptr;
        hipError_t e = dalloc(&g, Np, c->owned);
        if (e == hipSuccess) e = rc_launch_adj_forward(trav_fp(S->trav), &S->ds, &a->dst, &a->w, dI, dalb, a->K, a->B, a->seed, nullptr, S->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(g, a->w.SR, Np * 8, hipMemcpyDeviceToDevice, S->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
        rc_adjoint_destroy(a);
        if (e != hipSuccess) return fail(RC_ERR_HIP, "gradient masks: %s", hipGetErrorString(e));
        c->gmask.push_back(g);
    }
    return RC_OK;
}

// ------------------------------------------------------------------------------------ MLP cache
struct rc_mlp {
    rc_mlp_config cfg{};
    rc_mlpdev d{};
    rc_mlpbuf b{};
    int step = 0;
    float fallback = 0.f;   // per-query grid fallback threshold of the camera pass (0: none)
    hipEvent_t iev[2]{};    // around the batched inference of a timed rc_mlp_render
    float infer_ms = 0.f;   // its time
    // online-training workspace (max_batch samples)
    float4 *sP = nullptr, *sN = nullptr, *sD = nullptr, *hP = nullptr, *hN = nullptr, *hD = nullptr;
    float2* AE = nullptr;
    float *hy = nullptr, *tg = nullptr, *jacc = nullptr;
    int* ent = nullptr;
    std::vector<void*> owned;
    ~rc_mlp() {
        for (hipEvent_t e : iev) if (e) (void)hipEventDestroy(e);
        for (void* p : owned) (void)hipFree(p);
    }
};

extern "C" rc_status rc_mlp_create(const rc_mlp_config* c, rc_mlp** out) {
    if (!c || !out) return fail(RC_ERR_ARG, "NULL argument");
    if (c->log2_table < 10 || c->log2_table > 24 || c->base_res < 2 || c->max_res < c->base_res || c->max_res > (1 << 20) ||
        c->max_batch < 1 || c->max_batch > (1 << 24) || !(c->lr > 0.f) || !(c->ext[0] > 0.f) || !(c->ext[1] > 0.f) || !(c->ext[2] > 0.f))
        return fail(RC_ERR_ARG, "rc_mlp_create: bad configuration");
    rc_mlp* m = new rc_mlp();
    m->cfg = *c;
    rc_mlpdev& d = m->d;
    d.log2T = c->log2_table;
    const size_t T = (size_t)1 << c->log2_table, nt = (size_t)RC_MLP_LEVELS * 2 * T;
    const double gs = std::exp((std::log((double)c->max_res) - std::log((double)c->base_res)) / (RC_MLP_LEVELS - 1));
    d.dense = 0;
    for (int l = 0; l < RC_MLP_LEVELS; ++l) {
        d.res[l] = (int)std::floor(c->base_res * std::pow(gs, l) + 1e-9);
        const double r1 = d.res[l] + 1.0;
        if (r1 * r1 * r1 <= (double)T) d.dense |= 1u << l;
    }
    for (int a = 0; a < 3; ++a) { d.lo[a] = c->lo[a]; d.iext[a] = 1.f / c->ext[a]; }
    rc_mlpbuf& b = m->b;
    b.cap = rc_mlp_pad(c->max_batch);
    const size_t cap = b.cap, nsplit = (cap + 255) / 256;   // weight-gradient splits of 256 samples (rc_launch_mlp_grad)
#define MK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { delete m; return fail(RC_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); } } while (0)
    MK(dalloc(&d.table, nt, m->owned)); MK(dalloc(&d.gtable, nt, m->owned));
    MK(dalloc(&d.Wm, (size_t)RC_MLP_NW, m->owned)); MK(dalloc(&d.gW, (size_t)RC_MLP_NW, m->owned));
    uint16_t *wt = nullptr, *wk = nullptr;
    MK(dalloc(&wt, (size_t)RC_MLP_NW, m->owned)); MK(dalloc(&wk, (size_t)RC_MLP_NW, m->owned));
    d.Wt = wt; d.Wk = wk;
    uint16_t *act = nullptr, *gz = nullptr;
    MK(dalloc(&act, (size_t)(RC_MLP_HIDDEN + 1) * 64 * cap, m->owned)); MK(dalloc(&gz, (size_t)RC_MLP_HIDDEN * 64 * cap, m->owned));
    b.actT = act; b.gzT = gz;
    MK(dalloc(&b.y, cap, m->owned)); MK(dalloc(&b.dy, cap, m->owned)); MK(dalloc(&b.gX, cap * 2 * RC_MLP_LEVELS, m->owned));
    MK(dalloc(&b.part, (size_t)(RC_MLP_HIDDEN + 1) * nsplit * 64 * 64, m->owned));
    MK(dalloc(&b.mW, (size_t)RC_MLP_NW, m->owned)); MK(dalloc(&b.vW, (size_t)RC_MLP_NW, m->owned));
    MK(dalloc(&b.mT, nt, m->owned)); MK(dalloc(&b.vT, nt, m->owned)); MK(dalloc(&b.loss, 1, m->owned));
    {
        const size_t nbk = (size_t)RC_MLP_LEVELS * (T >> std::min(RC_EB_SB, c->log2_table));
        b.epairs = (int64_t)std::min<size_t>(cap, RC_EB_CHUNK) * RC_EB_PAIRS;
        MK(dalloc(&b.ecnt, nbk + 1, m->owned)); MK(dalloc(&b.ecur, nbk + 1, m->owned));
        MK(dalloc(&b.ebase, nbk + 1, m->owned)); MK(dalloc(&b.epart, nbk + 1, m->owned));
        MK(dalloc(&b.epmap, nbk + (size_t)(b.epairs + RC_EB_PMAX - 1) / RC_EB_PMAX, m->owned));
        MK(dalloc(&b.eslot, (size_t)b.epairs, m->owned)); MK(dalloc(&b.eval, 2 * (size_t)b.epairs, m->owned));
    }
    {
        const size_t nb = (size_t)c->max_batch;
        MK(dalloc(&m->sP, nb, m->owned)); MK(dalloc(&m->sN, nb, m->owned)); MK(dalloc(&m->sD, nb, m->owned));
        MK(dalloc(&m->hP, nb, m->owned)); MK(dalloc(&m->hN, nb, m->owned)); MK(dalloc(&m->hD, nb, m->owned));
        MK(dalloc(&m->AE, nb, m->owned)); MK(dalloc(&m->hy, nb, m->owned)); MK(dalloc(&m->tg, nb, m->owned));
        MK(dalloc(&m->ent, nb, m->owned)); MK(dalloc(&m->jacc, 2, m->owned));
    }
    // initialisation: weights uniform +-sqrt(6 / 64) (output layer x 0.1), table uniform +-1e-4
    std::vector<float> W(RC_MLP_NW), tb(nt);
    uint64_t r = (uint64_t)c->seed * 0x9E3779B97F4A7C15ull + 1;
    auto u = [&]() { return (float)(rc::mix64(r += 0x9E3779B97F4A7C15ull) >> 40) * (1.f / 16777216.f); };
    const float bw = std::sqrt(6.f / 64.f);
    for (int e = 0; e < RC_MLP_NW; ++e) {
        const bool out_layer = e >= RC_MLP_HIDDEN * 64 * 64;
        const bool used = !out_layer || ((e - RC_MLP_HIDDEN * 64 * 64) % 16) == 0;
        W[e] = used ? (2.f * u() - 1.f) * bw * (out_layer ? 0.1f : 1.f) : 0.f;
    }
    for (size_t e = 0; e < nt; ++e) tb[e] = (2.f * u() - 1.f) * 1e-4f;
    MK(hipMemcpy(d.Wm, W.data(), W.size() * 4, hipMemcpyHostToDevice));
    MK(hipMemcpy(d.table, tb.data(), tb.size() * 4, hipMemcpyHostToDevice));
    for (void* p : {(void*)d.gtable, (void*)d.gW, (void*)b.mW, (void*)b.vW, (void*)b.mT, (void*)b.vT})
        MK(hipMemset(p, 0, (p == d.gW || p == b.mW || p == b.vW) ? (size_t)RC_MLP_NW * 4 : nt * 4));
    MK(rc_launch_mlp_pack(&d, nullptr));
    MK(hipDeviceSynchronize());
#undef MK
    *out = m;
    return RC_OK;
}

extern "C" void rc_mlp_destroy(rc_mlp* m) { delete m; }

static rc_status mlp_check(rc_mlp* m, const float* P, const float* N, const float* D, int n) {
    if (!m) return fail(RC_ERR_ARG, "NULL mlp");
    if (n < 0 || n > m->cfg.max_batch) return fail(RC_ERR_ARG, "batch of %d outside 0..%d", n, m->cfg.max_batch);
    if (n > 0 && (!P || !N || !D)) return fail(RC_ERR_ARG, "NULL input array");
    return RC_OK;
}

extern "C" rc_status rc_mlp_infer(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, int n, float* d_y, void* stream) {
    rc_status r = mlp_check(m, d_P4, d_N4, d_D4, n);
    if (r != RC_OK) return r;
    if (n > 0 && !d_y) return fail(RC_ERR_ARG, "NULL output");
    HIPCK(rc_launch_mlp_infer(&m->d, (const float4*)d_P4, (const float4*)d_N4, (const float4*)d_D4, n, d_y, (hipStream_t)stream));
    return RC_OK;
}

extern "C" rc_status rc_mlp_backward(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, const float* d_dy, int n,
                                     void* stream) {
    rc_status r = mlp_check(m, d_P4, d_N4, d_D4, n);
    if (r != RC_OK) return r;
    if (n > 0 && !d_dy) return fail(RC_ERR_ARG, "NULL dy");
    hipStream_t st = (hipStream_t)stream;
    // the kernels read dy over the padded batch: stage it with zero padding
    HIPCK(rc_launch_mlp_zero(m->b.dy, m->b.cap, st));
    if (n > 0) HIPCK(hipMemcpyAsync(m->b.dy, d_dy, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    HIPCK(rc_launch_mlp_grad(&m->d, &m->b, (const float4*)d_P4, (const float4*)d_N4, (const float4*)d_D4, nullptr, m->b.dy, n, nullptr, st));
    return RC_OK;
}

extern "C" rc_status rc_mlp_step(rc_mlp* m, void* stream) {
    if (!m) return fail(RC_ERR_ARG, "NULL mlp");
    ++m->step;
    HIPCK(rc_launch_mlp_adam(&m->d, &m->b, m->cfg.lr, m->cfg.beta1, m->cfg.beta2, m->cfg.eps, m->step, (hipStream_t)stream));
    return RC_OK;
}

extern "C" rc_status rc_mlp_zero_grad(rc_mlp* m, void* stream) {
    if (!m) return fail(RC_ERR_ARG, "NULL mlp");
    const long long nt = (long long)RC_MLP_LEVELS * ((long long)2 << m->d.log2T);
    HIPCK(rc_launch_mlp_zero(m->d.gtable, nt, (hipStream_t)stream));
    HIPCK(rc_launch_mlp_zero(m->d.gW, RC_MLP_NW, (hipStream_t)stream));
    return RC_OK;
}

extern "C" rc_status rc_mlp_train(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, const float* d_t, int n,
                                  float* loss_out, void* stream) {
    rc_status r = mlp_check(m, d_P4, d_N4, d_D4, n);
    if (r != RC_OK) return r;
    if (n < 1 || !d_t) return fail(RC_ERR_ARG, "rc_mlp_train needs targets and n >= 1");
    hipStream_t st = (hipStream_t)stream;
    HIPCK(rc_launch_mlp_zero(m->b.loss, 1, st));
    HIPCK(rc_launch_mlp_grad(&m->d, &m->b, (const float4*)d_P4, (const float4*)d_N4, (const float4*)d_D4, d_t, nullptr, n, m->b.loss, st));
    ++m->step;
    HIPCK(rc_launch_mlp_adam(&m->d, &m->b, m->cfg.lr, m->cfg.beta1, m->cfg.beta2, m->cfg.eps, m->step, st));
    if (loss_out) {
        HIPCK(hipMemcpyAsync(loss_out, m->b.loss, 4, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
    }
    return RC_OK;
}

static bool mlp_array(const rc_mlp* m, int which, void** p, size_t* bytes) {
    const size_t nt = (size_t)RC_MLP_LEVELS * 2 * ((size_t)1 << m->d.log2T);
    switch (which) {
        case RC_MLP_W: *p = m->d.Wm; *bytes = (size_t)RC_MLP_NW * 4; return true;
        case RC_MLP_TABLE: *p = m->d.table; *bytes = nt * 4; return true;
        case RC_MLP_GW: *p = m->d.gW; *bytes = (size_t)RC_MLP_NW * 4; return true;
        case RC_MLP_GTABLE: *p = m->d.gtable; *bytes = nt * 4; return true;
        default: return false;
    }
}

extern "C" size_t rc_mlp_array_bytes(const rc_mlp* m, int which) {
    void* p; size_t b = 0;
    if (!m || !mlp_array(m, which, &p, &b)) return 0;
    return b;
}

extern "C" rc_status rc_mlp_get(const rc_mlp* m, int which, void* host_dst, size_t bytes) {
    void* p; size_t b = 0;
    if (!m || !host_dst || !mlp_array(m, which, &p, &b)) return fail(RC_ERR_ARG, "rc_mlp_get: bad arguments");
    if (bytes != b) return fail(RC_ERR_ARG, "rc_mlp_get: %zu bytes given, %zu expected", bytes, b);
    HIPCK(hipDeviceSynchronize());
    HIPCK(hipMemcpy(host_dst, p, b, hipMemcpyDeviceToHost));
    return RC_OK;
}

extern "C" rc_status rc_mlp_set(rc_mlp* m, int which, const void* host_src, size_t bytes) {
    void* p; size_t b = 0;
    if (!m || !host_src || !mlp_array(m, which, &p, &b)) return fail(RC_ERR_ARG, "rc_mlp_set: bad arguments");
    if (bytes != b) return fail(RC_ERR_ARG, "rc_mlp_set: %zu bytes given, %zu expected", bytes, b);
    HIPCK(hipDeviceSynchronize());
    HIPCK(hipMemcpy(p, host_src, b, hipMemcpyHostToDevice));
    if (which == RC_MLP_W) HIPCK(rc_launch_mlp_pack(&m->d, nullptr));
    HIPCK(hipDeviceSynchronize());
    return RC_OK;
}

extern "C" rc_status rc_mlp_train_scene(rc_mlp* m, rc_cache* c, int n, uint32_t seed, float* loss_out, float* judge_out) {
    if (!m || !c) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    if (n < 1 || n > m->cfg.max_batch) return fail(RC_ERR_ARG, "batch of %d outside 1..%d", n, m->cfg.max_batch);
    rc_scene* S = c->S;
    hipStream_t st = S->stream;
    const int si = c->prev_state;
    HIPCK(wait_query(c));
    if (!c->r_valid) {   // the judge reads the cache's resolved radiance
        HIPCK(rc_launch_resolve(&S->ds, &S->dst[si], c->c.Lc, c->c.R, c->dormant, st));
        c->r_valid = true;
    }
    rc_dstate dk;
    HIPCK(scene_dst(S, si, grid_threads(n), &dk));
    HIPCK(rc_launch_mlp_samples(trav_fp(S->trav), &S->ds, &dk, c->dormant, seed, n, m->sP, m->sN, m->sD, m->hP, m->hN, m->hD, m->AE, m->ent, st));
    HIPCK(rc_launch_mlp_infer(&m->d, m->hP, m->hN, m->hD, n, m->hy, st));
    HIPCK(rc_launch_mlp_targets(n, m->hP, m->hy, m->AE, m->tg, st));
    HIPCK(rc_launch_mlp_zero(m->b.loss, 1, st));
    HIPCK(rc_launch_mlp_zero(m->jacc, 2, st));
    HIPCK(rc_launch_mlp_grad(&m->d, &m->b, m->sP, m->sN, m->sD, m->tg, nullptr, n, m->b.loss, st));
    HIPCK(rc_launch_mlp_judge(n, m->b.y, c->c.R, m->ent, m->jacc, st));
    ++m->step;
    HIPCK(rc_launch_mlp_adam(&m->d, &m->b, m->cfg.lr, m->cfg.beta1, m->cfg.beta2, m->cfg.eps, m->step, st));
    float h[3];
    HIPCK(hipMemcpyAsync(h, m->b.loss, 4, hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(h + 1, m->jacc, 8, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (loss_out) *loss_out = h[0];
    if (judge_out) *judge_out = h[2] > 0.f ? std::sqrt(h[1] / h[2]) : -1.f;
    return RC_OK;
}

extern "C" rc_status rc_mlp_set_grid_fallback(rc_mlp* m, float thr) {
    if (!m || !(thr >= 0.f)) return fail(RC_ERR_ARG, "rc_mlp_set_grid_fallback: threshold must be >= 0");
    m->fallback = thr;
    return RC_OK;
}

extern "C" rc_status rc_mlp_render(rc_mlp* m, rc_cache* c, const rc_camera* cam, uint32_t seed, float* d_img, float* ms_out, void* stream) {
    if (!m || !c || !cam || !d_img) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    if (cam->W < 1 || cam->H < 1 || cam->spp < 1 || cam->bounces < 0 || cam->cache_bounce < 0 || (long long)cam->W * cam->H * cam->spp > (1ll << 30))
        return fail(RC_ERR_ARG, "camera: bad size / spp / bounces");
    rc_scene* S = c->S;
    if (S->trav < 2) return fail(RC_ERR_STATE, "the learned-cache camera pass runs on the 4-wide BVH (trav 2 or 3)");
    hipStream_t st = S->stream;
    const int si = c->prev_state;
    HIPCK(wait_query(c));   // the grid cache's camera pass may still use the staged workspace
    HIPCK(qwork_reserve(c, (long long)cam->W * cam->H * cam->spp));
    rc_dstate dk = S->dst[si];
    if (c->render_mode != RC_RENDER_STAGED) HIPCK(scene_dst(S, si, rc_query_mlp_threads(cam), &dk));
    if (stream) {   // after the caller's work on its stream
        HIPCK(hipEventRecord(c->ev_in, (hipStream_t)stream));
        HIPCK(hipStreamWaitEvent(st, c->ev_in, 0));
    }
    if (ms_out) HIPCK(hipEventRecord(c->ev[0], st));
    rc_mlpdev md = m->d;   // with the grid cache's resolved radiance as the per-query judge when a threshold is set
    md.jR = m->fallback > 0.f ? c->c.R : nullptr;
    md.jthr = m->fallback;
    if (m->fallback > 0.f && !c->r_valid) {   // no frame yet: resolve the warm-start channels
        HIPCK(rc_launch_resolve(&S->ds, &S->dst[si], c->c.Lc, c->c.R, c->dormant, st));
        c->r_valid = true;
    }
    if (ms_out && !m->iev[0])
        for (hipEvent_t& e : m->iev) HIPCK(hipEventCreate(&e));
    const hipEvent_t* iev = ms_out ? m->iev : nullptr;
    if (c->render_mode == RC_RENDER_STAGED)
        HIPCK(rc_launch_query_staged(&S->ds, &S->dst[si], nullptr, cam, seed, (uint32_t)c->f, c->dormant, d_img, nullptr, &c->qw, &md, st, iev));
    else
        HIPCK(rc_launch_query_mlp(&S->ds, &dk, cam, seed, (uint32_t)c->f, c->dormant, d_img, &c->qw, &md, st, iev));
    if (stream) {   // the caller's later work sees d_img
        HIPCK(hipEventRecord(c->ev_out, st));
        HIPCK(hipStreamWaitEvent((hipStream_t)stream, c->ev_out, 0));
    }
    if (ms_out) {
        HIPCK(hipEventRecord(c->ev[1], st));
        HIPCK(hipEventSynchronize(c->ev[1]));
        *ms_out = ms_between(c->ev[0], c->ev[1]);
        m->infer_ms = ms_between(m->iev[0], m->iev[1]);
    }
    return RC_OK;
}

extern "C" rc_status rc_mlp_last_infer_ms(const rc_mlp* m, float* ms_out) {
    if (!m || !ms_out) return fail(RC_ERR_ARG, "NULL argument");
    *ms_out = m->infer_ms;
    return RC_OK;
}

// ------------------------------------------------------------------------------------ compact snapshot wire format
static_assert(sizeof(rc_snap_rec) == 24 && sizeof(rc_snaprec) == 24, "snapshot records are 24 bytes");

extern "C" rc_status rc_cache_snapshot_save(rc_cache* c, rc_snap_rec* host_records, size_t capacity, size_t* count_out) {
    if (!c || !count_out || (capacity > 0 && !host_records)) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    const size_t Np = S->H.Np;
    HIPCK(wait_query(c));
    rc_snaprec* tmp = nullptr;
    HIPCK(hipMalloc((void**)&tmp, Np * sizeof(rc_snaprec)));
    int cnt = 0;
    hipError_t e = snap_pack(c, tmp, c->d_slotcnt + RC_KSNAP);
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, c->d_slotcnt + RC_KSNAP, sizeof(int), hipMemcpyDeviceToHost, S->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    const size_t n = std::min((size_t)cnt, capacity);
    if (e == hipSuccess && n) e = hipMemcpy(host_records, tmp, n * sizeof(rc_snaprec), hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(RC_ERR_HIP, "snapshot save: %s", hipGetErrorString(e));
    *count_out = n;
    return RC_OK;
}

// ---- allocation-free snapshot save
// The records are staged in a file-scope device array, `scratch`, instead of a buffer allocated per call.
// Host code never passes the array's name to a kernel or copy (on the host it is only a stand-in for the
// symbol): its device address on each GPU comes from hipGetSymbolAddress on the first call with that
// GPU current, and is kept per device. Every call on one GPU shares that GPU's array; the pack writes
// records [0, count) before they are read, so nothing in it needs clearing. Calls that overlap on one
// GPU race on it.
#define RC_NOALLOC_RECORDS (1u << 21)   // 2,097,152 records of 24 bytes: 48 MiB per GPU
#define RC_MAX_DEVICES 64
static __device__ __align__(256) rc_snaprec scratch[RC_NOALLOC_RECORDS];
static std::mutex g_scratch_mu;
static rc_snaprec* g_scratch_ptr[RC_MAX_DEVICES] = {};

static hipError_t scratch_records(rc_snaprec** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= RC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (rc_snaprec*)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" rc_status rc_cache_snapshot_save_noalloc(rc_cache* c, rc_snap_rec* host_records, size_t capacity, size_t* count_out) {
    if (!c || !count_out || (capacity > 0 && !host_records)) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    const size_t Np = S->H.Np;
    if (Np > RC_NOALLOC_RECORDS)
        return fail(RC_ERR_ARG, "rc_cache_snapshot_save_noalloc: the cache has %zu entries, at most %u fit the static record array "
                    "(use rc_cache_snapshot_save)", Np, RC_NOALLOC_RECORDS);
    rc_snaprec* rec = nullptr;
    HIPCK(scratch_records(&rec));
    HIPCK(wait_query(c));
    HIPCK(snap_pack(c, rec, c->d_slotcnt + RC_KSNAP));
    int cnt = 0;
    HIPCK(hipMemcpyAsync(&cnt, c->d_slotcnt + RC_KSNAP, sizeof(int), hipMemcpyDeviceToHost, S->stream));
    HIPCK(hipStreamSynchronize(S->stream));
    const size_t n = std::min((size_t)cnt, capacity);
    if (n) {
        HIPCK(hipMemcpyAsync(host_records, rec, n * sizeof(rc_snaprec), hipMemcpyDeviceToHost, S->stream));
        HIPCK(hipStreamSynchronize(S->stream));
    }
    *count_out = n;
    return RC_OK;
}

extern "C" rc_status rc_cache_snapshot_load(rc_cache* c, const rc_snap_rec* host_records, size_t count) {
    if (!c || (count > 0 && !host_records)) return fail(RC_ERR_ARG, "NULL argument");
    if (!c->begun) return fail(RC_ERR_STATE, "rc_cache_begin not called");
    rc_scene* S = c->S;
    const int Np = S->H.Np, nl = S->H.nl;
    if (count > (size_t)Np) return fail(RC_ERR_ARG, "more records (%zu) than entries (%d)", count, Np);
    hipStream_t st = S->stream;
    HIPCK(wait_query(c));
    rc_snaprec* tmp = nullptr;
    HIPCK(hipMalloc((void**)&tmp, std::max<size_t>(count, 1) * sizeof(rc_snaprec)));
    hipError_t e = count ? hipMemcpyAsync(tmp, host_records, count * sizeof(rc_snaprec), hipMemcpyHostToDevice, st) : hipSuccess;
    const int n = (int)count;
    if (e == hipSuccess) e = hipMemcpyAsync(c->d_slotcnt + RC_KSNAP, &n, sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = rc_launch_flush(&c->c, Np, nl, st);
    if (e == hipSuccess) e = rc_launch_snap_restore(&c->c, Np, nl, S->d_ekeys, tmp, c->d_slotcnt + RC_KSNAP, n, 0, c->dormant, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(RC_ERR_HIP, "snapshot load: %s", hipGetErrorString(e));
    c->r_valid = false;
    return RC_OK;
}

// Tracing helpers shared by the kernel translation units: block size, wave reductions, the
// full (dynamic + static) closest-hit trace per traversal mode, the dynamic-triangle LDS
// preload and the LDS stack layout, and the host-side grid / LDS sizes.
#pragma once
#include <map>
#include <mutex>
#include <tuple>
#include "rc_device.h"

#define RC_TB 256

// Reduction of a 64-bit integer over the 64 lanes of a wave (any order gives the same sum).
__device__ __forceinline__ long long rc_wave_sum_i64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double rc_wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ __forceinline__ long long rc_fx(float v, float scale) { return (long long)(v * scale); }

// Static geometry per traversal mode: 0 binary in the prototype's order, 1 binary closest-hit,
// 2 wide closest-hit (LDS stack), 3 the same over the quantized nodes.
template <int TRAV, class LS>
__device__ __forceinline__ void rc_trace_static(f3 o, f3 d, const rc_dscene& s, float& tmin, int& q, const LS& lstk) {
    if (TRAV == 0) rc_trace_bvh_ref(o, d, s, tmin, q);
    else if (TRAV == 1) rc_trace_bvh_closest(o, d, s, tmin, q);
    else rc_trace_wide_closest<TRAV == 3>(o, d, s, tmin, q, lstk);
}

// Dynamic triangles from LDS, then the static BVH carrying tmin (cuda/poc_gpu.cu:340-344).
template <int TRAV, class LS>
__device__ __forceinline__ bool rc_trace_full(f3 o, f3 d, const float* T, const int* TQ, int nt, const rc_dscene& s, float& tmin, int& q,
                                              const LS& lstk) {
    bool h = rc_trace_lds(o, d, T, TQ, nt, tmin, q);
    if (s.nst > 0) { rc_trace_static<TRAV>(o, d, s, tmin, q, lstk); h = q != -1; }
    return h;
}

// Dynamic LDS of the tracing kernels: nt triangles (9 floats) + their quad ids, then the
// traversal stack (wstack levels, or cap levels of a capped stack, x 256 lanes) and leaf queue
// (RC_WQ levels) at a 16-byte aligned offset.
__device__ __forceinline__ void rc_load_dyn(const rc_dstate& st, float* sh, int* shq) {
    for (int j = threadIdx.x; j < 9 * st.nt; j += blockDim.x) sh[j] = st.T[j];
    for (int j = threadIdx.x; j < st.nt; j += blockDim.x) shq[j] = st.TQ[j];
    __syncthreads();
}
__device__ __forceinline__ uint32_t* rc_lstack(float* sh, int nt) { return (uint32_t*)(sh + ((10 * nt + 3) & ~3)); }
// the whole stack in LDS (launched with rc_shm(..., false))
__device__ __forceinline__ rc_lstk_l rc_lstk_lds(float* sh, int nt, const rc_dscene& s) {
    rc_lstk_l k;
    k.col = rc_lstack(sh, nt) + (threadIdx.x & 255);
    k.cap = s.wstack;
    k.qwalk = RC_WQ - 4;
    return k;
}
// the state's split for this launch: the first min(wstack, ocap) levels in LDS, deeper ones in
// ovf's column of this thread (one column per thread of the grid); ocap 0: all levels in LDS
__device__ __forceinline__ rc_lstk_s rc_lstk_of(float* sh, const rc_dstate& st, const rc_dscene& s) {
    rc_lstk_s k;
    k.col = rc_lstack(sh, st.nt) + (threadIdx.x & 255);
    k.cap = st.ocap > 0 ? min(s.wstack, st.ocap) : s.wstack;
    k.ovf = st.ovf; k.ostride = gridDim.x * blockDim.x;
    k.qwalk = RC_WQ - 4;
    return k;
}

__device__ __forceinline__ bool rc_is_dormant(uint64_t dormant, int l) { return l < 64 && ((dormant >> l) & 1ull); }

// Facing normal and entry of a hit (rc_resolve_hit plus the normal the path continues with).
__device__ __forceinline__ bool rc_resolve_hit_n(const rc_dscene& s, const rc_dstate& st, f3 H, f3 D, int q, int& pi, f3& n) {
    if (q >= 0) {
        pi = rc_point_index(H, q, st);
        n = ld3(st.Q + 12 * q, 3);
        return dot3(n, D) > 0.f;
    }
    const float* n4 = s.STN + 4 * (int64_t)(-2 - q);
    n = mk3(n4[0], n4[1], n4[2]);
    if (dot3(n, D) > 0.f) {
        if (!(__float_as_int(n4[3]) & 1)) { pi = -1; return true; }
        n = n * -1.f;
    }
    pi = rc_hash_point(s, H, n);
    return pi < 0;
}

// ------------------------------------------------------------------------------ host-side launch sizes
static inline int rc_grid(long long n) { return (int)((n + RC_TB - 1) / RC_TB); }
// dynamic LDS of a tracing kernel: dynamic triangles + (wide traversal) the lane stacks
// Whether a wide-traversal launch takes its split-stack variant (the state's first ocap levels
// in LDS, the deeper ones in global memory) rather than the all-LDS one: when more of its blocks
// fit on a CU, or as many and tie_split (the split costs a branch per push and pop and a few
// VGPRs: at equal occupancy the adjoint's forward runs 10 % slower with it, the camera pass 1 %
// faster); decided once per kernel pair and LDS sizes.
static inline bool rc_use_split(const void* ks, const void* kl, size_t shm_s, size_t shm_l, bool tie_split) {
    if (shm_s >= shm_l) return false;
    static std::mutex mu;
    static std::map<std::tuple<const void*, size_t, size_t>, bool> memo;
    const auto key = std::make_tuple(ks, shm_s, shm_l);   // (ks fixes tie_split: one value per kernel)
    std::lock_guard<std::mutex> lock(mu);
    const auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    int bs = 0, bl = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bs, ks, RC_TB, shm_s) != hipSuccess) bs = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bl, kl, RC_TB, shm_l) != hipSuccess) bl = 0;
    return memo[key] = bs > bl || (tie_split && bs == bl && bs > 0);
}
// Launch of a wide-traversal kernel: KS / KL its split-stack and all-LDS instantiations (each in
// parentheses), the split one when the state forces it or rc_use_split picks it (TIE: its
// tie_split); returns
// hipErrorInvalidValue from the enclosing launcher when the split one runs and the state's global
// levels are too small for the grid.
#define RC_LAUNCH_WIDE(KS, KL, TIE, grid, st, s, stream, ...) RC_LAUNCH_WIDE_Q(KS, KL, TIE, RC_WQ, grid, st, s, stream, __VA_ARGS__)
// the same with WQ queue levels per lane
#define RC_LAUNCH_WIDE_Q(KS, KL, TIE, WQ, grid, st, s, stream, ...)                                             \
    do {                                                                                                    \
        const size_t shm_s_ = rc_shm(*(st), *(s), 2, true, WQ), shm_l_ = rc_shm(*(st), *(s), 2, false, WQ); \
        if (((st)->oforce && (st)->ocap > 0) || rc_use_split((const void*)KS, (const void*)KL, shm_s_, shm_l_, TIE)) { \
            if (!rc_ovf_ok(*(st), *(s), 2, (grid).x)) return hipErrorInvalidValue;                          \
            hipLaunchKernelGGL(KS, grid, dim3(RC_TB), shm_s_, stream, __VA_ARGS__);                         \
        } else {                                                                                            \
            hipLaunchKernelGGL(KL, grid, dim3(RC_TB), shm_l_, stream, __VA_ARGS__);                         \
        }                                                                                                   \
    } while (0)

// lane stacks of a kernel: SPLIT true the state's split (rc_lstk_of), false every level in LDS
template <bool SPLIT> struct rc_lstk_sel;
template <> struct rc_lstk_sel<true> {
    static __device__ __forceinline__ rc_lstk_s make(float* sh, const rc_dstate& st, const rc_dscene& s) { return rc_lstk_of(sh, st, s); }
};
template <> struct rc_lstk_sel<false> {
    static __device__ __forceinline__ rc_lstk_l make(float* sh, const rc_dstate& st, const rc_dscene& s) { return rc_lstk_lds(sh, st.nt, s); }
};

// (split: a kernel whose lane stacks follow the state's split, rc_lstk_of; else all levels in LDS)
// wq: the kernel's queue levels (RC_WQ, or RC_WQ_L for a kernel whose lane stacks set qwalk = RC_WQ_L - 4)
static inline size_t rc_shm(const rc_dstate& st, const rc_dscene& s, int trav, bool split = true, int wq = RC_WQ) {
    size_t tri = (((size_t)st.nt * 10 + 3) & ~(size_t)3) * sizeof(float);
    const int lv = split && st.ocap > 0 && st.ocap < s.wstack ? st.ocap : s.wstack;
    return tri + (trav >= 2 ? (size_t)(lv + wq) * 256 * sizeof(uint32_t) : 0);
}
// global stack words a launch of nthreads threads needs (0: every level in LDS)
static inline size_t rc_ovf_need(const rc_dstate& st, const rc_dscene& s, long long nthreads) {
    return st.ocap > 0 && st.ocap < s.wstack ? (size_t)nthreads * (size_t)(s.wstack - st.ocap) : 0;
}
// the split lane stacks address their global levels with 32-bit byte offsets (rc_lstk_s::off): a
// launch whose global levels span more than 4 GiB cannot take them
static inline bool rc_ovf_fits32(size_t words) { return words <= ((size_t)1 << 30); }
// false when the state's global stack levels are too small for a wide-traversal launch of
// nblocks blocks, or too large for 32-bit offsets (the launchers then fail with
// hipErrorInvalidValue instead of running)
static inline bool rc_ovf_ok(const rc_dstate& st, const rc_dscene& s, int trav, long long nblocks) {
    const size_t need = rc_ovf_need(st, s, nblocks * RC_TB);
    return trav != 2 || (need <= (st.ovf ? (size_t)st.ovf_words : 0) && rc_ovf_fits32(need));
}

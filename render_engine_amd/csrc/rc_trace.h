// Tracing helpers shared by the kernel translation units: block size, wave reductions, the
// full (dynamic + static) closest-hit trace per traversal mode, the dynamic-triangle LDS
// preload and the LDS stack layout, and the host-side grid / LDS sizes.
#pragma once
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
// 2 wide closest-hit (LDS stack).
template <int TRAV>
__device__ __forceinline__ void rc_trace_static(f3 o, f3 d, const rc_dscene& s, float& tmin, int& q, uint32_t* lstk) {
    if (TRAV == 0) rc_trace_bvh_ref(o, d, s, tmin, q);
    else if (TRAV == 1) rc_trace_bvh_closest(o, d, s, tmin, q);
    else rc_trace_wide_closest(o, d, s, tmin, q, lstk);
}

// Dynamic triangles from LDS, then the static BVH carrying tmin (cuda/poc_gpu.cu:340-344).
template <int TRAV>
__device__ __forceinline__ bool rc_trace_full(f3 o, f3 d, const float* T, const int* TQ, int nt, const rc_dscene& s, float& tmin, int& q,
                                              uint32_t* lstk) {
    bool h = rc_trace_lds(o, d, T, TQ, nt, tmin, q);
    if (s.nst > 0) { rc_trace_static<TRAV>(o, d, s, tmin, q, lstk); h = q != -1; }
    return h;
}

// Dynamic LDS of the tracing kernels: nt triangles (9 floats) + their quad ids, then the
// traversal stack (wstack levels x 256 lanes) and leaf queue (RC_WQ levels) at a 16-byte aligned offset.
__device__ __forceinline__ void rc_load_dyn(const rc_dstate& st, float* sh, int* shq) {
    for (int j = threadIdx.x; j < 9 * st.nt; j += blockDim.x) sh[j] = st.T[j];
    for (int j = threadIdx.x; j < st.nt; j += blockDim.x) shq[j] = st.TQ[j];
    __syncthreads();
}
__device__ __forceinline__ uint32_t* rc_lstack(float* sh, int nt) { return (uint32_t*)(sh + ((10 * nt + 3) & ~3)); }

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
static inline size_t rc_shm(const rc_dstate& st, const rc_dscene& s, int trav) {
    size_t tri = (((size_t)st.nt * 10 + 3) & ~(size_t)3) * sizeof(float);
    return tri + (trav == 2 ? (size_t)(s.wstack + RC_WQ) * 256 * sizeof(uint32_t) : 0);
}

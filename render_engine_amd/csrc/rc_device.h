// Device primitives of the cache frame: RNG, cosine sampling, ray/triangle and ray/box
// tests, BVH traversal, the 64-bit cell-key hash grid and the transit-grid DDA.
//
// Every function keeps the prototype's operation order so that a scalar CPU build of the
// same arithmetic (compiled with -ffp-contract=off, as this file is) gives identical bits.
// Two deliberate departures from the prototype, both needed for CPU/GPU bit equality:
//   * rsqrtf in cosine_dir (cuda/poc_gpu.cu:363) becomes 1/sqrtf (both correctly rounded);
//   * cosf/sinf (cuda/poc_gpu.cu:365) become rc_sincos below: Cody-Waite reduction by pi/2
//     (two-constant split) and Taylor polynomials of degree 9 (sin) / 10 (cos) on
//     [-pi/4, pi/4], evaluated in a fixed order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rc_types.h"

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 ld3(const float* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// ---------------------------------------------------------------- RNG (cuda/poc_gpu.cu:127-135)
__host__ __device__ __forceinline__ uint32_t rc_hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}
__device__ __forceinline__ float rc_rnd(uint32_t& s) {
    s = s * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    w = (w >> 22u) ^ w;
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}
// per-point / per-path seed chain (cuda/poc_gpu.cu:442)
__device__ __forceinline__ uint32_t rc_seed(uint32_t seed, uint32_t frame, uint32_t i) {
    return rc_hash32(seed * 0x9E3779B9u ^ rc_hash32(frame * 0x85EBCA6Bu ^ rc_hash32(i + 1u)));
}

// ---------------------------------------------------------------- sin / cos
__device__ __forceinline__ void rc_sincos(float x, float& s, float& c) {
    const float TWO_OVER_PI = 0.636619772367581f;
    const float PIO2_HI = 1.5703125f;              // 201/128, exact product with small q
    const float PIO2_LO = 4.83826794897e-4f;       // pi/2 - PIO2_HI
    float qf = floorf(x * TWO_OVER_PI + 0.5f);
    int q = (int)qf;
    float y = (x - qf * PIO2_HI) - qf * PIO2_LO;
    float y2 = y * y;
    float sp = y + y * y2 * (-1.66666667e-1f + y2 * (8.33333333e-3f + y2 * (-1.98412698e-4f + y2 * 2.75573192e-6f)));
    float cp = 1.0f + y2 * (-0.5f + y2 * (4.16666667e-2f + y2 * (-1.38888889e-3f + y2 * (2.48015873e-5f + y2 * -2.75573192e-7f))));
    switch (q & 3) {
        case 0: s = sp; c = cp; break;
        case 1: s = cp; c = -sp; break;
        case 2: s = -sp; c = -cp; break;
        default: s = -cp; c = sp; break;
    }
}

// cosine-weighted hemisphere direction around N (cuda/poc_gpu.cu:360-366)
__device__ __forceinline__ f3 rc_cosine_dir(f3 N, float u1, float u2) {
    float r = sqrtf(u1), phi = 6.28318530718f * u2;
    f3 a = fabsf(N.x) > 0.9f ? mk3(0.f, 1.f, 0.f) : mk3(1.f, 0.f, 0.f);
    f3 T = cross3(N, a);
    T = T * (1.0f / sqrtf(dot3(T, T)));
    f3 B = cross3(N, T);
    float sn, cs;
    rc_sincos(phi, sn, cs);
    return T * (r * cs) + B * (r * sn) + N * sqrtf(fmaxf(0.f, 1.f - u1));
}

// ---------------------------------------------------------------- triangles
// Dynamic triangles held in LDS: closest hit, strict '<' so the first of equal hits wins
// (cuda/poc_gpu.cu:200-220). q = quad id of the hit triangle, -1 if none.
__device__ __forceinline__ bool rc_trace_lds(f3 o, f3 d, const float* T, const int* TQ, int nt, float& tmin, int& q) {
    tmin = __int_as_float(0x7f800000); q = -1;
    for (int j = 0; j < nt; ++j) {
        f3 v0 = mk3(T[9 * j], T[9 * j + 1], T[9 * j + 2]);
        f3 e1 = mk3(T[9 * j + 3], T[9 * j + 4], T[9 * j + 5]);
        f3 e2 = mk3(T[9 * j + 6], T[9 * j + 7], T[9 * j + 8]);
        f3 p = cross3(d, e2);
        float det = dot3(e1, p);
        if (fabsf(det) <= 1e-9f) continue;
        float inv = 1.f / det;
        f3 tv = o - v0;
        float u = dot3(tv, p) * inv;
        if (u < 0.f) continue;
        f3 qv = cross3(tv, e1);
        float v = dot3(d, qv) * inv;
        if (v < 0.f || u + v > 1.f) continue;
        float t = dot3(e2, qv) * inv;
        if (t > 1e-4f && t < tmin) { tmin = t; q = TQ[j]; }
    }
    return q >= 0;
}

// slab entry distance of one child box of a wide node (same arithmetic as rc_aabb_enter)
__device__ __forceinline__ float rc_box_enter(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv, float tmax) {
    float tx1 = (lx - o.x) * inv.x, tx2 = (hx - o.x) * inv.x;
    float ty1 = (ly - o.y) * inv.y, ty2 = (hy - o.y) * inv.y;
    float tz1 = (lz - o.z) * inv.z, tz2 = (hz - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), 0.f));
    float tf = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fminf(fmaxf(tz1, tz2), tmax));
    return tn <= tf ? tn : 1e30f;
}

// Static-triangle test of the BVH leaves (cuda/poc_gpu.cu:305-321), reading the 48-byte
// aligned copy (v0, e1, e2 as three float4: three 16-byte loads). Returns t or -1.
__device__ __forceinline__ float rc_tri_static4(f3 o, f3 d, const float4* __restrict__ T4) {
    float4 a = T4[0], b = T4[1], c = T4[2];
    f3 v0 = mk3(a.x, a.y, a.z), e1 = mk3(b.x, b.y, b.z), e2 = mk3(c.x, c.y, c.z);
    f3 p = cross3(d, e2);
    float det = dot3(e1, p);
    if (fabsf(det) <= 1e-12f) return -1.f;
    float iv = 1.f / det;
    f3 tv = o - v0;
    float u = dot3(tv, p) * iv;
    if (u < 0.f || u > 1.f) return -1.f;
    f3 qv = cross3(tv, e1);
    float v = dot3(d, qv) * iv;
    if (v < 0.f || u + v > 1.f) return -1.f;
    return dot3(e2, qv) * iv;
}

// slab entry distance, 1e30 on a miss (cuda/poc_gpu.cu:288-295)
__device__ __forceinline__ float rc_aabb_enter(const float* lo, const float* hi, f3 o, f3 inv, float tmax) {
    float tx1 = (lo[0] - o.x) * inv.x, tx2 = (hi[0] - o.x) * inv.x;
    float ty1 = (lo[1] - o.y) * inv.y, ty2 = (hi[1] - o.y) * inv.y;
    float tz1 = (lo[2] - o.z) * inv.z, tz2 = (hi[2] - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), 0.f));
    float tf = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fminf(fmaxf(tz1, tz2), tmax));
    return tn <= tf ? tn : 1e30f;
}

__device__ __forceinline__ f3 rc_inv_dir(f3 d) {
    return mk3(fabsf(d.x) > 1e-12f ? 1.f / d.x : copysignf(1e30f, d.x),
               fabsf(d.y) > 1e-12f ? 1.f / d.y : copysignf(1e30f, d.y),
               fabsf(d.z) > 1e-12f ? 1.f / d.z : copysignf(1e30f, d.z));
}

// Binary-BVH traversal in the prototype's order: nearest child first (ties: left), a
// 64-entry stack whose overflowing pushes are dropped, strict '<' on t so the first
// triangle found among equal hits wins. q = -2 - triangle index (cuda/poc_gpu.cu:297-338).
__device__ __forceinline__ void rc_trace_bvh_ref(f3 o, f3 d, const rc_dscene& s, float& tmin, int& q) {
    f3 inv = rc_inv_dir(d);
    int stack[64]; int sp = 0, node = 0;
    while (true) {
        const rc_bvh_node nd = s.BN[node];
        if (nd.cnt > 0) {
            for (int j = nd.lf; j < nd.lf + nd.cnt; ++j) {
                float t = rc_tri_static4(o, d, (const float4*)s.ST4 + 3 * (int64_t)j);
                if (t > 1e-4f && t < tmin) { tmin = t; q = -2 - j; }
            }
        } else {
            const rc_bvh_node L = s.BN[nd.lf], R = s.BN[nd.lf + 1];
            float tl = rc_aabb_enter(L.lo, L.hi, o, inv, tmin), tr = rc_aabb_enter(R.lo, R.hi, o, inv, tmin);
            int first = nd.lf, second = nd.lf + 1;
            if (tr < tl) { float x = tl; tl = tr; tr = x; first = nd.lf + 1; second = nd.lf; }
            if (tl < 1e30f) {
                if (tr < 1e30f && sp < 64) stack[sp++] = second;
                node = first; continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
}

// Order-independent closest hit: smallest t, ties between static triangles go to the lowest
// triangle index, a dynamic-triangle hit keeps its ties with static ones. Any traversal order
// (binary or wide BVH) that never culls a box holding the closest hit returns the same (t, q).
__device__ __forceinline__ bool rc_closer(float t, int j, float tmin, int q) {
    return t > 1e-4f && (t < tmin || (t == tmin && q <= -2 && j < -2 - q));
}

__device__ __forceinline__ void rc_trace_bvh_closest(f3 o, f3 d, const rc_dscene& s, float& tmin, int& q) {
    f3 inv = rc_inv_dir(d);
    int stack[64]; int sp = 0, node = 0;
    while (true) {
        const rc_bvh_node nd = s.BN[node];
        if (nd.cnt > 0) {
            for (int j = nd.lf; j < nd.lf + nd.cnt; ++j) {
                float t = rc_tri_static4(o, d, (const float4*)s.ST4 + 3 * (int64_t)j);
                if (rc_closer(t, j, tmin, q)) { tmin = t; q = -2 - j; }
            }
        } else {
            const rc_bvh_node L = s.BN[nd.lf], R = s.BN[nd.lf + 1];
            float tl = rc_aabb_enter(L.lo, L.hi, o, inv, tmin), tr = rc_aabb_enter(R.lo, R.hi, o, inv, tmin);
            int first = nd.lf, second = nd.lf + 1;
            if (tr < tl) { float x = tl; tl = tr; tr = x; first = nd.lf + 1; second = nd.lf; }
            if (tl < 1e30f) {
                if (tr < 1e30f) stack[sp++] = second;   // depth <= 56 by construction
                node = first; continue;
            }
        }
        // pop, skipping entries whose box now starts beyond the current hit is left to the
        // box test of their parent: popped nodes are re-tested below when they are inner.
        if (sp == 0) break;
        node = stack[--sp];
    }
}

// Compare-exchange of (distance, node) pairs by selects (no branch: five VALU operations).
__device__ __forceinline__ void rc_cswap(float& ta, int& ia, float& tb, int& ib) {
    const bool c = tb < ta;
    const float lo = c ? tb : ta, hi = c ? ta : tb;
    const int il = c ? ib : ia, ih = c ? ia : ib;
    ta = lo; tb = hi; ia = il; ib = ih;
}

// Stack entries of the wide traversal are 32-bit: node index << 8 | an 8-bit code of the
// node's entry distance t. The code keeps t's exponent and top 3 mantissa bits (t >= 2^-15),
// and decodes to a lower bound of t (0 for code 0), so culling a popped node whose bound
// exceeds the current hit never drops a node the exact distance would have kept.
// Only called with t >= 0 (+0 or -0 included): the field extract drops the sign bit, so both
// zeros give code 0; then one subtract and one clamp.
__device__ __forceinline__ uint32_t rc_tcode(float t) {
    const int c = (int)__builtin_amdgcn_ubfe(__float_as_uint(t), 20, 11) - (112 << 3) + 1;
    return (uint32_t)min(max(c, 0), 255);
}
__device__ __forceinline__ float rc_tlb(uint32_t code) {
    return code == 0u ? 0.f : __uint_as_float(((code - 1u) + (112u << 3)) << 20);
}

// Per-ray constants of the wide traversal: origin, direction, reciprocal direction and the
// byte offsets, inside a 128-byte node, of the rows that hold each axis's near planes (near = lo
// when the direction is positive, hi when negative; node rows of 16 B: 0-2 lo x/y/z, 3-5 hi
// x/y/z, 6 children and leaf entries, 7 leaf ends). The far row of axis a is at (48 + 32 a) - near offset:
// node - near + constant, the constant going into the load's immediate offset (3 fewer VGPRs).
struct rc_wray { f3 o, d, inv; uint32_t nx, ny, nz; };
__device__ __forceinline__ rc_wray rc_wray_make(f3 o, f3 d) {
    rc_wray r;
    r.o = o; r.d = d; r.inv = rc_inv_dir(d);
    r.nx = r.inv.x < 0.f ? 48u : 0u; r.ny = r.inv.y < 0.f ? 64u : 16u; r.nz = r.inv.z < 0.f ? 80u : 32u;
    return r;
}

// One 16-byte row of a wide node: the scene's node array (uniform) plus a 32-bit byte offset,
// so the load takes the scalar-base + 32-bit vector-offset form (no 64-bit address math).
__device__ __forceinline__ float4 rc_wrow(const rc_dscene& s, uint32_t off) {
    return *(const float4*)((const char*)s.WN + off);
}
__device__ __forceinline__ int4 rc_wrow_i(const rc_dscene& s, uint32_t off) {
    return *(const int4*)((const char*)s.WN + off);
}

// Entry distances of a node's four child boxes against [0, tmax] (1e30: missed, empty slot or
// leaf), branch-free, and the leaf entries e of the entered leaf children (1 << 31 | first << 7 |
// count; 0 otherwise). With lo <= hi, (lo - o) * inv and (hi - o) * inv are ordered by the sign
// of inv, so the near-plane value is exactly the per-axis minimum of rc_box_enter and the result
// equals rc_box_enter's bit for bit. Empty slots hold lo = +inf, hi = -inf: their near value is
// +inf and their far value -inf whatever the direction, so they always miss.
__device__ __forceinline__ void rc_child_enter(const rc_wray& r, const rc_dscene& s, uint32_t node, float tmax, int4 ch,
                                               float& t0, float& t1, float& t2, float& t3, int& e0, int& e1, int& e2, int& e3) {
    const float4 nx = rc_wrow(s, node + r.nx), ny = rc_wrow(s, node + r.ny), nz = rc_wrow(s, node + r.nz);
    const float4 fx = rc_wrow(s, (node - r.nx) + 48u), fy = rc_wrow(s, (node - r.ny) + 80u), fz = rc_wrow(s, (node - r.nz) + 112u);
#define RC_SLAB(tc, c, e)                                                                                       \
    {                                                                                                           \
        const float tn = fmaxf(fmaxf((nx.c - r.o.x) * r.inv.x, (ny.c - r.o.y) * r.inv.y), fmaxf((nz.c - r.o.z) * r.inv.z, 0.f)); \
        const float tf = fminf(fminf((fx.c - r.o.x) * r.inv.x, (fy.c - r.o.y) * r.inv.y), fminf((fz.c - r.o.z) * r.inv.z, tmax)); \
        const bool in_ = tn <= tf, leaf_ = ch.c < 0;                                                            \
        tc = in_ && !leaf_ ? tn : 1e30f;                                                                        \
        e = in_ && leaf_ ? ch.c : 0;                                                                            \
    }
    RC_SLAB(t0, x, e0) RC_SLAB(t1, y, e1) RC_SLAB(t2, z, e2) RC_SLAB(t3, w, e3)
#undef RC_SLAB
}

// The same for a quantized node (rc_qnode, byte offset node into QN): per axis the base
// (o - ray origin) * inv and the step s * inv (exact: s is a power of two), the near and far
// plane words picked by the direction's sign, then per child and plane one byte conversion and
// one fma. The interval contains the fp32 box's (one step of slack on every side), so a box the
// fp32 test enters is entered here too.
__device__ __forceinline__ float4 rc_qrow(const rc_dscene& s, uint32_t off) {
    return *(const float4*)((const char*)s.QN + off);
}
__device__ __forceinline__ int4 rc_qrow_i(const rc_dscene& s, uint32_t off) {
    return *(const int4*)((const char*)s.QN + off);
}
__device__ __forceinline__ float rc_u8f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 255u); }
__device__ __forceinline__ void rc_child_enter_q(const rc_wray& r, const rc_dscene& s, uint32_t node, float tmax, int4 ch,
                                                 float& t0, float& t1, float& t2, float& t3, int& e0, int& e1, int& e2, int& e3) {
    const float4 a = rc_qrow(s, node), b = rc_qrow(s, node + 16u), c = rc_qrow(s, node + 32u);
    const float bx = (a.x - r.o.x) * r.inv.x, by = (a.y - r.o.y) * r.inv.y, bz = (a.z - r.o.z) * r.inv.z;
    const float sx = a.w * r.inv.x, sy = b.x * r.inv.y, sz = b.y * r.inv.z;
    const uint32_t lx = __float_as_uint(b.z), ly = __float_as_uint(b.w), lz = __float_as_uint(c.x);
    const uint32_t hx = __float_as_uint(c.y), hy = __float_as_uint(c.z), hz = __float_as_uint(c.w);
    const bool px = r.nx == 0u, py = r.ny == 16u, pz = r.nz == 32u;   // direction component >= 0
    const uint32_t nx = px ? lx : hx, fx = px ? hx : lx, ny = py ? ly : hy, fy = py ? hy : ly, nz = pz ? lz : hz, fz = pz ? hz : lz;
#define RC_QSLAB(tc, k, cc, e)                                                                                       \
    {                                                                                                               \
        const float tn = fmaxf(fmaxf(fmaf(rc_u8f(nx, k), sx, bx), fmaf(rc_u8f(ny, k), sy, by)), fmaxf(fmaf(rc_u8f(nz, k), sz, bz), 0.f)); \
        const float tf = fminf(fminf(fmaf(rc_u8f(fx, k), sx, bx), fmaf(rc_u8f(fy, k), sy, by)), fminf(fmaf(rc_u8f(fz, k), sz, bz), tmax)); \
        const bool in_ = tn <= tf, leaf_ = ch.cc < 0;                                                               \
        tc = in_ && !leaf_ ? tn : 1e30f;                                                                            \
        e = in_ && leaf_ ? ch.cc : 0;                                                                               \
    }
    RC_QSLAB(t0, 0, x, e0) RC_QSLAB(t1, 1, y, e1) RC_QSLAB(t2, 2, z, e2) RC_QSLAB(t3, 3, w, e3)
#undef RC_QSLAB
}

// four floats at a 4-byte aligned address, read by one 16-byte load
typedef float rc_f4u __attribute__((ext_vector_type(4), aligned(4)));

// Lane stack and leaf queue of the wide traversal: LDS levels one column per lane (256 entries
// apart), the queue levels (RC_WQ, or RC_WQ_L for the kernels launched with them) after the first
// cap levels. rc_lstk_l keeps every stack level in LDS (cap = wstack); rc_lstk_s keeps the first cap levels in LDS and the deeper ones in global
// memory (uniform base < 4 GiB; level cap + k of thread g at word k * ostride + g, g = the
// thread's index in the grid), which only a ray nested deeper than cap touches. The stack never
// holds more than wstack entries (at most 3 pushes per wide level).
struct rc_lstk_l {
    uint32_t* col;        // this lane's LDS column
    int cap;
    int qwalk;            // a lane holding more queued leaf ranges than this starts the wave's leaf walk (queue levels - 4)
    __device__ __forceinline__ void push(int sp, uint32_t e) const { col[sp * 256] = e; }
    __device__ __forceinline__ uint32_t pop(int sp) const { return col[sp * 256]; }
    __device__ __forceinline__ uint32_t* queue() const { return col + (size_t)cap * 256; }
};
struct rc_lstk_s {
    uint32_t* col;
    uint32_t* ovf;
    uint32_t ostride;
    int cap;
    int qwalk;            // as rc_lstk_l::qwalk
    // byte offset of global level sp (32-bit: a scalar base plus one vector offset per access)
    __device__ __forceinline__ uint32_t off(int sp) const {
        return ((uint32_t)(sp - cap) * ostride + blockIdx.x * blockDim.x + threadIdx.x) * 4u;
    }
    __device__ __forceinline__ void push(int sp, uint32_t e) const {
        if (sp < cap) col[sp * 256] = e;
        else *(uint32_t*)((char*)ovf + off(sp)) = e;
    }
    __device__ __forceinline__ uint32_t pop(int sp) const {
        // the global read names the global address space: otherwise the compiler merges the two
        // reads into one generic (flat) load through a select of addresses, slower than the LDS
        // read taken almost always
        typedef const __attribute__((address_space(1))) uint32_t gu32;
        uint32_t e;
        if (sp < cap) e = col[sp * 256];
        else e = *(gu32*)((const __attribute__((address_space(1))) char*)ovf + off(sp));
        return e;
    }
    __device__ __forceinline__ uint32_t* queue() const { return col + (size_t)cap * 256; }
};

// 4-wide BVH closest hit of the static geometry, order-independent rule, with postponed
// leaves. A node step tests the node's four child boxes, queues the triangle ranges of the
// entered leaf children (one LDS column per lane, RC_WQ or RC_WQ_L levels after the stack; entries
// first << 7 | count) and goes on with the nearest entered inner child, the others pushed
// far-to-near (LDS stack, 32-bit entries node << 8 | code(t)), or with a pop. The wave walks
// the queued triangles of all its lanes together once every lane still traversing has some
// queued (or one lane holds more than lstk.qwalk, the queue levels - 4, so that the next node step's
// four children still fit), so the triangle loop runs with most lanes busy
// instead of after each node step for the few lanes that entered a leaf. Culling uses the
// current tmin, which may lag the queued hits: more node steps at worst, never a missed hit,
// and the hit itself does not depend on the order. Leaf children hold at most RC_WLEAF_MAX
// triangles (the wide collapse splits longer binary leaves), so a node step only queues them: no
// triangle loop inside the node step, whose live values would raise the register pressure of the
// whole traversal. The stack depth is bounded by the scene's wstack levels (3 per wide level); the
// levels beyond lstk.cap live in global memory.
// QN: the node steps read the quantized nodes (rc_qnode) instead of the fp32 ones.
template <bool QN = false, class LS>
__device__ __forceinline__ void rc_trace_wide_closest(f3 o, f3 d, const rc_dscene& s, float& tmin, int& q, const LS& lstk) {
    uint32_t* que = lstk.queue();
    const rc_wray r = rc_wray_make(o, d);
    int sp = 0, cur = 0, nq = 0;   // cur < 0: this lane's traversal is over
    for (;;) {
        const bool trav = cur >= 0, pend = nq > 0;
        if (!__any(trav || pend)) break;
        if (__any(pend) && (__all(pend || !trav) || __any(nq > lstk.qwalk))) {
            int k = 0, a = 0, b = 0;
            for (;;) {
                if (a >= b) {
                    if (k >= nq) break;
                    const uint32_t e = que[k * 256];
                    ++k;
                    a = (int)(e >> 7); b = a + (int)(e & 127u);
                }
                const int j = a++;
                const float t = rc_tri_static4(r.o, r.d, (const float4*)s.ST4 + 3 * (int64_t)j);
                if (rc_closer(t, j, tmin, q)) { tmin = t; q = -2 - j; }
            }
            nq = 0;
            continue;
        }
        if (!trav) continue;
        const uint32_t node = QN ? (uint32_t)cur << 6 : (uint32_t)cur << 7;
        const int4 ch = QN ? rc_qrow_i(s, node + 48u) : rc_wrow_i(s, node + 96u);
        float t0, t1, t2, t3;
        int e0, e1, e2, e3;
        if (QN) rc_child_enter_q(r, s, node, tmin, ch, t0, t1, t2, t3, e0, e1, e2, e3);
        else rc_child_enter(r, s, node, tmin, ch, t0, t1, t2, t3, e0, e1, e2, e3);
        int i0 = ch.x, i1 = ch.y, i2 = ch.z, i3 = ch.w;
#define RC_QUEUE(e) \
        if (e != 0) { que[nq * 256] = (uint32_t)e & 0x7fffffffu; ++nq; }
        RC_QUEUE(e0) RC_QUEUE(e1) RC_QUEUE(e2) RC_QUEUE(e3)
#undef RC_QUEUE
        // sorting network on (t, node): ascending entry distance, misses (1e30) last
        rc_cswap(t0, i0, t1, i1); rc_cswap(t2, i2, t3, i3);
        rc_cswap(t0, i0, t2, i2); rc_cswap(t1, i1, t3, i3);
        rc_cswap(t1, i1, t2, i2);
        // the children still in front of the current hit are a prefix of the sorted list: push
        // all but the nearest far-to-near, visit the nearest next
#define RC_PUSH(tc, ic) \
        if (tc < 1e30f && tc <= tmin) { lstk.push(sp, ((uint32_t)ic << 8) | rc_tcode(tc)); ++sp; }
        RC_PUSH(t3, i3) RC_PUSH(t2, i2) RC_PUSH(t1, i1)
#undef RC_PUSH
        if (t0 < 1e30f && t0 <= tmin) { cur = i0; continue; }
        cur = -1;
        // pop culling on the codes: rc_tlb(c) <= tmin exactly when c <= rc_tcode(tmin) (tmin >= 0; the
        // code is monotone and rc_tlb(c) is the least value with code c), so no decode per entry
        const uint32_t ct = rc_tcode(tmin);
        while (sp > 0) {
            --sp;
            const uint32_t e = lstk.pop(sp);
            if ((e & 255u) <= ct) { cur = (int)(e >> 8); break; }
        }
    }
}

// Any-hit test of the static geometry for a shadow segment: a hit at 1e-4 < t < tlim blocks
// it (equivalent to the closest-hit test "closest t >= tlim" of the direct-light estimate, but
// stops at the first blocker). Boxes are culled against tlim; inner children are visited
// near-first, so blockers next to the shading point, the likely ones, are found early. Leaves
// are postponed as in the closest-hit walk (queued per lane, walked by the wave together); a
// lane stops at its first blocker. The answer does not depend on the order.
template <bool QN = false, class LS>
__device__ __forceinline__ bool rc_occluded_wide(f3 o, f3 d, const rc_dscene& s, float tlim, const LS& lstk) {
    uint32_t* que = lstk.queue();
    const rc_wray r = rc_wray_make(o, d);
    int sp = 0, cur = 0, nq = 0;   // cur < 0: this lane's traversal is over
    bool blocked = false;
    for (;;) {
        const bool trav = cur >= 0, pend = nq > 0;
        if (!__any(trav || pend)) break;
        if (__any(pend) && (__all(pend || !trav) || __any(nq > lstk.qwalk))) {
            int k = 0, a = 0, b = 0;
            for (;;) {
                if (a >= b) {
                    if (k >= nq) break;
                    const uint32_t e = que[k * 256];
                    ++k;
                    a = (int)(e >> 7); b = a + (int)(e & 127u);
                }
                const int j = a++;
                const float t = rc_tri_static4(r.o, r.d, (const float4*)s.ST4 + 3 * (int64_t)j);
                if (t > 1e-4f && t < tlim) { blocked = true; break; }
            }
            nq = 0;
            if (blocked) cur = -1;
            continue;
        }
        if (!trav) continue;
        const uint32_t node = QN ? (uint32_t)cur << 6 : (uint32_t)cur << 7;
        const int4 ch = QN ? rc_qrow_i(s, node + 48u) : rc_wrow_i(s, node + 96u);
        float t0, t1, t2, t3;
        int e0, e1, e2, e3;
        if (QN) rc_child_enter_q(r, s, node, tlim, ch, t0, t1, t2, t3, e0, e1, e2, e3);
        else rc_child_enter(r, s, node, tlim, ch, t0, t1, t2, t3, e0, e1, e2, e3);
        int i0 = ch.x, i1 = ch.y, i2 = ch.z, i3 = ch.w;
#define RC_QUEUE(e) \
        if (e != 0) { que[nq * 256] = (uint32_t)e & 0x7fffffffu; ++nq; }
        RC_QUEUE(e0) RC_QUEUE(e1) RC_QUEUE(e2) RC_QUEUE(e3)
#undef RC_QUEUE
        rc_cswap(t0, i0, t1, i1); rc_cswap(t2, i2, t3, i3);
        rc_cswap(t0, i0, t2, i2); rc_cswap(t1, i1, t3, i3);
        rc_cswap(t1, i1, t2, i2);
        // entered children are a prefix of the sorted list: push all but the nearest far-to-near
#define RC_ANYPUSH(tc, ic) \
        if (tc < 1e30f) { lstk.push(sp, (uint32_t)ic << 8); ++sp; }
        RC_ANYPUSH(t3, i3) RC_ANYPUSH(t2, i2) RC_ANYPUSH(t1, i1)
#undef RC_ANYPUSH
        if (t0 < 1e30f) { cur = i0; continue; }
        if (sp == 0) { cur = -1; continue; }
        --sp;
        cur = (int)(lstk.pop(sp) >> 8);
    }
    return blocked;
}

// the same test against the LDS-held dynamic triangles (cuda/poc_gpu.cu:200-220 arithmetic)
__device__ __forceinline__ bool rc_occluded_lds(f3 o, f3 d, const float* T, int nt, float tlim) {
    for (int j = 0; j < nt; ++j) {
        f3 v0 = mk3(T[9 * j], T[9 * j + 1], T[9 * j + 2]);
        f3 e1 = mk3(T[9 * j + 3], T[9 * j + 4], T[9 * j + 5]);
        f3 e2 = mk3(T[9 * j + 6], T[9 * j + 7], T[9 * j + 8]);
        f3 p = cross3(d, e2);
        float det = dot3(e1, p);
        if (fabsf(det) <= 1e-9f) continue;
        float inv = 1.f / det;
        f3 tv = o - v0;
        float u = dot3(tv, p) * inv;
        if (u < 0.f) continue;
        f3 qv = cross3(tv, e1);
        float v = dot3(d, qv) * inv;
        if (v < 0.f || u + v > 1.f) continue;
        float t = dot3(e2, qv) * inv;
        if (t > 1e-4f && t < tlim) return true;
    }
    return false;
}

// ---------------------------------------------------------------- hash grid (cuda/poc_gpu.cu:251-287)
__device__ __forceinline__ int rc_normal_bucket(float nx, float ny, float nz) {
    float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
    if (ax >= ay && ax >= az) return nx > 0.f ? 0 : 1;
    if (ay >= az) return ny > 0.f ? 2 : 3;
    return nz > 0.f ? 4 : 5;
}
__device__ __forceinline__ uint64_t rc_cell_key(int cx, int cy, int cz, int b) {
    return ((((uint64_t)(cx & 0xFFFFF) << 20 | (uint64_t)(cy & 0xFFFFF)) << 20 | (uint64_t)(cz & 0xFFFFF)) << 3 | (uint64_t)b) + 1ull;
}
__device__ __forceinline__ uint64_t rc_mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull; x ^= x >> 27; x *= 0x94d049bb133111ebull; x ^= x >> 31; return x;
}
__device__ __forceinline__ int rc_hash_find(const uint64_t* __restrict__ HK, const int* __restrict__ HV, uint32_t hmask, uint64_t key) {
    uint32_t h = (uint32_t)rc_mix64(key) & hmask;
    for (int probe = 0; probe < 64; ++probe) {
        uint64_t k = HK[h];
        if (k == key) return HV[h];
        if (k == 0ull) return -1;
        h = (h + 1) & hmask;
    }
    return -1;
}
__device__ __forceinline__ int rc_hash_point(const rc_dscene& s, f3 H, f3 n) {
    int b = rc_normal_bucket(n.x, n.y, n.z);
    float ih = 1.f / s.hcell;
    int cx = (int)floorf((H.x - s.lo[0]) * ih), cy = (int)floorf((H.y - s.lo[1]) * ih), cz = (int)floorf((H.z - s.lo[2]) * ih);
    int p = rc_hash_find(s.HK, s.HV, s.hmask, rc_cell_key(cx, cy, cz, b));
    if (p >= 0) return p;
    for (int dd = 0; dd < 27; ++dd) {
        int dz = dd / 9 - 1, dy = (dd / 3) % 3 - 1, dx = dd % 3 - 1;
        if (dx == 0 && dy == 0 && dz == 0) continue;
        p = rc_hash_find(s.HK, s.HV, s.hmask, rc_cell_key(cx + dx, cy + dy, cz + dz, b));
        if (p >= 0) return p;
    }
    return -1;
}

// cache point of a quad hit (cuda/poc_gpu.cu:239-246)
__device__ __forceinline__ int rc_point_index(f3 H, int q, const rc_dstate& st) {
    f3 o = ld3(st.Q + 12 * q, 0), u = ld3(st.Q + 12 * q, 1), v = ld3(st.Q + 12 * q, 2);
    f3 dd = H - o;
    float a = dot3(dd, u) / dot3(u, u), b = dot3(dd, v) / dot3(v, v);
    int nu = st.dims[2 * q], nv = st.dims[2 * q + 1];
    int iu = min(max((int)(a * nu), 0), nu - 1), iv = min(max((int)(b * nv), 0), nv - 1);
    return st.off[q] + iv * nu + iu;
}

// hit -> cache entry; true when the hit does not count (cuda/poc_gpu.cu:346-358)
__device__ __forceinline__ bool rc_resolve_hit(const rc_dscene& s, const rc_dstate& st, bool hit, f3 H, f3 D, int q, int i, int& pi) {
    if (!hit) { pi = i; return true; }
    if (q >= 0) { pi = rc_point_index(H, q, st); return dot3(ld3(st.Q + 12 * q, 3), D) > 0.f; }
    const float* n4 = s.STN + 4 * (int64_t)(-2 - q);
    f3 n = mk3(n4[0], n4[1], n4[2]);
    if (dot3(n, D) > 0.f) {
        if (!(__float_as_int(n4[3]) & 1)) { pi = i; return true; }
        n = n * -1.f;
    }
    pi = rc_hash_point(s, H, n);
    if (pi < 0) { pi = i; return true; }
    return false;
}

// ---------------------------------------------------------------- transit grid (cuda/poc_gpu.cu:137-197)
__device__ __forceinline__ int rc_vox_of(const rc_dscene& s, f3 p) {
    int i = min(max((int)((p.x - s.lo[0]) / s.ext[0] * s.vx), 0), s.vx - 1);
    int j = min(max((int)((p.y - s.lo[1]) / s.ext[1] * s.vy), 0), s.vy - 1);
    int k = min(max((int)((p.z - s.lo[2]) / s.ext[2] * s.vz), 0), s.vz - 1);
    return (k * s.vy + j) * s.vx + i;
}
__device__ __forceinline__ void rc_set_bit(uint32_t (&m)[RC_MAX_NW], int v) {
    uint32_t bit = 1u << (v & 31); int w0 = v >> 5;
#pragma unroll
    for (int w = 0; w < RC_MAX_NW; ++w) m[w] |= (w == w0) ? bit : 0u;
}
// Amanatides-Woo walk of segment a->b over the vx*vy*vz transit grid, OR-ing the voxels
// crossed into m (cuda/poc_gpu.cu:171-190).
__device__ __forceinline__ void rc_dda_reg(const rc_dscene& s, f3 a, f3 b, uint32_t (&m)[RC_MAX_NW]) {
    const int VX = s.vx, VY = s.vy, VZ = s.vz;
    const float sx = VX / s.ext[0], sy = VY / s.ext[1], sz = VZ / s.ext[2];
    float ax = (a.x - s.lo[0]) * sx, ay = (a.y - s.lo[1]) * sy, az = (a.z - s.lo[2]) * sz;
    float bx = (b.x - s.lo[0]) * sx, by = (b.y - s.lo[1]) * sy, bz = (b.z - s.lo[2]) * sz;
    int ix = min(max((int)ax, 0), VX - 1), iy = min(max((int)ay, 0), VY - 1), iz = min(max((int)az, 0), VZ - 1);
    int ex = min(max((int)bx, 0), VX - 1), ey = min(max((int)by, 0), VY - 1), ez = min(max((int)bz, 0), VZ - 1);
    float dx = bx - ax, dy = by - ay, dz = bz - az;
    int stx = dx > 0 ? 1 : -1, sty = dy > 0 ? 1 : -1, stz = dz > 0 ? 1 : -1;
    float tdx = dx != 0 ? fabsf(1.f / dx) : 1e30f, tdy = dy != 0 ? fabsf(1.f / dy) : 1e30f, tdz = dz != 0 ? fabsf(1.f / dz) : 1e30f;
    float tmx = dx != 0 ? ((dx > 0 ? ((float)(ix + 1) - ax) : (ax - (float)ix)) * tdx) : 1e30f;
    float tmy = dy != 0 ? ((dy > 0 ? ((float)(iy + 1) - ay) : (ay - (float)iy)) * tdy) : 1e30f;
    float tmz = dz != 0 ? ((dz > 0 ? ((float)(iz + 1) - az) : (az - (float)iz)) * tdz) : 1e30f;
    for (int g = 0; g < VX + VY + VZ; ++g) {
        rc_set_bit(m, (iz * VY + iy) * VX + ix);
        if (ix == ex && iy == ey && iz == ez) break;
        if (tmx < tmy && tmx < tmz) { ix += stx; tmx += tdx; if (ix < 0 || ix >= VX) break; }
        else if (tmy < tmz) { iy += sty; tmy += tdy; if (iy < 0 || iy >= VY) break; }
        else { iz += stz; tmz += tdz; if (iz < 0 || iz >= VZ) break; }
    }
}
// 20 point samples along the segment (cuda/poc_gpu.cu:192-197)
__device__ __forceinline__ void rc_rast_seg(const rc_dscene& s, f3 a, f3 b, uint32_t (&m)[RC_MAX_NW]) {
    for (int k = 0; k < 20; ++k) rc_set_bit(m, rc_vox_of(s, a + (b - a) * ((float)k / 19.f)));
}

// Amanatides-Woo walk of segment a->b over a g^3 voxel grid on the scene bounds (the stepping rule
// of rc_dda_reg), setting each crossed voxel's bit in V. Bits are gathered per 32-bit word while the
// walk stays in it (x steps within a word), then the word costs one load and, for bits not yet set,
// one atomicOr.
__device__ __forceinline__ void rc_dda_volume_mark(const rc_dscene& s, int gx, int gy, int gz, f3 a, f3 b, uint32_t* V) {
    const float sx = gx / s.ext[0], sy = gy / s.ext[1], sz = gz / s.ext[2];
    float ax = (a.x - s.lo[0]) * sx, ay = (a.y - s.lo[1]) * sy, az = (a.z - s.lo[2]) * sz;
    float bx = (b.x - s.lo[0]) * sx, by = (b.y - s.lo[1]) * sy, bz = (b.z - s.lo[2]) * sz;
    int ix = min(max((int)ax, 0), gx - 1), iy = min(max((int)ay, 0), gy - 1), iz = min(max((int)az, 0), gz - 1);
    int ex = min(max((int)bx, 0), gx - 1), ey = min(max((int)by, 0), gy - 1), ez = min(max((int)bz, 0), gz - 1);
    float dx = bx - ax, dy = by - ay, dz = bz - az;
    int stx = dx > 0 ? 1 : -1, sty = dy > 0 ? 1 : -1, stz = dz > 0 ? 1 : -1;
    float tdx = dx != 0 ? fabsf(1.f / dx) : 1e30f, tdy = dy != 0 ? fabsf(1.f / dy) : 1e30f, tdz = dz != 0 ? fabsf(1.f / dz) : 1e30f;
    float tmx = dx != 0 ? ((dx > 0 ? ((float)(ix + 1) - ax) : (ax - (float)ix)) * tdx) : 1e30f;
    float tmy = dy != 0 ? ((dy > 0 ? ((float)(iy + 1) - ay) : (ay - (float)iy)) * tdy) : 1e30f;
    float tmz = dz != 0 ? ((dz > 0 ? ((float)(iz + 1) - az) : (az - (float)iz)) * tdz) : 1e30f;
    int64_t cw = -1;     // word whose bits are being gathered
    uint32_t acc = 0u;
    for (int g = 0; g < gx + gy + gz; ++g) {
        const int64_t v = ((int64_t)iz * gy + iy) * gx + ix;
        if ((v >> 5) != cw) {
            if (cw >= 0 && (V[cw] & acc) != acc) atomicOr(V + cw, acc);
            cw = v >> 5; acc = 0u;
        }
        acc |= 1u << (v & 31);
        if (ix == ex && iy == ey && iz == ez) break;
        if (tmx < tmy && tmx < tmz) { ix += stx; tmx += tdx; if (ix < 0 || ix >= gx) break; }
        else if (tmy < tmz) { iy += sty; tmy += tdy; if (iy < 0 || iy >= gy) break; }
        else { iz += stz; tmz += tdz; if (iz < 0 || iz >= gz) break; }
    }
    if ((V[cw] & acc) != acc) atomicOr(V + cw, acc);
}

// Scene preparation on the GPU: the cache cells of a mesh (cuda/prep_mesh.cuh:129-157).
//
// The same cells as the host's sample_cells, bit for bit: every triangle draws its samples
// from its own host RNG chain (one thread per triangle, the chain is sequential), the samples
// are sorted by cell key with a stable radix sort (so the samples of a cell stay in the host's
// triangle / sample / side order), each cell is reduced by one thread in that order (the
// nearest-to-centre representative with the host's strict '<', the double-precision albedo
// and area sums in the host's addition order), and the cells are ordered by Morton code.
#include <hipcub/hipcub.hpp>

#include <numeric>

#include "rc_bvh.h"
#include "rc_device.h"
#include "rc_kernels.h"
#include "rc_scene.h"

namespace {

#define PK(x)                                                                                 \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) return std::string(#x ": ") + hipGetErrorString(e_);            \
    } while (0)

__device__ __forceinline__ uint64_t pr_spread3(uint64_t x) {   // 21 bits -> every third bit
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// triangle geometry of the host sampler: edges, unit normal, area, sample count (0: degenerate)
struct PrTri { float e1[3], e2[3], n[3], ar; int ns; };
__device__ __forceinline__ PrTri pr_tri(const float* V, int t, float h) {
    PrTri g;
    const float* v = V + 9 * (size_t)t;
    for (int a = 0; a < 3; ++a) { g.e1[a] = v[3 + a] - v[a]; g.e2[a] = v[6 + a] - v[a]; }
    const float cr[3] = {g.e1[1] * g.e2[2] - g.e1[2] * g.e2[1], g.e1[2] * g.e2[0] - g.e1[0] * g.e2[2],
                         g.e1[0] * g.e2[1] - g.e1[1] * g.e2[0]};
    const float l = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    g.ns = 0; g.ar = 0.f;
    if (l < 1e-12f) return g;
    g.ar = 0.5f * l;
    for (int a = 0; a < 3; ++a) g.n[a] = cr[a] / l;
    g.ns = (int)fmin(1e6, fmax(1.0, ceil((double)(g.ar / (h * h)) * 4.0)));
    return g;
}

__global__ void __launch_bounds__(256) k_cells_count(const float* V, const uint8_t* ds, int nt, float h, unsigned long long* cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const PrTri g = pr_tri(V, t, h);
    cnt[t] = (unsigned long long)g.ns * (ds[t] ? 2ull : 1ull);
}

// one thread per triangle: its samples in the host's order (sample k, then side), element
// g = off[t] + 2 k + side (or off[t] + k when single-sided)
__global__ void __launch_bounds__(256) k_cells_emit(const float* V, const uint8_t* ds, const float* alb, const int* mat, int nt, float h,
                                                   float lo0, float lo1, float lo2, const unsigned long long* off,
                                                   unsigned long long* key, unsigned* idx, float4* Pd, float4* Nn, double2* W, int* M) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const PrTri g = pr_tri(V, t, h);
    if (g.ns == 0) return;
    const float* v = V + 9 * (size_t)t;
    const float lo[3] = {lo0, lo1, lo2};
    const int sides = ds[t] ? 2 : 1;
    const double wa = (double)alb[t] * (double)g.ar / (double)g.ns, wr = (double)g.ar / (double)g.ns;
    const int m = mat ? mat[t] : -1;
    uint64_t s = rc_mix64((uint64_t)t * 0x9E3779B97F4A7C15ull + 1ull);
    unsigned long long e = off[t];
    for (int k = 0; k < g.ns; ++k) {
        float a = 1.f / 3.f, b = 1.f / 3.f;
        if (k > 0) {
            s = rc_mix64(s + 0x9E3779B97F4A7C15ull); a = (float)(s >> 40) * (1.0f / 16777216.0f);
            s = rc_mix64(s + 0x9E3779B97F4A7C15ull); b = (float)(s >> 40) * (1.0f / 16777216.0f);
            if (a + b > 1.f) { a = 1.f - a; b = 1.f - b; }
        }
        float P[3], d2 = 0.f;
        int c[3];
        for (int x = 0; x < 3; ++x) {
            P[x] = v[x] + g.e1[x] * a + g.e2[x] * b;
            c[x] = (int)floorf((P[x] - lo[x]) / h);
            const float cc = lo[x] + ((float)c[x] + 0.5f) * h;
            d2 += (P[x] - cc) * (P[x] - cc);
        }
        for (int sd = 0; sd < sides; ++sd, ++e) {
            const float nn[3] = {sd ? -g.n[0] : g.n[0], sd ? -g.n[1] : g.n[1], sd ? -g.n[2] : g.n[2]};
            key[e] = rc_cell_key(c[0], c[1], c[2], rc_normal_bucket(nn[0], nn[1], nn[2]));
            idx[e] = (unsigned)e;
            Pd[e] = make_float4(P[0], P[1], P[2], d2);
            Nn[e] = make_float4(nn[0], nn[1], nn[2], 0.f);
            W[e] = make_double2(wa, wr);
            M[e] = m;
        }
    }
}

__global__ void __launch_bounds__(256) k_cells_heads(const unsigned long long* skey, unsigned n, unsigned* head) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || skey[i] != skey[i - 1]) ? 1u : 0u;
}
__global__ void __launch_bounds__(256) k_cells_starts(const unsigned* head, const unsigned* cid, unsigned n, unsigned* start) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && head[i]) start[cid[i] - 1] = i;
    if (i == n - 1) start[cid[i]] = n;
}

// one thread per cell, its samples in the host's encounter order
__global__ void __launch_bounds__(256) k_cells_reduce(unsigned C, const unsigned* start, const unsigned long long* skey, const unsigned* sidx,
                                                     const float4* Pd, const float4* Nn, const double2* W, const int* M, float4* cp,
                                                     float4* cn, double2* cw, int* cm, unsigned long long* ckey, unsigned long long* mk,
                                                     unsigned* cidx) {
    const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float4 p = make_float4(0.f, 0.f, 0.f, 1e30f), n = make_float4(0.f, 0.f, 1.f, 0.f);
    double as = 0.0, ws = 0.0;
    int m = -1;
    for (unsigned i = start[c]; i < start[c + 1]; ++i) {
        const unsigned g = sidx[i];
        const float4 q = Pd[g];
        if (q.w < p.w) { p = q; n = Nn[g]; m = M[g]; }
        const double2 w = W[g];
        as += w.x; ws += w.y;
    }
    const uint64_t k = skey[start[c]];
    cp[c] = p; cn[c] = n; cw[c] = make_double2(as, ws); cm[c] = m; ckey[c] = k;
    const uint64_t kk = k - 1ull;
    const uint64_t b = kk & 7ull, cz = (kk >> 3) & 0xFFFFFull, cy = (kk >> 23) & 0xFFFFFull, cx = (kk >> 43) & 0xFFFFFull;
    mk[c] = (pr_spread3(cx) | pr_spread3(cy) << 1 | pr_spread3(cz) << 2) * 8ull + b;
    cidx[c] = c;
}

struct DevBuf {   // device allocations of one call, freed on every return path
    std::vector<void*> p;
    template <class T> hipError_t get(T** x, size_t n) { void* q = nullptr; hipError_t e = hipMalloc(&q, n * sizeof(T) + 16); if (e == hipSuccess) p.push_back(q); *x = (T*)q; return e; }
    ~DevBuf() { for (void* q : p) (void)hipFree(q); }
};


// ------------------------------------------------------------------------------ BVH build
// The host builder's breadth-first binned SAH (rc_bvh.h), level by level: items of a level
// with at most 64 triangles are decided by one wave each; larger ones are cut into chunks of
// RC_BCH triangles whose bounds and bins blocks reduce (exact min / max, integer counts), then
// one wave per item combines its chunks and runs the same sweep. A stable partition (scan of
// the per-triangle side flags) and a scan over the split items (children in item order) give
// the host's triangle order and node numbering.
#define RC_BCH 2048
struct BItem { int node, start, count, depth; };
struct BDec { int choice, bax, bs, nleft; float clo, chi, pad0, pad1; };

__device__ __forceinline__ uint32_t pr_ord(float f) { const uint32_t b = __float_as_uint(f); return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }
__device__ __forceinline__ float pr_unord(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }
__device__ __forceinline__ float pr_wmin(float v) { for (int o = 32; o > 0; o >>= 1) v = rc_fmin(v, __shfl_xor(v, o, 64)); return v; }
__device__ __forceinline__ float pr_wmax(float v) { for (int o = 32; o > 0; o >>= 1) v = rc_fmax(v, __shfl_xor(v, o, 64)); return v; }

__global__ void __launch_bounds__(256) k_bvh_tri(const float* V, int n, float4* tlo, float4* thi, float4* cen) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float* v = V + 9 * (size_t)t;
    float mn[3], mx[3], c[3];
    for (int a = 0; a < 3; ++a) {
        const float x0 = v[a] + 0.f, x1 = v[3 + a] + 0.f, x2 = v[6 + a] + 0.f;
        mn[a] = rc_fmin(rc_fmin(x0, x1), x2); mx[a] = rc_fmax(rc_fmax(x0, x1), x2);
        c[a] = 0.5f * (mn[a] + mx[a]) + 0.f;
    }
    tlo[t] = make_float4(mn[0], mn[1], mn[2], 0.f); thi[t] = make_float4(mx[0], mx[1], mx[2], 0.f); cen[t] = make_float4(c[0], c[1], c[2], 0.f);
}

// LDS bins of one axis as ordered integers (exact atomic min / max), then read back as floats
struct PrBinsU { int c[RC_BVH_NB]; uint32_t lo[RC_BVH_NB][3], hi[RC_BVH_NB][3]; };
__device__ __forceinline__ void pr_bins_clear(PrBinsU& B, int lane, int nl) {
    for (int b = lane; b < RC_BVH_NB; b += nl) { B.c[b] = 0; for (int a = 0; a < 3; ++a) { B.lo[b][a] = pr_ord(1e30f); B.hi[b][a] = pr_ord(-1e30f); } }
}
__device__ __forceinline__ void pr_bins_add(PrBinsU& B, int b, float4 lo, float4 hi) {
    atomicAdd(&B.c[b], 1);
    atomicMin(&B.lo[b][0], pr_ord(lo.x)); atomicMin(&B.lo[b][1], pr_ord(lo.y)); atomicMin(&B.lo[b][2], pr_ord(lo.z));
    atomicMax(&B.hi[b][0], pr_ord(hi.x)); atomicMax(&B.hi[b][1], pr_ord(hi.y)); atomicMax(&B.hi[b][2], pr_ord(hi.z));
}
__device__ __forceinline__ void pr_bins_float(const PrBinsU& U, rc_bins& B, int lane, int nl) {
    for (int b = lane; b < RC_BVH_NB; b += nl) {
        B.c[b] = U.c[b];
        for (int a = 0; a < 3; ++a) { B.lo[b][a] = pr_unord(U.lo[b][a]); B.hi[b][a] = pr_unord(U.hi[b][a]); }
    }
}

// shared tail of both paths (one wave, lane 0 decides): sweeps, choice, node bounds and decision
struct PrSweepLds { rc_bins B; float slo[RC_BVH_NB][3], shi[RC_BVH_NB][3]; int sc[RC_BVH_NB]; };

// items of at most 64 triangles: one wave (block of 64) per item, one triangle per lane
__global__ void __launch_bounds__(64) k_bvh_small(const BItem* items, int m, const int* ord, const float4* tlo, const float4* thi,
                                                 const float4* cen, rc_bvh_node* BN, BDec* dec) {
    __shared__ PrBinsU U;
    __shared__ PrSweepLds W;
    const int k = blockIdx.x, lane = threadIdx.x;
    if (k >= m) return;
    const BItem it = items[k];
    if (it.count > 64) return;   // chunked path
    const bool live = lane < it.count;
    const int t = live ? ord[it.start + lane] : 0;
    const float4 l4 = live ? tlo[t] : make_float4(1e30f, 1e30f, 1e30f, 0.f), h4 = live ? thi[t] : make_float4(-1e30f, -1e30f, -1e30f, 0.f);
    const float4 c4 = live ? cen[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    float nlo[3] = {pr_wmin(l4.x), pr_wmin(l4.y), pr_wmin(l4.z)}, nhi[3] = {pr_wmax(h4.x), pr_wmax(h4.y), pr_wmax(h4.z)};
    float clo[3] = {pr_wmin(live ? c4.x : 1e30f), pr_wmin(live ? c4.y : 1e30f), pr_wmin(live ? c4.z : 1e30f)};
    float chi[3] = {pr_wmax(live ? c4.x : -1e30f), pr_wmax(live ? c4.y : -1e30f), pr_wmax(live ? c4.z : -1e30f)};
    int ax = 0;
    for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
    int choice = 0, bs = -1, bax = -1;
    if (!rc_bvh_leaf_now(it.count, it.depth, chi[ax] - clo[ax])) {
        float best = 1e30f;
        for (int a0 = 0; a0 < 3; ++a0) {
            if (chi[a0] - clo[a0] < 1e-7f) continue;
            pr_bins_clear(U, lane, 64);
            __syncthreads();
            if (live) pr_bins_add(U, rc_bvh_bin(a0 == 0 ? c4.x : a0 == 1 ? c4.y : c4.z, clo[a0], chi[a0]), l4, h4);
            __syncthreads();
            pr_bins_float(U, W.B, lane, 64);
            __syncthreads();
            if (lane == 0) rc_bins_sweep(W.B, a0, best, bs, bax, W.slo, W.shi, W.sc);
            __syncthreads();
        }
        if (lane == 0) choice = rc_bvh_choice(it.count, best, bs, (float)it.count * rc_bvh_area(nlo, nhi));
    }
    if (lane == 0) {
        rc_bvh_node& nd = BN[it.node];
        for (int a = 0; a < 3; ++a) { nd.lo[a] = nlo[a]; nd.hi[a] = nhi[a]; }
        BDec d; d.choice = choice; d.bax = bax; d.bs = bs; d.nleft = it.count / 2;
        d.clo = bax >= 0 ? clo[bax] : 0.f; d.chi = bax >= 0 ? chi[bax] : 0.f; d.pad0 = d.pad1 = 0.f;
        dec[k] = d;
    }
}

// chunks of the larger items: chunk count per item, then the chunk table
__global__ void __launch_bounds__(256) k_bvh_nchunk(const BItem* items, int m, int* nch) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) nch[k] = items[k].count > 64 ? (items[k].count + RC_BCH - 1) / RC_BCH : 0;
    if (k == m) nch[k] = 0;
}
__global__ void __launch_bounds__(256) k_bvh_chunks(const BItem* items, int m, const int* choff, int2* chunk) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    for (int c = choff[k]; c < choff[k + 1]; ++c) chunk[c] = make_int2(k, c - choff[k]);
}

// per chunk: triangle box and centroid box (6 + 6 ordered integers)
__global__ void __launch_bounds__(256) k_bvh_cbounds(const BItem* items, const int2* chunk, const int* ord, const float4* tlo, const float4* thi,
                                                    const float4* cen, uint32_t* part) {
    __shared__ uint32_t s[12];
    const int2 ch = chunk[blockIdx.x];
    const BItem it = items[ch.x];
    const int j0 = it.start + ch.y * RC_BCH, j1 = min(it.start + it.count, j0 + RC_BCH);
    if (threadIdx.x < 12) s[threadIdx.x] = threadIdx.x % 6 < 3 ? pr_ord(1e30f) : pr_ord(-1e30f);
    __syncthreads();
    float v[12] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f, 1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const int t = ord[j];
        const float4 l = tlo[t], h = thi[t], c = cen[t];
        v[0] = rc_fmin(v[0], l.x); v[1] = rc_fmin(v[1], l.y); v[2] = rc_fmin(v[2], l.z);
        v[3] = rc_fmax(v[3], h.x); v[4] = rc_fmax(v[4], h.y); v[5] = rc_fmax(v[5], h.z);
        v[6] = rc_fmin(v[6], c.x); v[7] = rc_fmin(v[7], c.y); v[8] = rc_fmin(v[8], c.z);
        v[9] = rc_fmax(v[9], c.x); v[10] = rc_fmax(v[10], c.y); v[11] = rc_fmax(v[11], c.z);
    }
    for (int q = 0; q < 12; ++q) { if (q % 6 < 3) atomicMin(&s[q], pr_ord(v[q])); else atomicMax(&s[q], pr_ord(v[q])); }
    __syncthreads();
    if (threadIdx.x < 12) part[12 * (size_t)blockIdx.x + threadIdx.x] = s[threadIdx.x];
}

// per large item (one wave): combine the chunk boxes; leaf test; sets the item's centroid ranges
__global__ void __launch_bounds__(64) k_bvh_lbounds(const BItem* items, int m, const int* choff, const uint32_t* part, rc_bvh_node* BN,
                                                   float* crange, int* needbins, BDec* dec) {
    const int k = blockIdx.x, lane = threadIdx.x;
    if (k >= m) return;
    const BItem it = items[k];
    if (it.count <= 64) return;
    if (lane != 0) return;
    uint32_t u[12];
    for (int q = 0; q < 12; ++q) u[q] = q % 6 < 3 ? pr_ord(1e30f) : pr_ord(-1e30f);
    for (int c = choff[k]; c < choff[k + 1]; ++c)
        for (int q = 0; q < 12; ++q) { const uint32_t x = part[12 * (size_t)c + q]; u[q] = q % 6 < 3 ? min(u[q], x) : max(u[q], x); }
    float nlo[3], nhi[3], clo[3], chi[3];
    for (int a = 0; a < 3; ++a) { nlo[a] = pr_unord(u[a]); nhi[a] = pr_unord(u[3 + a]); clo[a] = pr_unord(u[6 + a]); chi[a] = pr_unord(u[9 + a]); }
    rc_bvh_node& nd = BN[it.node];
    for (int a = 0; a < 3; ++a) { nd.lo[a] = nlo[a]; nd.hi[a] = nhi[a]; crange[6 * (size_t)k + a] = clo[a]; crange[6 * (size_t)k + 3 + a] = chi[a]; }
    int ax = 0;
    for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
    const bool leaf = rc_bvh_leaf_now(it.count, it.depth, chi[ax] - clo[ax]);
    needbins[k] = leaf ? 0 : 1;
    if (leaf) { BDec d; d.choice = 0; d.bax = -1; d.bs = -1; d.nleft = it.count / 2; d.clo = d.chi = d.pad0 = d.pad1 = 0.f; dec[k] = d; }
}

// per chunk of an item that needs bins: the three axes' bins in LDS, then to global
__global__ void __launch_bounds__(256) k_bvh_cbins(const BItem* items, const int2* chunk, const int* needbins, const float* crange, const int* ord,
                                                  const float4* tlo, const float4* thi, const float4* cen, PrBinsU* part) {
    __shared__ PrBinsU U[3];
    const int2 ch = chunk[blockIdx.x];
    if (!needbins[ch.x]) return;
    const BItem it = items[ch.x];
    const float* cr = crange + 6 * (size_t)ch.x;
    for (int a0 = 0; a0 < 3; ++a0) pr_bins_clear(U[a0], threadIdx.x, blockDim.x);
    __syncthreads();
    const int j0 = it.start + ch.y * RC_BCH, j1 = min(it.start + it.count, j0 + RC_BCH);
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const int t = ord[j];
        const float4 l = tlo[t], h = thi[t], c = cen[t];
        const float cc[3] = {c.x, c.y, c.z};
        for (int a0 = 0; a0 < 3; ++a0)
            if (cr[3 + a0] - cr[a0] >= 1e-7f) pr_bins_add(U[a0], rc_bvh_bin(cc[a0], cr[a0], cr[3 + a0]), l, h);
    }
    __syncthreads();
    PrBinsU* out = part + 3 * (size_t)blockIdx.x;
    for (int e = threadIdx.x; e < 3 * RC_BVH_NB; e += blockDim.x) {
        const int a0 = e / RC_BVH_NB, b = e % RC_BVH_NB;
        out[a0].c[b] = U[a0].c[b];
        for (int a = 0; a < 3; ++a) { out[a0].lo[b][a] = U[a0].lo[b][a]; out[a0].hi[b][a] = U[a0].hi[b][a]; }
    }
}

// per large item needing bins (one wave): combine the chunks' bins, sweep, choose
__global__ void __launch_bounds__(64) k_bvh_ldecide(const BItem* items, int m, const int* choff, const int* needbins, const float* crange,
                                                   const PrBinsU* part, const rc_bvh_node* BN, BDec* dec) {
    __shared__ PrBinsU U;
    __shared__ PrSweepLds W;
    const int k = blockIdx.x, lane = threadIdx.x;
    if (k >= m) return;
    const BItem it = items[k];
    if (it.count <= 64 || !needbins[k]) return;
    const float* cr = crange + 6 * (size_t)k;
    float best = 1e30f;
    int bs = -1, bax = -1;
    for (int a0 = 0; a0 < 3; ++a0) {
        if (cr[3 + a0] - cr[a0] < 1e-7f) continue;
        if (lane < RC_BVH_NB) {
            int c = 0;
            uint32_t lo[3] = {pr_ord(1e30f), pr_ord(1e30f), pr_ord(1e30f)}, hi[3] = {pr_ord(-1e30f), pr_ord(-1e30f), pr_ord(-1e30f)};
            for (int q = choff[k]; q < choff[k + 1]; ++q) {
                const PrBinsU& P = part[3 * (size_t)q + a0];
                c += P.c[lane];
                for (int a = 0; a < 3; ++a) { lo[a] = min(lo[a], P.lo[lane][a]); hi[a] = max(hi[a], P.hi[lane][a]); }
            }
            U.c[lane] = c;
            for (int a = 0; a < 3; ++a) { U.lo[lane][a] = lo[a]; U.hi[lane][a] = hi[a]; }
        }
        __syncthreads();
        pr_bins_float(U, W.B, lane, 64);
        __syncthreads();
        if (lane == 0) rc_bins_sweep(W.B, a0, best, bs, bax, W.slo, W.shi, W.sc);
        __syncthreads();
    }
    if (lane == 0) {
        const rc_bvh_node nd = BN[it.node];
        BDec d; d.choice = rc_bvh_choice(it.count, best, bs, (float)it.count * rc_bvh_area(nd.lo, nd.hi));
        d.bax = bax; d.bs = bs; d.nleft = it.count / 2;
        d.clo = bax >= 0 ? cr[bax] : 0.f; d.chi = bax >= 0 ? cr[3 + bax] : 0.f; d.pad0 = d.pad1 = 0.f;
        dec[k] = d;
    }
}

// element -> item of the level: a head mark at each item's start, then a max-scan
__global__ void __launch_bounds__(256) k_bvh_marks(const BItem* items, int m, int* mark) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m && items[k].count > 0) mark[items[k].start] = k + 1;
}
// side flag of every element of a split item (1: left); 0 elsewhere
__global__ void __launch_bounds__(256) k_bvh_flags(int n, const int* eitem, const BItem* items, const BDec* dec, const int* ord,
                                                  const float4* cen, int* flag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    int f = 0;
    if (j < n) {
        const int k = eitem[j] - 1;
        if (k >= 0) {
            const BItem it = items[k];
            if (j < it.start + it.count) {
                const BDec d = dec[k];
                if (d.choice == 1) {
                    const float4 c = cen[ord[j]];
                    f = rc_bvh_bin(d.bax == 0 ? c.x : d.bax == 1 ? c.y : c.z, d.clo, d.chi) < d.bs ? 1 : 0;
                } else if (d.choice == 2) f = j - it.start < d.nleft ? 1 : 0;
            }
        }
    }
    flag[j] = f;
}
// stable partition inside each split item; other elements stay
__global__ void __launch_bounds__(256) k_bvh_scatter(int n, const int* eitem, const BItem* items, const BDec* dec, const int* flag,
                                                    const int* scan, const int* ord, int* ord2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int pos = j;
    const int k = eitem[j] - 1;
    if (k >= 0) {
        const BItem it = items[k];
        if (j < it.start + it.count && dec[k].choice != 0) {
            const int pre = scan[j] - scan[it.start], nleft = scan[it.start + it.count] - scan[it.start];
            pos = flag[j] ? it.start + pre : it.start + nleft + (j - it.start - pre);
        }
    }
    ord2[pos] = ord[j];
}
__global__ void __launch_bounds__(256) k_bvh_splitflag(const BDec* dec, int m, int* sf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) sf[k] = dec[k].choice != 0 ? 1 : 0;
    if (k == m) sf[k] = 0;
}
// children of the split items (in item order) and the nodes' child / leaf fields
__global__ void __launch_bounds__(256) k_bvh_children(const BItem* items, int m, const BDec* dec, const int* sfscan, const int* scan, int nodes0,
                                                     rc_bvh_node* BN, BItem* next) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const BItem it = items[k];
    const BDec d = dec[k];
    rc_bvh_node& nd = BN[it.node];
    if (d.choice == 0) { nd.lf = it.start; nd.cnt = it.count; return; }
    const int r = sfscan[k], left = nodes0 + 2 * r;
    const int nleft = scan[it.start + it.count] - scan[it.start];
    nd.lf = left; nd.cnt = 0;
    BItem a, b;
    a.node = left; a.start = it.start; a.count = nleft; a.depth = it.depth + 1;
    b.node = left + 1; b.start = it.start + nleft; b.count = it.count - nleft; b.depth = it.depth + 1;
    next[2 * r] = a; next[2 * r + 1] = b;
}

struct PrMax { __device__ __forceinline__ int operator()(int a, int b) const { return a > b ? a : b; } };

__global__ void __launch_bounds__(256) k_iota(int* v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

}  // namespace

namespace rc {

// The cache cells' hash table on the GPU (cuda/prep_mesh.cuh:356-366 builds it on the host): entry i's
// key with value i, inserted with 64-bit CAS linear probing from mix64(key) & (size - 1), the same slot
// function and probe order as the lookup. Slots may differ from the host's sequential insertion; every
// key looks up to the same entry, and empty slots hold value -1 as on the host.
std::string build_hash_device(const std::vector<uint64_t>& keys, int log2, std::vector<uint64_t>& HK, std::vector<int>& HV) {
    const size_t HS = (size_t)1 << log2, n = keys.size();
    if (HS > 0x80000000ull) return "hash table larger than 2^31 slots";
    hipStream_t st = nullptr;
    DevBuf B;
    unsigned long long *dK, *dHK; int *dV, *dHV, *dF;
    PK(B.get(&dK, std::max<size_t>(n, 1))); PK(B.get(&dV, std::max<size_t>(n, 1))); PK(B.get(&dHK, HS)); PK(B.get(&dHV, HS));
    PK(B.get(&dF, 1));
    PK(hipMemsetAsync(dHK, 0, HS * 8, st));
    PK(hipMemsetAsync(dHV, 0x7f, HS * 4, st));   // 0x7f7f7f7f: above every entry index, lowered by atomicMin
    PK(hipMemsetAsync(dF, 0, 4, st));
    if (n) {
        PK(hipMemcpyAsync(dK, keys.data(), n * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dV, (int)n);
        PK(hipGetLastError());
        PK(rc_launch_hash_insert(dHK, dHV, (uint32_t)(HS - 1), dK, dV, (int)n, dF, st));
    }
    PK(rc_launch_hash_finish(dHK, dHV, (uint32_t)HS, st));
    HK.resize(HS); HV.resize(HS);
    int failed = 0;
    PK(hipMemcpyAsync(HK.data(), dHK, HS * 8, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(HV.data(), dHV, HS * 4, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(&failed, dF, 4, hipMemcpyDeviceToHost, st));
    PK(hipStreamSynchronize(st));
    if (failed) return std::to_string(failed) + " cache keys found no hash slot within 64 probes";
    return "";
}

std::string sample_cells_device(const Mesh& M, float h, const float lo[3], std::vector<uint64_t>& keys_out, std::vector<CellRec>& recs_out) {
    const int nt = M.nt();
    keys_out.clear(); recs_out.clear();
    if (nt == 0) return "";
    hipStream_t st = nullptr;
    DevBuf B;
    float *dV, *dA; uint8_t* dD; int* dM = nullptr; unsigned long long *cnt, *off;
    PK(B.get(&dV, 9 * (size_t)nt)); PK(B.get(&dA, nt)); PK(B.get(&dD, nt)); PK(B.get(&cnt, nt + 1)); PK(B.get(&off, nt + 1));
    PK(hipMemcpyAsync(dV, M.V.data(), 36 * (size_t)nt, hipMemcpyHostToDevice, st));
    PK(hipMemcpyAsync(dA, M.alb.data(), 4 * (size_t)nt, hipMemcpyHostToDevice, st));
    PK(hipMemcpyAsync(dD, M.ds.data(), (size_t)nt, hipMemcpyHostToDevice, st));
    if ((int)M.mat.size() == nt) { PK(B.get(&dM, nt)); PK(hipMemcpyAsync(dM, M.mat.data(), 4 * (size_t)nt, hipMemcpyHostToDevice, st)); }
    const dim3 bt(256), gt((nt + 255) / 256);
    PK(hipMemsetAsync(cnt + nt, 0, 8, st));
    hipLaunchKernelGGL(k_cells_count, gt, bt, 0, st, dV, dD, nt, h, cnt);
    PK(hipGetLastError());
    size_t tb = 0;
    PK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, off, nt + 1, st));
    void* tmp; PK(B.get((char**)&tmp, tb));
    PK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, off, nt + 1, st));
    unsigned long long S64 = 0;
    PK(hipMemcpyAsync(&S64, off + nt, 8, hipMemcpyDeviceToHost, st));
    PK(hipStreamSynchronize(st));
    if (S64 == 0) return "";
    if (S64 >= (1ull << 31)) return "too many cache-cell samples for the device sampler (2^31)";
    const unsigned S = (unsigned)S64;
    unsigned long long *key, *skey; unsigned *idx, *sidx; float4 *Pd, *Nn; double2* W; int* Me;
    PK(B.get(&key, S)); PK(B.get(&skey, S)); PK(B.get(&idx, S)); PK(B.get(&sidx, S));
    PK(B.get(&Pd, S)); PK(B.get(&Nn, S)); PK(B.get(&W, S)); PK(B.get(&Me, S));
    hipLaunchKernelGGL(k_cells_emit, gt, bt, 0, st, dV, dD, dA, dM, nt, h, lo[0], lo[1], lo[2], off, key, idx, Pd, Nn, W, Me);
    PK(hipGetLastError());
    // stable radix sort by cell key: equal keys keep the host's encounter order
    tb = 0;
    PK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, idx, sidx, S, 0, 64, st));
    void* tmp2; PK(B.get((char**)&tmp2, tb));
    PK(hipcub::DeviceRadixSort::SortPairs(tmp2, tb, key, skey, idx, sidx, S, 0, 64, st));
    unsigned *head, *cid, *start;
    PK(B.get(&head, S)); PK(B.get(&cid, S)); PK(B.get(&start, (size_t)S + 1));
    const dim3 gs((S + 255) / 256);
    hipLaunchKernelGGL(k_cells_heads, gs, bt, 0, st, skey, S, head);
    tb = 0;
    PK(hipcub::DeviceScan::InclusiveSum(nullptr, tb, head, cid, S, st));
    void* tmp3; PK(B.get((char**)&tmp3, tb));
    PK(hipcub::DeviceScan::InclusiveSum(tmp3, tb, head, cid, S, st));
    hipLaunchKernelGGL(k_cells_starts, gs, bt, 0, st, head, cid, S, start);
    PK(hipGetLastError());
    unsigned C = 0;
    PK(hipMemcpyAsync(&C, cid + (S - 1), 4, hipMemcpyDeviceToHost, st));
    PK(hipStreamSynchronize(st));
    float4 *cp, *cn; double2* cw; int* cm; unsigned long long *ckey, *mk, *smk; unsigned *cidx, *order;
    PK(B.get(&cp, C)); PK(B.get(&cn, C)); PK(B.get(&cw, C)); PK(B.get(&cm, C)); PK(B.get(&ckey, C));
    PK(B.get(&mk, C)); PK(B.get(&smk, C)); PK(B.get(&cidx, C)); PK(B.get(&order, C));
    hipLaunchKernelGGL(k_cells_reduce, dim3((C + 255) / 256), bt, 0, st, C, start, skey, sidx, Pd, Nn, W, Me, cp, cn, cw, cm, ckey, mk, cidx);
    PK(hipGetLastError());
    // Morton order (morton * 8 + bucket is one-to-one with the key)
    tb = 0;
    PK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, mk, smk, cidx, order, C, 0, 64, st));
    void* tmp4; PK(B.get((char**)&tmp4, tb));
    PK(hipcub::DeviceRadixSort::SortPairs(tmp4, tb, mk, smk, cidx, order, C, 0, 64, st));
    std::vector<float4> hp(C), hn(C); std::vector<double2> hw(C); std::vector<int> hm(C), ho(C); std::vector<unsigned long long> hk(C);
    PK(hipMemcpyAsync(hp.data(), cp, 16 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(hn.data(), cn, 16 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(hw.data(), cw, 16 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(hm.data(), cm, 4 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(hk.data(), ckey, 8 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipMemcpyAsync(ho.data(), order, 4 * (size_t)C, hipMemcpyDeviceToHost, st));
    PK(hipStreamSynchronize(st));
    keys_out.resize(C); recs_out.resize(C);
    for (unsigned j = 0; j < C; ++j) {
        const int c = ho[j];
        CellRec& r = recs_out[j];
        r.p[0] = hp[c].x; r.p[1] = hp[c].y; r.p[2] = hp[c].z; r.d2 = hp[c].w;
        r.n[0] = hn[c].x; r.n[1] = hn[c].y; r.n[2] = hn[c].z;
        r.asum = hw[c].x; r.wsum = hw[c].y; r.mat = hm[c];
        keys_out[j] = hk[c];
    }
    return "";
}


std::string build_bvh_device(const Mesh& M, std::vector<rc_bvh_node>& nodes_out, std::vector<int>& order_out, int& depth_out) {
    const int n = M.nt();
    hipStream_t st = nullptr;
    DevBuf B;
    float* dV; float4 *tlo, *thi, *cen; int *ord, *ord2, *mark, *eitem, *flag, *scan, *nch, *choff, *sf, *sfscan, *needbins;
    BItem *items, *next; BDec* dec; rc_bvh_node* BN; float* crange;
    const size_t cap = (size_t)n + 2;
    PK(B.get(&dV, 9 * (size_t)std::max(n, 1))); PK(B.get(&tlo, cap)); PK(B.get(&thi, cap)); PK(B.get(&cen, cap));
    PK(B.get(&ord, cap)); PK(B.get(&ord2, cap)); PK(B.get(&mark, cap)); PK(B.get(&eitem, cap)); PK(B.get(&flag, cap)); PK(B.get(&scan, cap));
    PK(B.get(&nch, cap)); PK(B.get(&choff, cap)); PK(B.get(&sf, cap)); PK(B.get(&sfscan, cap)); PK(B.get(&needbins, cap));
    PK(B.get(&items, cap)); PK(B.get(&next, cap)); PK(B.get(&dec, cap)); PK(B.get(&BN, 2 * cap)); PK(B.get(&crange, 6 * cap));
    if (n > 0) PK(hipMemcpyAsync(dV, M.V.data(), 36 * (size_t)n, hipMemcpyHostToDevice, st));
    std::vector<int> iota(std::max(n, 1));
    std::iota(iota.begin(), iota.end(), 0);
    PK(hipMemcpyAsync(ord, iota.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    const dim3 b256(256), gn((n + 256) / 256);
    if (n > 0) hipLaunchKernelGGL(k_bvh_tri, gn, b256, 0, st, dV, n, tlo, thi, cen);
    BItem root{0, 0, n, 0};
    PK(hipMemcpyAsync(items, &root, sizeof root, hipMemcpyHostToDevice, st));
    // temp storage for the scans (sized for the largest)
    size_t tb1 = 0, tb2 = 0;
    PK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, flag, scan, n + 1, st));
    PK(hipcub::DeviceScan::InclusiveScan(nullptr, tb2, mark, eitem, PrMax(), n, st));
    void* tmp; PK(B.get((char**)&tmp, std::max(tb1, tb2) + 256));
    const size_t tbs = std::max(tb1, tb2) + 256;
    int m = 1, nodes = 1, depth = 0;
    uint32_t* cpart = nullptr; PrBinsU* bpart = nullptr; int2* chunk = nullptr; size_t ccap = 0;
    while (m > 0) {
        depth = std::max(depth, 0);
        const dim3 gm((m + 256) / 256);
        // decisions: small items by one wave each, large items by chunks
        hipLaunchKernelGGL(k_bvh_small, dim3(m), dim3(64), 0, st, items, m, ord, tlo, thi, cen, BN, dec);
        hipLaunchKernelGGL(k_bvh_nchunk, gm, b256, 0, st, items, m, nch);
        size_t tb = tbs;
        PK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, nch, choff, m + 1, st));
        int nck = 0;
        PK(hipMemcpyAsync(&nck, choff + m, 4, hipMemcpyDeviceToHost, st));
        PK(hipStreamSynchronize(st));
        if (nck > 0) {
            if ((size_t)nck > ccap) {
                ccap = (size_t)nck * 2;
                PK(B.get(&cpart, 12 * ccap)); PK(B.get(&bpart, 3 * ccap)); PK(B.get(&chunk, ccap));
            }
            hipLaunchKernelGGL(k_bvh_chunks, gm, b256, 0, st, items, m, choff, chunk);
            hipLaunchKernelGGL(k_bvh_cbounds, dim3(nck), b256, 0, st, items, chunk, ord, tlo, thi, cen, cpart);
            hipLaunchKernelGGL(k_bvh_lbounds, dim3(m), dim3(64), 0, st, items, m, choff, cpart, BN, crange, needbins, dec);
            hipLaunchKernelGGL(k_bvh_cbins, dim3(nck), b256, 0, st, items, chunk, needbins, crange, ord, tlo, thi, cen, bpart);
            hipLaunchKernelGGL(k_bvh_ldecide, dim3(m), dim3(64), 0, st, items, m, choff, needbins, crange, bpart, BN, dec);
        }
        // partition and children
        PK(hipMemsetAsync(mark, 0, 4 * (size_t)std::max(n, 1), st));
        hipLaunchKernelGGL(k_bvh_marks, gm, b256, 0, st, items, m, mark);
        if (n > 0) {
            tb = tbs;
            PK(hipcub::DeviceScan::InclusiveScan(tmp, tb, mark, eitem, PrMax(), n, st));
        }
        hipLaunchKernelGGL(k_bvh_flags, gn, b256, 0, st, n, eitem, items, dec, ord, cen, flag);
        tb = tbs;
        PK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, flag, scan, n + 1, st));
        if (n > 0) hipLaunchKernelGGL(k_bvh_scatter, gn, b256, 0, st, n, eitem, items, dec, flag, scan, ord, ord2);
        hipLaunchKernelGGL(k_bvh_splitflag, gm, b256, 0, st, dec, m, sf);
        tb = tbs;
        PK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, sf, sfscan, m + 1, st));
        hipLaunchKernelGGL(k_bvh_children, gm, b256, 0, st, items, m, dec, sfscan, scan, nodes, BN, next);
        PK(hipGetLastError());
        int nsplit = 0;
        PK(hipMemcpyAsync(&nsplit, sfscan + m, 4, hipMemcpyDeviceToHost, st));
        PK(hipStreamSynchronize(st));
        if (n > 0) std::swap(ord, ord2);
        std::swap(items, next);
        nodes += 2 * nsplit;
        if (nsplit > 0) ++depth;
        m = 2 * nsplit;
    }
    nodes_out.resize(nodes);
    order_out.resize(n);
    PK(hipMemcpyAsync(nodes_out.data(), BN, sizeof(rc_bvh_node) * (size_t)nodes, hipMemcpyDeviceToHost, st));
    if (n > 0) PK(hipMemcpyAsync(order_out.data(), ord, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    PK(hipStreamSynchronize(st));
    depth_out = depth;
    return "";
}

}  // namespace rc

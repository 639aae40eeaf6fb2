// Scene preparation on the GPU: the cache cells of a mesh (cuda/prep_mesh.cuh:129-157).
//
// The same cells as the host's sample_cells, bit for bit: every triangle draws its samples
// from its own host RNG chain (one thread per triangle, the chain is sequential), the samples
// are sorted by cell key with a stable radix sort (so the samples of a cell stay in the host's
// triangle / sample / side order), each cell is reduced by one thread in that order (the
// nearest-to-centre representative with the host's strict '<', the double-precision albedo
// and area sums in the host's addition order), and the cells are ordered by Morton code.
#include <hipcub/hipcub.hpp>

#include "rc_device.h"
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

}  // namespace

namespace rc {

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

}  // namespace rc

// Adjoint of the cached radiance with respect to per-light emission and per-material albedo.
//
// Differentiable model: the cache's own transport with fixed rays. With the unit-intensity
// direct light G[i][l] (visibility x cosine / distance^2: the k_build geometry at I = 1,
// cuda/poc_gpu.cu:369-391), E_i = sum_l I_l G_il and A_i = alb[mat_i], L^(0) = 0 and, for the
// bounces b = 0..B-1,
//     L^(b+1)_i = (1/K) sum over the K rays of entry i at bounce b that reach a valid entry h
//                 of A_h (E_h + L^(b)_h),
// and the cached radiance is R_i = A_i (E_i + L^(B)_i). The rays are those of the converged-
// reference iteration (cuda/prep_mesh.cuh:181-213): seed_b = hash32(seed + b * 0x9E3779B9),
// rs = hash32(seed_b ^ hash32(i + 1)), cosine directions from the entry.
//
// Reverse mode for J = sum_i g_i R_i keeps the per-bounce radiance L^(b) (one float per entry
// and bounce) and re-traces each bounce's rays from their seeds instead of storing paths; the
// weight of a ray's contribution is scattered to its hit entry with one hardware float atomic,
// and the per-entry factors are applied afterwards by a streaming pass.
// The final reductions (per light over entries, per material over its entries) run in a fixed
// order. The 64-bit support masks (bit min(l, 63) when dR_i/dI_l != 0) propagate with the
// forward: all terms are non-negative, so the nonzero pattern is the union over the rays.
#include "rc_device.h"
#include "rc_kernels.h"
#include "rc_trace.h"

// G[i][l] and its support SG[i]: the shadow-ray visibility of k_build (closest hit beyond
// r - 1e-3, equivalently no hit in (1e-4, r - 1e-3)), at unit intensity; rays are skipped
// where the cosine is not positive (G = 0 there).
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_adj_unit(rc_dscene s, rc_dstate st, float* G, uint64_t* SG) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    lstk.qwalk = RC_WQ_L - 4;   // launched with RC_WQ_L queue levels (RC_ADJ_T)
    rc_load_dyn(st, sh, shq);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.Np) return;
    const int nl = s.nl;
    const f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    uint64_t m = 0ull;
    for (int l = 0; l < nl; ++l) {
        float r; f3 D;
        if (s.ltype[l] == 1) { D = ld3(s.L, l); r = s.sun_len; }
        else { f3 dv = ld3(s.L, l) - P; r = sqrtf(dot3(dv, dv)); D = dv * (1.f / r); }
        const float c = dot3(N, D);
        float g = 0.f;
        if (c > 0.f) {
            bool vis;
            if (TRAV == 2) {
                const float tlim = r - 1e-3f;
                vis = !(rc_occluded_lds(O, D, sh, st.nt, tlim) || (s.nst > 0 && rc_occluded_wide(O, D, s, tlim, lstk)));
            } else {
                float t; int q;
                rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
                vis = t >= r - 1e-3f;
            }
            if (vis) {
                if (s.ltype[l] == 1) g = c;
                else { const float rr = fmaxf(r, 0.05f); g = c / (rr * rr); }
            }
        }
        G[(int64_t)nl * i + l] = g;
        if (g > 0.f) m |= 1ull << (l < 63 ? l : 63);
    }
    SG[i] = m;
}

__global__ void __launch_bounds__(RC_TB) k_adj_zero(int n, float* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

// A_i = alb[mat_i], E_i = sum_l I_l G_il (lights in order). A block owns RC_AS entries, whose
// G rows form one contiguous span: copied to LDS with coalesced loads (rows padded to nl + 1
// floats, so the per-entry walks along a row hit distinct banks), then summed per entry.
#define RC_AS 64
__global__ void __launch_bounds__(RC_TB) k_adj_setup(int Np, int nl, const int* mat, const float* alb, const float* I, const float* G,
                                                    float* A, float* E) {
    __shared__ float s_g[RC_AS * (RC_MAX_LIGHTS + 1)];
    const int base = blockIdx.x * RC_AS;
    const int T = min(RC_AS, Np - base), stride = nl + 1;
    const float* g = G + (int64_t)base * nl;
    for (int e = threadIdx.x; e < T * nl; e += blockDim.x) {
        const int j = e / nl;
        s_g[j * stride + (e - j * nl)] = g[e];
    }
    __syncthreads();
    if ((int)threadIdx.x < T) {
        const int i = base + threadIdx.x;
        A[i] = alb[mat[i]];
        const float* row = s_g + threadIdx.x * stride;
        float e = 0.f;
        for (int l = 0; l < nl; ++l) e += I[l] * row[l];
        E[i] = e;
    }
}

// the ray rule of the adjoint bounces: hit entry h of ray k of entry i, or -1
template <int TRAV, class LS>
__device__ __forceinline__ int rc_adj_ray(const rc_dscene& s, const rc_dstate& st, const float* sh, const int* shq, const LS& lstk,
                                          int i, f3 O, f3 N, uint32_t& rs) {
    const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
    const f3 D = rc_cosine_dir(N, u1, u2);
    float t; int q;
    const bool hit = rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
    const f3 H = hit ? O + D * t : O;
    int pi;
    return rc_resolve_hit(s, st, hit, H, D, q, i, pi) ? -1 : pi;
}

// The reverse bounce traces its rays in direction order within each block: the entry threads
// draw up to RC_ADJ_KS rays each (the RNG state before each ray is kept, so the ray is drawn again
// exactly by whichever thread traces it), a counting sort in LDS groups them by octant and
// dominant axis, and every thread then traces rays from the sorted list. The block's entries are
// neighbours (Morton order), so a wave's rays then share an origin region and a direction class
// and walk much the same nodes; entries of zero weight draw no rays, so the live rays are also
// compacted. The sums into S are float atomics either way: their order was never fixed. (The
// forward bounce keeps each entry's rays in the entry's thread: the same sort measured slower
// there, by the per-entry state it has to carry across the block's barriers.)
#define RC_ADJ_KS 4
#define RC_ADJ_NB 24
struct rc_adj_rays {
    uint32_t v[RC_TB * RC_ADJ_KS];   // RNG state before the ray
    uint16_t ord[RC_TB * RC_ADJ_KS]; // slots (thread * RC_ADJ_KS + ray) in direction order
    int cnt[RC_ADJ_NB];
    uint32_t rs[RC_TB];              // each thread's entry's RNG state between passes
};
__device__ __forceinline__ int rc_dir_class(f3 D) {
    const float ax = fabsf(D.x), ay = fabsf(D.y), az = fabsf(D.z);
    const int axis = ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
    return ((D.x < 0.f) | ((D.y < 0.f) << 1) | ((D.z < 0.f) << 2)) * 3 + axis;
}
// one pass over the next nk rays (nk <= RC_ADJ_KS) of the block's live entries; R.rs[t] is the
// RNG state of thread t's entry, advanced past the pass's rays; each hit h of a ray of entry i
// adds wgt[i] / K to S[h]. The state carried between passes lives in LDS, so nothing but the
// tracing loop's own state is held in registers across the traversal.
template <int TRAV, class LS>
__device__ __forceinline__ void rc_adj_pass(const rc_dscene& s, const rc_dstate& st, const float* sh, const int* shq, const LS& lstk,
                                            rc_adj_rays& R, bool live, int nk, const float* wgt, int K, float* S) {
    const int t = threadIdx.x;
    if (t < RC_ADJ_NB) R.cnt[t] = 0;
    __syncthreads();
    const f3 N = live ? ld3(st.N, blockIdx.x * RC_TB + t) : f3{0.f, 0.f, 1.f};
    int key[RC_ADJ_KS];
    uint32_t rs = R.rs[t];
#pragma unroll
    for (int j = 0; j < RC_ADJ_KS; ++j) {
        key[j] = -1;
        if (live && j < nk) {
            R.v[t * RC_ADJ_KS + j] = rs;
            const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
            key[j] = rc_dir_class(rc_cosine_dir(N, u1, u2));
            atomicAdd(&R.cnt[key[j]], 1);
        }
    }
    R.rs[t] = rs;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int b = 0; b < RC_ADJ_NB; ++b) { const int c = R.cnt[b]; R.cnt[b] = run; run += c; }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RC_ADJ_KS; ++j)
        if (key[j] >= 0) R.ord[atomicAdd(&R.cnt[key[j]], 1)] = (uint16_t)(t * RC_ADJ_KS + j);
    __syncthreads();
    const int n = R.cnt[RC_ADJ_NB - 1];   // every bucket's start moved to its end: the last end is the count
    for (int p = t; p < n; p += RC_TB) {
        const int r = R.ord[p];
        const int i = blockIdx.x * RC_TB + r / RC_ADJ_KS;
        uint32_t q = R.v[r];
        const f3 P = ld3(st.P, i), Ne = ld3(st.N, i), O = P + Ne * 1e-3f;
        const int h = rc_adj_ray<TRAV>(s, st, sh, shq, lstk, i, O, Ne, q);
        if (h >= 0) unsafeAtomicAdd(S + h, wgt[i] / (float)K);
    }
    __syncthreads();
}

// one forward bounce: Lout_i = (1/K) sum A_h (E_h + Lin_h), Sout_i = union of SG_h | Sin_h over
// hits with A_h > 0 (Lin, Sin NULL: zero). Five waves per SIMD: the split-stack build fits 96
// VGPRs (one spilled) and its smaller LDS lets five blocks reside, 3 % faster than four waves.
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) __attribute__((amdgpu_waves_per_eu(5))) k_adj_fwd(rc_dscene s, rc_dstate st, const float* A, const float* E, const uint64_t* SG,
                                                  const float* Lin, const uint64_t* Sin, float* Lout, uint64_t* Sout, uint32_t seed, int K) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    lstk.qwalk = RC_WQ_L - 4;   // launched with RC_WQ_L queue levels (RC_ADJ_T)
    rc_load_dyn(st, sh, shq);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.Np) return;
    uint32_t rs = rc_hash32(seed ^ rc_hash32((uint32_t)i + 1u));
    const f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    float acc = 0.f;
    uint64_t m = 0ull;
    for (int k = 0; k < K; ++k) {
        const int h = rc_adj_ray<TRAV>(s, st, sh, shq, lstk, i, O, N, rs);
        if (h < 0) continue;
        const float Ah = A[h];
        acc += Ah * (E[h] + (Lin ? Lin[h] : 0.f));
        if (Ah > 0.f) m |= SG[h] | (Sin ? Sin[h] : 0ull);
    }
    Lout[i] = acc / (float)K;
    Sout[i] = m;
}

// R_i = A_i (E_i + L_i), SR_i = support of R_i
__global__ void __launch_bounds__(RC_TB) k_adj_out(int Np, const float* A, const float* E, const float* L, const uint64_t* SG,
                                                  const uint64_t* S, float* R, uint64_t* SR) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const float a = A[i];
    if (R) R[i] = a * (E[i] + (L ? L[i] : 0.f));
    SR[i] = a > 0.f ? (SG[i] | (S ? S[i] : 0ull)) : 0ull;
}

// cotangent g on R: dA_i = g_i (E_i + L^(B)_i), dE_i = g_i A_i, adjoint of L^(B)_i = g_i A_i
__global__ void __launch_bounds__(RC_TB) k_adj_bwd_init(int Np, const float* g, const float* A, const float* E, const float* L,
                                                       float* dA, float* dE, float* lam) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    const float gi = g[i], a = A[i];
    dA[i] = gi * (E[i] + (L ? L[i] : 0.f));
    dE[i] = gi * a;
    lam[i] = gi * a;
}

// one reverse bounce: re-trace the rays of bounce b; each hit h collects the weight
// w = lam_i / K of the ray's entry i into S_h (one atomic per hit). Every term the hit feeds
// is a per-h factor times that weight, so k_adj_bwd_fin then adds (E_h + L^(b)_h) S_h to dA,
// A_h S_h to dE and sets the adjoint of L^(b) to A_h S_h.
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) __attribute__((amdgpu_waves_per_eu(5))) k_adj_bwd(rc_dscene s, rc_dstate st, const float* lam_in, float* S, uint32_t seed, int K) {
    extern __shared__ float sh[];
    __shared__ rc_adj_rays R;
    int* shq = (int*)(sh + 9 * st.nt);
    auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    lstk.qwalk = RC_WQ_L - 4;   // launched with RC_WQ_L queue levels (RC_ADJ_T)
    rc_load_dyn(st, sh, shq);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // an entry of zero weight contributes nothing: its rays are not traced
    const bool live = i < s.Np && lam_in[i] / (float)K != 0.f;
    R.rs[threadIdx.x] = rc_hash32(seed ^ rc_hash32((uint32_t)i + 1u));
    for (int k0 = 0; k0 < K; k0 += RC_ADJ_KS)
        rc_adj_pass<TRAV>(s, st, sh, shq, lstk, R, live, min(RC_ADJ_KS, K - k0), lam_in, K, S);
}

// per entry h after a reverse bounce: dA_h += (E_h + L^(b)_h) S_h, dE_h += A_h S_h, and the
// adjoint of L^(b) (lam_out, may alias S) = A_h S_h
__global__ void __launch_bounds__(RC_TB) k_adj_bwd_fin(int Np, const float* A, const float* E, const float* Lp, const float* S,
                                                      float* lam_out, float* dA, float* dE) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= Np) return;
    const float sw = S[h];
    if (sw == 0.f) { if (lam_out) lam_out[h] = 0.f; return; }
    const float a = A[h];
    dA[h] += (E[h] + (Lp ? Lp[h] : 0.f)) * sw;
    dE[h] += a * sw;
    if (lam_out) lam_out[h] = a * sw;
}

// block sum in a fixed order (wave shuffles, then the waves in order)
__device__ __forceinline__ float rc_block_sum(float v, float* s_w) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s_w[wv] = v;
    __syncthreads();
    float t = 0.f;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += s_w[k];
    return t;
}

// per-block partial sums of dE_i G_il for every light: part[block][l]. Lanes are lights (rows
// of G read whole, 64 lights at a time); each wave walks every fourth entry of the block's
// RC_TB, then the four waves' sums are added in wave order.
__global__ void __launch_bounds__(RC_TB) k_adj_dI_part(int Np, int nl, const float* dE, const float* G, float* part) {
    __shared__ float s_p[RC_TB / 64][RC_MAX_LIGHTS];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = blockIdx.x * RC_TB;
    for (int l = lane; l < nl; l += 64) {
        float acc = 0.f;
        for (int j = wv; j < RC_TB; j += RC_TB / 64) {
            const int i = base + j;
            if (i < Np) acc += dE[i] * G[(int64_t)i * nl + l];
        }
        s_p[wv][l] = acc;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nl; l += RC_TB) {
        float t = 0.f;
        for (int k = 0; k < RC_TB / 64; ++k) t += s_p[k][l];
        part[(int64_t)blockIdx.x * nl + l] = t;
    }
}

// dI_l = sum over blocks of part[block][l] (one block per light)
__global__ void __launch_bounds__(RC_TB) k_adj_dI_sum(int nblk, int nl, const float* part, float* dI) {
    __shared__ float s_w[RC_TB / 64];
    const int l = blockIdx.x;
    float v = 0.f;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) v += part[(int64_t)b * nl + l];
    const float t = rc_block_sum(v, s_w);
    if (threadIdx.x == 0) dI[l] = t;
}

// dalb_m = sum of dA over the entries of material m (perm: entries sorted by material, moff:
// offsets, pmat: material of each sorted position), balanced whatever the material sizes: each
// block scans one RC_TB tile of the sorted positions (segmented by material, fixed order) and
// leaves each segment's sum at the segment's last position; then one wave per material adds
// its tiles' segment sums in tile order.
__global__ void __launch_bounds__(RC_TB) k_adj_dalb_tile(int Np, const float* dA, const int* perm, const int* pmat, float* tail) {
    __shared__ float s_v[RC_TB];
    __shared__ int s_k[RC_TB];
    const int t = threadIdx.x, p = blockIdx.x * RC_TB + t;
    const bool live = p < Np;
    s_v[t] = live ? dA[perm[p]] : 0.f;
    s_k[t] = live ? pmat[p] : -1;
    __syncthreads();
    for (int d = 1; d < RC_TB; d <<= 1) {   // inclusive segmented scan (Hillis-Steele)
        const float add = (t >= d && s_k[t - d] == s_k[t]) ? s_v[t - d] : 0.f;
        __syncthreads();
        s_v[t] += add;
        __syncthreads();
    }
    if (live && (t == RC_TB - 1 || p == Np - 1 || s_k[t + 1] != s_k[t])) tail[p] = s_v[t];
}
__global__ void __launch_bounds__(RC_TB) k_adj_dalb(int nmat, const int* moff, const float* tail, float* dalb) {
    const int m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (m >= nmat) return;
    const int p0 = moff[m], p1 = moff[m + 1];
    float v = 0.f;
    if (p1 > p0)
        for (int tl = p0 / RC_TB + lane; tl <= (p1 - 1) / RC_TB; tl += 64) v += tail[min((tl + 1) * RC_TB, p1) - 1];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) dalb[m] = v;
}

// ------------------------------------------------------------------------------ launchers
#define RC_ADJ_T(KER, ...)                                                                         \
    do {                                                                                            \
        if (trav == 0) hipLaunchKernelGGL((KER<0, false>), g, b, shm, stream, __VA_ARGS__);        \
        else if (trav == 1) hipLaunchKernelGGL((KER<1, false>), g, b, shm, stream, __VA_ARGS__);   \
        else RC_LAUNCH_WIDE_Q((KER<2, true>), (KER<2, false>), false, RC_WQ_L, g, st, s, stream, __VA_ARGS__); \
    } while (0)

hipError_t rc_launch_adj_unit(int trav, const rc_dscene* s, const rc_dstate* st, float* G, uint64_t* SG, hipStream_t stream) {
    const dim3 g(rc_grid(s->Np)), b(RC_TB);
    const size_t shm = rc_shm(*st, *s, trav, false);
    RC_ADJ_T(k_adj_unit, *s, *st, G, SG);
    return hipGetLastError();
}

hipError_t rc_launch_adj_forward(int trav, const rc_dscene* s, const rc_dstate* st, const rc_adjbuf* w, const float* I,
                                 const float* alb, int K, int B, uint32_t seed, float* R, hipStream_t stream, hipEvent_t* ev) {
    const int Np = s->Np;
    const dim3 g(rc_grid(Np)), b(RC_TB);
    const size_t shm = rc_shm(*st, *s, trav, false);
    hipLaunchKernelGGL(k_adj_setup, dim3((Np + RC_AS - 1) / RC_AS), b, 0, stream, Np, s->nl, w->mat, alb, I, w->G, w->A, w->E);
    for (int k = 0; k < B; ++k) {
        const uint32_t sb = rc_hash32(seed + (uint32_t)k * 0x9E3779B9u);
        if (ev) { const hipError_t e = hipEventRecord(ev[2 * k], stream); if (e != hipSuccess) return e; }
        RC_ADJ_T(k_adj_fwd, *s, *st, w->A, w->E, w->SG, k > 0 ? w->L + (size_t)(k - 1) * Np : (const float*)nullptr,
                 k > 0 ? w->S[(k - 1) & 1] : (const uint64_t*)nullptr, w->L + (size_t)k * Np, w->S[k & 1], sb, K);
        if (ev) { const hipError_t e = hipEventRecord(ev[2 * k + 1], stream); if (e != hipSuccess) return e; }
    }
    hipLaunchKernelGGL(k_adj_out, g, b, 0, stream, Np, w->A, w->E, B > 0 ? w->L + (size_t)(B - 1) * Np : (const float*)nullptr, w->SG,
                       B > 0 ? w->S[(B - 1) & 1] : (const uint64_t*)nullptr, R, w->SR);
    return hipGetLastError();
}

hipError_t rc_launch_adj_backward(int trav, const rc_dscene* s, const rc_dstate* st, const rc_adjbuf* w, const float* gR, int K,
                                  int B, uint32_t seed, float* dI, float* dalb, hipStream_t stream, hipEvent_t* ev) {
    const int Np = s->Np, nl = s->nl;
    const dim3 g(rc_grid(Np)), b(RC_TB);
    const size_t shm = rc_shm(*st, *s, trav, false);
    hipLaunchKernelGGL(k_adj_bwd_init, g, b, 0, stream, Np, gR, w->A, w->E, B > 0 ? w->L + (size_t)(B - 1) * Np : (const float*)nullptr,
                       w->dA, w->dE, w->lam[B & 1]);
    for (int k = B - 1; k >= 0; --k) {
        const uint32_t sb = rc_hash32(seed + (uint32_t)k * 0x9E3779B9u);
        // the hit weights S collect in the buffer that then holds the adjoint of L^(k) (at k = 0,
        // where that adjoint is not needed, in the same buffer as scratch); it starts at zero
        // with a fill kernel (no memset node in a captured graph)
        float* S = w->lam[k & 1];
        hipLaunchKernelGGL(k_adj_zero, g, b, 0, stream, Np, S);
        if (ev) { const hipError_t e = hipEventRecord(ev[2 * k], stream); if (e != hipSuccess) return e; }
        RC_ADJ_T(k_adj_bwd, *s, *st, w->lam[(k + 1) & 1], S, sb, K);
        if (ev) { const hipError_t e = hipEventRecord(ev[2 * k + 1], stream); if (e != hipSuccess) return e; }
        hipLaunchKernelGGL(k_adj_bwd_fin, g, b, 0, stream, Np, w->A, w->E, k > 0 ? w->L + (size_t)(k - 1) * Np : (const float*)nullptr,
                           (const float*)S, k > 0 ? S : (float*)nullptr, w->dA, w->dE);
    }
    const int nblk = rc_grid(Np);
    hipLaunchKernelGGL(k_adj_dI_part, dim3(nblk), b, 0, stream, Np, nl, w->dE, w->G, w->part);
    hipLaunchKernelGGL(k_adj_dI_sum, dim3(nl), b, 0, stream, nblk, nl, w->part, dI);
    hipLaunchKernelGGL(k_adj_dalb_tile, dim3(nblk), b, 0, stream, Np, w->dA, w->perm, w->pmat, w->tail);
    hipLaunchKernelGGL(k_adj_dalb, dim3(rc_grid((long long)w->nmat * 64)), b, 0, stream, w->nmat, w->moff, w->tail, dalb);
    return hipGetLastError();
}
#undef RC_ADJ_T

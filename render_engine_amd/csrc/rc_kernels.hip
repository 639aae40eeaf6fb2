// Frame kernels of the dependency-aware radiance cache for gfx950 (wave64).
//
// One thread per cache entry for the entry passes (build / trace / update / event kernels),
// one thread per pixel for the camera query pass. Reductions that feed control flow use
// int64 fixed point so that the result does not depend on the order waves finish in;
// wave-level partial sums use the full 64-lane width.
#include "rc_device.h"
#include "rc_kernels.h"
#include "rc_trace.h"

// ------------------------------------------------------------------------------ rays (kernel ABI)
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_trace_rays(rc_dscene s, rc_dstate st, const float* __restrict__ O,
                                                     const float* __restrict__ D, int n, float* tout, int* qout) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t; int q;
    rc_trace_full<TRAV>(ld3(O, i), ld3(D, i), sh, shq, st.nt, s, t, q, lstk);
    tout[i] = t; qout[i] = q;
}

// ------------------------------------------------------------------------------ state build
// Direct light per entry and light, static transit mask of the shadow segments and the
// entry's light-visibility mask (cuda/poc_gpu.cu:369-391). Directional lights (the sun) use
// no distance falloff and a shadow segment of length s.sun_len.
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_build(rc_dscene s, rc_dstate st) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.Np) return;
    f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    uint32_t m[RC_MAX_NW] = {0u, 0u, 0u, 0u};
    uint64_t lm = 0ull;
    for (int l = 0; l < s.nl; ++l) {
        float r; f3 D;
        if (s.ltype[l] == 1) { D = ld3(s.L, l); r = s.sun_len; }
        else { f3 dv = ld3(s.L, l) - P; r = sqrtf(dot3(dv, dv)); D = dv * (1.f / r); }
        float c = dot3(N, D), t; int q;
        rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
        bool vis = t >= r - 1e-3f;
        float e;
        if (s.ltype[l] == 1) e = st.inten[l] * fmaxf(c, 0.f) * (vis ? 1.f : 0.f);
        else { float rr = fmaxf(r, 0.05f); e = st.inten[l] * fmaxf(c, 0.f) / (rr * rr) * (vis ? 1.f : 0.f); }
        st.Edir[(int64_t)s.nl * i + l] = e;
        f3 end = O + D * fminf(t, r);
        if (vis && c > 0.f) {
            lm |= 1ull << (l < 63 ? l : 63);
            if (s.fg > 0) rc_dda_volume_mark(s, s.fg, s.fg, s.fg, O, end, st.VOL + (int64_t)l * (((int64_t)s.fg * s.fg * s.fg + 31) >> 5));
        }
        rc_rast_seg(s, O, end, m);
    }
    for (int w = 0; w < s.nw; ++w) st.SV[(int64_t)s.nw * i + w] = m[w];
    st.LM[i] = lm;
}

// ------------------------------------------------------------------------------ frame: trace
__device__ __forceinline__ float rc_est_of(const float* Lc, int64_t i, int nl, uint64_t dormant) {
    float e = 0.f;
    for (int l = 0; l < nl; ++l) if (!(l < 64 && ((dormant >> l) & 1ull))) e += Lc[(int64_t)nl * i + l];
    return e;
}

// Per-entry sum of the light channels over the active lights, in light order (rc_est_of's sum, bit for
// bit) for every entry: a block of RC_TB threads stages the contiguous Lc span of 64 entries in LDS with
// coalesced 16-byte loads (the span starts at a multiple of 256 nl bytes; rows padded to an odd pitch:
// conflict-free row reads), then one thread per entry adds its row. Replaces one lane per entry walking
// its own row, whose loads touch 64 rows at once. (256 entries per block, every thread summing,
// measured 2.5 times slower: its 66 KB of LDS leave 8 waves per CU for the serial sums; 4-byte loads:
// 103 us against 1M entries x 65 lights.)
#define RC_RSE 64
__global__ void __launch_bounds__(RC_TB) k_rowsum(const float* __restrict__ Lc, int Np, int nl, uint64_t dormant, float* __restrict__ out) {
    extern __shared__ float s_r[];   // [RC_RSE][nl | 1]
    const int t = threadIdx.x, base = blockIdx.x * RC_RSE;
    const int T = min(RC_RSE, Np - base), pitch = nl | 1, nel = T * nl, nv = nel >> 2;
    const float* __restrict__ src = Lc + (int64_t)base * nl;
    {   // element 4 v + q of the span is light l of entry j (rows of nl)
        const int dq = 4 * RC_TB / nl, dr = 4 * RC_TB - dq * nl;
        int j = 4 * t / nl, l = 4 * t - (4 * t / nl) * nl;
        for (int v = t; v < nv; v += RC_TB) {
            const float4 x = ((const float4*)src)[v];
            int jj = j, ll = l;
            s_r[jj * pitch + ll] = x.x; if (++ll == nl) { ll = 0; ++jj; }
            s_r[jj * pitch + ll] = x.y; if (++ll == nl) { ll = 0; ++jj; }
            s_r[jj * pitch + ll] = x.z; if (++ll == nl) { ll = 0; ++jj; }
            s_r[jj * pitch + ll] = x.w;
            j += dq; l += dr;
            if (l >= nl) { l -= nl; ++j; }
        }
        const int e = 4 * nv + t;   // the last nel % 4 elements (a partial last block)
        if (e < nel) s_r[(e / nl) * pitch + e % nl] = src[e];
    }
    __syncthreads();
    if (t < T) {
        float e = 0.f;
        const float* row = s_r + t * pitch;
        for (int k = 0; k < nl; ++k) if (!rc_is_dormant(dormant, k)) e += row[k];
        out[base + t] = e;
    }
}

// k rays per entry, cache read at the hit (multi-bounce through the cache), per-light sample
// means in c.m, dependency masks (cuda/poc_gpu.cu:433-530). The per-light accumulator is the
// thread's own row of c.m (rays outer, lights inner) so any light count runs in one kernel.
// Light-dependency masks (deps on): own visible lights | hit's visible lights | hit's recorded
// light deps (cache-terminated path inherits them), staged in lnew and merged by k_update.
template <bool DEPS, int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_trace2(rc_dscene s, rc_dstate st, rc_dcache c, rc_params p, uint64_t* lnew) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.Np) return;
    const int nl = p.nl;
    uint32_t rs = rc_seed(p.seed, p.frame, (uint32_t)i);
    int k;
    bool pend = p.pend_left > 0 && c.pending[i];
    if (p.mode == 0) {
        float kb = floorf(p.budget);
        k = (int)kb + (rc_rnd(rs) < p.budget - kb ? 1 : 0);
    } else {
        float sumpr = (float)((double)c.sumpr[0] * (1.0 / 4294967296.0));
        float pr = 1.f / (c.n[i] + 1.f);
        float rrest = (float)(p.Np - c.npend[0]);
        float e = rrest * (0.5f / (float)p.Np + 0.5f * pr / sumpr) * p.rays;
        float eb = floorf(e);
        k = (pend ? 1 : 0) + (int)eb + (rc_rnd(rs) < e - eb ? 1 : 0);
    }
    f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    uint32_t m[RC_MAX_NW] = {0u, 0u, 0u, 0u};
    uint64_t lm = DEPS ? st.LM[i] : 0ull;
    float c0 = pend ? c.c0[i] : 0.f;
    int rg = pend ? s.region[i] : 0;
    float sd = 0.f, sq = 0.f;
    float* mcol = c.m + i;   // SoA: light l of entry i at mcol[l * mpitch], coalesced across lanes
    const int64_t ms = p.mpitch;
    for (int r = 0; r < k; ++r) {
        float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
        f3 D = rc_cosine_dir(N, u1, u2);
        float t; int q;
        bool hit = rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
        f3 H = hit ? O + D * t : O;
        int pi;
        bool back = rc_resolve_hit(s, st, hit, H, D, q, i, pi);
        float val = 0.f;
        if (!back) {
            float A = s.Alb[pi];
            const float* ed = st.Edir + (int64_t)nl * pi;
            const float* lc = c.Lc + (int64_t)nl * pi;
            // the hit entry's rows read 16 B at a time (rows are only 4 B aligned: nl floats each)
#define RC_ACC1(l_, e_, c_)                                                  \
            {                                                                \
                const float lv = rc_is_dormant(p.dormant, l_) ? 0.f : (c_);  \
                const float v = A * ((e_) + lv);                             \
                if (r == 0) mcol[(l_) * ms] = v; else mcol[(l_) * ms] += v;  \
                val += v;                                                    \
            }
            int l = 0;
            for (; l + 4 <= nl; l += 4) {
                const rc_f4u e4 = *(const rc_f4u*)(ed + l), c4 = *(const rc_f4u*)(lc + l);
                RC_ACC1(l, e4.x, c4.x) RC_ACC1(l + 1, e4.y, c4.y) RC_ACC1(l + 2, e4.z, c4.z) RC_ACC1(l + 3, e4.w, c4.w)
            }
            for (; l < nl; ++l) RC_ACC1(l, ed[l], lc[l])
#undef RC_ACC1
        } else if (r == 0) {
            for (int l = 0; l < nl; ++l) mcol[l * ms] = 0.f;
        }
        if (DEPS) {
            int src = back ? i : pi;
            const uint32_t* sv = st.SV + (int64_t)s.nw * src;
#pragma unroll
            for (int w = 0; w < RC_MAX_NW; ++w) if (w < s.nw) m[w] |= sv[w];
            rc_dda_reg(s, O, H, m);
            if (!back) lm |= st.LM[pi] | c.lcur[pi] | c.lprev[pi];
        }
        if (pend) { float dv = val - c0; sd += dv; sq += dv * dv; }
    }
    if (pend && k > 0) {
        atomicAdd((unsigned long long*)&c.rstat[4 * rg + 0], (unsigned long long)rc_fx(sd, RC_FX_STAT));
        atomicAdd((unsigned long long*)&c.rstat[4 * rg + 1], (unsigned long long)rc_fx(sq, RC_FX_STAT));
        atomicAdd((unsigned long long*)&c.rstat[4 * rg + 2], (unsigned long long)rc_fx(c0 * (float)k, RC_FX_STAT));
        atomicAdd((unsigned long long*)&c.rstat[4 * rg + 3], (unsigned long long)rc_fx((float)k, RC_FX_STAT));
    }
    c.kk[i] = k;
    float ik = k > 0 ? 1.f / (float)k : 0.f;
    float sa = 0.f;
    if (p.ddgi && k > 0)
        for (int l = 0; l < nl; ++l) sa += mcol[l * ms];   // sum of the per-light accumulators
    if (k == 0) {
        for (int l = 0; l < nl; ++l) mcol[l * ms] = 0.f;
    } else if (k > 1) {
        for (int l = 0; l < nl; ++l) mcol[l * ms] = mcol[l * ms] * ik;
    }
    if (p.ddgi && k > 0) {
        int rgd = s.region[i];
        atomicAdd((unsigned long long*)&c.dstat[3 * rgd + 0], (unsigned long long)rc_fx(sa * ik, RC_FX_STAT));
        atomicAdd((unsigned long long*)&c.dstat[3 * rgd + 1], (unsigned long long)rc_fx(rc_est_of(c.Lc, i, nl, p.dormant), RC_FX_STAT));
        atomicAdd((unsigned long long*)&c.dstat[3 * rgd + 2], (unsigned long long)rc_fx(1.f, RC_FX_STAT));
    }
    if (DEPS) {
        if (k > 0) {
#pragma unroll
            for (int w = 0; w < RC_MAX_NW; ++w) {
                if (w >= s.nw) break;
                uint32_t* cw = c.cur + (int64_t)w * p.Np + i;
                uint32_t old = *cw;
                if ((old | m[w]) != old) *cw = old | m[w];
            }
        }
        lnew[i] = k > 0 ? lm : 0ull;
    }
}

// Per-block counts of the event kernels (blkcnt[block]) added in block order by one block into
// *out (= the previous *out when add, else overwritten): no same-address atomics per entry.
__device__ __forceinline__ void rc_block_count(bool flag, int* blkcnt) {
    __shared__ int s_c[RC_TB / 64];
    const unsigned long long b = __ballot(flag);
    if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < RC_TB / 64; ++k) t += s_c[k];
        blkcnt[blockIdx.x] = t;
    }
}
// the same with a count per thread
__device__ __forceinline__ void rc_block_add(int v, int* blkcnt) {
    __shared__ int s_c[RC_TB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < RC_TB / 64; ++k) t += s_c[k];
        blkcnt[blockIdx.x] = t;
    }
}
#define RC_CSB 1024
__global__ void __launch_bounds__(RC_CSB) k_count_sum(const int* __restrict__ blkcnt, int nblk, int* out, int add) {
    __shared__ long long s_w[RC_CSB / 64];
    long long v = 0;
    for (int j = threadIdx.x; j < nblk; j += RC_CSB) v += blkcnt[j];
    v = rc_wave_sum_i64(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < RC_CSB / 64; ++k) t += s_w[k];
        *out = (int)(add ? *out + t : t);
    }
}

// verification decision per pending entry (cuda/poc_gpu.cu:532-555); entries still pending are counted
// per block (blkcnt), then added into npend[1] by k_count_sum
__global__ void __launch_bounds__(RC_TB) k_decide(rc_dscene s, rc_dcache c, rc_params p, int* blkcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool still = false;
    if (i < p.Np && c.pending[i]) {
        const long long* r4 = c.rstat + 4 * s.region[i];
        const float isc = 1.0f / RC_FX_STAT;
        float cnt = (float)r4[3] * isc;
        bool done = false;
        if (cnt >= 8.f) {
            float mean = ((float)r4[0] * isc) / cnt;
            float se = sqrtf(fmaxf(((float)r4[1] * isc) / cnt - mean * mean, 0.f) / cnt);
            float base = ((float)r4[2] * isc) / cnt;
            float rel = mean / base;
            if (fabsf(mean) > 2.f * se && fabsf(rel) > 0.03f) {
                float tgt = fminf(fmaxf(0.5f / fabsf(rel), 1.f) * p.rays, p.nmax);
                c.n[i] = fminf(c.n[i], tgt);
                c.pending[i] = 0;
                done = true;
            }
        }
        if (!done) {
            if (p.pend_left <= 1) c.pending[i] = 0;
            else still = true;
        }
    }
    rc_block_count(still, blkcnt);
}

#define RC_UT 64   // entries per k_update block
#define RC_UE ((RC_UT * RC_MAX_LIGHTS + RC_TB - 1) / RC_TB)   // channel elements per thread at most

// EMA update of the per-light channels + history count (cuda/poc_gpu.cu:562-602), with the
// light-dependency merge of this frame's samples, the next frame's sum of 1/(n+1), and the
// resolved radiance R = A * sum_l (Edir_l + Lc_l) of the updated entry (the value the query
// pass reads, one float instead of 2 NL). A block owns RC_UT entries, whose AoS channel rows
// form one contiguous span: every thread first issues the loads of its elements of that span
// (Lc, Edir) and of the tile's rows of the SoA sample means (rows of pitch p.mpitch, a multiple of
// RC_UT, so every row's tile is 256-byte aligned), so the block waits for memory once; the per-entry
// factors are computed by the first wave meanwhile, then the blend runs over the span in flat
// order (coalesced Lc2 stores). The means are staged in LDS at pitch RC_UT + 1: the flat order's
// lanes read consecutive lights of one entry, which then fall in distinct banks. The resolved terms
// A (Edir + Lc) then take the means' place in LDS (one more barrier) for the per-entry sums: half the
// LDS of separate arrays, so the registers (5 blocks per CU at 65 lights) rather than LDS (4) bound the
// blocks in flight (237 -> 231 us at 1M entries x 65 lights).
// UE: channel elements per thread (ceil(RC_UT * nl / RC_TB) at most), instantiated for a few
// light counts so that small ones keep few registers (more waves in flight)
template <int UE>
__global__ void __launch_bounds__(RC_TB) k_update(rc_dscene s, rc_dstate st, rc_dcache c, rc_params p, const uint64_t* lnew) {
    __shared__ float s_al[RC_UT], s_alr[RC_UT], s_A[RC_UT];
    extern __shared__ float sm[];                 // [nl][RC_UT + 1] sample means, then [RC_UT][nl] terms
    const int nl = p.nl, Np = p.Np;
    const int base = blockIdx.x * RC_UT;
    const int T = min(RC_UT, Np - base);
    const int t = threadIdx.x;
    const int nel = T * nl;
    float* s_m = sm;
    float* s_v = sm;   // after the blend
    const int64_t g0 = (int64_t)base * nl;
    const float* __restrict__ Lc = c.Lc + g0;
    const float* __restrict__ Ed = st.Edir + g0;
    float* __restrict__ Lc2 = c.Lc2 + g0;
    float ra[UE], re[UE], rm[UE];
#pragma unroll
    for (int u = 0; u < UE; ++u) {
        const int e = t + u * RC_TB;
        if (e < nel) { ra[u] = Lc[e]; re[u] = Ed[e]; }
    }
#pragma unroll
    for (int u = 0; u < UE; ++u) {
        const int e = t + u * RC_TB;   // s_m element: light e / RC_UT, entry e % RC_UT
        if (e < nl * RC_UT) rm[u] = c.m[(int64_t)(e / RC_UT) * p.mpitch + base + e % RC_UT];   // past Np: the padded row
    }
    if (t < RC_UT) {
        long long pr = 0;
        if (t < T) {
            const int i = base + t;
            int k = c.kk[i];
            float n = c.n[i], nn = n + (float)k;
            float nmx = p.nmax;
            if (p.ramp) { float a = c.age[i]; nmx = fminf(p.nmax, (4.f + 0.5f * a) * p.rays); c.age[i] = fminf(a + 1.f, 1e4f); }
            float alpha = k > 0 ? fminf((float)k / fminf(fmaxf(nn, 1.f), nmx), 1.f) : 0.f;
            float alpha_r = k > 0 ? fminf((float)k / fminf(fmaxf(nn, 1.f), p.nmax_resp), 1.f) : 0.f;
            if (p.ddgi) {
                const long long* d3 = c.dstat + 3 * s.region[i];
                const float isc = 1.0f / RC_FX_STAT;
                float scale = 1.f;
                float h = p.ddgi_h;
                float d2 = (float)d3[2] * isc;
                if (d2 > 0.f) {
                    float mn = ((float)d3[0] * isc) / d2, mc = ((float)d3[1] * isc) / d2;
                    float rel = fabsf(mn - mc) / fmaxf(mc, 1e-3f);
                    if (rel > p.ddgi_thr) h = fmaxf(0.f, h - 0.75f);
                    if (p.ddgi_clamp && rel > 0.10f) scale = 0.25f;
                }
                alpha = k > 0 ? (1.f - h) * scale : 0.f;   // DDGI blend weight
            }
            s_al[t] = alpha; s_alr[t] = alpha_r; s_A[t] = s.Alb[i];
            n = fminf(nn, nmx);
            c.n[i] = n;
            pr = rc_fx(1.f / (n + 1.f), RC_FX_PR);
            if (lnew) { uint64_t x = lnew[i]; if (x) c.lcur[i] |= x; }
        }
        pr = rc_wave_sum_i64(pr);
        if (t == 0) c.prpart[blockIdx.x] = pr;   // summed by k_prsum (no same-address atomics)
    }
#pragma unroll
    for (int u = 0; u < UE; ++u) {
        const int e = t + u * RC_TB;
        if (e < nl * RC_UT) s_m[e + e / RC_UT] = rm[u];
    }
    __syncthreads();
    // flat element e = j * nl + l of the span; (j, l) advanced by RC_TB per step
    const int dq = RC_TB / nl, dr = RC_TB - dq * nl;
    int j = t / nl, l = t - (t / nl) * nl;
#pragma unroll
    for (int u = 0; u < UE; ++u) {
        const int e = t + u * RC_TB;
        if (e < nel) {
            const float a = ra[u];
            const float mm = s_m[l * (RC_UT + 1) + j];
            const bool dm = rc_is_dormant(p.dormant, l);
            float o;
            if (p.ddgi) o = dm ? a : a + s_al[j] * (mm - a);
            else {
                float al = (l < 64 && ((p.resp_mask >> l) & 1ull)) ? s_alr[j] : s_al[j];
                if (dm) al = 0.f;
                o = a + al * (mm - a);
            }
            Lc2[e] = o;
            const float lv = dm ? 0.f : o;
            ra[u] = s_A[j] * (re[u] + lv);
        }
        j += dq; l += dr;
        if (l >= nl) { l -= nl; ++j; }
    }
    __syncthreads();   // every mean read: the terms overwrite them
#pragma unroll
    for (int u = 0; u < UE; ++u) {
        const int e = t + u * RC_TB;
        if (e < nel) s_v[e] = ra[u];
    }
    __syncthreads();
    if (t < T) {
        float val = 0.f;
        for (int k = 0; k < nl; ++k) val += s_v[t * nl + k];
        c.R[base + t] = val;
    }
}

// resolved radiance from the current channels (entries not yet updated this run)
__global__ void __launch_bounds__(RC_TB) k_resolve(rc_dscene s, rc_dstate st, const float* Lc, float* R, int nl, uint64_t dormant) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.Np) return;
    float A = s.Alb[i], val = 0.f;
    const float* ed = st.Edir + (int64_t)nl * i;
    const float* lc = Lc + (int64_t)nl * i;
    for (int l = 0; l < nl; ++l) {
        float lv = rc_is_dormant(dormant, l) ? 0.f : lc[l];
        val += A * (ed[l] + lv);
    }
    R[i] = val;
}

// Fixed-point sum of 1/(n+1) over all entries, in two passes: per-block partials, then one
// block adds them (k_prsum). Integer addition is exact, so the sum does not depend on the
// order; it replaces one same-address device atomic per wave, which this multi-die part
// serialises (~12 ns each: 100 us for 8k of them).
__global__ void __launch_bounds__(RC_TB) k_sumpr(rc_dcache c, int Np) {
    __shared__ long long s_w[RC_TB / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    long long pr = i < Np ? rc_fx(1.f / (c.n[i] + 1.f), RC_FX_PR) : 0;
    pr = rc_wave_sum_i64(pr);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = pr;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long v = 0;
        for (int k = 0; k < RC_TB / 64; ++k) v += s_w[k];
        c.prpart[blockIdx.x] = v;
    }
}

#define RC_PRB 1024
__global__ void __launch_bounds__(RC_PRB) k_prsum(const long long* __restrict__ part, int n, long long* out0, long long* out1) {
    __shared__ long long s_w[RC_PRB / 64];
    long long v = 0;
    for (int j = threadIdx.x; j < n; j += RC_PRB) v += part[j];
    v = rc_wave_sum_i64(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < RC_PRB / 64; ++k) t += s_w[k];
        *out0 = t;
        if (out1) *out1 = t;
    }
}

// restored snapshot: full confidence, optionally verified (cuda/poc_gpu.cu:612-617); with verify every
// entry becomes pending, so the pending count is Np
__global__ void __launch_bounds__(RC_TB) k_restored(rc_dcache c, int Np, int nl, float nmax, uint64_t dormant, int verify) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    c.n[i] = nmax; c.age[i] = 1e4f; c.pending[i] = 0;
    if (verify) {
        c.pending[i] = 1; c.c0[i] = c.rsum[i];   // k_rowsum of the restored channels
        if (i == 0) c.npend[1] = Np;
    }
}

// light-intensity change a -> b on channel l: linear rescale (cuda/poc_gpu.cu:637-640)
__global__ void __launch_bounds__(RC_TB) k_rescale(rc_dcache c, int Np, int nl, int l, float ratio) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Np) c.Lc[(int64_t)nl * i + l] *= ratio;
}

// Event invalidation from the transit-grid banks (cuda/poc_gpu.cu:642-660): an entry is touched when
// an event voxel's bit is set in cur | prev. The voxel list is given as a per-word mask vm (bits of the
// nw transit words), so each entry reads its nw words of both banks once (SoA, coalesced).
// kind 0: emitter appeared -> reset history; 1: occluder -> verification (count of pending entries into
// blkcnt); 2: stale marking only.
__global__ void __launch_bounds__(RC_TB) k_touched(rc_dcache c, int Np, int nl, rc_vmask vm, int nw, int kind, uint64_t dormant,
                                                  int* blkcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool t = false;
    if (i < Np) {
        uint32_t hv = 0;
#pragma unroll
        for (int w = 0; w < RC_MAX_NW; ++w)
            if (w < nw && vm.w[w]) hv |= (c.cur[(int64_t)w * Np + i] | c.prev[(int64_t)w * Np + i]) & vm.w[w];
        t = hv != 0;
        if (t) {
            if (kind == 0) { c.n[i] = 0.f; c.age[i] = 0.f; }
            else if (kind == 1) { c.pending[i] = 1; c.c0[i] = c.rsum[i]; }   // k_rowsum of the channels
            if (kind >= 1 && dormant) c.stale[i] |= dormant;
        }
    }
    if (kind == 1) rc_block_count(t, blkcnt);
}

// light wakes from dormancy (cuda/poc_gpu.cu:663-674); with verify the entries' channel sums come from
// k_rowsum (c.rsum) and the verified entries are counted per block
__global__ void __launch_bounds__(RC_TB) k_wake(rc_dcache c, int Np, int nl, int l, int verify, uint64_t dormant, int* blkcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool pend = false;
    if (i < Np && ((c.stale[i] >> l) & 1ull)) {
        if (verify) {
            c.pending[i] = 1;
            c.c0[i] = c.rsum[i];
            pend = true;
        } else { c.Lc[(int64_t)nl * i + l] = 0.f; c.n[i] = 0.f; c.age[i] = 0.f; }
        c.stale[i] &= ~(1ull << l);
    }
    if (verify) rc_block_count(pend, blkcnt);
}

// informed upper bound: shrink history by the true reference change (cuda/poc_gpu.cu:676-681)
__global__ void __launch_bounds__(RC_TB) k_oracle_method(rc_dcache c, int Np, const float* rold, const float* rnew) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Np) return;
    float rel = fabsf(rnew[i] - rold[i]) / (rold[i] + 1e-2f);
    c.n[i] *= 1.f - fminf(fmaxf(4.f * rel, 0.f), 1.f);
}

// Dependency-targeted invalidation (resolve-time, rtxgi_bench.patch:1028-1050): an entry whose
// recorded dependencies intersect the event (changed lights inv_lights | *inv_extra, or event voxels
// vm in the transit banks) restarts from zero; an entry with no recorded dependency is reset too
// (unknown -> conservative). ga / gb (may be NULL): per-entry gradient-support masks of the two
// states, used instead of the recorded light masks (an entry with no support depends on no light).
// A block owns RC_TB consecutive entries: one thread per entry decides (light masks, then the transit
// words only when needed), then the block zeroes the light rows of its reset entries over the tile's
// contiguous span of Lc in flat order (coalesced stores). Reset entries are counted per block.
__global__ void __launch_bounds__(RC_TB) k_invalidate(rc_dcache c, int Np, int nl, uint64_t inv_lights,
                                                     const unsigned long long* inv_extra, rc_vmask vm, int nw, const uint64_t* ga,
                                                     const uint64_t* gb, int* blkcnt) {
    __shared__ uint8_t s_hit[RC_TB];
    const int t = threadIdx.x, base = blockIdx.x * RC_TB, i = base + t;
    const uint64_t inv = inv_lights | (inv_extra ? (uint64_t)*inv_extra : 0ull);
    uint32_t vany = 0;
#pragma unroll
    for (int w = 0; w < RC_MAX_NW; ++w) vany |= w < nw ? vm.w[w] : 0u;
    bool hit = false;
    if (i < Np && (inv != 0ull || vany != 0u)) {   // an event with no changed light and no voxel resets nothing
        const uint64_t ld = ga ? (ga[i] | gb[i]) : (c.lcur[i] | c.lprev[i]);
        const bool known_l = ga != nullptr || ld != 0ull;
        hit = (ld & inv) != 0ull;
        if (!hit && (!known_l || vany)) {
            uint32_t any = 0, hv = 0;
#pragma unroll
            for (int w = 0; w < RC_MAX_NW; ++w) {
                if (w >= nw) break;
                const uint32_t x = c.cur[(int64_t)w * Np + i] | c.prev[(int64_t)w * Np + i];
                any |= x; hv |= x & vm.w[w];
            }
            hit = (!known_l && any == 0u) || hv != 0u;
        }
        if (hit) { c.n[i] = 0.f; c.age[i] = 0.f; }
    }
    s_hit[t] = hit ? 1 : 0;
    rc_block_count(hit, blkcnt);      // its barrier also publishes s_hit
    if (!__syncthreads_or(hit)) return;
    const int T = min(RC_TB, Np - base);
    float* __restrict__ L = c.Lc + (int64_t)base * nl;
    const int nel = T * nl, dq = RC_TB / nl, dr = RC_TB - dq * nl;
    int j = t / nl, l = t - (t / nl) * nl;
    for (int e = t; e < nel; e += RC_TB) {
        if (s_hit[j]) L[e] = 0.f;
        j += dq; l += dr;
        if (l >= nl) { l -= nl; ++j; }
    }
}

// whole-cache flush (event to a new state, rtxgi_bench.patch:363-366)
// Compact snapshots. Entries with samples (n > 0) become 24-byte records in entry order: per
// block count, host prefix, then each block writes its entries at its offset (wave ballots).
__global__ void __launch_bounds__(RC_TB) k_snap_count(rc_dcache c, int Np, int* blockcnt) {
    __shared__ int s_w[RC_TB / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long m = __ballot(i < Np && c.n[i] > 0.f);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < RC_TB / 64; ++k) t += s_w[k];
        blockcnt[blockIdx.x] = t;
    }
}
__global__ void __launch_bounds__(RC_TB) k_snap_pack(rc_dcache c, int Np, int nl, const uint64_t* ekeys, const int* blockoff, rc_snaprec* out) {
    __shared__ int s_w[RC_TB / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool v = i < Np && c.n[i] > 0.f;
    const unsigned long long m = __ballot(v);
    if (lane == 0) s_w[wv] = __popcll(m);
    __syncthreads();
    if (!v) return;
    int before = 0;
    for (int k = 0; k < wv; ++k) before += s_w[k];
    const int slot = blockoff[blockIdx.x] + before + __popcll(m & ((1ull << lane) - 1ull));
    rc_snaprec r;   // radiance: the sum of all channels in light order (k_rowsum with no dormant channel)
    r.index = (uint32_t)i; r.n = c.n[i]; r.key = ekeys[i]; r.radiance = c.rsum[i]; r.age = c.age[i];
    out[slot] = r;
}
// Exclusive prefix of n per-block counts in place, their total into *total (one block, any n): the
// record offsets of the snapshot packing stay on the device.
#define RC_SCB 1024
__global__ void __launch_bounds__(RC_SCB) k_scan_blocks(int* cnt, int n, int* total) {
    __shared__ int s_w[RC_SCB / 64];
    __shared__ int s_carry;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += RC_SCB) {
        const int j = base + threadIdx.x;
        const int v = j < n ? cnt[j] : 0;
        int x = v;   // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int before = s_carry;
        for (int k = 0; k < wv; ++k) before += s_w[k];
        if (j < n) cnt[j] = before + x - v;
        __syncthreads();
        if (threadIdx.x == RC_SCB - 1) s_carry = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// restore after a flush (rows already zero): record j < *countp, if its key matches its entry, gives the
// entry its radiance in channel 0, its history and age; verify opens the verification window on it (its
// channel sum is then the radiance, or 0 when channel 0 is dormant), counted per block. The record count
// is read on the device (the grid covers the most records). A thread takes RC_SRR records a block
// apart, their loads issued together: the key check waits on the record's index, so one record per
// thread left the kernel waiting on two dependent loads at a time.
#define RC_SRR 4
__global__ void __launch_bounds__(RC_TB) k_snap_restore(rc_dcache c, int Np, int nl, const uint64_t* ekeys, const rc_snaprec* rec,
                                                       const int* countp, int verify, uint64_t dormant, int* blkcnt) {
    const int cnt = *countp;
    const int j0 = blockIdx.x * (RC_TB * RC_SRR) + threadIdx.x;
    rc_snaprec r[RC_SRR];
    uint64_t key[RC_SRR];
#pragma unroll
    for (int u = 0; u < RC_SRR; ++u) {
        const int j = j0 + u * RC_TB;
        if (j < cnt) r[u] = rec[j];
        else r[u].index = 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < RC_SRR; ++u) key[u] = r[u].index < (uint32_t)Np ? ekeys[r[u].index] : ~r[u].key;
    int npend = 0;
#pragma unroll
    for (int u = 0; u < RC_SRR; ++u) {
        if (r[u].index < (uint32_t)Np && key[u] == r[u].key) {
            const int i = (int)r[u].index;
            c.Lc[(int64_t)nl * i] = r[u].radiance;
            c.n[i] = r[u].n; c.age[i] = r[u].age; c.pending[i] = 0;
            if (verify) { c.pending[i] = 1; c.c0[i] = rc_is_dormant(dormant, 0) ? 0.f : 0.f + r[u].radiance; ++npend; }
        }
    }
    if (verify) rc_block_add(npend, blkcnt);
}

// Whole-cache flush (event to a new state, rtxgi_bench.patch:363-366): the light rows as 16-byte stores
// over the whole Lc buffer (hipMalloc'd, so 16-byte aligned), then history and age.
__global__ void __launch_bounds__(RC_TB) k_flush(rc_dcache c, int Np, int nl) {
    const int64_t nf = (int64_t)Np * nl, n4 = nf >> 2;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n4) ((float4*)c.Lc)[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    else if (k < n4 + Np) { const int64_t i = k - n4; c.n[i] = 0.f; c.age[i] = 0.f; }
    if (k < (nf & 3)) c.Lc[(n4 << 2) + k] = 0.f;
}

// Relative RMS of the estimate against the converged reference (cuda/poc_gpu.cu:683-696), in two
// passes: each block writes its FP64 partial sums (squared error, reference, count) to part[3][blocks],
// then one block adds them in block order into err[0..2]. No same-address atomics (one per wave
// serialised ~12 ns each on this multi-die part), and the same sums in every run. An entry's light
// row is read 16 bytes at a time.
__global__ void __launch_bounds__(RC_TB) k_error(rc_dcache c, int Np, int nl, const float* __restrict__ ref,
                                                const uint8_t* __restrict__ valid, uint64_t dormant, double* part) {
    __shared__ double s_w[3][RC_TB / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0, rs = 0, cnt = 0;
    if (i < Np && valid[i]) {
        const float* row = c.Lc + (int64_t)nl * i;
        float e = 0.f;
        int l = 0;
        if (!dormant) {
            for (; l + 4 <= nl; l += 4) {
                const rc_f4u v = *(const rc_f4u*)(row + l);
                e += v.x; e += v.y; e += v.z; e += v.w;
            }
        }
        for (; l < nl; ++l) if (!rc_is_dormant(dormant, l)) e += row[l];
        const float r = ref[i];
        const float d = e - r;
        d2 = (double)d * d; rs = r; cnt = 1;
    }
    d2 = rc_wave_sum_f64(d2); rs = rc_wave_sum_f64(rs); cnt = rc_wave_sum_f64(cnt);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_w[0][w] = d2; s_w[1][w] = rs; s_w[2][w] = cnt; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0;
        for (int k = 0; k < RC_TB / 64; ++k) v += s_w[threadIdx.x][k];
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = v;
    }
}
#define RC_ERB 1024
__global__ void __launch_bounds__(RC_ERB) k_error_sum(const double* __restrict__ part, int nblk, double* err) {
    __shared__ double s_w[RC_ERB / 64];
    for (int q = 0; q < 3; ++q) {
        double v = 0;
        for (int j = threadIdx.x; j < nblk; j += RC_ERB) v += part[(int64_t)q * nblk + j];
        v = rc_wave_sum_f64(v);
        if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int k = 0; k < RC_ERB / 64; ++k) t += s_w[k];
            err[q] = t;
        }
        __syncthreads();
    }
}

// One fresh-ray Jacobi iteration Lout = mean_K A (Edir + Lin) of the converged reference
// (cuda/prep_mesh.cuh:182-213).
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_ref(rc_dscene s, rc_dstate st, const float* Lin, float* Lout, float* acc,
                                              float* closec, uint32_t seed, int K, int accumulate, int count_close) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.Np) return;
    const int nl = s.nl;
    uint32_t rs = rc_hash32(seed ^ rc_hash32((uint32_t)i + 1u));
    f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    float* out = Lout + (int64_t)nl * i;
    for (int l = 0; l < nl; ++l) out[l] = 0.f;
    int close = 0;
    for (int r = 0; r < K; ++r) {
        float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
        f3 D = rc_cosine_dir(N, u1, u2);
        float t; int q;
        bool hit = rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
        if (hit && t < 0.02f) close++;
        f3 H = hit ? O + D * t : O;
        int pi;
        if (rc_resolve_hit(s, st, hit, H, D, q, i, pi)) continue;
        float A = s.Alb[pi];
        const float* ed = st.Edir + (int64_t)nl * pi;
        const float* li = Lin + (int64_t)nl * pi;
        int l = 0;
        for (; l + 4 <= nl; l += 4) {   // 16-byte reads of the 4-byte aligned rows
            const rc_f4u e4 = *(const rc_f4u*)(ed + l), i4 = *(const rc_f4u*)(li + l);
            out[l] += A * (e4.x + i4.x); out[l + 1] += A * (e4.y + i4.y);
            out[l + 2] += A * (e4.z + i4.z); out[l + 3] += A * (e4.w + i4.w);
        }
        for (; l < nl; ++l) out[l] += A * (ed[l] + li[l]);
    }
    for (int l = 0; l < nl; ++l) {
        float v = out[l] / (float)K;
        out[l] = v;
        if (accumulate) acc[(int64_t)nl * i + l] += v;
    }
    if (count_close) closec[i] += (float)close;
}

// ------------------------------------------------------------------------------ camera query

// One-light next-event estimate at H (normal n, shadow-ray origin O = H + n 1e-3): light chosen
// uniformly, weight nl. When the path goes on (more), the next bounce's cosine direction is
// drawn right after the light into dn: the order of the path's draws (light, then bounce) is
// unchanged, and H and n need not stay live across the shadow ray.
template <int TRAV, class LS>
__device__ __forceinline__ float rc_nee(const rc_dscene& s, const rc_dstate& st, const float* sh, const int* shq,
                                        f3 H, f3 n, f3 O, uint32_t& rs, uint64_t dormant, int& shadow_rays, const LS& lstk,
                                        bool more, f3& dn) {
    const int nl = s.nl;
    float u = rc_rnd(rs);
    if (more) {
        const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
        dn = rc_cosine_dir(n, u1, u2);
    }
    int l = min((int)(u * (float)nl), nl - 1);
    float I = st.inten[l];
    if (!(I > 0.f) || rc_is_dormant(dormant, l)) return 0.f;
    float r; f3 D;
    if (s.ltype[l] == 1) { D = ld3(s.L, l); r = s.sun_len; }
    else { f3 dv = ld3(s.L, l) - H; r = sqrtf(dot3(dv, dv)); D = dv * (1.f / r); }
    float c = dot3(n, D);
    if (!(c > 0.f)) return 0.f;
    ++shadow_rays;
    // the light's contribution if unblocked, computed before the shadow ray (fewer values live across it)
    float val;
    if (s.ltype[l] == 1) val = (float)nl * (I * c);
    else { const float rr = fmaxf(r, 0.05f); val = (float)nl * (I * c / (rr * rr)); }
    const float tlim = r - 1e-3f;
    if (TRAV >= 2) {   // visibility only: any hit closer than r - 1e-3 blocks the light
        if (rc_occluded_lds(O, D, sh, st.nt, tlim) || (s.nst > 0 && rc_occluded_wide<TRAV == 3>(O, D, s, tlim, lstk))) return 0.f;
    } else {
        float t; int q;
        rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
        if (!(t >= tlim)) return 0.f;
    }
    return val;
}

// One camera path: primary hit x0, one-light NEE at every explicit vertex, cosine bounces; at
// bounce `cache_bounce` the path ends on the resolved radiance R of the hit entry (the value a
// cache-terminated ray reads, A (Edir + Lc) over the active lights, cuda/poc_gpu.cu:470-478);
// R == nullptr: no cache, the path ends after `bounces` bounces.
template <int TRAV, class LS>
__device__ __forceinline__ float rc_path(const rc_dscene& s, const rc_dstate& st, const float* sh, const int* shq, const float* R,
                                         const rc_camera& cam, int px, int py, uint32_t& rs, uint64_t dormant, int& ext_rays,
                                         int& sh_rays, const LS& lstk) {
    float ju = rc_rnd(rs), jv = rc_rnd(rs);
    float sx = (2.f * ((float)px + ju) / (float)cam.W - 1.f) * cam.tanx;
    float sy = (1.f - 2.f * ((float)py + jv) / (float)cam.H) * cam.tany;
    f3 d = mk3(cam.fwd[0], cam.fwd[1], cam.fwd[2]) + mk3(cam.right[0], cam.right[1], cam.right[2]) * sx +
           mk3(cam.up[0], cam.up[1], cam.up[2]) * sy;
    d = d * (1.f / sqrtf(dot3(d, d)));
    f3 o = mk3(cam.pos[0], cam.pos[1], cam.pos[2]);
    float thr = 1.f, rad = 0.f;
    for (int b = 0; ; ++b) {
        float t; int q;
        bool hit = rc_trace_full<TRAV>(o, d, sh, shq, st.nt, s, t, q, lstk);
        ++ext_rays;
        if (!hit) break;
        f3 H = o + d * t, n;
        int e;
        if (rc_resolve_hit_n(s, st, H, d, q, e, n)) break;
        if (R && b == cam.cache_bounce) { rad += thr * R[e]; break; }
        float A = s.Alb[e];
        const bool more = b < cam.bounces;
        const f3 on = H + n * 1e-3f;   // the shadow ray's and the next bounce's origin
        f3 dn = d;
        rad += thr * (A * rc_nee<TRAV>(s, st, sh, shq, H, n, on, rs, dormant, sh_rays, lstk, more, dn));
        if (!more) break;
        thr = thr * A;
        d = dn;
        o = on;
    }
    return rad;
}

// Pixel tile of one wave in the per-path camera pass: the 64 / spp pixels of a wave (spp divides
// 64, so a power of two) form a tw x th tile (4 x 4 at 4 spp, 8 x 8 at 1 spp), tiles in row-major
// order over the image, so a wave's primary rays and the points its bounces leave from are
// compact on screen. Paths keep their row-major ids (pixel * spp + sample) for seeds and output.
struct rc_qtile { int tw, th, tx, ty; };
__host__ __device__ inline rc_qtile rc_query_tile(int W, int H, int spp) {
    const int n = 64 / spp;
    int tw = 1;
    while (tw * tw < n) tw <<= 1;
    rc_qtile t;
    t.tw = tw; t.th = n / tw;
    t.tx = (W + 2 * tw - 1) / (2 * tw) * 2; t.ty = (H + 2 * t.th - 1) / (2 * t.th) * 2;   // even: blocks of 2 x 2 tiles
    return t;
}
// tile of wave w: the 4 waves of a 256-thread block take a 2 x 2 group of tiles
__host__ __device__ inline void rc_query_tile_of(const rc_qtile& t, int w, int& tx, int& ty) {
    const int b = w >> 2, q = w & 3, bx = t.tx >> 1;
    tx = (b % bx) * 2 + (q & 1);
    ty = (b / bx) * 2 + (q >> 1);
}

// Camera pass, one thread per path (spp divides 64: the spp paths of a pixel sit in adjacent
// lanes and are summed in sample order by the pixel's first lane; waves take pixel tiles), or
// one thread per pixel looping over its samples (any spp). Both sum ((r0 + r1) + r2) + ... and
// divide by spp. Six waves per SIMD (80 VGPRs, 1 register spilled to scratch): the traversal is
// latency bound as much as issue bound, and more waves hide more of its latency (1920x1080: 2.85 ms
// at five waves without spills, 2.70 ms at six).
template <int TRAV, bool PER_PATH, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) __attribute__((amdgpu_waves_per_eu(6))) k_query(rc_dscene s, rc_dstate st, const float* R, rc_camera cam, uint32_t seed,
                                                uint32_t frame, uint64_t dormant, float* img, unsigned long long* counters) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int ext_rays = 0, sh_rays = 0;
    const int npix = cam.W * cam.H;
    if (PER_PATH) {
        const int spp = cam.spp;
        const int lane = threadIdx.x & 63, w = tid >> 6, lp = lane / spp, first = lane - lane % spp;
        const rc_qtile qt = rc_query_tile(cam.W, cam.H, spp);
        int tx, ty;
        rc_query_tile_of(qt, w, tx, ty);
        const int px = tx * qt.tw + lp % qt.tw, py = ty * qt.th + lp / qt.tw;
        const bool live = px < cam.W && py < cam.H;
        const int pix = py * cam.W + px;
        float rad = 0.f;
        if (live) {
            uint32_t rs = rc_seed(seed, frame, (uint32_t)(pix * spp + lane - first));
            rad = rc_path<TRAV>(s, st, sh, shq, R, cam, px, py, rs, dormant, ext_rays, sh_rays, lstk);
        }
        float sum = rad;
        for (int k = 1; k < spp; ++k) {
            float rk = __shfl(rad, first + k, 64);
            if (lane == first) sum += rk;
        }
        if (live && lane == first) img[pix] = sum / (float)spp;
    } else if (tid < npix) {
        float sum = 0.f;
        for (int sp = 0; sp < cam.spp; ++sp) {
            uint32_t rs = rc_seed(seed, frame, (uint32_t)(tid * cam.spp + sp));
            sum += rc_path<TRAV>(s, st, sh, shq, R, cam, tid % cam.W, tid / cam.W, rs, dormant, ext_rays, sh_rays, lstk);
        }
        img[tid] = sum / (float)cam.spp;
    }
    if (counters) {
        long long a = rc_wave_sum_i64(ext_rays), b2 = rc_wave_sum_i64(sh_rays);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[0], (unsigned long long)a);
            atomicAdd(&counters[1], (unsigned long long)b2);
        }
    }
}

// ------------------------------------------------------------------------------ learned-cache query
// The fused camera pass for the learned cache: one thread per path (any spp) runs rc_path's
// arithmetic and RNG draws, and at the cache bounce records the query (hit point, normal,
// direction back along the ray) and the path's throughput in the staged pass's workspace
// instead of reading R; the radiance so far goes to w.rad. One batched inference, the
// prediction added to the querying paths and the per-pixel sums (the staged pass's last
// kernels) finish the image, bit for bit the staged learned-cache pass's.
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_query_mlp(rc_dscene s, rc_dstate st, rc_camera cam, uint32_t seed, uint32_t frame,
                                                    uint64_t dormant, rc_qwork w, int npath) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int p, px, py;
    if (64 % cam.spp == 0) {   // waves take pixel tiles, as in k_query; p stays the row-major path id
        const int lane = threadIdx.x & 63, lp = lane / cam.spp;
        const rc_qtile qt = rc_query_tile(cam.W, cam.H, cam.spp);
        int tx, ty;
        rc_query_tile_of(qt, tid >> 6, tx, ty);
        px = tx * qt.tw + lp % qt.tw;
        py = ty * qt.th + lp / qt.tw;
        if (px >= cam.W || py >= cam.H) return;
        p = (py * cam.W + px) * cam.spp + lane % cam.spp;
    } else {
        p = tid;
        if (p >= npath) return;
        const int pix = p / cam.spp;
        px = pix % cam.W; py = pix / cam.W;
    }
    uint32_t rs = rc_seed(seed, frame, (uint32_t)p);
    float ju = rc_rnd(rs), jv = rc_rnd(rs);
    float sx = (2.f * ((float)px + ju) / (float)cam.W - 1.f) * cam.tanx;
    float sy = (1.f - 2.f * ((float)py + jv) / (float)cam.H) * cam.tany;
    f3 d = mk3(cam.fwd[0], cam.fwd[1], cam.fwd[2]) + mk3(cam.right[0], cam.right[1], cam.right[2]) * sx +
           mk3(cam.up[0], cam.up[1], cam.up[2]) * sy;
    d = d * (1.f / sqrtf(dot3(d, d)));
    f3 o = mk3(cam.pos[0], cam.pos[1], cam.pos[2]);
    float thr = 1.f, rad = 0.f;
    float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0;
    int sh_rays = 0;
    for (int b = 0; ; ++b) {
        float t; int q;
        const bool hit = rc_trace_full<TRAV>(o, d, sh, shq, st.nt, s, t, q, lstk);
        if (!hit) break;
        f3 H = o + d * t, n;
        int e;
        if (rc_resolve_hit_n(s, st, H, d, q, e, n)) break;
        if (b == cam.cache_bounce) {
            q0 = make_float4(H.x, H.y, H.z, 1.f); q1 = make_float4(n.x, n.y, n.z, __int_as_float(e)); q2 = make_float4(-d.x, -d.y, -d.z, 0.f);
            break;
        }
        const float A = s.Alb[e];
        const f3 on = H + n * 1e-3f;
        f3 dn = d;
        rad += thr * (A * rc_nee<TRAV>(s, st, sh, shq, H, n, on, rs, dormant, sh_rays, lstk, false, dn));
        if (b >= cam.bounces) break;
        thr = thr * A;
        const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
        d = rc_cosine_dir(n, u1, u2);
        o = on;
    }
    w.rad[p] = rad; w.thr[p] = thr;
    ((float4*)w.mq0)[p] = q0;
    if (q0.w != 0.f) { ((float4*)w.mq1)[p] = q1; ((float4*)w.mq2)[p] = q2; }
}

// ------------------------------------------------------------------------------ wavefront query
// The camera pass as stages over compacted ray lists: generate primary rays, trace a batch,
// shade hits (cache read, next-event candidate, bounce), trace the shadow batch (any hit),
// add the visible next-event terms, repeat per bounce, then sum each pixel's samples. Every
// path performs exactly the fused path's operations in the same order (same RNG draws, same
// float expressions), so the image is bit-identical to k_query's; only the scheduling differs.

// Staged pass state is indexed by path id p (= pixel * spp + sample). The paths are split in
// segments of RC_QB consecutive ids; after each stage the rays still alive in a segment are
// compacted, in id order, into that segment's slice of an id list (idx: extension rays, sidx:
// shadow rays; counts in seg[0..nblk) and seg[nblk..2 nblk)). Compaction stays inside a block,
// so it needs no global atomics or scan; the trace kernels run one thread per list slot and
// waves past a segment's count exit at once. Only the scheduling differs from the fused pass.
#define RC_QB 1024

// Wave-ballot compaction inside a block: rank of each lane among the block's lanes with want
// (in thread order) and the block total.
__device__ __forceinline__ int rc_block_rank(bool want, int* s_w, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long m = __ballot(want);
    if (lane == 0) s_w[wv] = __popcll(m);
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < nw; ++k) { const int c = s_w[k]; before += k < wv ? c : 0; total += c; }
    __syncthreads();
    return before + __popcll(m & ((1ull << lane) - 1ull));
}

// primary rays of every path (the jitter draws of rc_path)
__global__ void __launch_bounds__(RC_TB) k_q_gen(rc_camera cam, uint32_t seed, uint32_t frame, int npath, rc_qwork w,
                                                unsigned long long* stat) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && stat) atomicAdd(stat, (unsigned long long)npath);
    if (p >= npath) return;
    ((float4*)w.mq0)[p] = make_float4(0.f, 0.f, 0.f, 0.f);   // no learned-cache query yet
    int pix = p / cam.spp;
    int px = pix % cam.W, py = pix / cam.W;
    uint32_t rs = rc_seed(seed, frame, (uint32_t)p);
    float ju = rc_rnd(rs), jv = rc_rnd(rs);
    float sx = (2.f * ((float)px + ju) / (float)cam.W - 1.f) * cam.tanx;
    float sy = (1.f - 2.f * ((float)py + jv) / (float)cam.H) * cam.tany;
    f3 d = mk3(cam.fwd[0], cam.fwd[1], cam.fwd[2]) + mk3(cam.right[0], cam.right[1], cam.right[2]) * sx +
           mk3(cam.up[0], cam.up[1], cam.up[2]) * sy;
    d = d * (1.f / sqrtf(dot3(d, d)));
    ((float4*)w.ro)[p] = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], 0.f);
    ((float4*)w.rd)[p] = make_float4(d.x, d.y, d.z, 0.f);
    w.rs[p] = rs; w.thr[p] = 1.f; w.rad[p] = 0.f;
}

// Trace one list slot per thread: ids == nullptr is the identity list of all npath paths.
// Closest: hit (t, q) of path p into ht / hq. ANY: the shadow ray of path p; when the light is
// visible the path's next-event term is added to its radiance (one shadow ray per path per
// bounce, after the bounce's shading, so each path's additions keep the fused order).
template <bool ANY>
__global__ void __launch_bounds__(RC_TB) k_q_trace(rc_dscene s, rc_dstate st, rc_qwork w, const int* __restrict__ ids,
                                                  const int* __restrict__ segcnt, int npath) {
    const int seg = blockIdx.x / (RC_QB / RC_TB);
    const int local = (blockIdx.x % (RC_QB / RC_TB)) * RC_TB + threadIdx.x;
    const int cnt = ids ? segcnt[seg] : min(RC_QB, npath - seg * RC_QB);
    if ((blockIdx.x % (RC_QB / RC_TB)) * RC_TB >= cnt) return;   // whole block past the segment's count
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_lds(sh, st.nt, s);
    rc_load_dyn(st, sh, shq);
    if (local >= cnt) return;
    const int p = ids ? ids[seg * RC_QB + local] : seg * RC_QB + local;
    if (ANY) {
        const float4 a = ((const float4*)w.so)[p], b = ((const float4*)w.sd)[p];
        const f3 o = mk3(a.x, a.y, a.z), d = mk3(b.x, b.y, b.z);
        const float tlim = a.w;
        if (rc_occluded_lds(o, d, sh, st.nt, tlim) || (s.nst > 0 && rc_occluded_wide(o, d, s, tlim, lstk))) return;
        w.rad[p] = w.rad[p] + w.sc[p];
    } else {
        const float4 a = ((const float4*)w.ro)[p], b = ((const float4*)w.rd)[p];
        float t; int q;
        rc_trace_full<2>(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), sh, shq, st.nt, s, t, q, lstk);
        w.ht[p] = t; w.hq[p] = q;
    }
}

// Shade the hits of bounce b over a segment's list: cache read at the cache bounce, else the
// next-event candidate (shadow ray + its term, rc_nee up to its shadow ray) and the bounce ray.
// The segment's new extension and shadow lists are written in place, in id order.
__global__ void __launch_bounds__(RC_QB) k_q_shade(rc_dscene s, rc_dstate st, const float* R, rc_camera cam, uint64_t dormant, int b,
                                                  int npath, int nblk, rc_qwork w, int mlp) {
    __shared__ int s_w[RC_QB / 64];
    const int seg = blockIdx.x, k = threadIdx.x;
    const int cnt = b == 0 ? min(RC_QB, npath - seg * RC_QB) : w.seg[seg];
    int p = -1;
    if (k < cnt) p = b == 0 ? seg * RC_QB + k : w.idx[seg * RC_QB + k];
    bool emit_sh = false, emit_b = false;
    if (p >= 0) {
        const float4 a = ((const float4*)w.ro)[p], c4 = ((const float4*)w.rd)[p];
        const f3 o = mk3(a.x, a.y, a.z), d = mk3(c4.x, c4.y, c4.z);
        const float t = w.ht[p];
        const int q = w.hq[p];
        if (q != -1) {
            f3 H = o + d * t, nrm;
            int e;
            if (!rc_resolve_hit_n(s, st, H, d, q, e, nrm)) {
                float thr = w.thr[p];
                if (mlp && b == cam.cache_bounce) {   // learned cache: query point, view direction back along the ray
                    ((float4*)w.mq0)[p] = make_float4(H.x, H.y, H.z, 1.f);
                    ((float4*)w.mq1)[p] = make_float4(nrm.x, nrm.y, nrm.z, __int_as_float(e));   // w: the query's grid entry
                    ((float4*)w.mq2)[p] = make_float4(-d.x, -d.y, -d.z, 0.f);
                } else if (R && b == cam.cache_bounce) {
                    w.rad[p] = w.rad[p] + thr * R[e];
                } else {
                    uint32_t rs = w.rs[p];
                    const int nl = s.nl;
                    const float A = s.Alb[e];
                    const float u = rc_rnd(rs);
                    const int l = min((int)(u * (float)nl), nl - 1);
                    const float I = st.inten[l];
                    if (I > 0.f && !rc_is_dormant(dormant, l)) {
                        float r; f3 D;
                        if (s.ltype[l] == 1) { D = ld3(s.L, l); r = s.sun_len; }
                        else { f3 dv = ld3(s.L, l) - H; r = sqrtf(dot3(dv, dv)); D = dv * (1.f / r); }
                        const float cth = dot3(nrm, D);
                        if (cth > 0.f) {
                            const f3 O = H + nrm * 1e-3f;
                            float v;
                            if (s.ltype[l] == 1) v = (float)nl * (I * cth);
                            else { float rr = fmaxf(r, 0.05f); v = (float)nl * (I * cth / (rr * rr)); }
                            w.sc[p] = thr * (A * v);
                            ((float4*)w.so)[p] = make_float4(O.x, O.y, O.z, r - 1e-3f);
                            ((float4*)w.sd)[p] = make_float4(D.x, D.y, D.z, 0.f);
                            emit_sh = true;
                        }
                    }
                    if (b < cam.bounces) {
                        thr = thr * A;
                        const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
                        const f3 nd = rc_cosine_dir(nrm, u1, u2);
                        const f3 no = H + nrm * 1e-3f;
                        ((float4*)w.ro)[p] = make_float4(no.x, no.y, no.z, 0.f);
                        ((float4*)w.rd)[p] = make_float4(nd.x, nd.y, nd.z, 0.f);
                        w.thr[p] = thr;
                        emit_b = true;
                    }
                    w.rs[p] = rs;
                }
            }
        }
    }
    int ns, nb2;
    const int rs_ = rc_block_rank(emit_sh, s_w, ns);
    const int rb_ = rc_block_rank(emit_b, s_w, nb2);   // ranks read the old list first: p is in a register
    if (emit_sh) w.sidx[seg * RC_QB + rs_] = p;
    if (emit_b) w.idx[seg * RC_QB + rb_] = p;
    if (k == 0) { w.seg[nblk + seg] = ns; w.seg[seg] = nb2; }
}

// the learned cache's prediction ends the paths that queried it (their last addition). With a judge
// (jR), the grid cache's resolved radiance at the query's entry (stored in the query's normal row, w)
// replaces a prediction that differs from it by more than jthr times it: the grid is the fallback
// wherever it disagrees, e.g. in the region a light-off event just changed.
__global__ void __launch_bounds__(RC_TB) k_q_mlp_add(rc_qwork w, int npath, const float* __restrict__ jR, float jthr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npath || ((const float4*)w.mq0)[p].w == 0.f) return;
    float y = fmaxf(w.mpred[p], 0.f);
    if (jR) {
        const float r = jR[__float_as_int(((const float4*)w.mq1)[p].w)];
        if (fabsf(y - r) > jthr * r) y = r;
    }
    w.rad[p] = w.rad[p] + w.thr[p] * y;
}

// rays in the lists (ray counters): sum of the segment counts
__global__ void __launch_bounds__(RC_TB) k_q_count(const int* segcnt, int nblk, unsigned long long* stat) {
    __shared__ long long s_p[RC_TB / 64];
    long long v = 0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) v += segcnt[i];
    v = rc_wave_sum_i64(v);
    if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int k = 0; k < RC_TB / 64; ++k) t += s_p[k];
        atomicAdd(stat, (unsigned long long)t);
    }
}

// pixel = (((r0 + r1) + r2) + ...) / spp over the pixel's paths in sample order
__global__ void __launch_bounds__(RC_TB) k_q_pixels(rc_camera cam, rc_qwork w, float* img) {
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= cam.W * cam.H) return;
    const float* r = w.rad + (int64_t)pix * cam.spp;
    // the fused pass starts from the first sample when spp divides 64, else from 0
    const bool per_path = cam.spp <= 64 && 64 % cam.spp == 0;
    float sum = per_path ? r[0] : 0.f + r[0];
    for (int k = 1; k < cam.spp; ++k) sum += r[k];
    img[pix] = sum / (float)cam.spp;
}

// ------------------------------------------------------------------------------ hash grid insert
// Insert keys with 64-bit CAS and linear probing (same home slot mix64(key) & hmask and probe
// order as the lookup, cuda/poc_gpu.cu:263-272). Wave-level dedup by a ballot loop over the wave's
// distinct keys: the lowest lane still to be grouped names a key (a uniform shuffle), one ballot
// gives the lanes holding it, and only a group of more than one lane reduces its values; the group's
// lowest lane then inserts the group's minimum. Distinct keys cost one ballot each, duplicates cost
// nothing more to probe. The value stored for a key is the minimum inserted (atomicMin), whatever the
// arrival order; slots may differ from a sequential insertion, lookups do not. fail[0] counts keys
// that found no slot within 64 probes.
__global__ void __launch_bounds__(RC_TB) k_hash_insert(unsigned long long* HK, int* HV, uint32_t hmask,
                                                      const unsigned long long* keys, const int* vals, int n, int* fail) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool active = i < n;
    const unsigned long long key = active ? keys[i] : 0ull;
    const int val = active ? vals[i] : 0x7fffffff;
    active = active && key != 0ull;
    bool leader = false;
    int best = val;
    unsigned long long todo = __ballot(active);
    while (todo) {
        const int src = __ffsll((long long)todo) - 1;
        const unsigned long long k = __shfl(key, src, 64);
        const unsigned long long grp = __ballot(active && key == k);
        if (__popcll(grp) > 1) {   // duplicates of k in this wave: their minimum value
            int v = ((grp >> lane) & 1ull) ? val : 0x7fffffff;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
            if (lane == src) best = v;
        }
        if (lane == src) leader = true;
        todo &= ~grp;
    }
    if (!leader) return;
    uint32_t h = (uint32_t)rc_mix64(key) & hmask;
    for (int probe = 0; probe < 64; ++probe) {
        unsigned long long prev = atomicCAS(HK + h, 0ull, key);
        if (prev == 0ull || prev == key) { atomicMin(HV + h, best); return; }
        h = (h + 1) & hmask;
    }
    atomicAdd(fail, 1);
}

// slots left empty by the insert: value -1 (the host table's empty value)
__global__ void __launch_bounds__(RC_TB) k_hash_finish(const unsigned long long* HK, int* HV, uint32_t hsize) {
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < hsize && HK[h] == 0ull) HV[h] = -1;
}

__global__ void __launch_bounds__(RC_TB) k_hash_query(const unsigned long long* HK, const int* HV, uint32_t hmask,
                                                     const unsigned long long* keys, int n, int* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = rc_hash_find((const uint64_t*)HK, HV, hmask, (uint64_t)keys[i]);
}

// cache entry of hit points (position + facing normal), the lookup of the cache-read path
__global__ void __launch_bounds__(RC_TB) k_hash_point(rc_dscene s, const float* H, const float* Nn, int n, int* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = rc_hash_point(s, ld3(H, i), ld3(Nn, i));
}

// ------------------------------------------------------------------------------ DDA kernels
// Transit-grid masks of segments (a_i -> b_i), one thread per segment.
__global__ void __launch_bounds__(RC_TB) k_dda_segments(rc_dscene s, const float* A, const float* B, int n, uint32_t* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t m[RC_MAX_NW] = {0u, 0u, 0u, 0u};
    rc_dda_reg(s, ld3(A, i), ld3(B, i), m);
    for (int w = 0; w < s.nw; ++w) out[(int64_t)s.nw * i + w] = m[w];
}

// Fine voxel volume: Amanatides-Woo over a gx*gy*gz bit volume (stepping rule of the transit
// grid, cuda/poc_gpu.cu:171-190), OR-ing each crossed voxel's bit with a read-before-atomic
// so a voxel already set costs one load. vol has ceil(gx*gy*gz/32) words per volume; seg_vol
// selects the volume of each segment (e.g. its light). Steps are bounded by gx+gy+gz.
__global__ void __launch_bounds__(RC_TB) k_dda_volume(rc_dscene s, int gx, int gy, int gz, const float* A, const float* B,
                                                     const int* seg_vol, int n, uint32_t* vol, long long words_per_vol) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* V = vol + (seg_vol ? (int64_t)seg_vol[i] * words_per_vol : 0);
    rc_dda_volume_mark(s, gx, gy, gz, ld3(A, i), ld3(B, i), V);
}

// Lights affected by a geometry event: light l is affected when its fine volume in the state
// before (Va) or after (Vb) the event has a bit among the event's fine voxels. One block per
// light; the result is ORed into out[0] as bit min(l, 63).
__global__ void __launch_bounds__(RC_TB) k_vol_lights(const uint32_t* Va, const uint32_t* Vb, long long words_per_vol,
                                                     const int* vox, int nvox, unsigned long long* out) {
    const int l = blockIdx.x;
    const uint32_t* a = Va + (int64_t)l * words_per_vol;
    const uint32_t* b = Vb + (int64_t)l * words_per_vol;
    int hit = 0;
    for (int j = threadIdx.x; j < nvox && !hit; j += blockDim.x) {
        int v = vox[j];
        hit = ((a[v >> 5] | b[v >> 5]) >> (v & 31)) & 1u;
    }
    if (__syncthreads_or(hit) && threadIdx.x == 0) atomicOr(out, 1ull << (l < 63 ? l : 63));
}

// ------------------------------------------------------------------------------ launchers

hipError_t rc_launch_trace_rays(int trav, const rc_dscene* s, const rc_dstate* st, const float* O, const float* D, int n,
                                float* t, int* q, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    dim3 g(rc_grid(n)), b(RC_TB);
    size_t shm = rc_shm(*st, *s, trav, false);
    if (trav == 0) hipLaunchKernelGGL((k_trace_rays<0, false>), g, b, shm, stream, *s, *st, O, D, n, t, q);
    else if (trav == 1) hipLaunchKernelGGL((k_trace_rays<1, false>), g, b, shm, stream, *s, *st, O, D, n, t, q);
    else if (trav == 2) RC_LAUNCH_WIDE((k_trace_rays<2, true>), (k_trace_rays<2, false>), false, g, st, s, stream, *s, *st, O, D, n, t, q);
    else RC_LAUNCH_WIDE((k_trace_rays<3, true>), (k_trace_rays<3, false>), false, g, st, s, stream, *s, *st, O, D, n, t, q);
    return hipGetLastError();
}

hipError_t rc_launch_build(int trav, const rc_dscene* s, const rc_dstate* st, hipStream_t stream) {
    if (s->Np <= 0) return hipSuccess;
    dim3 g(rc_grid(s->Np)), b(RC_TB);
    size_t shm = rc_shm(*st, *s, trav, false);
    if (trav == 0) hipLaunchKernelGGL((k_build<0, false>), g, b, shm, stream, *s, *st);
    else if (trav == 1) hipLaunchKernelGGL((k_build<1, false>), g, b, shm, stream, *s, *st);
    else RC_LAUNCH_WIDE((k_build<2, true>), (k_build<2, false>), false, g, st, s, stream, *s, *st);
    return hipGetLastError();
}

hipError_t rc_launch_trace2(int trav, int deps, const rc_dscene* s, const rc_dstate* st, const rc_dcache* c, const rc_params* p,
                            uint64_t* lnew, hipStream_t stream) {
    if (p->Np <= 0) return hipSuccess;
    dim3 g(rc_grid(p->Np)), b(RC_TB);
    size_t shm = rc_shm(*st, *s, trav, false);
#define RC_T2(D_, T_) hipLaunchKernelGGL((k_trace2<D_, T_, false>), g, b, shm, stream, *s, *st, *c, *p, lnew)
    if (deps) {
        if (trav == 0) RC_T2(true, 0); else if (trav == 1) RC_T2(true, 1);
        else RC_LAUNCH_WIDE((k_trace2<true, 2, true>), (k_trace2<true, 2, false>), false, g, st, s, stream, *s, *st, *c, *p, lnew);
    } else {
        if (trav == 0) RC_T2(false, 0); else if (trav == 1) RC_T2(false, 1);
        else RC_LAUNCH_WIDE((k_trace2<false, 2, true>), (k_trace2<false, 2, false>), false, g, st, s, stream, *s, *st, *c, *p, lnew);
    }
#undef RC_T2
    return hipGetLastError();
}

static void rc_rowsum(const rc_dcache* c, int Np, int nl, uint64_t dormant, hipStream_t stream) {
    hipLaunchKernelGGL(k_rowsum, dim3((Np + RC_RSE - 1) / RC_RSE), dim3(RC_TB), (size_t)RC_RSE * (nl | 1) * sizeof(float), stream,
                       (const float*)c->Lc, Np, nl, dormant, c->rsum);
}
hipError_t rc_launch_decide(const rc_dscene* s, const rc_dcache* c, const rc_params* p, hipStream_t stream) {
    const int nblk = rc_grid(p->Np);
    hipLaunchKernelGGL(k_decide, dim3(nblk), dim3(RC_TB), 0, stream, *s, *c, *p, c->blkcnt);
    hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(RC_CSB), 0, stream, (const int*)c->blkcnt, nblk, c->npend + 1, 1);
    return hipGetLastError();
}
hipError_t rc_launch_update(const rc_dscene* s, const rc_dstate* st, const rc_dcache* c, const rc_params* p, const uint64_t* lnew,
                            hipStream_t stream) {
    if (p->Np <= 0) return hipSuccess;
    dim3 g((p->Np + RC_UT - 1) / RC_UT), b(RC_TB);
    size_t shm = (size_t)p->nl * (RC_UT + 1) * sizeof(float);
    const int ue = (RC_UT * p->nl + RC_TB - 1) / RC_TB;
    if (ue <= 4) hipLaunchKernelGGL(k_update<4>, g, b, shm, stream, *s, *st, *c, *p, lnew);
    else if (ue <= 9) hipLaunchKernelGGL(k_update<9>, g, b, shm, stream, *s, *st, *c, *p, lnew);
    else hipLaunchKernelGGL(k_update<RC_UE>, g, b, shm, stream, *s, *st, *c, *p, lnew);
    hipLaunchKernelGGL(k_prsum, dim3(1), dim3(RC_PRB), 0, stream, (const long long*)c->prpart, (int)g.x, c->sumpr + 1, c->sumpr);
    return hipGetLastError();
}
hipError_t rc_launch_resolve(const rc_dscene* s, const rc_dstate* st, const float* Lc, float* R, uint64_t dormant, hipStream_t stream) {
    if (s->Np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_resolve, dim3(rc_grid(s->Np)), dim3(RC_TB), 0, stream, *s, *st, Lc, R, s->nl, dormant);
    return hipGetLastError();
}
hipError_t rc_launch_sumpr(const rc_dcache* c, int Np, hipStream_t stream) {
    if (Np <= 0) return hipSuccess;
    const int nb = rc_grid(Np);
    hipLaunchKernelGGL(k_sumpr, dim3(nb), dim3(RC_TB), 0, stream, *c, Np);
    hipLaunchKernelGGL(k_prsum, dim3(1), dim3(RC_PRB), 0, stream, (const long long*)c->prpart, nb, c->sumpr, (long long*)nullptr);
    return hipGetLastError();
}
hipError_t rc_launch_restored(const rc_dcache* c, int Np, int nl, float nmax, uint64_t dormant, int verify, hipStream_t stream) {
    if (verify) rc_rowsum(c, Np, nl, dormant, stream);
    hipLaunchKernelGGL(k_restored, dim3(rc_grid(Np)), dim3(RC_TB), 0, stream, *c, Np, nl, nmax, dormant, verify);
    return hipGetLastError();
}
hipError_t rc_launch_rescale(const rc_dcache* c, int Np, int nl, int l, float ratio, hipStream_t stream) {
    hipLaunchKernelGGL(k_rescale, dim3(rc_grid(Np)), dim3(RC_TB), 0, stream, *c, Np, nl, l, ratio);
    return hipGetLastError();
}
hipError_t rc_launch_touched(const rc_dcache* c, int Np, int nl, rc_vmask vm, int nw, int kind, uint64_t dormant, hipStream_t stream) {
    const int nblk = rc_grid(Np);
    if (kind == 1) rc_rowsum(c, Np, nl, dormant, stream);
    hipLaunchKernelGGL(k_touched, dim3(nblk), dim3(RC_TB), 0, stream, *c, Np, nl, vm, nw, kind, dormant, c->blkcnt);
    if (kind == 1) hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(RC_CSB), 0, stream, (const int*)c->blkcnt, nblk, c->npend + 1, 1);
    return hipGetLastError();
}
hipError_t rc_launch_wake(const rc_dcache* c, int Np, int nl, int l, int verify, uint64_t dormant, hipStream_t stream) {
    const int nblk = rc_grid(Np);
    if (verify) rc_rowsum(c, Np, nl, dormant, stream);
    hipLaunchKernelGGL(k_wake, dim3(nblk), dim3(RC_TB), 0, stream, *c, Np, nl, l, verify, dormant, c->blkcnt);
    if (verify) hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(RC_CSB), 0, stream, (const int*)c->blkcnt, nblk, c->npend + 1, 1);
    return hipGetLastError();
}
hipError_t rc_launch_oracle_method(const rc_dcache* c, int Np, const float* rold, const float* rnew, hipStream_t stream) {
    hipLaunchKernelGGL(k_oracle_method, dim3(rc_grid(Np)), dim3(RC_TB), 0, stream, *c, Np, rold, rnew);
    return hipGetLastError();
}
hipError_t rc_launch_invalidate(const rc_dcache* c, int Np, int nl, uint64_t inv_lights, const unsigned long long* inv_extra, rc_vmask vm,
                                int nw, const uint64_t* ga, const uint64_t* gb, int* count, hipStream_t stream) {
    const int nblk = rc_grid(Np);
    hipLaunchKernelGGL(k_invalidate, dim3(nblk), dim3(RC_TB), 0, stream, *c, Np, nl, inv_lights, inv_extra, vm, nw, ga, gb, c->blkcnt);
    hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(RC_CSB), 0, stream, (const int*)c->blkcnt, nblk, count, 0);
    return hipGetLastError();
}
hipError_t rc_launch_flush(const rc_dcache* c, int Np, int nl, hipStream_t stream) {
    const long long n = (((long long)Np * nl) >> 2) + Np;
    hipLaunchKernelGGL(k_flush, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, *c, Np, nl);
    return hipGetLastError();
}
hipError_t rc_launch_error(const rc_dcache* c, int Np, int nl, const float* ref, const uint8_t* valid, uint64_t dormant, hipStream_t stream) {
    const int nblk = rc_grid(Np);
    hipLaunchKernelGGL(k_error, dim3(nblk), dim3(RC_TB), 0, stream, *c, Np, nl, ref, valid, dormant, c->errpart);
    hipLaunchKernelGGL(k_error_sum, dim3(1), dim3(RC_ERB), 0, stream, (const double*)c->errpart, nblk, c->err);
    return hipGetLastError();
}
hipError_t rc_launch_ref(int trav, const rc_dscene* s, const rc_dstate* st, const float* Lin, float* Lout, float* acc, float* closec,
                         uint32_t seed, int K, int accumulate, int count_close, hipStream_t stream) {
    dim3 g(rc_grid(s->Np)), b(RC_TB);
    size_t shm = rc_shm(*st, *s, trav, false);
    if (trav == 0) hipLaunchKernelGGL((k_ref<0, false>), g, b, shm, stream, *s, *st, Lin, Lout, acc, closec, seed, K, accumulate, count_close);
    else if (trav == 1) hipLaunchKernelGGL((k_ref<1, false>), g, b, shm, stream, *s, *st, Lin, Lout, acc, closec, seed, K, accumulate, count_close);
    else RC_LAUNCH_WIDE((k_ref<2, true>), (k_ref<2, false>), false, g, st, s, stream, *s, *st, Lin, Lout, acc, closec, seed, K, accumulate, count_close);
    return hipGetLastError();
}
size_t rc_ovf_words(const rc_dstate* st, const rc_dscene* s, long long nthreads) { return rc_ovf_need(*st, *s, nthreads); }
long long rc_grid_threads(long long n) { return (long long)rc_grid(n) * RC_TB; }
long long rc_query_mlp_threads(const rc_camera* cam) {
    if (64 % cam->spp == 0) {
        const rc_qtile t = rc_query_tile(cam->W, cam->H, cam->spp);
        return (long long)rc_grid((long long)t.tx * t.ty * 64) * RC_TB;
    }
    return (long long)rc_grid((long long)cam->W * cam->H * cam->spp) * RC_TB;
}
long long rc_query_threads(const rc_camera* cam) {
    const bool per_path = cam->spp <= 64 && 64 % cam->spp == 0;
    if (per_path) {   // one wave per pixel tile
        const rc_qtile t = rc_query_tile(cam->W, cam->H, cam->spp);
        return (long long)rc_grid((long long)t.tx * t.ty * 64) * RC_TB;
    }
    return (long long)rc_grid((long long)cam->W * cam->H) * RC_TB;
}
hipError_t rc_launch_query(int trav, const rc_dscene* s, const rc_dstate* st, const float* R, const rc_camera* cam, uint32_t seed,
                           uint32_t frame, uint64_t dormant, float* img, unsigned long long* counters, hipStream_t stream) {
    const bool per_path = cam->spp <= 64 && 64 % cam->spp == 0;
    dim3 g((unsigned)(rc_query_threads(cam) / RC_TB)), b(RC_TB);
    size_t shm = rc_shm(*st, *s, trav, false);
#define RC_Q(T_, P_) hipLaunchKernelGGL((k_query<T_, P_, false>), g, b, shm, stream, *s, *st, R, *cam, seed, frame, dormant, img, counters)
#define RC_QW(P_) RC_LAUNCH_WIDE((k_query<2, P_, true>), (k_query<2, P_, false>), true, g, st, s, stream, *s, *st, R, *cam, seed, frame, dormant, \
                                 img, counters)
#define RC_QW3(P_) RC_LAUNCH_WIDE((k_query<3, P_, true>), (k_query<3, P_, false>), true, g, st, s, stream, *s, *st, R, *cam, seed, frame, \
                                  dormant, img, counters)
    if (per_path) { if (trav == 0) RC_Q(0, true); else if (trav == 1) RC_Q(1, true); else if (trav == 2) RC_QW(true); else RC_QW3(true); }
    else { if (trav == 0) RC_Q(0, false); else if (trav == 1) RC_Q(1, false); else if (trav == 2) RC_QW(false); else RC_QW3(false); }
#undef RC_QW3
#undef RC_QW
#undef RC_Q
    return hipGetLastError();
}
hipError_t rc_launch_hash_finish(const unsigned long long* HK, int* HV, uint32_t hsize, hipStream_t stream) {
    hipLaunchKernelGGL(k_hash_finish, dim3(rc_grid(hsize)), dim3(RC_TB), 0, stream, HK, HV, hsize);
    return hipGetLastError();
}
hipError_t rc_launch_hash_insert(unsigned long long* HK, int* HV, uint32_t hmask, const unsigned long long* keys, const int* vals,
                                 int n, int* fail, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_insert, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, HK, HV, hmask, keys, vals, n, fail);
    return hipGetLastError();
}
hipError_t rc_launch_hash_query(const unsigned long long* HK, const int* HV, uint32_t hmask, const unsigned long long* keys, int n,
                                int* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_query, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, HK, HV, hmask, keys, n, out);
    return hipGetLastError();
}
hipError_t rc_launch_hash_point(const rc_dscene* s, const float* H, const float* Nn, int n, int* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_point, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, *s, H, Nn, n, out);
    return hipGetLastError();
}
hipError_t rc_launch_dda_segments(const rc_dscene* s, const float* A, const float* B, int n, uint32_t* out, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dda_segments, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, *s, A, B, n, out);
    return hipGetLastError();
}
hipError_t rc_launch_dda_volume(const rc_dscene* s, int gx, int gy, int gz, const float* A, const float* B, const int* seg_vol, int n,
                                uint32_t* vol, long long words_per_vol, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dda_volume, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, *s, gx, gy, gz, A, B, seg_vol, n, vol, words_per_vol);
    return hipGetLastError();
}

hipError_t rc_launch_vol_lights(int nl, const uint32_t* Va, const uint32_t* Vb, long long words_per_vol, const int* vox, int nvox,
                                unsigned long long* out, hipStream_t stream) {
    if (nl <= 0 || nvox <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_vol_lights, dim3(nl), dim3(RC_TB), 0, stream, Va, Vb, words_per_vol, vox, nvox, out);
    return hipGetLastError();
}

hipError_t rc_launch_query_mlp(const rc_dscene* s, const rc_dstate* st, const rc_camera* cam, uint32_t seed, uint32_t frame, uint64_t dormant,
                               float* img, const rc_qwork* w, const rc_mlpdev* mlp, hipStream_t stream, const hipEvent_t* infer_ev) {
    const int npath = cam->W * cam->H * cam->spp;
    if (npath <= 0) return hipSuccess;
    const dim3 g((unsigned)(rc_query_mlp_threads(cam) / RC_TB));
    RC_LAUNCH_WIDE((k_query_mlp<2, true>), (k_query_mlp<2, false>), false, g, st, s, stream, *s, *st, *cam, seed, frame, dormant, *w, npath);
    if (infer_ev) (void)hipEventRecord(infer_ev[0], stream);
    hipError_t e = rc_launch_mlp_infer(mlp, (const float4*)w->mq0, (const float4*)w->mq1, (const float4*)w->mq2, npath, w->mpred, stream);
    if (e != hipSuccess) return e;
    if (infer_ev) (void)hipEventRecord(infer_ev[1], stream);
    hipLaunchKernelGGL(k_q_mlp_add, dim3(rc_grid(npath)), dim3(RC_TB), 0, stream, *w, npath, mlp->jR, mlp->jthr);
    hipLaunchKernelGGL(k_q_pixels, dim3(rc_grid((long long)cam->W * cam->H)), dim3(RC_TB), 0, stream, *cam, *w, img);
    return hipGetLastError();
}
hipError_t rc_launch_query_staged(const rc_dscene* s, const rc_dstate* st, const float* R, const rc_camera* cam, uint32_t seed,
                                  uint32_t frame, uint64_t dormant, float* img, unsigned long long* counters, const rc_qwork* w,
                                  const rc_mlpdev* mlp, hipStream_t stream, const hipEvent_t* infer_ev) {
    const int npath = cam->W * cam->H * cam->spp;
    const size_t shm = rc_shm(*st, *s, 2, false);   // the stage kernels keep every stack level in LDS
    const int nblk = (npath + RC_QB - 1) / RC_QB;
    const dim3 tgrid(nblk * (RC_QB / RC_TB));
    hipLaunchKernelGGL(k_q_gen, dim3(rc_grid(npath)), dim3(RC_TB), 0, stream, *cam, seed, frame, npath, *w, counters);
    const int last = (R || mlp) ? min(cam->cache_bounce, cam->bounces) : cam->bounces;
    for (int b = 0; b <= last; ++b) {
        if (b > 0 && counters) hipLaunchKernelGGL(k_q_count, dim3(1), dim3(RC_TB), 0, stream, (const int*)w->seg, nblk, counters);
        hipLaunchKernelGGL(k_q_trace<false>, tgrid, dim3(RC_TB), shm, stream, *s, *st, *w, b == 0 ? (const int*)nullptr : (const int*)w->idx,
                           (const int*)w->seg, npath);
        hipLaunchKernelGGL(k_q_shade, dim3(nblk), dim3(RC_QB), 0, stream, *s, *st, R, *cam, dormant, b, npath, nblk, *w, mlp ? 1 : 0);
        if ((R || mlp) && b == cam->cache_bounce) break;   // cache read: no shadow or bounce rays follow
        if (counters) hipLaunchKernelGGL(k_q_count, dim3(1), dim3(RC_TB), 0, stream, (const int*)w->seg + nblk, nblk, counters + 1);
        hipLaunchKernelGGL(k_q_trace<true>, tgrid, dim3(RC_TB), shm, stream, *s, *st, *w, (const int*)w->sidx, (const int*)w->seg + nblk, npath);
    }
    if (mlp) {
        if (infer_ev) (void)hipEventRecord(infer_ev[0], stream);
        hipError_t e = rc_launch_mlp_infer(mlp, (const float4*)w->mq0, (const float4*)w->mq1, (const float4*)w->mq2, npath, w->mpred, stream);
        if (e != hipSuccess) return e;
        if (infer_ev) (void)hipEventRecord(infer_ev[1], stream);
        hipLaunchKernelGGL(k_q_mlp_add, dim3(rc_grid(npath)), dim3(RC_TB), 0, stream, *w, npath, mlp->jR, mlp->jthr);
    }
    hipLaunchKernelGGL(k_q_pixels, dim3(rc_grid((long long)cam->W * cam->H)), dim3(RC_TB), 0, stream, *cam, *w, img);
    return hipGetLastError();
}

hipError_t rc_launch_snap_count(const rc_dcache* c, int Np, int* blockcnt, hipStream_t stream) {
    hipLaunchKernelGGL(k_snap_count, dim3(rc_grid(Np)), dim3(RC_TB), 0, stream, *c, Np, blockcnt);
    return hipGetLastError();
}
hipError_t rc_launch_snap_pack(const rc_dcache* c, int Np, int nl, const uint64_t* ekeys, const int* blockoff, rc_snaprec* out,
                               hipStream_t stream) {
    rc_rowsum(c, Np, nl, 0ull, stream);
    hipLaunchKernelGGL(k_snap_pack, dim3(rc_grid(Np)), dim3(RC_TB), 0, stream, *c, Np, nl, ekeys, blockoff, out);
    return hipGetLastError();
}
hipError_t rc_launch_snap_restore(const rc_dcache* c, int Np, int nl, const uint64_t* ekeys, const rc_snaprec* rec, const int* count,
                                  int max_count, int verify, uint64_t dormant, hipStream_t stream) {
    if (max_count <= 0) return hipSuccess;
    const int nblk = (max_count + RC_TB * RC_SRR - 1) / (RC_TB * RC_SRR);
    hipLaunchKernelGGL(k_snap_restore, dim3(nblk), dim3(RC_TB), 0, stream, *c, Np, nl, ekeys, rec, count, verify, dormant, c->blkcnt);
    if (verify) hipLaunchKernelGGL(k_count_sum, dim3(1), dim3(RC_CSB), 0, stream, (const int*)c->blkcnt, nblk, c->npend + 1, 1);
    return hipGetLastError();
}
hipError_t rc_launch_scan_blocks(int* cnt, int n, int* total, hipStream_t stream) {
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(RC_SCB), 0, stream, cnt, n, total);
    return hipGetLastError();
}

// Hash-encoded MLP radiance cache (extension) for gfx950.
//
// Input encoding: 16-level multiresolution hash grid over the position normalised by the scene
// bounds (2 features per level, 2^log2T entries per level; levels whose (res+1)^3 grid fits the
// table are indexed densely, finer ones by the xor of the corner coordinates times 1,
// 2654435761, 805459861), trilinear interpolation; then the normal and the view direction
// mapped to [0, 1], a constant 1, zero padding: 64 inputs. Network: 5 hidden layers of 64 ReLU
// units, scalar output, no biases. Weights and activations are BF16 MFMA operands
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation); Adam keeps FP32 master weights and table.
//
// MFMA lane layout (16x16x32): lane l holds A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16]
// (j = 0..7), and D[4 (l / 16) + i][l % 16] (i = 0..3). Every product below is D = P Q^T with P and
// Q stored K-contiguous, so both operands are 16-byte loads.
#include "rc_device.h"
#include "rc_kernels.h"
#include "rc_trace.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MW 64                   // layer width
#define MLD 72                  // LDS row stride in bf16 (144-byte rows spread 16-byte reads over the banks)
#define MROWS 64                // samples per wave
#define MWAVES 4                // waves per block

__device__ __forceinline__ __bf16 rc_bf(float x) { return (__bf16)x; }
__device__ __forceinline__ float rc_fb(__bf16 x) { return (float)x; }

__device__ __forceinline__ f32x4 rc_mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
// 16 rows x 32 K of a K-contiguous bf16 matrix (row stride ld) as this lane's operand
__device__ __forceinline__ bf16x8 rc_opnd(const __bf16* M, int ld, int r0, int k0) {
    const int l = threadIdx.x & 63;
    return *(const bf16x8*)(M + (int64_t)(r0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}

// ------------------------------------------------------------------------------ encoding
__device__ __forceinline__ uint32_t rc_mlp_index(const rc_mlpdev& m, int lv, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t r1 = (uint32_t)m.res[lv] + 1u;
    if ((m.dense >> lv) & 1u) return x + r1 * (y + r1 * z);
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & ((1u << m.log2T) - 1u);
}

// Orders one wave's LDS writes before its later LDS reads. The forward and backward passes keep
// each wave's activations in a tile of its own, so no other wave waits: a wave's LDS instructions
// complete in issue order, and the fences keep the compiler from moving accesses across this point.
__device__ __forceinline__ void rc_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// corner grid of level lv for the normalised position x: base corner and fractions
__device__ __forceinline__ void rc_mlp_cell(const rc_mlpdev& m, int lv, f3 x, uint32_t& ix, uint32_t& iy, uint32_t& iz, f3& fr) {
    const float r = (float)m.res[lv];
    const float sx = x.x * r, sy = x.y * r, sz = x.z * r;
    const float fx = fminf(floorf(sx), r - 1.f), fy = fminf(floorf(sy), r - 1.f), fz = fminf(floorf(sz), r - 1.f);
    ix = (uint32_t)fx; iy = (uint32_t)fy; iz = (uint32_t)fz;
    fr = mk3(sx - fx, sy - fy, sz - fz);
}

__device__ __forceinline__ f3 rc_mlp_norm(const rc_mlpdev& m, float4 p) {
    return mk3(fminf(fmaxf((p.x - m.lo[0]) * m.iext[0], 0.f), 1.f), fminf(fmaxf((p.y - m.lo[1]) * m.iext[1], 0.f), 1.f),
               fminf(fmaxf((p.z - m.lo[2]) * m.iext[2], 0.f), 1.f));
}

// the 64 inputs of one sample into row[0..63]
__device__ void rc_mlp_encode(const rc_mlpdev& m, float4 p, float4 n, float4 v, __bf16* row) {
    const f3 x = rc_mlp_norm(m, p);
    const size_t T = (size_t)1 << m.log2T;
#pragma unroll 4   // 32 corner gathers in flight per lane (each level's sum keeps its own order)
    for (int lv = 0; lv < RC_MLP_LEVELS; ++lv) {
        uint32_t ix, iy, iz; f3 fr;
        rc_mlp_cell(m, lv, x, ix, iy, iz, fr);
        float f0 = 0.f, f1 = 0.f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const uint32_t cx = ix + (c & 1), cy = iy + ((c >> 1) & 1), cz = iz + (c >> 2);
            const float w = ((c & 1) ? fr.x : 1.f - fr.x) * (((c >> 1) & 1) ? fr.y : 1.f - fr.y) * ((c >> 2) ? fr.z : 1.f - fr.z);
            const float2 t = *(const float2*)(m.table + 2 * ((size_t)lv * T + rc_mlp_index(m, lv, cx, cy, cz)));
            f0 += w * t.x;
            f1 += w * t.y;
        }
        row[2 * lv] = rc_bf(f0);
        row[2 * lv + 1] = rc_bf(f1);
    }
    row[32] = rc_bf(0.5f * (n.x + 1.f)); row[33] = rc_bf(0.5f * (n.y + 1.f)); row[34] = rc_bf(0.5f * (n.z + 1.f));
    row[35] = rc_bf(0.5f * (v.x + 1.f)); row[36] = rc_bf(0.5f * (v.y + 1.f)); row[37] = rc_bf(0.5f * (v.z + 1.f));
    row[38] = rc_bf(1.f);
    for (int k = 39; k < MW; ++k) row[k] = rc_bf(0.f);
}

// ------------------------------------------------------------------------------ forward
// One wave = 64 samples; activations in the wave's LDS tile, updated in place layer by layer
// (a row tile is read for all its products before it is overwritten). TRAIN also stores every
// layer's activations transposed, actT[layer][feature][sample] (row stride ldT), for the backward.
template <bool TRAIN>
__global__ void __launch_bounds__(MROWS * MWAVES) k_mlp_fwd(rc_mlpdev m, const float4* P, const float4* N, const float4* D, int n,
                                                            float* y, __bf16* actT, int ldT) {
    __shared__ __align__(16) __bf16 s_act[MWAVES][MROWS * MLD];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    const int base = (blockIdx.x * MWAVES + wv) * MROWS;
    __bf16* act = s_act[wv];
    {
        const int s = base + lane;
        __bf16* row = act + lane * MLD;
        if (s < n) rc_mlp_encode(m, P[s], N[s], D[s], row);
        else for (int k = 0; k < MW; ++k) row[k] = rc_bf(0.f);
        if (TRAIN) for (int k = 0; k < MW; ++k) actT[(int64_t)k * ldT + base + lane] = row[k];
    }
    rc_wave_sync();
    // Each layer's weight operands are loaded once into registers (8 x 16 bytes per lane) and serve
    // the wave's 4 row tiles; the products run as D = W A^T, so a lane's 4 results are 4 consecutive
    // output features of one sample: one 8-byte LDS store per column tile. (Operands reloaded from
    // memory for every product, and D = A W^T with 2-byte stores: 1.7x slower.)
    const __bf16* Wt = (const __bf16*)m.Wt;
    for (int L = 0; L < RC_MLP_HIDDEN; ++L) {
        const __bf16* W = Wt + (size_t)L * MW * MW;
        bf16x8 wop[2][4];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) wop[ks][ct] = rc_opnd(W, MW, ct * 16, ks * 32);
#pragma unroll
        for (int rt = 0; rt < MROWS / 16; ++rt) {
            const bf16x8 a0 = rc_opnd(act, MLD, rt * 16, 0), a1 = rc_opnd(act, MLD, rt * 16, 32);
            f32x4 acc[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = rc_mfma(wop[1][ct], a1, rc_mfma(wop[0][ct], a0, f32x4{0.f, 0.f, 0.f, 0.f}));
            // lane: features ct * 16 + 4 g + i of sample rt * 16 + c16
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                __bf16 h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) h[i] = rc_bf(fmaxf(acc[ct][i], 0.f));
                typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                *(bf16x4*)(act + (rt * 16 + c16) * MLD + ct * 16 + 4 * g) = bf16x4{h[0], h[1], h[2], h[3]};
                if (TRAIN) {
                    __bf16* dst = actT + ((int64_t)(L + 1) * MW + ct * 16 + 4 * g) * ldT + base + rt * 16 + c16;
#pragma unroll
                    for (int i = 0; i < 4; ++i) dst[(int64_t)i * ldT] = h[i];
                }
            }
        }
        rc_wave_sync();
    }
    const __bf16* W5 = Wt + (size_t)RC_MLP_HIDDEN * MW * MW;   // [16][64], output column 0
    const bf16x8 w50 = rc_opnd(W5, MW, 0, 0), w51 = rc_opnd(W5, MW, 0, 32);
#pragma unroll
    for (int rt = 0; rt < MROWS / 16; ++rt) {
        f32x4 acc = rc_mfma(rc_opnd(act, MLD, rt * 16, 0), w50, f32x4{0.f, 0.f, 0.f, 0.f});
        acc = rc_mfma(rc_opnd(act, MLD, rt * 16, 32), w51, acc);
        if (c16 == 0)
            for (int i = 0; i < 4; ++i) {
                const int s = base + rt * 16 + 4 * g + i;
                if (s < n) y[s] = acc[i];
            }
    }
}

// relative L2 loss (y - t)^2 / (y^2 + 0.01) averaged over n (the denominator is held constant in
// the gradient); dy over the padded batch (zero past n); loss sum accumulated in *loss
__global__ void __launch_bounds__(RC_TB) k_mlp_loss(const float* y, const float* t, int n, int npad, float* dy, float* loss) {
    __shared__ float s_w[RC_TB / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < npad) {
        float d = 0.f;
        if (i < n) {
            const float yi = y[i], den = yi * yi + 0.01f, e = yi - t[i];
            d = 2.f * e / den / (float)n;
            l = e * e / den / (float)n;
        }
        dy[i] = d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t2 = 0.f;
        for (int k = 0; k < RC_TB / 64; ++k) t2 += s_w[k];
        unsafeAtomicAdd(loss, t2);
    }
}

// ------------------------------------------------------------------------------ backward
// Gradients w.r.t. the pre-activations of every layer, gzT[L][feature][sample] for z_{L+1}
// (L = 0..4), and w.r.t. the 32 hash features of the input (gX[sample][32], FP32).
//   g_z5 = dy W5[:, 0] (.) [a5 > 0];  g_a_L = g_z_{L+1} W_L^T;  g_z_L = g_a_L (.) [a_L > 0] (L >= 1);  gX = g_a_0
__global__ void __launch_bounds__(MROWS * MWAVES) k_mlp_bwd(rc_mlpdev m, const float* dy, const __bf16* actT, int ldT, __bf16* gzT,
                                                            float* gX) {
    __shared__ __align__(16) __bf16 s_g[MWAVES][MROWS * MLD];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    const int base = (blockIdx.x * MWAVES + wv) * MROWS;
    __bf16* gs = s_g[wv];
    const __bf16* Wk = (const __bf16*)m.Wk;
    const __bf16* W5 = Wk + (size_t)RC_MLP_HIDDEN * MW * MW;   // [64][16]
    {
        const int s = base + lane;
        const float d = dy[s];
        for (int k = 0; k < MW; ++k) {
            const float a5 = rc_fb(actT[((int64_t)RC_MLP_HIDDEN * MW + k) * ldT + s]);
            const __bf16 v = rc_bf(a5 > 0.f ? d * rc_fb(W5[k * 16]) : 0.f);
            gs[lane * MLD + k] = v;
            gzT[((int64_t)(RC_MLP_HIDDEN - 1) * MW + k) * ldT + s] = v;
        }
    }
    rc_wave_sync();
    for (int L = RC_MLP_HIDDEN - 1; L >= 0; --L) {
        const __bf16* W = Wk + (size_t)L * MW * MW;   // W_L[k][j]
        bf16x8 wop[2][4];   // the layer's weight operands, loaded once for the wave's 4 row tiles
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) wop[ks][ct] = rc_opnd(W, MW, ct * 16, ks * 32);
        for (int rt = 0; rt < MROWS / 16; ++rt) {
            f32x4 acc[4];
            for (int ct = 0; ct < 4; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < 2; ++ks) {
                const bf16x8 a = rc_opnd(gs, MLD, rt * 16, ks * 32);
                for (int ct = 0; ct < 4; ++ct) acc[ct] = rc_mfma(a, wop[ks][ct], acc[ct]);
            }
            const int s0 = base + rt * 16 + 4 * g;
            for (int ct = 0; ct < 4; ++ct) {
                const int k = ct * 16 + c16;
                if (L > 0) {
                    const __bf16* ap = actT + ((int64_t)L * MW + k) * ldT + s0;
                    __bf16* dst = gzT + ((int64_t)(L - 1) * MW + k) * ldT + s0;
                    for (int i = 0; i < 4; ++i) {
                        const __bf16 v = rc_bf(rc_fb(ap[i]) > 0.f ? acc[ct][i] : 0.f);
                        gs[(rt * 16 + 4 * g + i) * MLD + k] = v;
                        dst[i] = v;
                    }
                } else if (k < 2 * RC_MLP_LEVELS) {
                    for (int i = 0; i < 4; ++i) gX[(int64_t)(s0 + i) * (2 * RC_MLP_LEVELS) + k] = acc[ct][i];
                }
            }
        }
        rc_wave_sync();
    }
}

// Weight gradients, split over the batch: dW_L[k][j] = sum_b a_L[b][k] g_z_{L+1}[b][j] for L = 0..4
// (P = actT[L], Q = gzT[L], both sample-contiguous), dW_5[k][0] = sum_b a_5[b][k] dy[b].
// Block (split, L): wave w computes rows k = 16w..16w+15; partial sums part[L][split][k][j]. For
// dW_5 the block's 4 waves sum quarters of the split (thread k + 64 q: samples q, q + 4, ...) and
// wave 0 adds the quarters in order (a fixed order: the same sums every run).
__global__ void __launch_bounds__(MROWS * MWAVES) k_mlp_dw(const __bf16* actT, const __bf16* gzT, const float* dy, int ldT, int chunk,
                                                           float* part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    const int L = blockIdx.y, split = blockIdx.x, nsplit = gridDim.x;
    const int b0 = split * chunk, b1 = min(ldT, b0 + chunk);
    float* out = part + ((size_t)L * nsplit + split) * MW * MW;
    if (L < RC_MLP_HIDDEN) {
        const __bf16* Pm = actT + (size_t)L * MW * ldT;
        const __bf16* Qm = gzT + (size_t)L * MW * ldT;
        f32x4 acc[4];
        for (int ct = 0; ct < 4; ++ct) acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int bs = b0; bs < b1; bs += 32) {
            const bf16x8 a = rc_opnd(Pm, ldT, wv * 16, bs);
            for (int ct = 0; ct < 4; ++ct) acc[ct] = rc_mfma(a, rc_opnd(Qm, ldT, ct * 16, bs), acc[ct]);
        }
        for (int ct = 0; ct < 4; ++ct)
            for (int i = 0; i < 4; ++i) out[(wv * 16 + 4 * g + i) * MW + ct * 16 + c16] = acc[ct][i];
    } else {
        __shared__ float s_q[4][MW];
        const int k = threadIdx.x & 63, q = threadIdx.x >> 6;
        const __bf16* a5 = actT + ((size_t)RC_MLP_HIDDEN * MW + k) * ldT;
        float s = 0.f;
        for (int b = b0 + q; b < b1; b += 4) s += rc_fb(a5[b]) * dy[b];
        s_q[q][k] = s;
        __syncthreads();
        if (q == 0) {
            const float t = ((s_q[0][k] + s_q[1][k]) + s_q[2][k]) + s_q[3][k];
            for (int j = 0; j < MW; ++j) out[k * MW + j] = j == 0 ? t : 0.f;
        }
    }
}

// gW (master layout: layers 0..4 [64][64], layer 5 [64][16]) = sum over splits, in split order
// (a block = 64 weights x 4 quarters of the splits: each thread adds its quarter in split order, then
// the quarters are added in order; the same sums every run)
__global__ void __launch_bounds__(RC_TB) k_mlp_dw_sum(const float* part, int nsplit, float* gW) {
    __shared__ float s_q[4][64];
    const int e = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
    float s = 0.f;
    if (e < RC_MLP_NW) {
        int L, k, j;
        if (e < RC_MLP_HIDDEN * MW * MW) { L = e / (MW * MW); k = (e / MW) % MW; j = e % MW; }
        else { L = RC_MLP_HIDDEN; k = (e - RC_MLP_HIDDEN * MW * MW) / 16; j = (e - RC_MLP_HIDDEN * MW * MW) % 16; }
        const int q0 = q * nsplit / 4, q1 = (q + 1) * nsplit / 4;
        for (int sp = q0; sp < q1; ++sp) s += part[((size_t)L * nsplit + sp) * MW * MW + k * MW + j];
    }
    s_q[q][threadIdx.x & 63] = s;
    __syncthreads();
    if (q == 0 && e < RC_MLP_NW) gW[e] = ((s_q[0][threadIdx.x] + s_q[1][threadIdx.x]) + s_q[2][threadIdx.x]) + s_q[3][threadIdx.x];
}

// ------------------------------------------------------------------------------ table gradient
// The gradient of each level's table is every sample's two feature gradients spread over its 8
// corners with the trilinear weights. One device-scope float atomic per corner into the 64 MB
// gradient (random addresses) runs at the atomic rate past the L2s; instead the updates are grouped
// by table slice (2^RC_EB_SB slots) and summed in LDS:
//   k_eb_count    per block (one level, RC_EB_SPT x 256 samples): corner pairs per slice, counted in
//                 LDS and added to the per-slice totals;
//   k_eb_scan     slice starts, and the parts (at most RC_EB_PMAX pairs each) each slice is summed in;
//   k_eb_scatter  the same pairs again, each block reserving its range of every slice with one
//                 atomic per slice, stored as (local slot, g0, g1);
//   k_eb_sum      one block per part: its pairs added into the slice's sums in LDS, then the slice's
//                 gradient read, added to and written back with 16-byte accesses; a slice summed in
//                 several parts (coarse levels, where many samples share few corners) adds its
//                 nonzero sums with atomics instead.
// Within a wave, lanes hold consecutive samples of one level. Training samples come in entry order,
// which is spatially sorted, so at the coarse levels neighbouring lanes share corners: lanes holding
// the same slot in a run of consecutive lanes add their values first (a segmented scan by shuffles)
// and only the run's last lane emits the pair. Batches above RC_EB_CHUNK samples run in rounds.
#define RC_EB_SPT 4

// a block's sweep over its samples of the round at level lv; emit(slot, v0, v1) for every pair
template <class F>
__device__ __forceinline__ void rc_eb_sweep(const rc_mlpdev& m, int lv, const float4* P, const float* gX, int s0, int nc, F&& emit) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int k = 0; k < RC_EB_SPT; ++k) {
        const int s = (blockIdx.x * RC_EB_SPT + k) * RC_TB + threadIdx.x;   // sample of the round
        const bool live = s < nc;
        float g0 = 0.f, g1 = 0.f;
        if (live) {
            const float* g = gX + (int64_t)(s0 + s) * (2 * RC_MLP_LEVELS) + 2 * lv;
            g0 = g[0]; g1 = g[1];
        }
        const bool act = live && (g0 != 0.f || g1 != 0.f);
        if (!__any(act)) continue;   // wave-uniform
        uint32_t ix = 0, iy = 0, iz = 0;
        f3 fr = mk3(0.f, 0.f, 0.f);
        if (act) rc_mlp_cell(m, lv, rc_mlp_norm(m, P[s0 + s]), ix, iy, iz, fr);
        for (int c = 0; c < 8; ++c) {
            const uint32_t cx = ix + (c & 1), cy = iy + ((c >> 1) & 1), cz = iz + (c >> 2);
            const float w = ((c & 1) ? fr.x : 1.f - fr.x) * (((c >> 1) & 1) ? fr.y : 1.f - fr.y) * ((c >> 2) ? fr.z : 1.f - fr.z);
            const int idx = act ? (int)rc_mlp_index(m, lv, cx, cy, cz) : -1 - lane;   // slots < 2^24; inactive: a run of its own
            float v0 = act ? w * g0 : 0.f, v1 = act ? w * g1 : 0.f;
            const int prev = __shfl_up(idx, 1, 64);
            const unsigned long long heads = __ballot(lane == 0 || prev != idx);
            const int seg = __popcll(heads & (below | (1ull << lane)));   // run number of this lane
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const float o0 = __shfl_up(v0, d, 64), o1 = __shfl_up(v1, d, 64);
                const int os = __shfl_up(seg, d, 64);
                if (lane >= d && os == seg) { v0 += o0; v1 += o1; }
            }
            const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
            if (act && tail) emit((uint32_t)idx, v0, v1);
        }
    }
}

__global__ void __launch_bounds__(RC_TB) k_eb_zero(uint32_t* x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0u;
}

__global__ void __launch_bounds__(RC_TB) k_eb_count(rc_mlpdev m, const float4* P, const float* gX, int s0, int nc, uint32_t* cnt) {
    extern __shared__ uint32_t s_ecnt[];   // [nsl] pairs per slice
    const int lv = blockIdx.y, sb = min(RC_EB_SB, m.log2T), nsl = 1 << (m.log2T - sb);
    for (int j = threadIdx.x; j < nsl; j += RC_TB) s_ecnt[j] = 0u;
    __syncthreads();
    rc_eb_sweep(m, lv, P, gX, s0, nc, [&](uint32_t slot, float, float) { atomicAdd(&s_ecnt[slot >> sb], 1u); });
    __syncthreads();
    for (int j = threadIdx.x; j < nsl; j += RC_TB)
        if (s_ecnt[j]) atomicAdd(&cnt[lv * nsl + j], s_ecnt[j]);
}

// exclusive scans of the slice counts (pair starts, base[nbk] = total) and of their part counts
// (part[nbk] = total parts), the slice of every part (pmap); clears the cursors of the scatter
__global__ void __launch_bounds__(1024) k_eb_scan(const uint32_t* cnt, int nbk, uint32_t* base, uint32_t* part, uint32_t* cur,
                                                  uint32_t* pmap) {
    __shared__ uint32_t s_a[16], s_p[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (nbk + 1023) / 1024, j0 = min(nbk, t * per), j1 = min(nbk, j0 + per);
    uint32_t a = 0u, pc = 0u;
    for (int j = j0; j < j1; ++j) { const uint32_t c = cnt[j]; a += c; pc += (c + RC_EB_PMAX - 1) / RC_EB_PMAX; }
    uint32_t ia = a, ip = pc;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t oa = __shfl_up(ia, d, 64), op = __shfl_up(ip, d, 64);
        if (lane >= d) { ia += oa; ip += op; }
    }
    if (lane == 63) { s_a[w] = ia; s_p[w] = ip; }
    __syncthreads();
    if (t == 0) {
        uint32_t x = 0u, y = 0u;
        for (int k = 0; k < 16; ++k) { const uint32_t u = s_a[k], v = s_p[k]; s_a[k] = x; s_p[k] = y; x += u; y += v; }
    }
    __syncthreads();
    uint32_t ea = s_a[w] + ia - a, ep = s_p[w] + ip - pc;
    for (int j = j0; j < j1; ++j) {
        const uint32_t c = cnt[j], np = (c + RC_EB_PMAX - 1) / RC_EB_PMAX;
        base[j] = ea; part[j] = ep; cur[j] = 0u;
        for (uint32_t k = 0; k < np; ++k) pmap[ep + k] = (uint32_t)j;   // the slice of every part
        ea += c; ep += np;
    }
    if (t == 1023) { base[nbk] = ea; part[nbk] = ep; }
}

__global__ void __launch_bounds__(RC_TB) k_eb_scatter(rc_mlpdev m, const float4* P, const float* gX, int s0, int nc, const uint32_t* base,
                                                      uint32_t* cur, uint16_t* eslot, float2* eval) {
    extern __shared__ uint32_t s_erank[];   // [nsl] counts, then ranks; [nsl] the block's start in every slice
    const int lv = blockIdx.y, sb = min(RC_EB_SB, m.log2T), nsl = 1 << (m.log2T - sb);
    uint32_t* s_off = s_erank + nsl;
    const uint32_t* bb = base + lv * nsl;
    for (int j = threadIdx.x; j < nsl; j += RC_TB) s_erank[j] = 0u;
    __syncthreads();
    rc_eb_sweep(m, lv, P, gX, s0, nc, [&](uint32_t slot, float, float) { atomicAdd(&s_erank[slot >> sb], 1u); });
    __syncthreads();
    for (int j = threadIdx.x; j < nsl; j += RC_TB) {
        const uint32_t c = s_erank[j];
        s_off[j] = c ? bb[j] + atomicAdd(&cur[lv * nsl + j], c) : 0u;
        s_erank[j] = 0u;
    }
    __syncthreads();
    const uint32_t lmask = (1u << sb) - 1u;
    rc_eb_sweep(m, lv, P, gX, s0, nc, [&](uint32_t slot, float v0, float v1) {
        const uint32_t sl = slot >> sb, pos = s_off[sl] + atomicAdd(&s_erank[sl], 1u);
        if (pos < bb[sl + 1]) {   // the two sweeps emit the same pairs: always inside the slice's range
            eslot[pos] = (uint16_t)(slot & lmask);
            eval[pos] = make_float2(v0, v1);
        }
    });
}

__global__ void __launch_bounds__(RC_TB) k_eb_sum(rc_mlpdev m, int nbk, const uint32_t* base, const uint32_t* part, const uint32_t* pmap,
                                                  const uint16_t* eslot, const float2* eval) {
    extern __shared__ float4 s_eacc4[];    // [2^sb] x 2 sums
    float* s_acc = (float*)s_eacc4;
    const int sb = min(RC_EB_SB, m.log2T), nsl = 1 << (m.log2T - sb), S = 1 << sb;
    const uint32_t blk = blockIdx.x;
    if (blk >= part[nbk]) return;          // the grid is an upper bound on the parts
    const int b = (int)pmap[blk];
    for (int j = threadIdx.x; j < 2 * S; j += RC_TB) s_acc[j] = 0.f;
    __syncthreads();
    const uint32_t np = part[b + 1] - part[b], k = blk - part[b];
    const uint32_t p0 = base[b] + k * RC_EB_PMAX, p1 = min(base[b + 1], p0 + RC_EB_PMAX);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += RC_TB) {
        const uint32_t loc = eslot[p];
        const float2 v = eval[p];
        atomicAdd(&s_acc[2 * loc], v.x);
        atomicAdd(&s_acc[2 * loc + 1], v.y);
    }
    __syncthreads();
    const int lv = b / nsl, sl = b - lv * nsl;
    float* g = m.gtable + (((size_t)lv << m.log2T) + ((size_t)sl << sb)) * 2;
    if (np == 1) {
        float4* g4 = (float4*)g;
        for (int j = threadIdx.x; j < S / 2; j += RC_TB) {
            float4 x = g4[j];
            const float4 a = s_eacc4[j];
            x.x += a.x; x.y += a.y; x.z += a.z; x.w += a.w;
            g4[j] = x;
        }
    } else {
        for (int j = threadIdx.x; j < 2 * S; j += RC_TB) {
            const float a = s_acc[j];
            if (a != 0.f) unsafeAtomicAdd(&g[j], a);
        }
    }
}


// Adam (torch.optim.Adam's formula): m, v moments, bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t;
// the gradient is cleared for the next step
__global__ void __launch_bounds__(RC_TB) k_mlp_adam(float* p, float* gr, float* mo, float* vo, long long n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mm = b1 * mo[i] + (1.f - b1) * g;
    const float vv = b2 * vo[i] + (1.f - b2) * g * g;
    mo[i] = mm; vo[i] = vv;
    p[i] -= (lr / c1) * mm / (sqrtf(vv) / sqrtf(c2) + eps);
    gr[i] = 0.f;
}

// BF16 operand copies of the master weights: Wt[L][j][k] (forward), Wk[L][k][j] (backward)
__global__ void __launch_bounds__(RC_TB) k_mlp_pack(const float* Wm, __bf16* Wt, __bf16* Wk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= RC_MLP_NW) return;
    const __bf16 v = rc_bf(Wm[e]);
    Wk[e] = v;
    if (e < RC_MLP_HIDDEN * MW * MW) {
        const int L = e / (MW * MW), k = (e / MW) % MW, j = e % MW;
        Wt[(size_t)L * MW * MW + j * MW + k] = v;
    } else {
        const int r = e - RC_MLP_HIDDEN * MW * MW, k = r / 16, j = r % 16;
        Wt[(size_t)RC_MLP_HIDDEN * MW * MW + j * MW + k] = v;
    }
}

__global__ void __launch_bounds__(RC_TB) k_mlp_zero(float* x, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

// ------------------------------------------------------------------------------ online training from the scene
// Sample s: cache entry i = hash(seed, s) mod Np of the state, one cosine ray from it. Training
// input (P_i, N_i, ray direction); the hit's query (point, facing normal, direction back along the
// ray; w = 1 when the hit is a valid entry); A_i and the direct light E_i = sum of the active
// lights' Edir (the cache's resolved radiance is R_i = A_i (E_i + indirect)).
template <int TRAV, bool SPLIT>
__global__ void __launch_bounds__(RC_TB) k_mlp_samples(rc_dscene s, rc_dstate st, uint64_t dormant, uint32_t seed, int n, float4* sP,
                                                      float4* sN, float4* sD, float4* hP, float4* hN, float4* hD, float2* AE, int* ent) {
    extern __shared__ float sh[];
    int* shq = (int*)(sh + 9 * st.nt);
    const auto lstk = rc_lstk_sel<SPLIT>::make(sh, st, s);
    rc_load_dyn(st, sh, shq);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t rs = rc_hash32(seed ^ rc_hash32((uint32_t)k + 1u));
    // stratified over the entries in entry order (one uniform pick in each of n equal strata):
    // every entry equally likely, and consecutive samples spatially close (entries are stored
    // in Morton order), which the table-gradient scatter exploits
    const float u0 = rc_rnd(rs);
    const int i = min(s.Np - 1, (int)(((double)k + (double)u0) * (double)s.Np / (double)n));
    const f3 P = ld3(st.P, i), N = ld3(st.N, i), O = P + N * 1e-3f;
    const float u1 = rc_rnd(rs), u2 = rc_rnd(rs);
    const f3 D = rc_cosine_dir(N, u1, u2);
    float t; int q;
    const bool hit = rc_trace_full<TRAV>(O, D, sh, shq, st.nt, s, t, q, lstk);
    float valid = 0.f;
    f3 H = O, nh = N;
    if (hit) {
        H = O + D * t;
        int e;
        valid = rc_resolve_hit_n(s, st, H, D, q, e, nh) ? 0.f : 1.f;
    }
    sP[k] = make_float4(P.x, P.y, P.z, 0.f);
    sN[k] = make_float4(N.x, N.y, N.z, 0.f);
    sD[k] = make_float4(D.x, D.y, D.z, 0.f);
    hP[k] = make_float4(H.x, H.y, H.z, valid);
    hN[k] = make_float4(nh.x, nh.y, nh.z, 0.f);
    hD[k] = make_float4(-D.x, -D.y, -D.z, 0.f);
    const float* ed = st.Edir + (int64_t)s.nl * i;
    float E = 0.f;
    for (int l = 0; l < s.nl; ++l) E += rc_is_dormant(dormant, l) ? 0.f : ed[l];
    AE[k] = make_float2(s.Alb[i], E);
    ent[k] = i;
}

// self-training targets t = A_i (E_i + max(y(hit), 0)) (0 for the indirect part on a miss)
__global__ void __launch_bounds__(RC_TB) k_mlp_targets(int n, const float4* hP, const float* hy, const float2* AE, float* t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float2 ae = AE[k];
    t[k] = ae.x * (ae.y + (hP[k].w != 0.f ? fmaxf(hy[k], 0.f) : 0.f));
}

// judge: sums of (y - R_i)^2 and R_i^2 over the samples (the grid cache's resolved radiance)
__global__ void __launch_bounds__(RC_TB) k_mlp_judge(int n, const float* y, const float* R, const int* ent, float* acc) {
    __shared__ float s_w[2][RC_TB / 64];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.f, r2 = 0.f;
    if (k < n) {
        const float r = R[ent[k]], e = y[k] - r;
        d2 = e * e; r2 = r * r;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { d2 += __shfl_down(d2, o, 64); r2 += __shfl_down(r2, o, 64); }
    if ((threadIdx.x & 63) == 0) { s_w[0][threadIdx.x >> 6] = d2; s_w[1][threadIdx.x >> 6] = r2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.f, b = 0.f;
        for (int w = 0; w < RC_TB / 64; ++w) { a += s_w[0][w]; b += s_w[1][w]; }
        unsafeAtomicAdd(acc, a);
        unsafeAtomicAdd(acc + 1, b);
    }
}

// ------------------------------------------------------------------------------ launchers
static inline int rc_mlp_blocks(int n) { return (n + MROWS * MWAVES - 1) / (MROWS * MWAVES); }
int rc_mlp_pad(int n) { return rc_mlp_blocks(n) * MROWS * MWAVES; }

hipError_t rc_launch_mlp_infer(const rc_mlpdev* m, const float4* P, const float4* N, const float4* D, int n, float* y, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_fwd<false>, dim3(rc_mlp_blocks(n)), dim3(MROWS * MWAVES), 0, st, *m, P, N, D, n, y, (__bf16*)nullptr, 0);
    return hipGetLastError();
}

// table gradient of n samples added to m->gtable, in rounds of RC_EB_CHUNK samples
static hipError_t rc_mlp_table_grad(const rc_mlpdev* m, const rc_mlpbuf* b, const float4* P, int n, hipStream_t st) {
    const int sb = std::min(RC_EB_SB, m->log2T), nsl = 1 << (m->log2T - sb), nbk = RC_MLP_LEVELS * nsl;
    const size_t lds_bucket = 2 * (size_t)nsl * sizeof(uint32_t), lds_sum = ((size_t)2 << sb) * sizeof(float);
    for (int s0 = 0; s0 < n; s0 += RC_EB_CHUNK) {
        const int nc = std::min(RC_EB_CHUNK, n - s0);
        const dim3 g((nc + RC_EB_SPT * RC_TB - 1) / (RC_EB_SPT * RC_TB), RC_MLP_LEVELS);
        hipLaunchKernelGGL(k_eb_zero, dim3((nbk + RC_TB - 1) / RC_TB), dim3(RC_TB), 0, st, b->ecnt, nbk);
        hipLaunchKernelGGL(k_eb_count, g, dim3(RC_TB), lds_bucket / 2, st, *m, P, (const float*)b->gX, s0, nc, b->ecnt);
        hipLaunchKernelGGL(k_eb_scan, dim3(1), dim3(1024), 0, st, (const uint32_t*)b->ecnt, nbk, b->ebase, b->epart, b->ecur, b->epmap);
        hipLaunchKernelGGL(k_eb_scatter, g, dim3(RC_TB), lds_bucket, st, *m, P, (const float*)b->gX, s0, nc, (const uint32_t*)b->ebase,
                           b->ecur, b->eslot, (float2*)b->eval);
        const long long maxparts = nbk + ((long long)nc * RC_EB_PAIRS + RC_EB_PMAX - 1) / RC_EB_PMAX;
        hipLaunchKernelGGL(k_eb_sum, dim3((unsigned)maxparts), dim3(RC_TB), lds_sum, st, *m, nbk, (const uint32_t*)b->ebase,
                           (const uint32_t*)b->epart, (const uint32_t*)b->epmap, (const uint16_t*)b->eslot, (const float2*)b->eval);
    }
    return hipGetLastError();
}

// forward (train) + backward: gradients of sum_i dy_i y_i added to m->gW (overwritten) and m->gtable (accumulated)
hipError_t rc_launch_mlp_grad(const rc_mlpdev* m, const rc_mlpbuf* b, const float4* P, const float4* N, const float4* D, const float* t,
                              const float* dy_in, int n, float* loss, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int npad = rc_mlp_pad(n), nb = rc_mlp_blocks(n);
    __bf16* actT = (__bf16*)b->actT;
    __bf16* gzT = (__bf16*)b->gzT;
    hipLaunchKernelGGL(k_mlp_fwd<true>, dim3(nb), dim3(MROWS * MWAVES), 0, st, *m, P, N, D, n, b->y, actT, npad);
    if (t) hipLaunchKernelGGL(k_mlp_loss, dim3(rc_grid(npad)), dim3(RC_TB), 0, st, (const float*)b->y, t, n, npad, b->dy, loss);
    const float* dy = t ? (const float*)b->dy : dy_in;
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nb), dim3(MROWS * MWAVES), 0, st, *m, dy, (const __bf16*)actT, npad, gzT, b->gX);
    const int chunk = 256, nsplit = (npad + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_mlp_dw, dim3(nsplit, RC_MLP_HIDDEN + 1), dim3(MROWS * MWAVES), 0, st, (const __bf16*)actT, (const __bf16*)gzT, dy,
                       npad, chunk, b->part);
    hipLaunchKernelGGL(k_mlp_dw_sum, dim3((RC_MLP_NW + 63) / 64), dim3(RC_TB), 0, st, (const float*)b->part, nsplit, m->gW);
    return rc_mlp_table_grad(m, b, P, n, st);
}

hipError_t rc_launch_mlp_adam(const rc_mlpdev* m, const rc_mlpbuf* b, float lr, float b1, float b2, float eps, int step, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    const long long nt = (long long)RC_MLP_LEVELS * ((long long)2 << m->log2T);
    hipLaunchKernelGGL(k_mlp_adam, dim3(rc_grid(RC_MLP_NW)), dim3(RC_TB), 0, st, m->Wm, m->gW, b->mW, b->vW, (long long)RC_MLP_NW, lr, b1, b2,
                       eps, c1, c2);
    hipLaunchKernelGGL(k_mlp_adam, dim3(rc_grid(nt)), dim3(RC_TB), 0, st, m->table, m->gtable, b->mT, b->vT, nt, lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_mlp_pack, dim3(rc_grid(RC_MLP_NW)), dim3(RC_TB), 0, st, (const float*)m->Wm, (__bf16*)m->Wt, (__bf16*)m->Wk);
    return hipGetLastError();
}

hipError_t rc_launch_mlp_pack(const rc_mlpdev* m, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(rc_grid(RC_MLP_NW)), dim3(RC_TB), 0, st, (const float*)m->Wm, (__bf16*)m->Wt, (__bf16*)m->Wk);
    return hipGetLastError();
}

hipError_t rc_launch_mlp_zero(float* x, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_zero, dim3(rc_grid(n)), dim3(RC_TB), 0, st, x, n);
    return hipGetLastError();
}

hipError_t rc_launch_mlp_samples(int trav, const rc_dscene* s, const rc_dstate* st, uint64_t dormant, uint32_t seed, int n, float4* sP,
                                 float4* sN, float4* sD, float4* hP, float4* hN, float4* hD, float2* AE, int* ent, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const dim3 g(rc_grid(n)), b(RC_TB);
    const size_t shm = rc_shm(*st, *s, trav, false);
    if (trav == 0) hipLaunchKernelGGL((k_mlp_samples<0, false>), g, b, shm, stream, *s, *st, dormant, seed, n, sP, sN, sD, hP, hN, hD, AE, ent);
    else if (trav == 1) hipLaunchKernelGGL((k_mlp_samples<1, false>), g, b, shm, stream, *s, *st, dormant, seed, n, sP, sN, sD, hP, hN, hD, AE, ent);
    else RC_LAUNCH_WIDE((k_mlp_samples<2, true>), (k_mlp_samples<2, false>), false, g, st, s, stream, *s, *st, dormant, seed, n, sP, sN, sD, hP, hN, hD,
                        AE, ent);
    return hipGetLastError();
}

hipError_t rc_launch_mlp_targets(int n, const float4* hP, const float* hy, const float2* AE, float* t, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_targets, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, n, hP, hy, AE, t);
    return hipGetLastError();
}

hipError_t rc_launch_mlp_judge(int n, const float* y, const float* R, const int* ent, float* acc, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_judge, dim3(rc_grid(n)), dim3(RC_TB), 0, stream, n, y, R, ent, acc);
    return hipGetLastError();
}

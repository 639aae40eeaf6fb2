// Plain-old-data layouts shared by the host C++ (g++) and the HIP kernels (hipcc).
// Field meanings follow the prototype's DevState / Cache / Params structs
// (cuda/poc_gpu.cu:222-237, :394-409, :411-422); the light count, transit-grid
// size and channel count are runtime values here instead of -DNLIGHTS/-DGX
// compile-time constants (cuda/build.sh:12).
#pragma once
#include <stdint.h>

#define RC_MAX_LIGHTS 72      // light channels per cache entry (64 lamps + sun fit)
#define RC_MAX_NW 4           // transit-grid mask words per entry (128 bits, cuda/build.sh:12)
#define RC_KSNAP 6            // snapshot slots (cuda/poc_gpu.cu:423)
#define RC_VERIF_FRAMES 4     // verification window (cuda/poc_gpu.cu:115)

// Fixed-point scales of the order-independent reductions. Float sums that feed
// control flow (ray allocation, verification, DDGI blend) are accumulated as
// int64 so that every summation order gives the same bits.
#define RC_FX_PR 4294967296.0f    // sum of 1/(n+1)           (2^32)
#define RC_FX_STAT 1048576.0f     // per-region sample stats   (2^20)

// One BVH node of the binary tree: the prototype's two float4 {lo.xyz, lf} {hi.xyz, cnt}
// (cuda/prep_mesh.cuh:39-40). cnt > 0: leaf with triangles lf..lf+cnt-1; else children lf, lf+1.
typedef struct rc_bvh_node {
    float lo[3]; int32_t lf;
    float hi[3]; int32_t cnt;
} rc_bvh_node;

// 4-wide BVH node (128 B = two 64-byte lines): child boxes as fp32 SoA so the four slab
// tests of a node read whole float4s. child[c] < 0: leaf, child[c] = 1 << 31 | first << 7 |
// count (triangles first .. first + count - 1, same order and indices as the binary BVH, count
// 1..RC_WLEAF_MAX: a longer binary leaf is split over a wide node of its own whose children are
// chunks of its range with its box; cend[c] = first + count); child[c] > 0: inner node child[c]
// (cend 0); child[c] == 0: empty slot (box lo = +inf, hi = -inf: never entered). The traversal
// reads rows 0-6 only.
#define RC_WLEAF_MAX 127
typedef struct rc_wnode {
    float lox[4], loy[4], loz[4], hix[4];   // 64 B
    float hiy[4], hiz[4];                   // 32 B
    int32_t child[4];                       // 16 B
    int32_t cend[4];                        // 16 B
} rc_wnode;

// The same 4-wide node with quantized child boxes (64 B = one line; node i of QN is node i of WN):
// per axis an origin o and a power-of-two step s, and per child and axis the planes o + q s with
// 8-bit q (byte c of each q word = child c). The quantized box contains the fp32 box with at
// least one step to spare on every side, and the step is at least 2^-18 of the scene's largest
// coordinate, so the slab test on it is conservative with respect to the fp32 test of the exact
// box despite both tests' rounding (the closest hit is never culled: same hits). Empty slots
// hold q lo = 255, q hi = 0 on every axis (never entered). child[] as in rc_wnode.
typedef struct rc_qnode {
    float ox, oy, oz, sx;                   // row 0
    float sy, sz; uint32_t qlox, qloy;      // row 1
    uint32_t qloz, qhix, qhiy, qhiz;        // row 2
    int32_t child[4];                       // row 3
} rc_qnode;
#define RC_WSTACK 48          // most LDS stack levels per lane (48 KB per 256-lane block)
// Queued leaf ranges per lane of the wide traversals (LDS levels after the stack): RC_WQ, the wave
// walks the queued leaves once a lane holds 3 (1920x1080 frame +0.8 % against 7 levels walked at 4);
// the adjoint's kernels keep RC_WQ_L = 7 walked at 4 (their short rays lose 0.5-1 % walking at 3)
#define RC_WQ 6
#define RC_WQ_L 7
#define RC_LCAP 16            // LDS levels of the camera pass's lane stack (deeper levels: global overflow)

// Per-state device view of the scene (cuda/poc_gpu.cu:222-237).
typedef struct rc_dstate {
    const float* P;          // [3 Np] cache point positions (quad points move with the state)
    const float* N;          // [3 Np] cache point normals
    const float* Q;          // [12 Nq] quads: o, u, v, n
    const int32_t* dims;     // [2 Nq] points per quad side
    const int32_t* off;      // [Nq] first cache point of the quad
    const float* T;          // [9 nt] dynamic triangles v0, e1, e2
    const int32_t* TQ;       // [nt] quad of each dynamic triangle
    float* Edir;             // [NL Np] direct light per entry and light
    uint32_t* SV;            // [NW Np] static shadow-ray transit mask (AoS)
    uint64_t* LM;            // [Np] light-visibility mask of the entry (bit min(l,63))
    const float* ref;        // [Np] converged reference (optional)
    const uint8_t* valid;    // [Np]
    const float* Pw;         // [NL Np] converged per-light reference (optional)
    const float* inten;      // [NL] light intensities of the state
    uint32_t* VOL;           // [NL][fg^3 / 32] per-light fine voxel volumes of the lit shadow segments
    uint32_t* ovf;           // [ovf_words] global levels of the wide traversal's lane stacks (levels >= ocap),
    unsigned long long ovf_words;   //   one column per thread of the launch; set per launch by its owner
    int32_t nt;              // dynamic triangles in this state
    int32_t ocap;            // LDS levels of the split lane stacks (0: all wstack levels in LDS, ovf unused)
    int32_t oforce;          // 1: wide-traversal launches take the split stacks; 0: only where that raises occupancy
    int32_t pad0;
} rc_dstate;

// Static scene data on the device (format-3 part of cuda/poc_gpu.cu:231-236).
typedef struct rc_dscene {
    const float* ST;         // [9 nst] static triangles v0, e1, e2 (BVH leaf order)
    const float* ST4;        // [12 nst] the same triangles as 48-byte rows (v0,0 e1,0 e2,0)
    const float* STN;        // [4 nst] geometric normal + flags (w bit 0 = double sided)
    const rc_bvh_node* BN;   // binary BVH nodes
    const rc_wnode* WN;      // 4-wide BVH nodes
    const rc_qnode* QN;      // the same nodes with quantized child boxes (traversal mode 3)
    const uint64_t* HK;      // hash keys (0 = empty)
    const int32_t* HV;       // entry of each key
    const float* L;          // [3 NL] light positions (point) or unit direction to the light (sun)
    const int32_t* ltype;    // [NL] 0 = point, 1 = directional
    const float* Alb;        // [Np] entry albedo
    const int32_t* region;   // [Np] verification zone
    uint32_t hmask;          // table size - 1
    float hcell;             // cache cell size (m)
    float lo[3], ext[3];     // transit-grid / hash-grid bounds
    float sun_len;           // shadow-ray length toward a directional light
    int32_t nst, nl, Np, NR;
    int32_t vx, vy, vz, nw;  // transit grid (vx*vy*vz = 32 nw bits per entry)
    int32_t wstack;          // LDS stack levels of the wide traversal (3 per wide-BVH level)
    int32_t fg;              // fine voxel grid fg^3 over the bounds (0: no fine volumes)
} rc_dscene;

// Cache buffers (cuda/poc_gpu.cu:394-409) without the v5-B projection buffers (out of scope).
typedef struct rc_dcache {
    float *Lc, *Lc2, *n, *c0;
    float* m;                    // [NL][Np] SoA: this frame's per-light sample means
    float* R;                    // [Np] resolved radiance A (sum Edir + sum Lc) read by the query pass
    int32_t* kk;
    uint8_t* pending;
    uint32_t *cur, *prev;        // [NW][Np] SoA transit-grid dependency banks
    uint64_t *lcur, *lprev;      // [Np] light-dependency banks
    long long* rstat;            // [4 NR] fixed point (RC_FX_STAT)
    long long* dstat;            // [3 NR] fixed point (RC_FX_STAT)
    uint64_t* stale;             // [Np] bit l = dormant channel l is out of date here (l < 64)
    float* snap;                 // [KSNAP NL Np]
    void* csnap;                 // [KSNAP][Np] compact snapshot records (24 B)
    float* age;                  // [Np]
    long long* sumpr;            // [2] fixed point (RC_FX_PR): current, next
    long long* prpart;           // [ceil(Np / 64)] per-block partial sums of the sumpr terms
    int32_t* npend;              // [2]
    double* err;                 // [3] squared error, reference, valid count of the error pass
    double* errpart;             // [3][ceil(Np / 256)] per-block partials of the error pass
    int32_t* blkcnt;             // [ceil(Np / 256) + 1] per-block counts of the event kernels, then their total
    float* rsum;                 // [Np] per-entry channel sums of the event kernels (k_rowsum)
} rc_dcache;

// Event voxels of the transit grid as a per-word bit mask (bit v & 31 of word v >> 5), built on the host
// from the event's voxel list (cuda/poc_gpu.cu:899-911 lists; the kernels test cur | prev against it).
typedef struct rc_vmask { uint32_t w[RC_MAX_NW]; } rc_vmask;

// Per-frame parameters (cuda/poc_gpu.cu:411-422).
typedef struct rc_params {
    int32_t Np, NR, mode;        // mode 0 = fixed budget, 1 = adaptive allocation
    float budget, nmax;
    uint32_t seed, frame;
    int32_t pend_left, deps;
    uint64_t resp_mask;          // lights blended with the responsive history cap
    float nmax_resp;
    int32_t ddgi; float ddgi_h, ddgi_thr; int32_t ddgi_clamp;
    uint64_t dormant;            // dormant light channels (frozen, excluded from estimates)
    int32_t ramp;
    float rays;
    int32_t nl;
    int32_t mpitch;              // row pitch (entries) of the SoA sample means: Np rounded up to a multiple of 64
} rc_params;

// Pinhole camera of the query pass. Paths: primary hit, one-light NEE per explicit vertex,
// cosine bounces; with a cache the path ends on the cached radiance at bounce cache_bounce,
// without one it ends after `bounces` bounces.
#ifndef RC_CAMERA_DEFINED
#define RC_CAMERA_DEFINED
typedef struct rc_camera {
    float pos[3], fwd[3], right[3], up[3];
    float tanx, tany;
    int32_t W, H, spp, bounces, cache_bounce, pad0;
} rc_camera;
#endif

// Workspace of the staged camera pass, indexed by path id (W*H*spp paths, segments of 1024).
typedef struct rc_qwork {
    void* ro; void* rd;            // float4 extension ray of each path: origin, direction
    void* so; void* sd;            // float4 shadow ray: origin + tlim, direction
    float* sc;                     // next-event term of the shadow ray
    float* ht; int32_t* hq;        // hit of the extension ray
    uint32_t* rs; float* thr; float* rad;   // RNG state, throughput, radiance
    int32_t* idx; int32_t* sidx;   // per-segment compacted id lists: extension rays, shadow rays
    int32_t* seg;                  // [2 * segments]: extension-list counts, then shadow-list counts
    void* mq0; void* mq1; void* mq2;   // float4 learned-cache query of each path: (point, queried), normal, view direction
    float* mpred;                  // the learned cache's prediction per path
} rc_qwork;

// Device buffers of the cached-radiance adjoint (Np entries, nl lights, B bounces).
typedef struct rc_adjbuf {
    float* G;                      // [Np][nl] unit-intensity direct light
    uint64_t* SG;                  // [Np] support of G
    int32_t* mat;                  // [Np] material of each entry
    float* A; float* E;            // [Np] albedo, emission-weighted direct light of the last forward
    float* L;                      // [B][Np] L^(1..B) of the last forward
    uint64_t* S[2];                // [Np] support ping-pong of the forward bounces
    uint64_t* SR;                  // [Np] support of R
    float* dA; float* dE;          // [Np] adjoints of A and E
    float* lam[2];                 // [Np] adjoint of L^(b) ping-pong
    float* part;                   // [blocks][nl] partial sums of dI
    int32_t* perm; int32_t* moff;  // entries sorted by material, offsets [nmat + 1]
    int32_t* pmat;                 // [Np] material of each sorted position
    float* tail;                   // [Np] per-tile segment sums of dA (at each segment's last position)
    int32_t nmat, pad0;
} rc_adjbuf;

// Hash-encoded MLP cache: 16 levels x 2 features, 5 hidden layers of 64, scalar output.
#define RC_MLP_LEVELS 16
#define RC_MLP_HIDDEN 5
#define RC_MLP_NW (RC_MLP_HIDDEN * 64 * 64 + 64 * 16)   // master weights: [5][64][64] + [64][16]
typedef struct rc_mlpdev {
    float* table; float* gtable;   // [16][T][2] hash-grid features and their gradient
    float* Wm; float* gW;          // FP32 master weights and their gradient
    void* Wt; void* Wk;            // BF16 operand copies: Wt[L][out][in], Wk[L][in][out]
    int32_t log2T; uint32_t dense; // table size 2^log2T per level; bit lv: level lv indexed densely
    int32_t res[RC_MLP_LEVELS];    // grid resolution per level
    float lo[3], iext[3];          // position bounds (lo, 1 / extent)
    const float* jR;               // camera pass: the grid cache's resolved radiance judging each query (NULL: none)
    float jthr;                    // a prediction off the grid's value at its entry by more than jthr times it -> the grid's value
    int32_t pad1;
} rc_mlpdev;
typedef struct rc_mlpbuf {
    void* actT; void* gzT;         // BF16 [6][64][cap] activations, [5][64][cap] pre-activation gradients
    float* y; float* dy; float* gX; float* part;   // outputs, output gradients, input-feature gradients [cap][32], dW partials
    float* mW; float* vW; float* mT; float* vT;    // Adam moments (weights, table)
    float* loss;                   // [1]
    int32_t cap, pad0;             // padded batch capacity
    // table-gradient scatter by table slice (2^RC_EB_SB slots each): per-slice pair counts, cursors,
    // starts (nbk + 1 each, nbk = 16 levels x slices), part starts, the slice of every part (nbk +
    // epairs / RC_EB_PMAX at most), and the pairs of one round (local slot, the two feature
    // gradients), RC_EB_PAIRS per sample of the round
    uint32_t* ecnt; uint32_t* ecur; uint32_t* ebase; uint32_t* epart; uint32_t* epmap;
    uint16_t* eslot; float* eval;
    int64_t epairs;                // pair capacity of one round
} rc_mlpbuf;
// table slices of 4096 slots: a slice's 2 x 4096 float sums fit in 32 KB of LDS
#define RC_EB_SB 12
#define RC_EB_CHUNK (1 << 16)      // samples per round of the scatter
#define RC_EB_PAIRS (8 * 16)       // at most 8 corners x 16 levels per sample
#define RC_EB_PMAX 16384           // pairs summed by one block

// compact snapshot record, 24 bytes (the C ABI's rc_snap_rec)
typedef struct rc_snaprec { uint32_t index; float n; uint64_t key; float radiance; float age; } rc_snaprec;

/* CPU oracle of the radiance-cache hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of the prototype's device functions and frame kernels
 * (cuda/poc_gpu.cu, cuda/prep_mesh.cuh) and of the SDK bench's resolve-time invalidation
 * (windows/patch/rtxgi_bench.patch), used by tests/, the smoke check and the CPU-baseline leg
 * of bench.py to check and time the HIP library. Nothing in the product links or calls it.
 *
 * FP32 functions follow the same operation order as the HIP kernels and are compiled with
 * -ffp-contract=off, so hit ids, keys, masks and cache values are bit-equal to a correct
 * GPU run. The *_f64 functions recompute radiance in double precision from the discrete
 * decisions (rays, hits) of the FP32 path: the FP64 reference of the radiance tolerance.
 *
 * Parity pins (see tests/): SplitMix64 / hash constants against published outputs, sin/cos
 * against libm, cosine sampling against its closed-form moments, ray hits against brute
 * force over every triangle, DDA voxel sets against exact slab-walk sets, hash tables
 * against their probing invariants, EMA / rescale / invalidation against closed forms.
 */
#ifndef RC_ORACLE_H
#define RC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_node { float lo[3]; int32_t lf; float hi[3]; int32_t cnt; } orc_node;
typedef struct orc_wnode {
    float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
    int32_t child[4], cend[4];   /* leaf: child = 1 << 31 | first << 7 | (count <= 127 ? count : 0), cend = first + count;
                                    inner: child = node, cend 0; empty: both 0 */
} orc_wnode;

/* static scene + one state (inputs); Edir/SV/LM are the state's build outputs */
typedef struct orc_scene {
    const float* ST; const float* STN; const orc_node* BN; const orc_wnode* WN;
    const uint64_t* HK; const int32_t* HV; uint32_t hmask; float hcell;
    const float* L; const int32_t* ltype; const float* Alb; const int32_t* region;
    float lo[3], ext[3]; float sun_len;
    int32_t nst, nl, Np, NR, vx, vy, vz, nw;
    const float* P; const float* N; const float* Q; const int32_t* dims; const int32_t* off;
    const float* T; const int32_t* TQ; int32_t nt;
    const float* Edir; const uint32_t* SV; const uint64_t* LM; const float* inten;
    uint32_t* VOL; int32_t fg;       /* per-light fine volumes [nl][ceil(fg^3/32)] (fg = 0: none) */
} orc_scene;

typedef struct orc_cache {
    float *Lc, *Lc2, *n, *c0;
    float* m;        /* [nl][Np] SoA sample means */
    float* R;        /* [Np] resolved radiance A * sum_l (Edir_l + Lc_l) after the update */
    int32_t* kk;
    uint8_t* pending;
    uint32_t *cur, *prev;
    uint64_t *lcur, *lprev;
    long long *rstat, *dstat;
    uint64_t* stale;
    float* age;
    long long* sumpr;
    int32_t* npend;
    double* err;
} orc_cache;

typedef struct orc_params {
    int32_t Np, NR, mode; float budget, nmax; uint32_t seed, frame; int32_t pend_left, deps;
    uint64_t resp_mask; float nmax_resp; int32_t ddgi; float ddgi_h, ddgi_thr; int32_t ddgi_clamp;
    uint64_t dormant; int32_t ramp; float rays; int32_t nl; int32_t pad0;
} orc_params;

typedef struct orc_camera {
    float pos[3], fwd[3], right[3], up[3];
    float tanx, tany;
    int32_t W, H, spp, bounces, cache_bounce, pad0;
} orc_camera;

/* ---- primitives ---- */
uint32_t orc_hash32(uint32_t x);
float orc_rnd(uint32_t* s);
uint32_t orc_seed(uint32_t seed, uint32_t frame, uint32_t i);
void orc_sincos(float x, float* s, float* c);
void orc_cosine_dir(const float N[3], float u1, float u2, float D[3]);
uint64_t orc_mix64(uint64_t x);
uint64_t orc_cell_key(int cx, int cy, int cz, int b);
void orc_decode_key(uint64_t k, int* cx, int* cy, int* cz, int* b);
int orc_normal_bucket(float nx, float ny, float nz);
uint64_t orc_spread3(uint64_t x);
float orc_hostrng_u(uint64_t* state);          /* SplitMix64 step, top 24 bits */
uint64_t orc_hostrng_init(uint64_t x);

/* ---- hash grid ---- */
int orc_hash_find(const uint64_t* HK, const int32_t* HV, uint32_t hmask, uint64_t key);
int orc_hash_point(const orc_scene* s, const float H[3], const float n[3]);
void orc_hash_build(const uint64_t* keys, int n, int log2, uint64_t* HK, int32_t* HV);
/* cell sampling of a mesh (V: 9 floats per triangle, alb, ds) with cell h from lo:
 * returns the number of cells; when keys != NULL fills (up to cap) the sorted keys and
 * their point / normal / albedo (cuda/prep_mesh.cuh:129-157, :349-370). */
int orc_sample_cells(const float* V, const float* alb, const uint8_t* ds, int nt, float h, const float lo[3],
                     uint64_t* keys, float* P, float* N, float* A, int cap);

/* ---- traversal ---- */
int orc_trace_lds(const float o[3], const float d[3], const float* T, const int32_t* TQ, int nt, float* tmin, int* q);
void orc_trace_bvh_ref(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q);
void orc_trace_bvh_closest(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q);
void orc_trace_brute(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q);
/* trav: 0 binary reference order, 1 binary closest, 3 brute force (closest) */
void orc_trace_rays(const orc_scene* s, int trav, const float* O, const float* D, int n, float* t, int32_t* q);

/* ---- transit grid ---- */
int orc_vox_of(const orc_scene* s, const float p[3]);
void orc_dda_reg(const orc_scene* s, const float a[3], const float b[3], uint32_t* m);
void orc_dda_segments(const orc_scene* s, const float* A, const float* B, int n, uint32_t* out);
void orc_dda_volume(const orc_scene* s, int gx, int gy, int gz, const float* A, const float* B, const int32_t* seg_vol, int n,
                    uint32_t* vol, long long words_per_vol);

/* ---- frame kernels (one call = the whole grid) ---- */
/* also marks each lit shadow segment into s->VOL[l] when s->fg > 0 */
void orc_k_build(const orc_scene* s, int trav, float* Edir, uint32_t* SV, uint64_t* LM);
/* the same for entries [i0, i1) only, into full-size Edir / SV / LM; fine-volume bits of those entries'
 * lit segments are ORed into VOL (may be NULL) */
void orc_k_build_rows(const orc_scene* s, int trav, int i0, int i1, float* Edir, uint32_t* SV, uint64_t* LM, uint32_t* VOL);
void orc_k_trace2(const orc_scene* s, int trav, int deps, const orc_cache* c, const orc_params* p, uint64_t* lnew, int32_t* hits, int kmax);
void orc_k_decide(const orc_scene* s, const orc_cache* c, const orc_params* p);
void orc_k_update(const orc_scene* s, const orc_cache* c, const orc_params* p, const uint64_t* lnew);
void orc_k_sumpr(const orc_cache* c, int Np);
void orc_k_restored(const orc_cache* c, int Np, int nl, float nmax, uint64_t dormant, int verify);
void orc_k_rescale(const orc_cache* c, int Np, int nl, int l, float ratio);
void orc_k_touched(const orc_cache* c, int Np, int nl, const int32_t* V, int nv, int kind, uint64_t dormant);
void orc_k_wake(const orc_cache* c, int Np, int nl, int l, int verify, uint64_t dormant);
void orc_k_oracle_method(const orc_cache* c, int Np, const float* rold, const float* rnew);
void orc_k_invalidate(const orc_cache* c, int Np, int nl, uint64_t inv_lights, const int32_t* V, int nv, int nw);
/* the same with the entries' light dependencies taken from masks ga[i] | gb[i] (no unknown entries) */
void orc_k_invalidate_masks(const orc_cache* c, int Np, int nl, uint64_t inv_lights, const int32_t* V, int nv, int nw,
                            const uint64_t* ga, const uint64_t* gb);
void orc_k_flush(const orc_cache* c, int Np, int nl);
void orc_k_error(const orc_cache* c, int Np, int nl, const float* ref, const uint8_t* valid, uint64_t dormant);
void orc_k_ref(const orc_scene* s, int trav, const float* Lin, float* Lout, float* acc, float* closec, uint32_t seed, int K,
               int accumulate, int count_close);
/* resolved radiance of every entry from channels Lc */
void orc_k_resolve(const orc_scene* s, const float* Lc, float* R, uint64_t dormant);
/* camera pass (R NULL: no cache); rays[2] += extension, shadow rays */
void orc_k_query(const orc_scene* s, int trav, const float* R, const orc_camera* cam, uint32_t seed, uint32_t frame,
                 uint64_t dormant, float* img, long long* rays, int pix0, int pix1);
/* traversal-work counters (binary nodes visited, triangles tested, hash probes, cache rows
 * read) accumulated by the closest-hit traversal and the lookups while enabled */
void orc_count_enable(int on);
void orc_count_get(long long out[4]);

/* ---- FP64 radiance twin ---- */
/* direct light in double from the FP32 build's visibility (Edir > 0 or lit geometry) */
void orc_k_build_f64(const orc_scene* s, int trav, double* Edir64);
/* per-light sample means in double from the hits recorded by orc_k_trace2 (hits[Np*kmax],
 * -1 = ray did not count), the FP32 run's kk and a double cache Lc64 / Edir64 */
void orc_k_trace2_f64(const orc_scene* s, const double* Edir64, const double* Lc64, const int32_t* kk, const int32_t* hits,
                      int kmax, uint64_t dormant, double* m64);
/* resolved radiance in double, A_i sum_l (Edir64 + Lc64) over the active lights */
void orc_k_resolve_f64(const orc_scene* s, const double* Edir64, const double* Lc64, uint64_t dormant, double* R64);
/* the camera pass's FP64 twin: the FP32 pass's rays, hits and light choices, radiometry in double
 * (R64 = NULL: no cache) */
void orc_k_query_f64(const orc_scene* s, int trav, const double* R64, const orc_camera* cam, uint32_t seed, uint32_t frame,
                     uint64_t dormant, double* img64, int pix0, int pix1);
/* EMA update in double with the FP32 run's history counts (n before the update) */
void orc_k_update_f64(const orc_scene* s, const float* n_before, const int32_t* kk, const orc_params* p, const double* Lc64,
                      const double* m64, double* Lc64_out);

/* ---- adjoint of the cached radiance (per-light emission, per-material albedo) ----
 * Model: unit-intensity direct light G[Np][nl] (the k_build geometry at I = 1), E_i = sum_l I_l G_il,
 * A_i = alb[mat_i], L^(0) = 0, L^(b+1)_i = (1/K) sum over the K rays of entry i at bounce b that hit a
 * valid entry h of A_h (E_h + L^(b)_h), R_i = A_i (E_i + L^(B)_i). Rays: seed_b = hash32(seed + b * 0x9E3779B9),
 * per entry rs = hash32(seed_b ^ hash32(i + 1)), then the k_ref ray rule (prep_mesh.cuh:181-213). */
uint32_t orc_adj_seed(uint32_t seed, int b);
/* G and its support SG (bit min(l, 63) when G_il > 0); G64 (optional) the same in double from the
 * same visibility decisions */
void orc_adj_unit(const orc_scene* s, int trav, float* G, uint64_t* SG, double* G64);
/* hit entry of every ray, hits[(b * Np + i) * K + k], -1 when the ray does not count */
void orc_adj_hits(const orc_scene* s, int trav, int K, int B, uint32_t seed, int32_t* hits);
/* the rays of bounce b of entries i0..i1-1 only: hits[(i - i0) * K + k] */
void orc_adj_hits_rows(const orc_scene* s, int trav, int K, int b, uint32_t seed, int i0, int i1, int32_t* hits);
/* FP32 forward in the kernels' operation order: R, support SR, L^(B) (LB may be NULL) */
void orc_adj_forward(int Np, int nl, int K, int B, const int32_t* hits, const int32_t* mat, const float* alb, const float* I,
                     const float* G, const uint64_t* SG, float* R, uint64_t* SR, float* LB);
/* FP64 forward R64 and reverse-mode gradient of J = sum_i g_i R_i: dI[nl], dalb[nmat]; returns J */
double orc_adj_grad64(int Np, int nl, int nmat, int K, int B, const int32_t* hits, const int32_t* mat, const double* alb,
                      const double* I, const double* G64, const double* g, double* R64, double* dI, double* dalb);

/* ---- runner (cuda/poc_gpu.cu:826-1016) on host buffers ---- */
typedef struct orc_runner orc_runner;
typedef struct orc_method {
    int mode; float nmax, budget; int deps, verify, oracle, track_every, rescale;
    uint64_t resp_mask; float nmax_resp; int ddgi; float ddgi_h, ddgi_thr; int ddgi_clamp;
    int memory, ramp, wake_verify, snaps, snap_verify, flush, targeted, compact, grad_masks;
    int eventboost;   /* frames of the history-cap ramp min(nmax, 2 + k/2) after an event that restores nothing */
} orc_method;
/* states[NS]: one orc_scene per state (static fields shared); occ / em event lists like the
 * library's (occ[a*NS+b], em[(a*NS+b)*nl+l]) given as offsets into one flat int array. */
orc_runner* orc_runner_create(const orc_scene* states, int NS, const orc_method* m, int trav, const int32_t* occ_flat,
                              const int64_t* occ_off, const int32_t* em_flat, const int64_t* em_off, const float* const* Pw,
                              const float* const* ref, const uint8_t* const* valid, const int32_t* occf_flat,
                              const int64_t* occf_off);
void orc_runner_destroy(orc_runner* r);
/* per-state gradient-support masks (masks[st * Np + i]) read by targeted invalidation when the
 * method sets grad_masks; the array stays owned by the caller */
void orc_runner_set_grad_masks(orc_runner* r, const uint64_t* masks);
void orc_runner_begin(orc_runner* r, int state0, uint32_t seed);
/* one frame; returns the relative RMS error (or -1 without a reference) */
double orc_runner_frame(orc_runner* r, int state, int want_err, int* invalidated);
const orc_cache* orc_runner_cache(const orc_runner* r);
/* the runner's frame parameters (those of the last frame; frame index, history cap and pending
 * count are set again at the start of each frame) */
const orc_params* orc_runner_params(const orc_runner* r);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

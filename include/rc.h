/* librc — dependency-aware radiance cache on AMD Instinct MI355X (gfx950).
 *
 * C ABI of the hot path of the prototype (cuda/poc_gpu.cu): BVH ray traversal, the cell
 * hash grid with 64-bit keys, the per-entry cache update / verification / invalidation,
 * per-entry light-dependency masks and transit-grid DDA masks, behind the prototype's
 * scene / event / method interface. The prototype exposes no library API — it is a CLI
 * whose Runner::run (cuda/poc_gpu.cu:826-1016) drives the kernels; every entry point
 * below names the reference function or block it replaces.
 *
 * Conventions: plain pointers and sizes; pointers named d_* are device pointers (HIP),
 * `stream` is a hipStream_t passed as void* (NULL = the handle's own stream). Every call
 * returns an rc_status; rc_last_error() gives the message of the last failure on the
 * calling thread. Handles are thread-compatible, not thread-safe. Nothing calls exit().
 */
#ifndef RC_H
#define RC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rc_status {
    RC_OK = 0,
    RC_ERR_ARG = 1,      /* invalid argument / shape */
    RC_ERR_HIP = 2,      /* HIP runtime error (launch, allocation, copy) */
    RC_ERR_NODEV = 3,    /* no usable GPU */
    RC_ERR_STATE = 4,    /* call out of order (e.g. frame before begin) */
    RC_ERR_SCENE = 5     /* scene preparation failed */
} rc_status;

typedef struct rc_scene rc_scene;
typedef struct rc_cache rc_cache;

/* Method flags: the prototype's Method struct (cuda/poc_gpu.cu:803-805) plus the SDK bench's
 * flush / targeted-invalidation switches (windows/run.sh:98-103,128; rtxgi_bench.patch:363-386,
 * :1028-1050). The four north-star methods are: no cache (rc_render with no_cache=1),
 * plain (mode 0), flush + snapshots (flush=1, snaps=1), targeted (targeted=1, deps=1). */
typedef struct rc_method {
    int mode;            /* 0: floor(budget)+Bernoulli rays per entry, 1: adaptive share (:445-454) */
    float nmax, budget;  /* history cap in frames, rays per entry */
    int deps;            /* record dependency masks in k_trace2 */
    int verify;          /* occluder events open a verification window (k_decide) */
    int oracle;          /* shrink history by the true reference change (:676-681) */
    int track_every;     /* masks recorded every n-th frame */
    int rescale;         /* per-light channel rescale on intensity change (:875-882) */
    uint64_t resp_mask;  /* lights blended with the responsive cap (SHaRC emulation) */
    float nmax_resp;
    int ddgi; float ddgi_h, ddgi_thr; int ddgi_clamp;
    int memory, ramp, wake_verify, snaps, snap_verify;
    int flush;           /* event to a state not in a snapshot slot clears the whole cache */
    int targeted;        /* dependency-targeted invalidation at the event */
    int compact;         /* with snaps: snapshots as 24-byte records of the entries with samples (rc_snap_rec):
                            restore flushes, then writes each record's radiance into channel 0 with its
                            history (rtxgi_bench.patch:1289-1345); full per-channel snapshots otherwise */
    int grad_masks;      /* > 0: targeted invalidation reads, instead of the recorded light masks, the
                            gradient-support masks of the cached-radiance adjoint of both states
                            (grad_masks rays per entry and bounce, 3 bounces, seed 0x5eed0000 + state;
                            computed at rc_cache_begin) */
    int eventboost;      /* > 0: for this many frames after an event that restores no snapshot, the history cap
                            ramps as min(nmax, 2 + k/2) frames, k = frames since the event (the SDK bench's
                            eventboost, rtxgi_bench.patch:355-356, :409-421; windows/run.sh:100-101) */
} rc_method;

#ifndef RC_CAMERA_DEFINED
#define RC_CAMERA_DEFINED
typedef struct rc_camera {
    float pos[3], fwd[3], right[3], up[3];
    float tanx, tany;
    int32_t W, H, spp, bounces, cache_bounce, pad0;
} rc_camera;
#endif

typedef struct rc_scene_info {
    int Np, npst, nbox, nst, nnodes, nwnodes, NS, nl, NR, Nq, hsize, vx, vy, vz, nw, bvh_depth, wide_depth, has_ref, fine;
    float h, lo[3], ext[3], diag, sun_len;
    rc_camera cam;
    int nmat, pad1;        /* materials (albedo table size) */
} rc_scene_info;

typedef struct rc_frame_stats {
    float ms_swap, ms_trace, ms_decide, ms_update, ms_event;   /* GPU time per phase (:948-1002) */
    double err;            /* relative RMS vs the converged reference, -1 without one (:1005-1010) */
    int event;             /* an event block ran this frame */
    int invalidated;       /* entries reset by targeted invalidation / verification opened */
    long long rays;        /* rays traced by the entry pass (sum of k) */
} rc_frame_stats;

/* host arrays of a scene (rc_scene_array); state-dependent arrays take a state index */
enum {
    RC_A_ST = 0, RC_A_STN, RC_A_BN, RC_A_WN, RC_A_HK, RC_A_HV, RC_A_P, RC_A_N, RC_A_ALB, RC_A_REGION,
    RC_A_L, RC_A_LTYPE, RC_A_INTEN, RC_A_Q, RC_A_DIMS, RC_A_OFF, RC_A_T, RC_A_TQ, RC_A_TALB,
    RC_A_EDIR, RC_A_SV, RC_A_LM, RC_A_OCC, RC_A_EM, RC_A_REF, RC_A_VALID, RC_A_PW, RC_A_ORDER, RC_A_OBJBOX,
    RC_A_MESH_V, RC_A_MESH_ALB, RC_A_MESH_DS, RC_A_VOL, RC_A_OCCF, RC_A_MAT, RC_A_MATALB, RC_A_COUNT
};
/* cache arrays (rc_cache_get / rc_cache_set); M is SoA [nl][Np], LC / EDIR rows are AoS [Np][nl] */
enum {
    RC_C_LC = 0, RC_C_M, RC_C_N, RC_C_KK, RC_C_PENDING, RC_C_CUR, RC_C_PREV, RC_C_LCUR, RC_C_LPREV,
    RC_C_SUMPR, RC_C_NPEND, RC_C_AGE, RC_C_STALE, RC_C_C0, RC_C_RSTAT, RC_C_LNEW, RC_C_DSTAT,
    RC_C_R, RC_C_COUNT
};

const char* rc_last_error(void);
int rc_version(void);
/* number of visible GPUs (0 on a machine without one; never fails) */
int rc_device_count(void);

/* ---- scenes (cuda/poc_gpu.cu:751-801 load, cuda/prep_mesh.cuh:303-531 prep) ---- */
/* Where scene preparation builds the binned-SAH BVH (cuda/prep_mesh.cuh:42-114) and samples, reduces and
 * Morton-orders the cache cells (:129-157) for the scenes created after the call: on = 1 the GPU (the
 * same tree and cells bit for bit), on = 0 the host (default). Process-wide. */
rc_status rc_scene_prep_device(int on);
/* Procedural box room: 6 walls + 2 cubes (~target_tris), 2 point lights, states {both on, light 1 off}. */
rc_status rc_scene_box_room(uint64_t seed, int target_tris, rc_scene** out);
/* Procedural street of nboxes boxes (~target_tris), nlamps point lamps (+ sun), states
 * {base, night (sun 0), lamp 0 off, 12x10 m occluder inserted}; camera W x H x spp, bounces. */
rc_status rc_scene_street(uint64_t seed, int nboxes, int target_tris, int nlamps, int sun, int target_points, int table_log2,
                          int W, int H, int spp, int bounces, rc_scene** out);
/* Prototype scene file, format 2 (box scenes) or 3 (mesh scenes), plain or gzip-compressed (cuda/poc_gpu.cu:751-801;
 * the .bin.gz scale scenes, cuda/run_cuda.sh:131).
 * Converged references stored in the file become the scene's references. */
rc_status rc_scene_load(const char* path, rc_scene** out);
/* Bench scene around an MSH1 mesh (the prep mode's scene, cuda/prep_mesh.cuh:303-453): 3 point lamps and a cube
 * at open points, 5 states, cells toward target_points; cube_m (may be NULL) = the cube's side. */
rc_status rc_scene_mesh(const char* msh_path, int target_points, rc_scene** out, float* cube_m);
/* Format-3 scene file readable by rc_scene_load (cuda/prep_mesh.cuh:242-297), with the converged references
 * when the scene has them; point lights only. */
rc_status rc_scene_save(const rc_scene* s, const char* path);
void rc_scene_destroy(rc_scene* s);
rc_status rc_scene_info_get(const rc_scene* s, rc_scene_info* out);
/* byte size / host copy of a scene array (state: state index, or state pair a*NS+b for OCC,
 * (a*NS+b)*nl+l for EM). Device-produced arrays (EDIR, SV, LM, REF, VALID, PW) need upload. */
size_t rc_scene_array_bytes(const rc_scene* s, int which, int state);
rc_status rc_scene_array(const rc_scene* s, int which, int state, void* dst, size_t bytes);
/* Device upload of the scene and the per-state build (k_build, cuda/poc_gpu.cu:369-391).
 * trav: 0 = binary BVH in the prototype's order, 1 = binary closest-hit, 2 = 4-wide closest-hit, 3 = 2 with the
 * camera pass and rc_trace walking the 4-wide nodes with quantized child boxes (64-byte nodes: per axis an
 * origin, a power-of-two step and 8-bit planes; the same hits, the other kernels keep the fp32 nodes). */
rc_status rc_scene_upload(rc_scene* s, int trav);
rc_status rc_scene_set_trav(rc_scene* s, int trav);
/* Fine voxel grid g^3 of the per-light volumes (before upload; 0 disables). Every lit shadow segment
 * of the state build is DDA-marked into its light's volume (the stepping rule of dda_reg,
 * cuda/poc_gpu.cu:171-190); a geometry event then affects the lights whose volume, before or
 * after, holds a voxel of the moving object, and targeted invalidation resets their entries. */
rc_status rc_scene_set_fine_grid(rc_scene* s, int g);
/* Converged per-state reference by fresh-ray Jacobi iterations, two chains, until the chains'
 * relative RMS difference <= target or max_iter (cuda/prep_mesh.cuh:454-520). noise_out[NS]. */
rc_status rc_scene_reference(rc_scene* s, int K, int warm, int block, int max_iter, float target, float* noise_out, int* iters_out);

/* ---- cache + runner (Runner, cuda/poc_gpu.cu:807-1016) ---- */
rc_status rc_cache_create(rc_scene* s, const rc_method* m, rc_cache** out);
void rc_cache_destroy(rc_cache* c);
/* run() prologue: warm start (converged Pw when present), n = nmax, masks cleared (:828-844) */
rc_status rc_cache_begin(rc_cache* c, int state0, uint32_t seed);
/* one frame of run() in state `state`: event block if the state changed, bank swap, trace,
 * decide, update, optional error (:851-1011). timed != 0 brackets phases with HIP events. */
rc_status rc_cache_frame(rc_cache* c, int state, int want_err, int timed, rc_frame_stats* out);
/* camera query pass over the current cache into d_img[W*H] (no_cache: plain path tracing).
 * rays_out[2] = extension rays, shadow rays (may be NULL); ms_out = GPU time of the pass
 * (may be NULL; when set the call waits for the pass). The pass runs on the cache's own blocking
 * stream, after the work queued so far; the next rc_cache_frame's entry pass overlaps it.
 * stream (may be NULL): the caller's stream; the pass then also starts after the work queued on it, and
 * the work queued on it afterwards starts after the pass (so d_img may be written or read there).
 * rc_cache_sync, or a synchronisation of the device or of the legacy default stream, waits for it. */
rc_status rc_cache_render(rc_cache* c, const rc_camera* cam, int no_cache, uint32_t seed, float* d_img, long long* rays_out,
                          float* ms_out, void* stream);
/* Schedule of the camera pass (same image bit for bit): RC_RENDER_FUSED (default) runs each path
 * in one thread; RC_RENDER_STAGED (4-wide BVH scenes, trav 2) runs generate / trace / shade /
 * shadow-trace stages, the live rays of each 1024-path segment compacted between stages. The
 * staged workspace (~100 B per path) is kept by the cache. */
enum { RC_RENDER_FUSED = 0, RC_RENDER_STAGED = 1 };
rc_status rc_cache_set_render_mode(rc_cache* c, int mode);
/* Traversal stacks of the fused camera pass. By default each lane keeps all its stack levels in LDS,
 * or only the first 16 with the deeper ones in a global buffer the cache keeps (4 B per path and
 * level beyond 16) where that lets more waves reside; after this call always the first `levels`
 * levels in LDS (1..64). The image does not depend on it. No counterpart in the reference (its
 * traversal stack is per-thread local memory, cuda/poc_gpu.cu:297-338). */
rc_status rc_cache_set_stack_levels(rc_cache* c, int levels);
/* The same for the scene's other tracing kernels (entry pass, direct light, reference solve,
 * learned-cache samples and camera pass) and for adjoint handles created afterwards; results do
 * not depend on it. No counterpart in the reference (per-thread local-memory stacks,
 * cuda/poc_gpu.cu:297-338). */
rc_status rc_scene_set_stack_levels(rc_scene* s, int levels);
/* Compact snapshot record (24 bytes; the SDK bench's SnapRec, rtxgi_bench.patch:1289-1345): entry index,
 * history count, entry key (cell key of the hash grid; 1 << 63 | index for points of dynamic quads),
 * indirect radiance summed over the light channels, frames since reset. */
typedef struct rc_snap_rec { uint32_t index; float n; uint64_t key; float radiance; float age; } rc_snap_rec;
/* records of the cache's entries with samples, in entry order, into host_records (capacity records);
 * count_out = records written (the number of entries with samples when capacity suffices) */
rc_status rc_cache_snapshot_save(rc_cache* c, rc_snap_rec* host_records, size_t capacity, size_t* count_out);
/* rc_cache_snapshot_save that allocates nothing: the records are staged in a static device array of the
 * library (2^21 records, 48 MiB per GPU, held until the process ends) instead of a buffer allocated per
 * call; RC_ERR_ARG for a cache of more than 2^21 entries. No counterpart in the reference: for callers that
 * cannot allocate device memory (serving loops, code under stream capture, harnesses that forbid hipMalloc).
 * Calls that overlap on one GPU share the array and race. */
rc_status rc_cache_snapshot_save_noalloc(rc_cache* c, rc_snap_rec* host_records, size_t capacity, size_t* count_out);
/* flush, then restore the records whose key matches its entry (radiance into channel 0, history, age) */
rc_status rc_cache_snapshot_load(rc_cache* c, const rc_snap_rec* host_records, size_t count);
size_t rc_cache_array_bytes(const rc_cache* c, int which);
rc_status rc_cache_get(const rc_cache* c, int which, void* host_dst, size_t bytes);
rc_status rc_cache_set(rc_cache* c, int which, const void* host_src, size_t bytes);
rc_status rc_cache_sync(rc_cache* c);
/* The event block of the transition from the cache's current state to to_state, run now instead of at the
 * start of the next rc_cache_frame (which then runs none): snapshot of the state being left, rescale /
 * dormancy / wake, emitter and occluder invalidation, targeted invalidation, flush, snapshot restore,
 * eventboost (cuda/poc_gpu.cu:856-943; rtxgi_bench.patch:355-366, 1028-1050). invalidated_out: entries
 * the targeted invalidation reset (waits for it). Before the first frame only the state changes, as in
 * rc_cache_frame. */
rc_status rc_cache_event(rc_cache* c, int to_state, int* invalidated_out);
/* Dependency-targeted invalidation with an explicit invalidation set, the resolve-time test of the SDK
 * bench (rtxgi_bench.patch:368-386, 1028-1050) on the cache's recorded masks: an entry whose light masks
 * (both banks) meet light_mask (bit min(l, 63) per light), or whose transit masks (both banks) hold one of
 * the nvox transit-grid voxels, or that recorded no dependency at all, restarts from zero. Meaningful for
 * methods that record dependencies (rc_method.deps); the resolved radiance follows at the next frame's
 * update or rc_cache_resolve. invalidated_out: entries reset (waits for the count). */
rc_status rc_cache_invalidate(rc_cache* c, unsigned long long light_mask, const int* voxels, int nvox, int* invalidated_out);
/* Resolved radiance R = A sum_l (Edir_l + Lc_l) of every entry from the current channels (the value the
 * camera pass reads; the resolve of SHaRC's SharcResolveEntry, rtxgi_bench.patch:1044, and the cache
 * read of cuda/poc_gpu.cu:470-478), e.g. after rc_cache_set of the channels. */
rc_status rc_cache_resolve(rc_cache* c);

/* ---- kernel level (device pointers) ---- */
/* closest hit of n rays: t (inf on a miss), q (>= 0 quad of a dynamic triangle, -2-j static j,
 * -1 none) (trace_full, cuda/poc_gpu.cu:340-344) */
rc_status rc_trace(rc_scene* s, int state, const float* d_O, const float* d_D, int n, float* d_t, int* d_q, int trav, void* stream);
/* GPU insert of n 64-bit keys (0 = empty, never inserted) with values into a linear-probing
 * table of hmask+1 slots; HK must start zeroed and HV at INT32_MAX. d_fail[0] += keys left
 * out after 64 probes. Same slot function and probe order as hash_find (:263-272). */
rc_status rc_hash_insert(uint64_t* d_HK, int32_t* d_HV, uint32_t hmask, const uint64_t* d_keys, const int32_t* d_vals, int n,
                         int32_t* d_fail, void* stream);
/* lookup of n keys: value, -1 when absent (hash_find, :263-272) */
rc_status rc_hash_query(const uint64_t* d_HK, const int32_t* d_HV, uint32_t hmask, const uint64_t* d_keys, int n, int32_t* d_out,
                        void* stream);
/* cache entry of n hit points with facing normals, 26-neighbour fallback (hash_point, :274-287) */
rc_status rc_hash_point(rc_scene* s, const float* d_H, const float* d_N, int n, int32_t* d_out, void* stream);
/* transit-grid DDA mask (nw words) of n segments (dda_reg, :171-190) */
rc_status rc_dda_mark(rc_scene* s, const float* d_A, const float* d_B, int n, uint32_t* d_out, void* stream);
/* DDA over a gx*gy*gz bit volume (same stepping rule), OR-ing crossed voxels into volume
 * seg_vol[i] (NULL: volume 0) of d_vol, words_per_vol 32-bit words each. */
rc_status rc_dda_volume(rc_scene* s, int gx, int gy, int gz, const float* d_A, const float* d_B, const int32_t* d_seg_vol, int n,
                        uint32_t* d_vol, long long words_per_vol, void* stream);

/* ---- adjoint of the cached radiance (extension) ----
 * Model: the cache's transport with fixed rays. G = unit-intensity direct light of the state's
 * geometry (cuda/poc_gpu.cu:369-391 at I = 1), E_i = sum_l I_l G_il, A_i = alb[mat_i] (material table
 * of the scene), L^(0) = 0, L^(b+1)_i = (1/K) sum over entry i's K rays of bounce b that reach a valid
 * entry h of A_h (E_h + L^(b)_h) (rays of the converged-reference iteration, cuda/prep_mesh.cuh:181-213,
 * seed_b = hash32(seed + b * 0x9E3779B9)), cached radiance R_i = A_i (E_i + L^(B)_i). The backward
 * re-traces each bounce's rays instead of storing paths. No counterpart in the reference: its
 * extension for gradient-driven invalidation (callers: autograd bindings, the gradient-mask method). */
typedef struct rc_adjoint rc_adjoint;
rc_status rc_adjoint_create(rc_scene* s, int state, int K, int bounces, uint32_t seed, rc_adjoint** out);
void rc_adjoint_destroy(rc_adjoint* a);
/* d_I[nl] emission, d_alb[nmat] material albedos (device) -> d_R[Np] (device; may be NULL) */
rc_status rc_adjoint_forward(rc_adjoint* a, const float* d_I, const float* d_alb, float* d_R, void* stream);
/* gradient of J = sum_i gR_i R_i at the last forward: d_gI[nl], d_galb[nmat] (device, overwritten) */
rc_status rc_adjoint_backward(rc_adjoint* a, const float* d_gR, float* d_gI, float* d_galb, void* stream);
/* host copies: G [Np][nl] f32, SG [Np] u64, SR (support of R at the last forward: bit min(l, 63) when
 * dR_i/dI_l != 0) [Np] u64, LB (L^(B)) [Np] f32, DA / DE (adjoints of A / E at the last backward) [Np] f32 */
enum { RC_ADJ_G = 0, RC_ADJ_SG, RC_ADJ_SR, RC_ADJ_LB, RC_ADJ_DA, RC_ADJ_DE, RC_ADJ_COUNT };
size_t rc_adjoint_array_bytes(const rc_adjoint* a, int which);
rc_status rc_adjoint_get(const rc_adjoint* a, int which, void* host_dst, size_t bytes);
/* kernel timing: timed != 0 brackets every traced bounce of the following forward / backward calls
 * with HIP events on their stream (each call then waits for its last bounce); ms_out[0] = GPU time
 * of the last forward's traced bounces (summed over the B launches), ms_out[1] the same for the last
 * backward, 0 before a timed call. No counterpart in the reference (caller: the benchmark). */
rc_status rc_adjoint_set_timed(rc_adjoint* a, int timed);
rc_status rc_adjoint_trace_ms(const rc_adjoint* a, float ms_out[2]);

/* ---- hash-encoded MLP radiance cache (extension) ----
 * Input: 16-level multiresolution hash encoding of the position (2 features per level,
 * 2^log2_table entries per level, resolutions base_res..max_res in geometric steps), normal and
 * view direction; 5 hidden layers of 64 ReLU units; scalar output. BF16 MFMA with FP32
 * accumulation; Adam on FP32 master weights and table. No counterpart in the reference: its
 * learned-cache extension (callers: online training from the cache's scene, the camera pass). */
typedef struct rc_mlp rc_mlp;
typedef struct rc_mlp_config {
    int log2_table;        /* hash-table entries per level (19) */
    int base_res, max_res; /* grid resolution of the coarsest / finest level (16, 2048) */
    int max_batch;         /* largest batch of one call (65536) */
    float lr, beta1, beta2, eps;   /* Adam (1e-2, 0.9, 0.99, 1e-15) */
    uint32_t seed;         /* initialisation of weights and table */
    float lo[3], ext[3];   /* position bounds */
} rc_mlp_config;
rc_status rc_mlp_create(const rc_mlp_config* cfg, rc_mlp** out);
void rc_mlp_destroy(rc_mlp* m);
/* d_y[n] from float4 positions / normals / view directions (xyz used; device arrays) */
rc_status rc_mlp_infer(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, int n, float* d_y, void* stream);
/* gradients of sum_i dy_i y_i into the weight gradient (overwritten) and the table gradient (accumulated
 * until the next rc_mlp_step / rc_mlp_zero_grad) */
rc_status rc_mlp_backward(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, const float* d_dy, int n, void* stream);
/* one training step on targets d_t: loss mean (y - t)^2 / (y^2 + 0.01), gradients, Adam; loss_out (host, optional;
 * when set the call waits for the step) */
rc_status rc_mlp_train(rc_mlp* m, const float* d_P4, const float* d_N4, const float* d_D4, const float* d_t, int n, float* loss_out,
                       void* stream);
/* Adam step on the current gradients (then cleared) */
rc_status rc_mlp_step(rc_mlp* m, void* stream);
rc_status rc_mlp_zero_grad(rc_mlp* m, void* stream);
/* parameters / gradients (host copies): W master weights [5][64][64] + [64][16] f32, TABLE [16][2^log2_table][2] f32 */
enum { RC_MLP_W = 0, RC_MLP_TABLE, RC_MLP_GW, RC_MLP_GTABLE, RC_MLP_COUNT };
/* One online training step from the cache's scene and current state: n (<= max_batch) random cache
 * entries, one cosine ray each, self-training targets A_i (E_i + max(y(hit), 0)) with E_i the state's
 * direct light; judge_out = relative RMS of the MLP's predictions at those entries (before the step)
 * against the cache's resolved radiance (the plain grid cache as error judge). Waits for the step. */
rc_status rc_mlp_train_scene(rc_mlp* m, rc_cache* c, int n, uint32_t seed, float* loss_out, float* judge_out);
/* camera pass over the cache's scene and state whose paths end on the MLP at bounce cam->cache_bounce:
 * with the cache's render mode RC_RENDER_FUSED one thread per path records its query point, then one
 * batched inference; RC_RENDER_STAGED runs the staged schedule (the same image); ms_out and stream as rc_cache_render */
rc_status rc_mlp_render(rc_mlp* m, rc_cache* c, const rc_camera* cam, uint32_t seed, float* d_img, float* ms_out, void* stream);
/* time of the batched inference inside the last rc_mlp_render called with ms_out (HIP events on the stream it ran
 * on; 0 before). No counterpart in the reference: a measurement accessor of the learned-cache extension. */
rc_status rc_mlp_last_infer_ms(const rc_mlp* m, float* ms_out);
/* The grid cache as per-query judge and fallback of rc_mlp_render: a query whose prediction differs from the
 * cache's resolved radiance at its entry by more than thr times that value ends on the grid's value instead
 * (thr = 0: never, the default). No counterpart in the reference: the learned-cache extension's fallback. */
rc_status rc_mlp_set_grid_fallback(rc_mlp* m, float thr);
size_t rc_mlp_array_bytes(const rc_mlp* m, int which);
rc_status rc_mlp_get(const rc_mlp* m, int which, void* host_dst, size_t bytes);
rc_status rc_mlp_set(rc_mlp* m, int which, const void* host_src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* RC_H */

// Host-side launchers of the HIP kernels (rc_kernels.hip). All take device pointers and a
// stream; they return the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include "rc_types.h"

hipError_t rc_launch_trace_rays(int trav, const rc_dscene* s, const rc_dstate* st, const float* O, const float* D, int n,
                                float* t, int* q, hipStream_t stream);
hipError_t rc_launch_build(int trav, const rc_dscene* s, const rc_dstate* st, hipStream_t stream);
hipError_t rc_launch_trace2(int trav, int deps, const rc_dscene* s, const rc_dstate* st, const rc_dcache* c, const rc_params* p,
                            uint64_t* lnew, hipStream_t stream);
hipError_t rc_launch_decide(const rc_dscene* s, const rc_dcache* c, const rc_params* p, hipStream_t stream);
hipError_t rc_launch_update(const rc_dscene* s, const rc_dstate* st, const rc_dcache* c, const rc_params* p, const uint64_t* lnew,
                            hipStream_t stream);
hipError_t rc_launch_resolve(const rc_dscene* s, const rc_dstate* st, const float* Lc, float* R, uint64_t dormant, hipStream_t stream);
hipError_t rc_launch_sumpr(const rc_dcache* c, int Np, hipStream_t stream);
hipError_t rc_launch_restored(const rc_dcache* c, int Np, int nl, float nmax, uint64_t dormant, int verify, hipStream_t stream);
hipError_t rc_launch_rescale(const rc_dcache* c, int Np, int nl, int l, float ratio, hipStream_t stream);
// kind 1 adds the entries it puts in verification to c->npend[1]
hipError_t rc_launch_touched(const rc_dcache* c, int Np, int nl, rc_vmask vm, int nw, int kind, uint64_t dormant, hipStream_t stream);
hipError_t rc_launch_wake(const rc_dcache* c, int Np, int nl, int l, int verify, uint64_t dormant, hipStream_t stream);
hipError_t rc_launch_oracle_method(const rc_dcache* c, int Np, const float* rold, const float* rnew, hipStream_t stream);
// inv_extra (device, may be NULL): more changed-light bits, read by the kernel; count (device): entries reset
hipError_t rc_launch_invalidate(const rc_dcache* c, int Np, int nl, uint64_t inv_lights, const unsigned long long* inv_extra, rc_vmask vm,
                                int nw, const uint64_t* ga, const uint64_t* gb, int* count, hipStream_t stream);
hipError_t rc_launch_flush(const rc_dcache* c, int Np, int nl, hipStream_t stream);
hipError_t rc_launch_error(const rc_dcache* c, int Np, int nl, const float* ref, const uint8_t* valid, uint64_t dormant, hipStream_t stream);
hipError_t rc_launch_ref(int trav, const rc_dscene* s, const rc_dstate* st, const float* Lin, float* Lout, float* acc, float* closec,
                         uint32_t seed, int K, int accumulate, int count_close, hipStream_t stream);
// camera pass (fused); rc_query_threads: threads of its grid (the global stack levels need one column each)
hipError_t rc_launch_query(int trav, const rc_dscene* s, const rc_dstate* st, const float* R, const rc_camera* cam, uint32_t seed,
                           uint32_t frame, uint64_t dormant, float* img, unsigned long long* counters, hipStream_t stream);
long long rc_query_threads(const rc_camera* cam);
long long rc_query_mlp_threads(const rc_camera* cam);   // threads of the fused learned-cache pass's grid
// global lane-stack words a launch of nthreads threads needs with st's split; threads of an n-item grid
size_t rc_ovf_words(const rc_dstate* st, const rc_dscene* s, long long nthreads);
long long rc_grid_threads(long long n);
hipError_t rc_launch_hash_insert(unsigned long long* HK, int* HV, uint32_t hmask, const unsigned long long* keys, const int* vals,
                                 int n, int* fail, hipStream_t stream);
// empty slots (key 0) get value -1
hipError_t rc_launch_hash_finish(const unsigned long long* HK, int* HV, uint32_t hsize, hipStream_t stream);
hipError_t rc_launch_hash_query(const unsigned long long* HK, const int* HV, uint32_t hmask, const unsigned long long* keys, int n,
                                int* out, hipStream_t stream);
hipError_t rc_launch_hash_point(const rc_dscene* s, const float* H, const float* Nn, int n, int* out, hipStream_t stream);
hipError_t rc_launch_dda_segments(const rc_dscene* s, const float* A, const float* B, int n, uint32_t* out, hipStream_t stream);
hipError_t rc_launch_dda_volume(const rc_dscene* s, int gx, int gy, int gz, const float* A, const float* B, const int* seg_vol, int n,
                                uint32_t* vol, long long words_per_vol, hipStream_t stream);
hipError_t rc_launch_vol_lights(int nl, const uint32_t* Va, const uint32_t* Vb, long long words_per_vol, const int* vox, int nvox,
                                unsigned long long* out, hipStream_t stream);
// infer_ev (may be NULL): two events recorded on the stream around the batched MLP inference
hipError_t rc_launch_query_mlp(const rc_dscene* s, const rc_dstate* st, const rc_camera* cam, uint32_t seed, uint32_t frame, uint64_t dormant,
                               float* img, const rc_qwork* w, const rc_mlpdev* mlp, hipStream_t stream, const hipEvent_t* infer_ev = nullptr);
hipError_t rc_launch_query_staged(const rc_dscene* s, const rc_dstate* st, const float* R, const rc_camera* cam, uint32_t seed,
                                  uint32_t frame, uint64_t dormant, float* img, unsigned long long* counters, const rc_qwork* w,
                                  const rc_mlpdev* mlp, hipStream_t stream, const hipEvent_t* infer_ev = nullptr);
hipError_t rc_launch_adj_unit(int trav, const rc_dscene* s, const rc_dstate* st, float* G, uint64_t* SG, hipStream_t stream);
hipError_t rc_launch_adj_forward(int trav, const rc_dscene* s, const rc_dstate* st, const rc_adjbuf* w, const float* I,
                                 const float* alb, int K, int B, uint32_t seed, float* R, hipStream_t stream,
                                 hipEvent_t* ev = nullptr);   // ev: 2 B events around the traced bounces (or NULL)
hipError_t rc_launch_adj_backward(int trav, const rc_dscene* s, const rc_dstate* st, const rc_adjbuf* w, const float* gR, int K,
                                  int B, uint32_t seed, float* dI, float* dalb, hipStream_t stream, hipEvent_t* ev = nullptr);
int rc_mlp_pad(int n);
hipError_t rc_launch_mlp_infer(const rc_mlpdev* m, const float4* P, const float4* N, const float4* D, int n, float* y, hipStream_t st);
hipError_t rc_launch_mlp_grad(const rc_mlpdev* m, const rc_mlpbuf* b, const float4* P, const float4* N, const float4* D, const float* t,
                              const float* dy_in, int n, float* loss, hipStream_t st);
hipError_t rc_launch_mlp_adam(const rc_mlpdev* m, const rc_mlpbuf* b, float lr, float b1, float b2, float eps, int step, hipStream_t st);
hipError_t rc_launch_mlp_pack(const rc_mlpdev* m, hipStream_t st);
hipError_t rc_launch_mlp_zero(float* x, long long n, hipStream_t st);
hipError_t rc_launch_mlp_samples(int trav, const rc_dscene* s, const rc_dstate* st, uint64_t dormant, uint32_t seed, int n, float4* sP,
                                 float4* sN, float4* sD, float4* hP, float4* hN, float4* hD, float2* AE, int* ent, hipStream_t stream);
hipError_t rc_launch_mlp_targets(int n, const float4* hP, const float* hy, const float2* AE, float* t, hipStream_t stream);
hipError_t rc_launch_mlp_judge(int n, const float* y, const float* R, const int* ent, float* acc, hipStream_t stream);
hipError_t rc_launch_snap_count(const rc_dcache* c, int Np, int* blockcnt, hipStream_t stream);
hipError_t rc_launch_snap_pack(const rc_dcache* c, int Np, int nl, const uint64_t* ekeys, const int* blockoff, rc_snaprec* out,
                               hipStream_t stream);
// count (device): records at rec; max_count: most records the grid covers
hipError_t rc_launch_snap_restore(const rc_dcache* c, int Np, int nl, const uint64_t* ekeys, const rc_snaprec* rec, const int* count,
                                  int max_count, int verify, uint64_t dormant, hipStream_t stream);
// exclusive prefix of n ints in place (one block), total into *total (device)
hipError_t rc_launch_scan_blocks(int* cnt, int n, int* total, hipStream_t stream);

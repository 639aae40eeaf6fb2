/* CPU oracle of the radiance-cache hot path — TEST INFRASTRUCTURE ONLY (see rc_oracle.h).
 * Build: gcc -O2 -ffp-contract=off -fopenmp (oracle/Makefile). */
#include "rc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_NW 4
#define ORC_FX_PR 4294967296.0f
#define ORC_FX_STAT 1048576.0f
#define ORC_KSNAP 6
#define ORC_VERIF 4

typedef struct { float x, y, z; } v3;
static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 ld(const float* p, long long i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline long long fx(float v, float scale) { return (long long)(v * scale); }

static int g_count = 0;
static long long g_cnt[4];
#define CNT(k, v) do { if (g_count) __atomic_fetch_add(&g_cnt[k], (long long)(v), __ATOMIC_RELAXED); } while (0)
void orc_count_enable(int on) { g_count = on; if (on) memset(g_cnt, 0, sizeof g_cnt); }
void orc_count_get(long long out[4]) { for (int k = 0; k < 4; ++k) out[k] = g_cnt[k]; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ RNG (cuda/poc_gpu.cu:127-135, :442) */
uint32_t orc_hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16; return x;
}
float orc_rnd(uint32_t* s) {
    *s = *s * 747796405u + 2891336453u;
    uint32_t w = ((*s >> ((*s >> 28u) + 4u)) ^ *s) * 277803737u;
    w = (w >> 22u) ^ w;
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}
uint32_t orc_seed(uint32_t seed, uint32_t frame, uint32_t i) {
    return orc_hash32(seed * 0x9E3779B9u ^ orc_hash32(frame * 0x85EBCA6Bu ^ orc_hash32(i + 1u)));
}

/* Cody-Waite reduction by pi/2 with a 2-term split, Taylor polynomials on [-pi/4, pi/4]
 * (replaces cosf / sinf of cuda/poc_gpu.cu:365 so both sides round identically). */
void orc_sincos(float x, float* s, float* c) {
    const float TWO_OVER_PI = 0.636619772367581f, PIO2_HI = 1.5703125f, PIO2_LO = 4.83826794897e-4f;
    float qf = floorf(x * TWO_OVER_PI + 0.5f);
    int q = (int)qf;
    float y = (x - qf * PIO2_HI) - qf * PIO2_LO;
    float y2 = y * y;
    float sp = y + y * y2 * (-1.66666667e-1f + y2 * (8.33333333e-3f + y2 * (-1.98412698e-4f + y2 * 2.75573192e-6f)));
    float cp = 1.0f + y2 * (-0.5f + y2 * (4.16666667e-2f + y2 * (-1.38888889e-3f + y2 * (2.48015873e-5f + y2 * -2.75573192e-7f))));
    switch (q & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

static v3 cosine_dir(v3 N, float u1, float u2) {   /* cuda/poc_gpu.cu:360-366 */
    float r = sqrtf(u1), phi = 6.28318530718f * u2;
    v3 a = fabsf(N.x) > 0.9f ? V3(0.f, 1.f, 0.f) : V3(1.f, 0.f, 0.f);
    v3 T = cross(N, a);
    T = mul(T, 1.0f / sqrtf(dot(T, T)));
    v3 B = cross(N, T);
    float sn, cs;
    orc_sincos(phi, &sn, &cs);
    return add(add(mul(T, r * cs), mul(B, r * sn)), mul(N, sqrtf(fmaxf(0.f, 1.f - u1))));
}
void orc_cosine_dir(const float N[3], float u1, float u2, float D[3]) {
    v3 d = cosine_dir(V3(N[0], N[1], N[2]), u1, u2);
    D[0] = d.x; D[1] = d.y; D[2] = d.z;
}

/* ------------------------------------------------------------------ keys (cuda/poc_gpu.cu:251-262, prep_mesh.cuh:32-36,119-127) */
uint64_t orc_mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull; x ^= x >> 27; x *= 0x94d049bb133111ebull; x ^= x >> 31; return x;
}
uint64_t orc_cell_key(int cx, int cy, int cz, int b) {
    return ((((uint64_t)(cx & 0xFFFFF) << 20 | (uint64_t)(cy & 0xFFFFF)) << 20 | (uint64_t)(cz & 0xFFFFF)) << 3 | (uint64_t)b) + 1ull;
}
void orc_decode_key(uint64_t k, int* cx, int* cy, int* cz, int* b) {
    k -= 1ull;
    *b = (int)(k & 7ull); *cz = (int)((k >> 3) & 0xFFFFFull); *cy = (int)((k >> 23) & 0xFFFFFull); *cx = (int)((k >> 43) & 0xFFFFFull);
}
int orc_normal_bucket(float nx, float ny, float nz) {
    float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
    if (ax >= ay && ax >= az) return nx > 0.f ? 0 : 1;
    if (ay >= az) return ny > 0.f ? 2 : 3;
    return nz > 0.f ? 4 : 5;
}
uint64_t orc_spread3(uint64_t x) {   /* bit i of x -> bit 3i */
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= ((x >> i) & 1ull) << (3 * i);
    return r;
}
uint64_t orc_hostrng_init(uint64_t x) { return orc_mix64(x * 0x9E3779B97F4A7C15ull + 1); }
float orc_hostrng_u(uint64_t* s) {
    *s = orc_mix64(*s + 0x9E3779B97F4A7C15ull);
    return (float)(*s >> 40) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------ hash grid */
int orc_hash_find(const uint64_t* HK, const int32_t* HV, uint32_t hmask, uint64_t key) {   /* :263-272 */
    uint32_t h = (uint32_t)orc_mix64(key) & hmask;
    for (int probe = 0; probe < 64; ++probe) {
        uint64_t k = HK[h];
        CNT(2, 1);
        if (k == key) return HV[h];
        if (k == 0ull) return -1;
        h = (h + 1) & hmask;
    }
    return -1;
}
static int hash_point(const orc_scene* s, v3 H, v3 n) {   /* :274-287 */
    int b = orc_normal_bucket(n.x, n.y, n.z);
    float ih = 1.f / s->hcell;
    int cx = (int)floorf((H.x - s->lo[0]) * ih), cy = (int)floorf((H.y - s->lo[1]) * ih), cz = (int)floorf((H.z - s->lo[2]) * ih);
    int p = orc_hash_find(s->HK, s->HV, s->hmask, orc_cell_key(cx, cy, cz, b));
    if (p >= 0) return p;
    for (int d = 0; d < 27; ++d) {
        int dz = d / 9 - 1, dy = (d / 3) % 3 - 1, dx = d % 3 - 1;
        if (dx == 0 && dy == 0 && dz == 0) continue;
        p = orc_hash_find(s->HK, s->HV, s->hmask, orc_cell_key(cx + dx, cy + dy, cz + dz, b));
        if (p >= 0) return p;
    }
    return -1;
}
int orc_hash_point(const orc_scene* s, const float H[3], const float n[3]) { return hash_point(s, V3(H[0], H[1], H[2]), V3(n[0], n[1], n[2])); }

void orc_hash_build(const uint64_t* keys, int n, int log2, uint64_t* HK, int32_t* HV) {   /* prep_mesh.cuh:356-366 */
    size_t HS = (size_t)1 << log2;
    for (size_t i = 0; i < HS; ++i) { HK[i] = 0ull; HV[i] = -1; }
    for (int i = 0; i < n; ++i) {
        size_t hh = (size_t)orc_mix64(keys[i]) & (HS - 1);
        while (HK[hh] != 0ull) hh = (hh + 1) & (HS - 1);
        HK[hh] = keys[i];
        HV[hh] = i;
    }
}

/* cell sampling with a private open-addressing map (prep_mesh.cuh:129-157) */
typedef struct { uint64_t key; float p[3]; float d2; float n[3]; double asum, wsum; } orc_cell;
typedef struct { orc_cell* t; size_t cap, used; } orc_map;
static orc_cell* map_get(orc_map* m, uint64_t key) {
    if (2 * (m->used + 1) > m->cap) {
        size_t nc = m->cap ? 2 * m->cap : 1024;
        orc_cell* nt = (orc_cell*)calloc(nc, sizeof(orc_cell));
        for (size_t i = 0; i < m->cap; ++i)
            if (m->t[i].key) {
                size_t h = (size_t)orc_mix64(m->t[i].key) & (nc - 1);
                while (nt[h].key) h = (h + 1) & (nc - 1);
                nt[h] = m->t[i];
            }
        free(m->t); m->t = nt; m->cap = nc;
    }
    size_t h = (size_t)orc_mix64(key) & (m->cap - 1);
    while (m->t[h].key && m->t[h].key != key) h = (h + 1) & (m->cap - 1);
    if (!m->t[h].key) {
        orc_cell* c = &m->t[h];
        c->key = key; c->p[0] = c->p[1] = c->p[2] = 0.f; c->d2 = 1e30f; c->n[0] = 0.f; c->n[1] = 0.f; c->n[2] = 1.f;
        c->asum = 0.0; c->wsum = 0.0;
        m->used++;
    }
    return &m->t[h];
}
typedef struct { uint64_t mort, key; size_t slot; } orc_mk;
static int mk_cmp(const void* a, const void* b) {
    const orc_mk *x = (const orc_mk*)a, *y = (const orc_mk*)b;
    if (x->mort != y->mort) return x->mort < y->mort ? -1 : 1;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return 0;
}
int orc_sample_cells(const float* V, const float* alb, const uint8_t* ds, int nt, float h, const float lo[3],
                     uint64_t* keys, float* P, float* N, float* A, int cap) {
    orc_map m = {NULL, 0, 0};
    for (int t = 0; t < nt; ++t) {
        const float* v = V + 9 * (size_t)t;
        float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float l = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        if (l < 1e-12f) continue;
        float ar = 0.5f * l, n[3] = {cr[0] / l, cr[1] / l, cr[2] / l};
        double nsd = ceil(ar / (h * h) * 4.0);
        if (nsd < 1.0) nsd = 1.0;
        if (nsd > 1e6) nsd = 1e6;
        int ns = (int)nsd;
        uint64_t rs = orc_hostrng_init((uint64_t)t);
        int sides = ds[t] ? 2 : 1;
        for (int k = 0; k < ns; ++k) {
            float a = 1.f / 3.f, b = 1.f / 3.f;
            if (k > 0) { a = orc_hostrng_u(&rs); b = orc_hostrng_u(&rs); if (a + b > 1.f) { a = 1.f - a; b = 1.f - b; } }
            float Pp[3] = {v[0] + e1[0] * a + e2[0] * b, v[1] + e1[1] * a + e2[1] * b, v[2] + e1[2] * a + e2[2] * b};
            int c[3]; float d2 = 0.f;
            for (int x = 0; x < 3; ++x) {
                c[x] = (int)floorf((Pp[x] - lo[x]) / h);
                float cc = lo[x] + ((float)c[x] + 0.5f) * h;
                d2 += (Pp[x] - cc) * (Pp[x] - cc);
            }
            for (int sd = 0; sd < sides; ++sd) {
                float nn[3] = {sd ? -n[0] : n[0], sd ? -n[1] : n[1], sd ? -n[2] : n[2]};
                orc_cell* rec = map_get(&m, orc_cell_key(c[0], c[1], c[2], orc_normal_bucket(nn[0], nn[1], nn[2])));
                if (d2 < rec->d2) { rec->d2 = d2; for (int x = 0; x < 3; ++x) { rec->p[x] = Pp[x]; rec->n[x] = nn[x]; } }
                rec->asum += (double)alb[t] * ar / ns;
                rec->wsum += (double)ar / ns;
            }
        }
    }
    int count = (int)m.used;
    if (keys) {
        orc_mk* mk = (orc_mk*)malloc(sizeof(orc_mk) * (m.used ? m.used : 1));
        size_t j = 0;
        for (size_t i = 0; i < m.cap; ++i)
            if (m.t[i].key) {
                int cx, cy, cz, b;
                orc_decode_key(m.t[i].key, &cx, &cy, &cz, &b);
                mk[j].mort = (orc_spread3((uint64_t)cx) | orc_spread3((uint64_t)cy) << 1 | orc_spread3((uint64_t)cz) << 2) * 8 + (uint64_t)b;
                mk[j].key = m.t[i].key; mk[j].slot = i; ++j;
            }
        qsort(mk, j, sizeof(orc_mk), mk_cmp);
        for (size_t i = 0; i < j && (int)i < cap; ++i) {
            const orc_cell* c = &m.t[mk[i].slot];
            keys[i] = c->key;
            for (int x = 0; x < 3; ++x) { P[3 * i + x] = c->p[x]; N[3 * i + x] = c->n[x]; }
            A[i] = (float)(c->wsum > 0 ? c->asum / c->wsum : 0.5);
        }
        free(mk);
    }
    free(m.t);
    return count;
}

/* ------------------------------------------------------------------ traversal */
int orc_trace_lds(const float o_[3], const float d_[3], const float* T, const int32_t* TQ, int nt, float* tmin, int* q) {   /* :200-220 */
    v3 o = V3(o_[0], o_[1], o_[2]), d = V3(d_[0], d_[1], d_[2]);
    *tmin = INFINITY; *q = -1;
    for (int j = 0; j < nt; ++j) {
        v3 v0 = ld(T, 3 * j), e1 = ld(T, 3 * j + 1), e2 = ld(T, 3 * j + 2);
        v3 p = cross(d, e2);
        float det = dot(e1, p);
        if (fabsf(det) <= 1e-9f) continue;
        float inv = 1.f / det;
        v3 tv = sub(o, v0);
        float u = dot(tv, p) * inv;
        if (u < 0.f) continue;
        v3 qv = cross(tv, e1);
        float v = dot(d, qv) * inv;
        if (v < 0.f || u + v > 1.f) continue;
        float t = dot(e2, qv) * inv;
        if (t > 1e-4f && t < *tmin) { *tmin = t; *q = TQ[j]; }
    }
    return *q >= 0;
}
static float tri_static(v3 o, v3 d, const float* Tp) {   /* :305-321 */
    v3 v0 = V3(Tp[0], Tp[1], Tp[2]), e1 = V3(Tp[3], Tp[4], Tp[5]), e2 = V3(Tp[6], Tp[7], Tp[8]);
    v3 p = cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det) <= 1e-12f) return -1.f;
    float iv = 1.f / det;
    v3 tv = sub(o, v0);
    float u = dot(tv, p) * iv;
    if (u < 0.f || u > 1.f) return -1.f;
    v3 qv = cross(tv, e1);
    float v = dot(d, qv) * iv;
    if (v < 0.f || u + v > 1.f) return -1.f;
    return dot(e2, qv) * iv;
}
static float aabb_enter(const float* lo, const float* hi, v3 o, v3 inv, float tmax) {   /* :288-295 */
    float tx1 = (lo[0] - o.x) * inv.x, tx2 = (hi[0] - o.x) * inv.x;
    float ty1 = (lo[1] - o.y) * inv.y, ty2 = (hi[1] - o.y) * inv.y;
    float tz1 = (lo[2] - o.z) * inv.z, tz2 = (hi[2] - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fmaxf(fminf(tz1, tz2), 0.f));
    float tf = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fminf(fmaxf(tz1, tz2), tmax));
    return tn <= tf ? tn : 1e30f;
}
static v3 inv_dir(v3 d) {
    return V3(fabsf(d.x) > 1e-12f ? 1.f / d.x : copysignf(1e30f, d.x), fabsf(d.y) > 1e-12f ? 1.f / d.y : copysignf(1e30f, d.y),
              fabsf(d.z) > 1e-12f ? 1.f / d.z : copysignf(1e30f, d.z));
}
static void bvh_ref(const orc_scene* s, v3 o, v3 d, float* tmin, int* q) {   /* :297-338 */
    v3 inv = inv_dir(d);
    int stack[64]; int sp = 0, node = 0;
    for (;;) {
        const orc_node* nd = &s->BN[node];
        if (nd->cnt > 0) {
            for (int j = nd->lf; j < nd->lf + nd->cnt; ++j) {
                float t = tri_static(o, d, s->ST + 9 * (size_t)j);
                if (t > 1e-4f && t < *tmin) { *tmin = t; *q = -2 - j; }
            }
        } else {
            const orc_node *L = &s->BN[nd->lf], *R = &s->BN[nd->lf + 1];
            float tl = aabb_enter(L->lo, L->hi, o, inv, *tmin), tr = aabb_enter(R->lo, R->hi, o, inv, *tmin);
            int first = nd->lf, second = nd->lf + 1;
            if (tr < tl) { float x = tl; tl = tr; tr = x; first = nd->lf + 1; second = nd->lf; }
            if (tl < 1e30f) {
                if (tr < 1e30f && sp < 64) stack[sp++] = second;
                node = first; continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
}
/* order-independent closest hit: smallest t, static ties -> lowest triangle index, a
 * dynamic hit keeps its ties with static triangles */
static inline int closer(float t, int j, float tmin, int q) { return t > 1e-4f && (t < tmin || (t == tmin && q <= -2 && j < -2 - q)); }
static void bvh_closest(const orc_scene* s, v3 o, v3 d, float* tmin, int* q) {
    v3 inv = inv_dir(d);
    int stack[128]; int sp = 0, node = 0;
    for (;;) {
        const orc_node* nd = &s->BN[node];
        CNT(0, 1);
        if (nd->cnt > 0) {
            CNT(1, nd->cnt);
            for (int j = nd->lf; j < nd->lf + nd->cnt; ++j) {
                float t = tri_static(o, d, s->ST + 9 * (size_t)j);
                if (closer(t, j, *tmin, *q)) { *tmin = t; *q = -2 - j; }
            }
        } else {
            const orc_node *L = &s->BN[nd->lf], *R = &s->BN[nd->lf + 1];
            float tl = aabb_enter(L->lo, L->hi, o, inv, *tmin), tr = aabb_enter(R->lo, R->hi, o, inv, *tmin);
            int first = nd->lf, second = nd->lf + 1;
            if (tr < tl) { float x = tl; tl = tr; tr = x; first = nd->lf + 1; second = nd->lf; }
            if (tl < 1e30f) {
                if (tr < 1e30f) stack[sp++] = second;
                node = first; continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
}
static void brute_closest(const orc_scene* s, v3 o, v3 d, float* tmin, int* q) {
    for (int j = 0; j < s->nst; ++j) {
        float t = tri_static(o, d, s->ST + 9 * (size_t)j);
        if (closer(t, j, *tmin, *q)) { *tmin = t; *q = -2 - j; }
    }
}
void orc_trace_bvh_ref(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q) { bvh_ref(s, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmin, q); }
void orc_trace_bvh_closest(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q) { bvh_closest(s, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmin, q); }
void orc_trace_brute(const orc_scene* s, const float o[3], const float d[3], float* tmin, int* q) { brute_closest(s, V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), tmin, q); }

/* dynamic triangles then static geometry (:340-344); trav 2 (wide) has the same closest-hit
 * semantics as 1, so the oracle answers it with the binary closest traversal */
static int trace_full(const orc_scene* s, int trav, v3 o, v3 d, float* tmin, int* q) {
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    int h = orc_trace_lds(oo, dd, s->T, s->TQ, s->nt, tmin, q);
    if (s->nst > 0) {
        if (trav == 0) bvh_ref(s, o, d, tmin, q);
        else if (trav == 3) brute_closest(s, o, d, tmin, q);
        else bvh_closest(s, o, d, tmin, q);
        h = *q != -1;
    }
    return h;
}
void orc_trace_rays(const orc_scene* s, int trav, const float* O, const float* D, int n, float* t, int32_t* q) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        float tt; int qq;
        trace_full(s, trav, ld(O, i), ld(D, i), &tt, &qq);
        t[i] = tt; q[i] = qq;
    }
}

/* ------------------------------------------------------------------ hits -> entries */
static int point_index(v3 H, int q, const orc_scene* s) {   /* :239-246 */
    v3 o = ld(s->Q + 12 * q, 0), u = ld(s->Q + 12 * q, 1), v = ld(s->Q + 12 * q, 2);
    v3 dd = sub(H, o);
    float a = dot(dd, u) / dot(u, u), b = dot(dd, v) / dot(v, v);
    int nu = s->dims[2 * q], nv = s->dims[2 * q + 1];
    int iu = imin(imax((int)(a * nu), 0), nu - 1), iv = imin(imax((int)(b * nv), 0), nv - 1);
    return s->off[q] + iv * nu + iu;
}
static int resolve_hit(const orc_scene* s, int hit, v3 H, v3 D, int q, int i, int* pi) {   /* :346-358 */
    if (!hit) { *pi = i; return 1; }
    if (q >= 0) { *pi = point_index(H, q, s); return dot(ld(s->Q + 12 * q, 3), D) > 0.f; }
    const float* n4 = s->STN + 4 * (size_t)(-2 - q);
    v3 n = V3(n4[0], n4[1], n4[2]);
    if (dot(n, D) > 0.f) {
        int32_t fl; memcpy(&fl, &n4[3], 4);
        if (!(fl & 1)) { *pi = i; return 1; }
        n = mul(n, -1.f);
    }
    *pi = hash_point(s, H, n);
    if (*pi < 0) { *pi = i; return 1; }
    return 0;
}
static int resolve_hit_n(const orc_scene* s, v3 H, v3 D, int q, int* pi, v3* n) {
    if (q >= 0) { *pi = point_index(H, q, s); *n = ld(s->Q + 12 * q, 3); return dot(*n, D) > 0.f; }
    const float* n4 = s->STN + 4 * (size_t)(-2 - q);
    *n = V3(n4[0], n4[1], n4[2]);
    if (dot(*n, D) > 0.f) {
        int32_t fl; memcpy(&fl, &n4[3], 4);
        if (!(fl & 1)) { *pi = -1; return 1; }
        *n = mul(*n, -1.f);
    }
    *pi = hash_point(s, H, *n);
    return *pi < 0;
}

/* ------------------------------------------------------------------ transit grid (:137-197) */
int orc_vox_of(const orc_scene* s, const float p[3]) {
    int i = imin(imax((int)((p[0] - s->lo[0]) / s->ext[0] * s->vx), 0), s->vx - 1);
    int j = imin(imax((int)((p[1] - s->lo[1]) / s->ext[1] * s->vy), 0), s->vy - 1);
    int k = imin(imax((int)((p[2] - s->lo[2]) / s->ext[2] * s->vz), 0), s->vz - 1);
    return (k * s->vy + j) * s->vx + i;
}
/* Amanatides-Woo walk over a gx*gy*gz grid; visit(v) for each voxel crossed */
typedef void (*orc_visit)(void* ctx, long long v);
static void dda_walk(const orc_scene* s, int gx, int gy, int gz, v3 a, v3 b, orc_visit visit, void* ctx) {
    const float sx = gx / s->ext[0], sy = gy / s->ext[1], sz = gz / s->ext[2];
    float ax = (a.x - s->lo[0]) * sx, ay = (a.y - s->lo[1]) * sy, az = (a.z - s->lo[2]) * sz;
    float bx = (b.x - s->lo[0]) * sx, by = (b.y - s->lo[1]) * sy, bz = (b.z - s->lo[2]) * sz;
    int ix = imin(imax((int)ax, 0), gx - 1), iy = imin(imax((int)ay, 0), gy - 1), iz = imin(imax((int)az, 0), gz - 1);
    int ex = imin(imax((int)bx, 0), gx - 1), ey = imin(imax((int)by, 0), gy - 1), ez = imin(imax((int)bz, 0), gz - 1);
    float dx = bx - ax, dy = by - ay, dz = bz - az;
    int stx = dx > 0 ? 1 : -1, sty = dy > 0 ? 1 : -1, stz = dz > 0 ? 1 : -1;
    float tdx = dx != 0 ? fabsf(1.f / dx) : 1e30f, tdy = dy != 0 ? fabsf(1.f / dy) : 1e30f, tdz = dz != 0 ? fabsf(1.f / dz) : 1e30f;
    float tmx = dx != 0 ? ((dx > 0 ? ((float)(ix + 1) - ax) : (ax - (float)ix)) * tdx) : 1e30f;
    float tmy = dy != 0 ? ((dy > 0 ? ((float)(iy + 1) - ay) : (ay - (float)iy)) * tdy) : 1e30f;
    float tmz = dz != 0 ? ((dz > 0 ? ((float)(iz + 1) - az) : (az - (float)iz)) * tdz) : 1e30f;
    for (int g = 0; g < gx + gy + gz; ++g) {
        visit(ctx, ((long long)iz * gy + iy) * gx + ix);
        if (ix == ex && iy == ey && iz == ez) break;
        if (tmx < tmy && tmx < tmz) { ix += stx; tmx += tdx; if (ix < 0 || ix >= gx) break; }
        else if (tmy < tmz) { iy += sty; tmy += tdy; if (iy < 0 || iy >= gy) break; }
        else { iz += stz; tmz += tdz; if (iz < 0 || iz >= gz) break; }
    }
}
static void set_mask_bit(void* ctx, long long v) { uint32_t* m = (uint32_t*)ctx; m[v >> 5] |= 1u << (v & 31); }
static void set_mask_bit_atomic(void* ctx, long long v) { uint32_t* m = (uint32_t*)ctx; __atomic_fetch_or(&m[v >> 5], 1u << (v & 31), __ATOMIC_RELAXED); }
static long long fine_words(int g) { return ((long long)g * g * g + 31) / 32; }
void orc_dda_reg(const orc_scene* s, const float a[3], const float b[3], uint32_t* m) {
    dda_walk(s, s->vx, s->vy, s->vz, V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]), set_mask_bit, m);
}
void orc_dda_segments(const orc_scene* s, const float* A, const float* B, int n, uint32_t* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        uint32_t m[ORC_MAX_NW] = {0, 0, 0, 0};
        dda_walk(s, s->vx, s->vy, s->vz, ld(A, i), ld(B, i), set_mask_bit, m);
        for (int w = 0; w < s->nw; ++w) out[(size_t)s->nw * i + w] = m[w];
    }
}
void orc_dda_volume(const orc_scene* s, int gx, int gy, int gz, const float* A, const float* B, const int32_t* seg_vol, int n,
                    uint32_t* vol, long long words_per_vol) {
    for (int i = 0; i < n; ++i) {
        uint32_t* V = vol + (seg_vol ? (long long)seg_vol[i] * words_per_vol : 0);
        dda_walk(s, gx, gy, gz, ld(A, i), ld(B, i), set_mask_bit, V);
    }
}
static void rast_seg(const orc_scene* s, v3 a, v3 b, uint32_t* m) {   /* 20 samples, :192-197 */
    for (int k = 0; k < 20; ++k) {
        v3 p = add(a, mul(sub(b, a), (float)k / 19.f));
        float pp[3] = {p.x, p.y, p.z};
        int v = orc_vox_of(s, pp);
        m[v >> 5] |= 1u << (v & 31);
    }
}

/* ------------------------------------------------------------------ state build (:369-391) */
static void build_entry(const orc_scene* s, int trav, int i, float* Edir, uint32_t* SV, uint64_t* LM, uint32_t* VOL) {
    v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
    uint32_t m[ORC_MAX_NW] = {0, 0, 0, 0};
    uint64_t lm = 0;
    for (int l = 0; l < s->nl; ++l) {
        float r; v3 D;
        if (s->ltype[l] == 1) { D = ld(s->L, l); r = s->sun_len; }
        else { v3 dv = sub(ld(s->L, l), P); r = sqrtf(dot(dv, dv)); D = mul(dv, 1.f / r); }
        float c = dot(N, D), t; int q;
        trace_full(s, trav, O, D, &t, &q);
        int vis = t >= r - 1e-3f;
        float e;
        if (s->ltype[l] == 1) e = s->inten[l] * fmaxf(c, 0.f) * (vis ? 1.f : 0.f);
        else { float rr = fmaxf(r, 0.05f); e = s->inten[l] * fmaxf(c, 0.f) / (rr * rr) * (vis ? 1.f : 0.f); }
        Edir[(size_t)s->nl * i + l] = e;
        v3 end = add(O, mul(D, fminf(t, r)));
        if (vis && c > 0.f) {
            lm |= 1ull << (l < 63 ? l : 63);
            if (s->fg > 0 && VOL) dda_walk(s, s->fg, s->fg, s->fg, O, end, set_mask_bit_atomic, VOL + (long long)l * fine_words(s->fg));
        }
        rast_seg(s, O, end, m);
    }
    for (int w = 0; w < s->nw; ++w) SV[(size_t)s->nw * i + w] = m[w];
    LM[i] = lm;
}
void orc_k_build(const orc_scene* s, int trav, float* Edir, uint32_t* SV, uint64_t* LM) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s->Np; ++i) build_entry(s, trav, i, Edir, SV, LM, s->VOL);
}
void orc_k_build_rows(const orc_scene* s, int trav, int i0, int i1, float* Edir, uint32_t* SV, uint64_t* LM, uint32_t* VOL) {
    if (i1 > s->Np) i1 = s->Np;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = i0; i < i1; ++i) build_entry(s, trav, i, Edir, SV, LM, VOL);
}

static inline int dormant_l(uint64_t dm, int l) { return l < 64 && ((dm >> l) & 1ull); }
static float est_of(const float* Lc, long long i, int nl, uint64_t dm) {
    float e = 0.f;
    for (int l = 0; l < nl; ++l) if (!dormant_l(dm, l)) e += Lc[(long long)nl * i + l];
    return e;
}

/* ------------------------------------------------------------------ trace (:433-530) */
void orc_k_trace2(const orc_scene* s, int trav, int deps, const orc_cache* c, const orc_params* p, uint64_t* lnew, int32_t* hits, int kmax) {
    const int nl = p->nl;
    float sumpr = (float)((double)c->sumpr[0] * (1.0 / 4294967296.0));
    int npend0 = c->npend[0];
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < p->Np; ++i) {
        uint32_t rs = orc_seed(p->seed, p->frame, (uint32_t)i);
        int k;
        int pend = p->pend_left > 0 && c->pending[i];
        if (p->mode == 0) {
            float kb = floorf(p->budget);
            k = (int)kb + (orc_rnd(&rs) < p->budget - kb ? 1 : 0);
        } else {
            float pr = 1.f / (c->n[i] + 1.f);
            float rrest = (float)(p->Np - npend0);
            float e = rrest * (0.5f / (float)p->Np + 0.5f * pr / sumpr) * p->rays;
            float eb = floorf(e);
            k = (pend ? 1 : 0) + (int)eb + (orc_rnd(&rs) < e - eb ? 1 : 0);
        }
        v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
        uint32_t m[ORC_MAX_NW] = {0, 0, 0, 0};
        uint64_t lm = deps ? s->LM[i] : 0ull;
        float c0 = pend ? c->c0[i] : 0.f;
        int rg = pend ? s->region[i] : 0;
        float sd = 0.f, sq = 0.f;
        float* mcol = c->m + i;              /* SoA: light l at mcol[l * Np] */
        const size_t ms = (size_t)p->Np;
        for (int r = 0; r < k; ++r) {
            float u1 = orc_rnd(&rs), u2 = orc_rnd(&rs);
            v3 D = cosine_dir(N, u1, u2);
            float t; int q;
            int hit = trace_full(s, trav, O, D, &t, &q);
            v3 H = hit ? add(O, mul(D, t)) : O;
            int pi;
            int back = resolve_hit(s, hit, H, D, q, i, &pi);
            if (hits && r < kmax) hits[(size_t)kmax * i + r] = back ? -1 : pi;
            float val = 0.f;
            if (!back) {
                float A = s->Alb[pi];
                const float* ed = s->Edir + (size_t)nl * pi;
                const float* lc = c->Lc + (size_t)nl * pi;
                for (int l = 0; l < nl; ++l) {
                    float lv = dormant_l(p->dormant, l) ? 0.f : lc[l];
                    float v = A * (ed[l] + lv);
                    if (r == 0) mcol[l * ms] = v; else mcol[l * ms] += v;
                    val += v;
                }
            } else if (r == 0) {
                for (int l = 0; l < nl; ++l) mcol[l * ms] = 0.f;
            }
            if (deps) {
                int src = back ? i : pi;
                const uint32_t* sv = s->SV + (size_t)s->nw * src;
                for (int w = 0; w < s->nw; ++w) m[w] |= sv[w];
                dda_walk(s, s->vx, s->vy, s->vz, O, H, set_mask_bit, m);
                if (!back) lm |= s->LM[pi] | c->lcur[pi] | c->lprev[pi];
            }
            if (pend) { float dv = val - c0; sd += dv; sq += dv * dv; }
        }
        if (hits) for (int r = k; r < kmax; ++r) hits[(size_t)kmax * i + r] = -2;
        if (pend && k > 0) {
            long long a0 = fx(sd, ORC_FX_STAT), a1 = fx(sq, ORC_FX_STAT), a2 = fx(c0 * (float)k, ORC_FX_STAT), a3 = fx((float)k, ORC_FX_STAT);
#pragma omp atomic
            c->rstat[4 * rg + 0] += a0;
#pragma omp atomic
            c->rstat[4 * rg + 1] += a1;
#pragma omp atomic
            c->rstat[4 * rg + 2] += a2;
#pragma omp atomic
            c->rstat[4 * rg + 3] += a3;
        }
        c->kk[i] = k;
        float ik = k > 0 ? 1.f / (float)k : 0.f;
        float sa = 0.f;
        if (p->ddgi && k > 0)
            for (int l = 0; l < nl; ++l) sa += mcol[l * ms];
        if (k == 0) { for (int l = 0; l < nl; ++l) mcol[l * ms] = 0.f; }
        else if (k > 1) { for (int l = 0; l < nl; ++l) mcol[l * ms] = mcol[l * ms] * ik; }
        if (p->ddgi && k > 0) {
            int rgd = s->region[i];
            long long d0 = fx(sa * ik, ORC_FX_STAT), d1 = fx(est_of(c->Lc, i, nl, p->dormant), ORC_FX_STAT), d2 = fx(1.f, ORC_FX_STAT);
#pragma omp atomic
            c->dstat[3 * rgd + 0] += d0;
#pragma omp atomic
            c->dstat[3 * rgd + 1] += d1;
#pragma omp atomic
            c->dstat[3 * rgd + 2] += d2;
        }
        if (deps) {
            if (k > 0)
                for (int w = 0; w < s->nw; ++w) {
                    uint32_t* cw = c->cur + (size_t)w * p->Np + i;
                    uint32_t old = *cw;
                    if ((old | m[w]) != old) *cw = old | m[w];
                }
            lnew[i] = k > 0 ? lm : 0ull;
        }
    }
}

void orc_k_decide(const orc_scene* s, const orc_cache* c, const orc_params* p) {   /* :532-555 */
    for (int i = 0; i < p->Np; ++i) {
        if (!c->pending[i]) continue;
        const long long* r4 = c->rstat + 4 * s->region[i];
        const float isc = 1.0f / ORC_FX_STAT;
        float cnt = (float)r4[3] * isc;
        int done = 0;
        if (cnt >= 8.f) {
            float mean = ((float)r4[0] * isc) / cnt;
            float se = sqrtf(fmaxf(((float)r4[1] * isc) / cnt - mean * mean, 0.f) / cnt);
            float base = ((float)r4[2] * isc) / cnt;
            float rel = mean / base;
            if (fabsf(mean) > 2.f * se && fabsf(rel) > 0.03f) {
                float tgt = fminf(fmaxf(0.5f / fabsf(rel), 1.f) * p->rays, p->nmax);
                c->n[i] = fminf(c->n[i], tgt);
                c->pending[i] = 0;
                done = 1;
            }
        }
        if (!done) {
            if (p->pend_left <= 1) c->pending[i] = 0;
            else c->npend[1] += 1;
        }
    }
}

void orc_k_update(const orc_scene* s, const orc_cache* c, const orc_params* p, const uint64_t* lnew) {   /* :562-602 */
    const int nl = p->nl;
    const size_t Np = (size_t)p->Np;
    long long tot = 0;
#pragma omp parallel for schedule(static) reduction(+ : tot)
    for (int i = 0; i < p->Np; ++i) {
        int k = c->kk[i];
        float n = c->n[i], nn = n + (float)k;
        float nmx = p->nmax;
        if (p->ramp) { float a = c->age[i]; nmx = fminf(p->nmax, (4.f + 0.5f * a) * p->rays); c->age[i] = fminf(a + 1.f, 1e4f); }
        float alpha = k > 0 ? fminf((float)k / fminf(fmaxf(nn, 1.f), nmx), 1.f) : 0.f;
        float alpha_r = k > 0 ? fminf((float)k / fminf(fmaxf(nn, 1.f), p->nmax_resp), 1.f) : 0.f;
        const float* a_ = c->Lc + (size_t)nl * i;
        float* o_ = c->Lc2 + (size_t)nl * i;
        const float* ed = s->Edir + (size_t)nl * i;
        float A = s->Alb[i], val = 0.f;
        if (p->ddgi) {
            const long long* d3 = c->dstat + 3 * s->region[i];
            const float isc = 1.0f / ORC_FX_STAT;
            float scale = 1.f, h = p->ddgi_h, d2 = (float)d3[2] * isc;
            if (d2 > 0.f) {
                float mn = ((float)d3[0] * isc) / d2, mc = ((float)d3[1] * isc) / d2;
                float rel = fabsf(mn - mc) / fmaxf(mc, 1e-3f);
                if (rel > p->ddgi_thr) h = fmaxf(0.f, h - 0.75f);
                if (p->ddgi_clamp && rel > 0.10f) scale = 0.25f;
            }
            float w = k > 0 ? (1.f - h) * scale : 0.f;
            for (int l = 0; l < nl; ++l) {
                float a = a_[l], mm = c->m[l * Np + i];
                int dm = dormant_l(p->dormant, l);
                float o = dm ? a : a + w * (mm - a);
                o_[l] = o;
                val += A * (ed[l] + (dm ? 0.f : o));
            }
        } else {
            for (int l = 0; l < nl; ++l) {
                float a = a_[l], mm = c->m[l * Np + i];
                float al = (l < 64 && ((p->resp_mask >> l) & 1ull)) ? alpha_r : alpha;
                int dm = dormant_l(p->dormant, l);
                if (dm) al = 0.f;
                float o = a + al * (mm - a);
                o_[l] = o;
                val += A * (ed[l] + (dm ? 0.f : o));
            }
        }
        c->R[i] = val;
        n = fminf(nn, nmx);
        c->n[i] = n;
        tot += fx(1.f / (n + 1.f), ORC_FX_PR);
        if (lnew) { uint64_t x = lnew[i]; if (x) c->lcur[i] |= x; }
    }
    c->sumpr[1] += tot;
}

void orc_k_resolve(const orc_scene* s, const float* Lc, float* R, uint64_t dormant) {
    const int nl = s->nl;
    for (int i = 0; i < s->Np; ++i) {
        float A = s->Alb[i], val = 0.f;
        for (int l = 0; l < nl; ++l) {
            float lv = dormant_l(dormant, l) ? 0.f : Lc[(size_t)nl * i + l];
            val += A * (s->Edir[(size_t)nl * i + l] + lv);
        }
        R[i] = val;
    }
}

void orc_k_sumpr(const orc_cache* c, int Np) {
    long long tot = 0;
    for (int i = 0; i < Np; ++i) tot += fx(1.f / (c->n[i] + 1.f), ORC_FX_PR);
    c->sumpr[0] += tot;
}
void orc_k_restored(const orc_cache* c, int Np, int nl, float nmax, uint64_t dormant, int verify) {   /* :612-617 */
    for (int i = 0; i < Np; ++i) {
        c->n[i] = nmax; c->age[i] = 1e4f; c->pending[i] = 0;
        if (verify) { c->pending[i] = 1; c->c0[i] = est_of(c->Lc, i, nl, dormant); c->npend[1] += 1; }
    }
}
void orc_k_rescale(const orc_cache* c, int Np, int nl, int l, float ratio) {   /* :637-640 */
    for (int i = 0; i < Np; ++i) c->Lc[(size_t)nl * i + l] *= ratio;
}
void orc_k_touched(const orc_cache* c, int Np, int nl, const int32_t* V, int nv, int kind, uint64_t dormant) {   /* :642-660 */
    for (int i = 0; i < Np; ++i) {
        int t = 0;
        for (int j = 0; j < nv && !t; ++j) {
            int v = V[j];
            size_t o = (size_t)(v >> 5) * Np + i;
            t = ((c->cur[o] | c->prev[o]) >> (v & 31)) & 1u;
        }
        if (!t) continue;
        if (kind == 0) { c->n[i] = 0.f; c->age[i] = 0.f; }
        else if (kind == 1) { c->pending[i] = 1; c->c0[i] = est_of(c->Lc, i, nl, dormant); c->npend[1] += 1; }
        if (kind >= 1 && dormant) c->stale[i] |= dormant;
    }
}
void orc_k_wake(const orc_cache* c, int Np, int nl, int l, int verify, uint64_t dormant) {   /* :663-674 */
    for (int i = 0; i < Np; ++i) {
        if (!((c->stale[i] >> l) & 1ull)) continue;
        if (verify) { c->pending[i] = 1; c->c0[i] = est_of(c->Lc, i, nl, dormant); c->npend[1] += 1; }
        else { c->Lc[(size_t)nl * i + l] = 0.f; c->n[i] = 0.f; c->age[i] = 0.f; }
        c->stale[i] &= ~(1ull << l);
    }
}
void orc_k_oracle_method(const orc_cache* c, int Np, const float* rold, const float* rnew) {   /* :676-681 */
    for (int i = 0; i < Np; ++i) {
        float rel = fabsf(rnew[i] - rold[i]) / (rold[i] + 1e-2f);
        c->n[i] *= 1.f - fminf(fmaxf(4.f * rel, 0.f), 1.f);
    }
}
/* resolve-time targeted invalidation (rtxgi_bench.patch:1028-1050) */
void orc_k_invalidate(const orc_cache* c, int Np, int nl, uint64_t inv_lights, const int32_t* V, int nv, int nw) {
    orc_k_invalidate_masks(c, Np, nl, inv_lights, V, nv, nw, NULL, NULL);
}
void orc_k_invalidate_masks(const orc_cache* c, int Np, int nl, uint64_t inv_lights, const int32_t* V, int nv, int nw,
                            const uint64_t* ga, const uint64_t* gb) {
    for (int i = 0; i < Np; ++i) {
        uint64_t ldp = ga ? (ga[i] | gb[i]) : (c->lcur[i] | c->lprev[i]);
        int known = ga != NULL || ldp != 0ull;
        for (int w = 0; w < nw && !known; ++w) known = (c->cur[(size_t)w * Np + i] | c->prev[(size_t)w * Np + i]) != 0u;
        int t = !known || (ldp & inv_lights) != 0ull;
        for (int j = 0; j < nv && !t; ++j) {
            int v = V[j];
            size_t o = (size_t)(v >> 5) * Np + i;
            t = ((c->cur[o] | c->prev[o]) >> (v & 31)) & 1u;
        }
        if (t) {
            c->n[i] = 0.f; c->age[i] = 0.f;
            for (int l = 0; l < nl; ++l) c->Lc[(size_t)nl * i + l] = 0.f;
            c->npend[1] += 1;
        }
    }
}
void orc_k_flush(const orc_cache* c, int Np, int nl) {
    for (int i = 0; i < Np; ++i) {
        c->n[i] = 0.f; c->age[i] = 0.f;
        for (int l = 0; l < nl; ++l) c->Lc[(size_t)nl * i + l] = 0.f;
    }
}
void orc_k_error(const orc_cache* c, int Np, int nl, const float* ref, const uint8_t* valid, uint64_t dormant) {   /* :683-696 */
    double d2 = 0, rs = 0, cnt = 0;
    for (int i = 0; i < Np; ++i) {
        if (!valid[i]) continue;
        float e = est_of(c->Lc, i, nl, dormant) - ref[i];
        d2 += (double)e * e; rs += ref[i]; cnt += 1;
    }
    c->err[0] += d2; c->err[1] += rs; c->err[2] += cnt;
}

/* ------------------------------------------------------------------ reference iteration (prep_mesh.cuh:182-213) */
void orc_k_ref(const orc_scene* s, int trav, const float* Lin, float* Lout, float* acc, float* closec, uint32_t seed, int K,
               int accumulate, int count_close) {
    const int nl = s->nl;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s->Np; ++i) {
        uint32_t rs = orc_hash32(seed ^ orc_hash32((uint32_t)i + 1u));
        v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
        float* out = Lout + (size_t)nl * i;
        for (int l = 0; l < nl; ++l) out[l] = 0.f;
        int close = 0;
        for (int r = 0; r < K; ++r) {
            float u1 = orc_rnd(&rs), u2 = orc_rnd(&rs);
            v3 D = cosine_dir(N, u1, u2);
            float t; int q;
            int hit = trace_full(s, trav, O, D, &t, &q);
            if (hit && t < 0.02f) close++;
            v3 H = hit ? add(O, mul(D, t)) : O;
            int pi;
            if (resolve_hit(s, hit, H, D, q, i, &pi)) continue;
            float A = s->Alb[pi];
            const float* ed = s->Edir + (size_t)nl * pi;
            const float* li = Lin + (size_t)nl * pi;
            for (int l = 0; l < nl; ++l) out[l] += A * (ed[l] + li[l]);
        }
        for (int l = 0; l < nl; ++l) {
            float v = out[l] / (float)K;
            out[l] = v;
            if (accumulate) acc[(size_t)nl * i + l] += v;
        }
        if (count_close) closec[i] += (float)close;
    }
}

/* ------------------------------------------------------------------ camera query */
static float nee(const orc_scene* s, int trav, v3 H, v3 n, uint32_t* rs, uint64_t dormant, long long* shadow) {
    const int nl = s->nl;
    float u = orc_rnd(rs);
    int l = imin((int)(u * (float)nl), nl - 1);
    float I = s->inten[l];
    if (!(I > 0.f) || dormant_l(dormant, l)) return 0.f;
    float r; v3 D;
    if (s->ltype[l] == 1) { D = ld(s->L, l); r = s->sun_len; }
    else { v3 dv = sub(ld(s->L, l), H); r = sqrtf(dot(dv, dv)); D = mul(dv, 1.f / r); }
    float c = dot(n, D);
    if (!(c > 0.f)) return 0.f;
    v3 O = add(H, mul(n, 1e-3f));
    float t; int q;
    trace_full(s, trav, O, D, &t, &q);
    ++*shadow;
    if (!(t >= r - 1e-3f)) return 0.f;
    if (s->ltype[l] == 1) return (float)nl * (I * c);
    float rr = fmaxf(r, 0.05f);
    return (float)nl * (I * c / (rr * rr));
}
void orc_k_query(const orc_scene* s, int trav, const float* R, const orc_camera* cam, uint32_t seed, uint32_t frame,
                 uint64_t dormant, float* img, long long* rays, int pix0, int pix1) {
    long long ext_tot = 0, sh_tot = 0;
    if (pix1 > cam->W * cam->H) pix1 = cam->W * cam->H;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ext_tot, sh_tot)
    for (int pix = pix0; pix < pix1; ++pix) {
        long long ext = 0, sh = 0;
        int px = pix % cam->W, py = pix / cam->W;
        v3 pos = V3(cam->pos[0], cam->pos[1], cam->pos[2]), fw = V3(cam->fwd[0], cam->fwd[1], cam->fwd[2]);
        v3 rt = V3(cam->right[0], cam->right[1], cam->right[2]), up = V3(cam->up[0], cam->up[1], cam->up[2]);
        float sum = 0.f;
        for (int sp = 0; sp < cam->spp; ++sp) {
            uint32_t rs = orc_seed(seed, frame, (uint32_t)(pix * cam->spp + sp));
            float ju = orc_rnd(&rs), jv = orc_rnd(&rs);
            float sx = (2.f * ((float)px + ju) / (float)cam->W - 1.f) * cam->tanx;
            float sy = (1.f - 2.f * ((float)py + jv) / (float)cam->H) * cam->tany;
            v3 d = add(add(fw, mul(rt, sx)), mul(up, sy));
            d = mul(d, 1.f / sqrtf(dot(d, d)));
            v3 o = pos;
            float thr = 1.f, rad = 0.f;
            for (int b = 0;; ++b) {
                float t; int q;
                int hit = trace_full(s, trav, o, d, &t, &q);
                ++ext;
                if (!hit) break;
                v3 H = add(o, mul(d, t)), n;
                int e;
                if (resolve_hit_n(s, H, d, q, &e, &n)) break;
                if (R && b == cam->cache_bounce) { CNT(3, 1); rad += thr * R[e]; break; }
                float A = s->Alb[e];
                rad += thr * (A * nee(s, trav, H, n, &rs, dormant, &sh));
                if (b >= cam->bounces) break;
                thr = thr * A;
                float u1 = orc_rnd(&rs), u2 = orc_rnd(&rs);
                d = cosine_dir(n, u1, u2);
                o = add(H, mul(n, 1e-3f));
            }
            sum += rad;
        }
        img[pix] = sum / (float)cam->spp;
        ext_tot += ext; sh_tot += sh;
    }
    if (rays) { rays[0] += ext_tot; rays[1] += sh_tot; }
}

/* ------------------------------------------------------------------ FP64 radiance twin */
void orc_k_build_f64(const orc_scene* s, int trav, double* Edir64) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s->Np; ++i) {
        v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
        for (int l = 0; l < s->nl; ++l) {
            float r; v3 D;
            if (s->ltype[l] == 1) { D = ld(s->L, l); r = s->sun_len; }
            else { v3 dv = sub(ld(s->L, l), P); r = sqrtf(dot(dv, dv)); D = mul(dv, 1.f / r); }
            float t; int q;
            trace_full(s, trav, O, D, &t, &q);
            int vis = t >= r - 1e-3f;   /* visibility decision of the FP32 path */
            double e = 0.0;
            if (vis) {
                double px = P.x, py = P.y, pz = P.z, nx = N.x, ny = N.y, nz = N.z;
                if (s->ltype[l] == 1) {
                    double c = nx * s->L[3 * l] + ny * s->L[3 * l + 1] + nz * s->L[3 * l + 2];
                    e = s->inten[l] * (c > 0 ? c : 0.0);
                } else {
                    double dx = s->L[3 * l] - px, dy = s->L[3 * l + 1] - py, dz = s->L[3 * l + 2] - pz;
                    double rd = sqrt(dx * dx + dy * dy + dz * dz);
                    double c = (nx * dx + ny * dy + nz * dz) / rd;
                    double rr = rd > 0.05 ? rd : 0.05;
                    e = s->inten[l] * (c > 0 ? c : 0.0) / (rr * rr);
                }
            }
            Edir64[(size_t)s->nl * i + l] = e;
        }
    }
}
void orc_k_trace2_f64(const orc_scene* s, const double* Edir64, const double* Lc64, const int32_t* kk, const int32_t* hits,
                      int kmax, uint64_t dormant, double* m64) {
    const int nl = s->nl;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->Np; ++i) {
        int k = kk[i];
        double* row = m64 + (size_t)nl * i;
        for (int l = 0; l < nl; ++l) row[l] = 0.0;
        int kr = k < kmax ? k : kmax;
        for (int r = 0; r < kr; ++r) {
            int pi = hits[(size_t)kmax * i + r];
            if (pi < 0) continue;
            double A = s->Alb[pi];
            for (int l = 0; l < nl; ++l) {
                double lv = dormant_l(dormant, l) ? 0.0 : Lc64[(size_t)nl * pi + l];
                row[l] += A * (Edir64[(size_t)nl * pi + l] + lv);
            }
        }
        if (k > 0) for (int l = 0; l < nl; ++l) row[l] /= (double)k;
    }
}
void orc_k_update_f64(const orc_scene* s, const float* n_before, const int32_t* kk, const orc_params* p, const double* Lc64,
                      const double* m64, double* Lc64_out) {
    const int nl = p->nl;
    (void)s;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < p->Np; ++i) {
        int k = kk[i];
        double nn = (double)n_before[i] + k;
        double nmx = p->nmax;
        double den = nn > 1.0 ? nn : 1.0;
        if (den > nmx) den = nmx;
        double alpha = k > 0 ? (k / den < 1.0 ? k / den : 1.0) : 0.0;
        double den_r = nn > 1.0 ? nn : 1.0;
        if (den_r > p->nmax_resp) den_r = p->nmax_resp;
        double alpha_r = k > 0 ? (k / den_r < 1.0 ? k / den_r : 1.0) : 0.0;
        for (int l = 0; l < nl; ++l) {
            double a = Lc64[(size_t)nl * i + l];
            double al = (l < 64 && ((p->resp_mask >> l) & 1ull)) ? alpha_r : alpha;
            if (dormant_l(p->dormant, l)) al = 0.0;
            Lc64_out[(size_t)nl * i + l] = a + al * (m64[(size_t)nl * i + l] - a);
        }
    }
}

/* resolved radiance in double: R64_i = A_i sum_l (Edir64_il + Lc64_il) over the active lights */
void orc_k_resolve_f64(const orc_scene* s, const double* Edir64, const double* Lc64, uint64_t dormant, double* R64) {
    const int nl = s->nl;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->Np; ++i) {
        double v = 0.0;
        for (int l = 0; l < nl; ++l) v += Edir64[(size_t)nl * i + l] + (dormant_l(dormant, l) ? 0.0 : Lc64[(size_t)nl * i + l]);
        R64[i] = (double)s->Alb[i] * v;
    }
}

/* FP64 twin of orc_k_query: each path takes the FP32 pass's rays, hits and light choices (the same
 * draws and traversals), and its radiometry (next-event terms, throughput, the cached value R64 at the
 * cache bounce, the pixel mean) is evaluated in double from the FP32 geometry. */
void orc_k_query_f64(const orc_scene* s, int trav, const double* R64, const orc_camera* cam, uint32_t seed, uint32_t frame,
                     uint64_t dormant, double* img64, int pix0, int pix1) {
    if (pix1 > cam->W * cam->H) pix1 = cam->W * cam->H;
    const int nl = s->nl;
#pragma omp parallel for schedule(dynamic, 64)
    for (int pix = pix0; pix < pix1; ++pix) {
        int px = pix % cam->W, py = pix / cam->W;
        v3 pos = V3(cam->pos[0], cam->pos[1], cam->pos[2]), fw = V3(cam->fwd[0], cam->fwd[1], cam->fwd[2]);
        v3 rt = V3(cam->right[0], cam->right[1], cam->right[2]), up = V3(cam->up[0], cam->up[1], cam->up[2]);
        double sum = 0.0;
        for (int sp = 0; sp < cam->spp; ++sp) {
            uint32_t rs = orc_seed(seed, frame, (uint32_t)(pix * cam->spp + sp));
            float ju = orc_rnd(&rs), jv = orc_rnd(&rs);
            float sx = (2.f * ((float)px + ju) / (float)cam->W - 1.f) * cam->tanx;
            float sy = (1.f - 2.f * ((float)py + jv) / (float)cam->H) * cam->tany;
            v3 d = add(add(fw, mul(rt, sx)), mul(up, sy));
            d = mul(d, 1.f / sqrtf(dot(d, d)));
            v3 o = pos;
            double thr = 1.0, rad = 0.0;
            for (int b = 0;; ++b) {
                float t; int q;
                int hit = trace_full(s, trav, o, d, &t, &q);
                if (!hit) break;
                v3 H = add(o, mul(d, t)), n;
                int e;
                if (resolve_hit_n(s, H, d, q, &e, &n)) break;
                if (R64 && b == cam->cache_bounce) { rad += thr * R64[e]; break; }
                const double A = s->Alb[e];
                /* next-event term: the FP32 light choice and visibility, the value in double */
                float u = orc_rnd(&rs);
                int l = imin((int)(u * (float)nl), nl - 1);
                float I = s->inten[l];
                if (I > 0.f && !dormant_l(dormant, l)) {
                    float r; v3 D;
                    if (s->ltype[l] == 1) { D = ld(s->L, l); r = s->sun_len; }
                    else { v3 dv = sub(ld(s->L, l), H); r = sqrtf(dot(dv, dv)); D = mul(dv, 1.f / r); }
                    float c = dot(n, D);
                    if (c > 0.f) {
                        v3 O = add(H, mul(n, 1e-3f));
                        float ts; int qs;
                        trace_full(s, trav, O, D, &ts, &qs);
                        if (ts >= r - 1e-3f) {
                            double val;
                            if (s->ltype[l] == 1) {
                                double cd = (double)n.x * s->L[3 * l] + (double)n.y * s->L[3 * l + 1] + (double)n.z * s->L[3 * l + 2];
                                val = (double)nl * I * cd;
                            } else {
                                double dx = (double)s->L[3 * l] - H.x, dy = (double)s->L[3 * l + 1] - H.y, dz = (double)s->L[3 * l + 2] - H.z;
                                double rd = sqrt(dx * dx + dy * dy + dz * dz);
                                double cd = ((double)n.x * dx + (double)n.y * dy + (double)n.z * dz) / rd;
                                double rr = rd > 0.05 ? rd : 0.05;
                                val = (double)nl * I * cd / (rr * rr);
                            }
                            rad += thr * (A * val);
                        }
                    }
                }
                if (b >= cam->bounces) break;
                thr = thr * A;
                float u1 = orc_rnd(&rs), u2 = orc_rnd(&rs);
                d = cosine_dir(n, u1, u2);
                o = add(H, mul(n, 1e-3f));
            }
            sum += rad;
        }
        img64[pix] = sum / cam->spp;
    }
}

/* ------------------------------------------------------------------ runner (cuda/poc_gpu.cu:826-1016) */
struct orc_runner {
    const orc_scene* st; int NS; orc_method M; int trav;
    const int32_t* occ; const int64_t* occ_off; const int32_t* em; const int64_t* em_off;
    const int32_t* occf; const int64_t* occf_off;
    const float* const* Pw; const float* const* ref; const uint8_t* const* valid;
    orc_cache c; orc_params p;
    uint64_t* lnew; float* snap;
    int prev_state, f, pend_left, nslots, slot_next, last_event_f;
    int slot_state[ORC_KSNAP];
    uint64_t dormant; float ilearn[128];
    float nmax;
    int boost;        /* frames left of the post-event history-cap ramp (rtxgi_bench.patch:355-356, :409-421) */
    const uint64_t* gmask;
    /* compact snapshots: records (index, n, key, radiance, age) of the entries with samples */
    struct orc_rec { uint32_t index; float n; uint64_t key; float radiance; float age; } * csnap;
    int slot_count[ORC_KSNAP];
    uint64_t* ekeys;
};

void orc_runner_set_grad_masks(orc_runner* r, const uint64_t* masks) { r->gmask = masks; }

orc_runner* orc_runner_create(const orc_scene* states, int NS, const orc_method* m, int trav, const int32_t* occ_flat,
                              const int64_t* occ_off, const int32_t* em_flat, const int64_t* em_off, const float* const* Pw,
                              const float* const* ref, const uint8_t* const* valid, const int32_t* occf_flat,
                              const int64_t* occf_off) {
    orc_runner* r = (orc_runner*)calloc(1, sizeof(orc_runner));
    r->occf = occf_flat; r->occf_off = occf_off;
    r->st = states; r->NS = NS; r->M = *m; r->trav = trav;
    r->occ = occ_flat; r->occ_off = occ_off; r->em = em_flat; r->em_off = em_off;
    r->Pw = Pw; r->ref = ref; r->valid = valid;
    const orc_scene* s = &states[0];
    size_t Np = s->Np, nl = s->nl, nw = s->nw, NR = s->NR;
    orc_cache* c = &r->c;
    c->Lc = (float*)calloc(nl * Np, 4); c->Lc2 = (float*)calloc(nl * Np, 4); c->m = (float*)calloc(nl * Np, 4);
    c->R = (float*)calloc(Np, 4);
    c->n = (float*)calloc(Np, 4); c->c0 = (float*)calloc(Np, 4); c->kk = (int32_t*)calloc(Np, 4);
    c->pending = (uint8_t*)calloc(Np, 1);
    c->cur = (uint32_t*)calloc(nw * Np, 4); c->prev = (uint32_t*)calloc(nw * Np, 4);
    c->lcur = (uint64_t*)calloc(Np, 8); c->lprev = (uint64_t*)calloc(Np, 8);
    c->rstat = (long long*)calloc(4 * NR + 1, 8); c->dstat = (long long*)calloc(3 * NR + 1, 8);
    c->stale = (uint64_t*)calloc(Np, 8); c->age = (float*)calloc(Np, 4);
    c->sumpr = (long long*)calloc(2, 8); c->npend = (int32_t*)calloc(2, 4); c->err = (double*)calloc(3, 8);
    r->lnew = (uint64_t*)calloc(Np, 8);
    if (m->snaps && !m->compact) r->snap = (float*)calloc((size_t)ORC_KSNAP * nl * Np, 4);
    if (m->snaps && m->compact) r->csnap = (struct orc_rec*)calloc((size_t)ORC_KSNAP * Np, sizeof(struct orc_rec));
    r->ekeys = (uint64_t*)malloc(sizeof(uint64_t) * Np);   /* entry keys from the hash grid (quad points: 1 << 63 | index) */
    for (size_t i = 0; i < Np; ++i) r->ekeys[i] = (1ull << 63) | (uint64_t)i;
    if (s->HK && s->hmask < 0x80000000u)   /* scenes without a hash table (box scenes) have hmask = size - 1 = ~0 */
        for (size_t k = 0; k <= (size_t)s->hmask; ++k)
            if (s->HK[k] != 0ull && s->HV[k] >= 0 && s->HV[k] < (int)Np) r->ekeys[s->HV[k]] = s->HK[k];
    return r;
}
void orc_runner_destroy(orc_runner* r) {
    if (!r) return;
    orc_cache* c = &r->c;
    free(c->Lc); free(c->Lc2); free(c->m); free(c->R); free(c->n); free(c->c0); free(c->kk); free(c->pending); free(c->cur); free(c->prev);
    free(c->lcur); free(c->lprev); free(c->rstat); free(c->dstat); free(c->stale); free(c->age); free(c->sumpr); free(c->npend);
    free(c->err); free(r->lnew); free(r->snap); free(r->csnap); free(r->ekeys); free(r);
}
const orc_cache* orc_runner_cache(const orc_runner* r) { return &r->c; }
const orc_params* orc_runner_params(const orc_runner* r) { return &r->p; }

void orc_runner_begin(orc_runner* r, int state0, uint32_t seed) {
    const orc_scene* s = &r->st[0];
    size_t Np = s->Np, nl = s->nl, nw = s->nw;
    orc_cache* c = &r->c;
    if (r->Pw && r->Pw[state0]) memcpy(c->Lc, r->Pw[state0], nl * Np * 4);
    else memset(c->Lc, 0, nl * Np * 4);
    r->nmax = r->M.nmax;
    for (size_t i = 0; i < Np; ++i) { c->n[i] = r->nmax; c->age[i] = 1e4f; }
    memset(c->pending, 0, Np); memset(c->cur, 0, nw * Np * 4); memset(c->prev, 0, nw * Np * 4);
    memset(c->lcur, 0, Np * 8); memset(c->lprev, 0, Np * 8); memset(r->lnew, 0, Np * 8); memset(c->stale, 0, Np * 8);
    c->npend[0] = c->npend[1] = 0; c->sumpr[0] = c->sumpr[1] = 0;
    orc_k_sumpr(c, (int)Np);
    r->prev_state = state0; r->f = 0; r->pend_left = 0; r->dormant = 0; r->boost = 0;
    memset(r->ilearn, 0, sizeof r->ilearn);
    r->nslots = 0; r->slot_next = 0; r->last_event_f = 0;
    for (int k = 0; k < ORC_KSNAP; ++k) r->slot_state[k] = -1;
    orc_params* p = &r->p;
    memset(p, 0, sizeof *p);
    p->Np = (int)Np; p->NR = s->NR; p->mode = r->M.mode; p->budget = r->M.budget; p->nmax = r->nmax; p->seed = seed;
    p->deps = r->M.deps; p->resp_mask = r->M.resp_mask; p->nmax_resp = r->M.nmax_resp;
    p->ddgi = r->M.ddgi; p->ddgi_h = r->M.ddgi_h; p->ddgi_thr = r->M.ddgi_thr; p->ddgi_clamp = r->M.ddgi_clamp;
    p->ramp = r->M.ramp; p->rays = 1.f; p->nl = (int)nl;
}

static void event_block(orc_runner* r, int e, int si, int* invalidated) {
    const orc_scene* s = &r->st[0];
    const orc_method* M = &r->M;
    orc_cache* c = &r->c;
    const int Np = s->Np, nl = s->nl, NS = r->NS, NR = s->NR;
    const int pe = e * NS + si;
    const float* ie = r->st[e].inten; const float* is = r->st[si].inten;
    if (M->snaps && r->f - r->last_event_f >= 20) {
        int k = -1;
        for (int a = 0; a < r->nslots; ++a) if (r->slot_state[a] == e) k = a;
        if (k < 0) { if (r->nslots < ORC_KSNAP) k = r->nslots++; else { k = r->slot_next; r->slot_next = (r->slot_next + 1) % ORC_KSNAP; } }
        r->slot_state[k] = e;
        if (M->compact) {   /* records of the entries with samples, entry order (rtxgi_bench.patch:1312-1334) */
            struct orc_rec* out = r->csnap + (size_t)k * Np;
            int cnt = 0;
            for (int i = 0; i < Np; ++i) {
                if (!(c->n[i] > 0.f)) continue;
                float L = 0.f;
                for (int l = 0; l < nl; ++l) L += c->Lc[(size_t)nl * i + l];
                out[cnt].index = (uint32_t)i; out[cnt].n = c->n[i]; out[cnt].key = r->ekeys[i]; out[cnt].radiance = L; out[cnt].age = c->age[i];
                ++cnt;
            }
            r->slot_count[k] = cnt;
        } else {
            memcpy(r->snap + (size_t)k * Np * nl, c->Lc, (size_t)Np * nl * 4);
        }
    }
    if (M->oracle && r->ref) orc_k_oracle_method(c, Np, r->ref[e], r->ref[si]);
    if (M->rescale && !M->deps) {
        for (int l = 0; l < nl; ++l) {
            float a = ie[l], b = is[l];
            if (a == b) continue;
            int dm = dormant_l(r->dormant, l);
            if (M->memory && l < 64 && a > 0 && b == 0) { r->dormant |= 1ull << l; r->ilearn[l] = a; }
            else if (M->memory && a == 0 && b > 0 && dm) { orc_k_rescale(c, Np, nl, l, b / r->ilearn[l]); r->dormant &= ~(1ull << l); }
            else if (a > 0) orc_k_rescale(c, Np, nl, l, b / a);
        }
    }
    if (M->deps && !M->targeted) {
        uint64_t woke = 0;
        for (int l = 0; l < nl; ++l) {
            float a = ie[l], b = is[l];
            if (a == b) continue;
            int dm = dormant_l(r->dormant, l);
            if (M->memory && l < 64 && a > 0 && b == 0) { r->dormant |= 1ull << l; r->ilearn[l] = a; }
            else if (M->memory && a == 0 && b > 0 && dm) {
                orc_k_rescale(c, Np, nl, l, b / r->ilearn[l]);
                r->dormant &= ~(1ull << l);
                woke |= 1ull << l;
                if (M->wake_verify) {
                    c->npend[0] = c->npend[1] = 0;
                    memset(c->rstat, 0, 4 * (size_t)NR * 8);
                    orc_k_wake(c, Np, nl, l, 1, r->dormant);
                    c->npend[0] = c->npend[1];
                    r->pend_left = ORC_VERIF;
                } else orc_k_wake(c, Np, nl, l, 0, r->dormant);
            } else if (a > 0) orc_k_rescale(c, Np, nl, l, b / a);
        }
        for (int l = 0; l < nl; ++l) {
            long long ix = (long long)pe * nl + l;
            int ne = (int)(r->em_off[ix + 1] - r->em_off[ix]);
            if (ne > 0 && !(l < 64 && ((woke >> l) & 1ull))) orc_k_touched(c, Np, nl, r->em + r->em_off[ix], ne, 0, r->dormant);
        }
        int no = (int)(r->occ_off[pe + 1] - r->occ_off[pe]);
        if (no > 0) {
            if (M->verify) {
                c->npend[0] = c->npend[1] = 0;
                memset(c->rstat, 0, 4 * (size_t)NR * 8);
                orc_k_touched(c, Np, nl, r->occ + r->occ_off[pe], no, 1, r->dormant);
                c->npend[0] = c->npend[1];
                r->pend_left = ORC_VERIF;
            } else if (r->dormant) orc_k_touched(c, Np, nl, r->occ + r->occ_off[pe], no, 2, r->dormant);
        }
    }
    int hit = -1;
    if (M->snaps) for (int a = 0; a < r->nslots; ++a) if (r->slot_state[a] == si) hit = a;
    if (M->targeted) {
        uint64_t inv = 0;
        for (int l = 0; l < nl; ++l) if (ie[l] != is[l]) inv |= 1ull << (l < 63 ? l : 63);
        int nv = (int)(r->occ_off[pe + 1] - r->occ_off[pe]);
        const orc_scene *sa = &r->st[e], *sb = &r->st[si];
        int nf = r->occf ? (int)(r->occf_off[pe + 1] - r->occf_off[pe]) : 0;
        if (s->fg > 0 && nf > 0 && sa->VOL && sb->VOL) {   /* lights whose lit segments cross the object */
            const long long wpv = fine_words(s->fg);
            for (int l = 0; l < nl; ++l)
                for (int j = 0; j < nf; ++j) {
                    int v = r->occf[r->occf_off[pe] + j];
                    uint32_t w = sa->VOL[l * wpv + (v >> 5)] | sb->VOL[l * wpv + (v >> 5)];
                    if ((w >> (v & 31)) & 1u) { inv |= 1ull << (l < 63 ? l : 63); break; }
                }
        }
        if (inv || nv) {
            c->npend[1] = 0;
            const int gm = M->grad_masks > 0 && r->gmask;
            orc_k_invalidate_masks(c, Np, nl, inv, r->occ + r->occ_off[pe], nv, s->nw, gm ? r->gmask + (size_t)e * Np : NULL,
                                   gm ? r->gmask + (size_t)si * Np : NULL);
            *invalidated = c->npend[1];
            c->npend[1] = 0;
        }
    }
    if (M->flush && hit < 0) orc_k_flush(c, Np, nl);
    if (M->snaps && M->compact && hit >= 0) {   /* flush, then the records (rtxgi_bench.patch:1337-1345) */
        orc_k_flush(c, Np, nl);
        c->npend[0] = c->npend[1] = 0;
        if (M->snap_verify) memset(c->rstat, 0, 4 * (size_t)NR * 8);
        const struct orc_rec* rec = r->csnap + (size_t)hit * Np;
        for (int j = 0; j < r->slot_count[hit]; ++j) {
            const uint32_t i = rec[j].index;
            if (i >= (uint32_t)Np || r->ekeys[i] != rec[j].key) continue;
            float* row = c->Lc + (size_t)nl * i;
            row[0] = rec[j].radiance;
            for (int l = 1; l < nl; ++l) row[l] = 0.f;
            c->n[i] = rec[j].n; c->age[i] = rec[j].age; c->pending[i] = 0;
            if (M->snap_verify) { c->pending[i] = 1; c->c0[i] = est_of(c->Lc, i, nl, r->dormant); c->npend[1] += 1; }
        }
        if (M->snap_verify) { c->npend[0] = c->npend[1]; r->pend_left = ORC_VERIF; }
        else r->pend_left = 0;
    } else if (M->snaps && hit >= 0) {
        memcpy(c->Lc, r->snap + (size_t)hit * Np * nl, (size_t)Np * nl * 4);
        if (M->snap_verify) {
            c->npend[0] = c->npend[1] = 0;
            memset(c->rstat, 0, 4 * (size_t)NR * 8);
            orc_k_restored(c, Np, nl, r->nmax, r->dormant, 1);
            c->npend[0] = c->npend[1];
            r->pend_left = ORC_VERIF;
        } else {
            orc_k_restored(c, Np, nl, r->nmax, r->dormant, 0);
            c->npend[0] = c->npend[1] = 0;
            r->pend_left = 0;
        }
    }
    r->last_event_f = r->f;
    if (M->eventboost > 0 && !(M->snaps && hit >= 0)) r->boost = M->eventboost;
    if (M->mode == 1) { c->sumpr[0] = 0; orc_k_sumpr(c, Np); }
}

double orc_runner_frame(orc_runner* r, int si, int want_err, int* invalidated) {
    const orc_scene* s = &r->st[si];
    const orc_method* M = &r->M;
    orc_cache* c = &r->c;
    orc_params* p = &r->p;
    const int Np = s->Np, nl = s->nl, nw = s->nw;
    int inv = 0;
    p->frame = (uint32_t)r->f;
    if (r->f > 0 && r->prev_state != si) event_block(r, r->prev_state, si, &inv);
    if (invalidated) *invalidated = inv;
    p->pend_left = r->pend_left;
    p->dormant = r->dormant;
    if (r->boost > 0) {   /* SHaRC accumulation frames min(default, 2 + k/2), k frames after the event */
        const int k = M->eventboost - r->boost;
        p->nmax = fminf(r->nmax, (2.f + (float)(k / 2)) * p->rays);
        r->boost--;
    } else {
        p->nmax = r->nmax;
    }
    if (M->ddgi) memset(c->dstat, 0, 3 * (size_t)s->NR * 8);
    if (M->deps && r->f % 32 == 0) {
        uint32_t* t = c->cur; c->cur = c->prev; c->prev = t;
        uint64_t* u = c->lcur; c->lcur = c->lprev; c->lprev = u;
        memset(c->cur, 0, (size_t)nw * Np * 4);
        memset(c->lcur, 0, (size_t)Np * 8);
    }
    int track = M->deps && (r->f % M->track_every == 0);
    orc_k_trace2(s, r->trav, track, c, p, r->lnew, NULL, 0);
    if (r->pend_left > 0) {
        c->npend[1] = 0;
        orc_k_decide(s, c, p);
        c->npend[0] = c->npend[1];
        r->pend_left--;
    }
    c->sumpr[1] = 0;
    orc_k_update(s, c, p, track ? r->lnew : NULL);
    c->sumpr[0] = c->sumpr[1];
    { float* t = c->Lc; c->Lc = c->Lc2; c->Lc2 = t; }
    double err = -1.0;
    if (want_err && r->ref && r->valid) {
        c->err[0] = c->err[1] = c->err[2] = 0.0;
        orc_k_error(c, Np, nl, r->ref[si], r->valid[si], r->dormant);
        err = c->err[2] > 0 ? sqrt(c->err[0] / c->err[2]) / (c->err[1] / c->err[2]) : 0.0;
    }
    r->prev_state = si;
    r->f++;
    return err;
}

/* ------------------------------------------------------------------ adjoint of the cached radiance */
uint32_t orc_adj_seed(uint32_t seed, int b) { return orc_hash32(seed + (uint32_t)b * 0x9E3779B9u); }

void orc_adj_unit(const orc_scene* s, int trav, float* G, uint64_t* SG, double* G64) {
    const int nl = s->nl;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s->Np; ++i) {
        v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
        uint64_t m = 0;
        for (int l = 0; l < nl; ++l) {
            float r; v3 D;
            if (s->ltype[l] == 1) { D = ld(s->L, l); r = s->sun_len; }
            else { v3 dv = sub(ld(s->L, l), P); r = sqrtf(dot(dv, dv)); D = mul(dv, 1.f / r); }
            float c = dot(N, D), g = 0.f;
            double g64 = 0.0;
            if (c > 0.f) {
                float t; int q;
                trace_full(s, trav, O, D, &t, &q);
                if (t >= r - 1e-3f) {   /* visible: closest hit beyond the light (k_build) */
                    if (s->ltype[l] == 1) g = c;
                    else { float rr = fmaxf(r, 0.05f); g = c / (rr * rr); }
                    if (G64) {   /* the same geometry in double */
                        double d[3], r64, c64;
                        if (s->ltype[l] == 1) {
                            double n = sqrt((double)s->L[3 * l] * s->L[3 * l] + (double)s->L[3 * l + 1] * s->L[3 * l + 1] +
                                            (double)s->L[3 * l + 2] * s->L[3 * l + 2]);
                            for (int a = 0; a < 3; ++a) d[a] = s->L[3 * l + a] / n;
                            c64 = N.x * d[0] + N.y * d[1] + N.z * d[2];
                            g64 = c64 > 0 ? c64 : 0.0;
                        } else {
                            const float* pp = s->P + 3 * (size_t)i;
                            for (int a = 0; a < 3; ++a) d[a] = (double)s->L[3 * l + a] - (double)pp[a];
                            r64 = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            c64 = (N.x * d[0] + N.y * d[1] + N.z * d[2]) / r64;
                            double rr = r64 > 0.05 ? r64 : 0.05;
                            g64 = c64 > 0 ? c64 / (rr * rr) : 0.0;
                        }
                    }
                }
            }
            G[(size_t)nl * i + l] = g;
            if (G64) G64[(size_t)nl * i + l] = g64;
            if (g > 0.f) m |= 1ull << (l < 63 ? l : 63);
        }
        SG[i] = m;
    }
}

void orc_adj_hits_rows(const orc_scene* s, int trav, int K, int b, uint32_t seed, int i0, int i1, int32_t* hits) {
    const uint32_t sb = orc_adj_seed(seed, b);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = i0; i < i1; ++i) {
        uint32_t rs = orc_hash32(sb ^ orc_hash32((uint32_t)i + 1u));
        v3 P = ld(s->P, i), N = ld(s->N, i), O = add(P, mul(N, 1e-3f));
        for (int k = 0; k < K; ++k) {
            float u1 = orc_rnd(&rs), u2 = orc_rnd(&rs);
            v3 D = cosine_dir(N, u1, u2);
            float t; int q;
            int hit = trace_full(s, trav, O, D, &t, &q);
            v3 H = hit ? add(O, mul(D, t)) : O;
            int pi;
            hits[(size_t)(i - i0) * K + k] = resolve_hit(s, hit, H, D, q, i, &pi) ? -1 : pi;
        }
    }
}

void orc_adj_hits(const orc_scene* s, int trav, int K, int B, uint32_t seed, int32_t* hits) {
    for (int b = 0; b < B; ++b) orc_adj_hits_rows(s, trav, K, b, seed, 0, s->Np, hits + (size_t)b * s->Np * K);
}

void orc_adj_forward(int Np, int nl, int K, int B, const int32_t* hits, const int32_t* mat, const float* alb, const float* I,
                     const float* G, const uint64_t* SG, float* R, uint64_t* SR, float* LB) {
    float* A = (float*)malloc(sizeof(float) * Np);
    float* E = (float*)malloc(sizeof(float) * Np);
    float* L0 = (float*)calloc(Np, sizeof(float));
    float* L1 = (float*)calloc(Np, sizeof(float));
    uint64_t* S0 = (uint64_t*)calloc(Np, sizeof(uint64_t));
    uint64_t* S1 = (uint64_t*)calloc(Np, sizeof(uint64_t));
    for (int i = 0; i < Np; ++i) {
        A[i] = alb[mat[i]];
        float e = 0.f;
        for (int l = 0; l < nl; ++l) e += I[l] * G[(size_t)nl * i + l];
        E[i] = e;
    }
    for (int b = 0; b < B; ++b) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < Np; ++i) {
            float acc = 0.f;
            uint64_t m = 0;
            for (int k = 0; k < K; ++k) {
                int h = hits[((size_t)b * Np + i) * K + k];
                if (h < 0) continue;
                acc += A[h] * (E[h] + (b > 0 ? L0[h] : 0.f));
                if (A[h] > 0.f) m |= SG[h] | (b > 0 ? S0[h] : 0ull);
            }
            L1[i] = acc / (float)K;
            S1[i] = m;
        }
        float* tl = L0; L0 = L1; L1 = tl;
        uint64_t* ts = S0; S0 = S1; S1 = ts;
    }
    for (int i = 0; i < Np; ++i) {
        float L = B > 0 ? L0[i] : 0.f;
        R[i] = A[i] * (E[i] + L);
        SR[i] = A[i] > 0.f ? (SG[i] | (B > 0 ? S0[i] : 0ull)) : 0ull;
        if (LB) LB[i] = L;
    }
    free(A); free(E); free(L0); free(L1); free(S0); free(S1);
}

double orc_adj_grad64(int Np, int nl, int nmat, int K, int B, const int32_t* hits, const int32_t* mat, const double* alb,
                      const double* I, const double* G64, const double* g, double* R64, double* dI, double* dalb) {
    double* A = (double*)malloc(sizeof(double) * Np);
    double* E = (double*)malloc(sizeof(double) * Np);
    double* L = (double*)calloc((size_t)(B + 1) * Np, sizeof(double));   /* L^(0..B) */
    for (int i = 0; i < Np; ++i) {
        A[i] = alb[mat[i]];
        double e = 0.0;
        for (int l = 0; l < nl; ++l) e += I[l] * G64[(size_t)nl * i + l];
        E[i] = e;
    }
    for (int b = 0; b < B; ++b) {
        const double* Lp = L + (size_t)b * Np;
        double* Ln = L + (size_t)(b + 1) * Np;
        for (int i = 0; i < Np; ++i) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k) {
                int h = hits[((size_t)b * Np + i) * K + k];
                if (h >= 0) acc += A[h] * (E[h] + Lp[h]);
            }
            Ln[i] = acc / K;
        }
    }
    const double* LBp = L + (size_t)B * Np;
    double J = 0.0;
    for (int i = 0; i < Np; ++i) { R64[i] = A[i] * (E[i] + LBp[i]); J += g[i] * R64[i]; }
    /* reverse mode */
    double* dA = (double*)malloc(sizeof(double) * Np);
    double* dE = (double*)malloc(sizeof(double) * Np);
    double* lam = (double*)malloc(sizeof(double) * Np);
    double* lam2 = (double*)malloc(sizeof(double) * Np);
    for (int i = 0; i < Np; ++i) { dA[i] = g[i] * (E[i] + LBp[i]); dE[i] = g[i] * A[i]; lam[i] = g[i] * A[i]; }
    for (int b = B - 1; b >= 0; --b) {
        const double* Lp = L + (size_t)b * Np;
        for (int i = 0; i < Np; ++i) lam2[i] = 0.0;
        for (int i = 0; i < Np; ++i) {
            double w = lam[i] / K;
            if (w == 0.0) continue;
            for (int k = 0; k < K; ++k) {
                int h = hits[((size_t)b * Np + i) * K + k];
                if (h < 0) continue;
                dA[h] += w * (E[h] + Lp[h]);
                dE[h] += w * A[h];
                lam2[h] += w * A[h];
            }
        }
        double* t = lam; lam = lam2; lam2 = t;
    }
    for (int l = 0; l < nl; ++l) dI[l] = 0.0;
    for (int m = 0; m < nmat; ++m) dalb[m] = 0.0;
    for (int i = 0; i < Np; ++i) {
        for (int l = 0; l < nl; ++l) dI[l] += dE[i] * G64[(size_t)nl * i + l];
        dalb[mat[i]] += dA[i];
    }
    free(A); free(E); free(L); free(dA); free(dE); free(lam); free(lam2);
    return J;
}

// Split rule of the binned-SAH BVH build, shared by the host builder (rc_scene.cpp) and the
// device builder (rc_prep.hip) so that both make the same decisions from the same numbers:
// 64 centroid bins on each of the three axes, the cheapest split over all of them (axes in
// order, bins in order, strict '<'), leaves of at most RC_BVH_LEAF triangles or when the SAH
// says so, a middle split (no reordering) for clusters of equal centroids, depth <= 56
// (cuda/prep_mesh.cuh:42-114 bins 16 along the longest axis only).
#pragma once
#ifdef __HIPCC__
#define RC_HD __host__ __device__
#else
#define RC_HD
#endif

#define RC_BVH_NB 64
#define RC_BVH_LEAF 2

struct rc_bins {   // one axis: triangle count and box of each centroid bin
    int c[RC_BVH_NB];
    float lo[RC_BVH_NB][3], hi[RC_BVH_NB][3];
};

RC_HD inline float rc_fmin(float a, float b) { return b < a ? b : a; }
RC_HD inline float rc_fmax(float a, float b) { return b > a ? b : a; }

RC_HD inline float rc_bvh_area(const float* lo, const float* hi) {
    const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    return d0 < 0 ? 0.f : 2.f * (d0 * d1 + d1 * d2 + d2 * d0);
}

// bin of a centroid coordinate c on an axis whose centroid range is [clo, chi] (chi - clo >= 1e-7)
RC_HD inline int rc_bvh_bin(float c, float clo, float chi) {
    const int b = (int)((c - clo) / (chi - clo) * (float)RC_BVH_NB);
    return b < RC_BVH_NB - 1 ? b : RC_BVH_NB - 1;
}

RC_HD inline void rc_bins_clear(rc_bins& B) {
    for (int b = 0; b < RC_BVH_NB; ++b) {
        B.c[b] = 0;
        for (int a = 0; a < 3; ++a) { B.lo[b][a] = 1e30f; B.hi[b][a] = -1e30f; }
    }
}

// leaf before binning: few triangles, the depth limit, or equal centroids on the longest axis
// (then at most 64 triangles, else a middle split)
RC_HD inline bool rc_bvh_leaf_now(int count, int depth, float ext_longest) {
    return (count <= RC_BVH_LEAF || depth >= 56 || ext_longest < 1e-7f) && (count <= 64 || depth >= 56);
}

// sweep of one axis's bins: suffix boxes and counts first (into the caller's slo / shi / sc,
// RC_BVH_NB each), then the prefix on the fly; a cheaper split (strict '<', bins in order)
// updates best, bs, bax
RC_HD inline void rc_bins_sweep(const rc_bins& B, int axis, float& best, int& bs, int& bax, float (*slo)[3], float (*shi)[3], int* sc) {
    for (int b = RC_BVH_NB - 1; b >= 0; --b) {
        for (int a = 0; a < 3; ++a) {
            slo[b][a] = rc_fmin(b < RC_BVH_NB - 1 ? slo[b + 1][a] : 1e30f, B.lo[b][a]);
            shi[b][a] = rc_fmax(b < RC_BVH_NB - 1 ? shi[b + 1][a] : -1e30f, B.hi[b][a]);
        }
        sc[b] = (b < RC_BVH_NB - 1 ? sc[b + 1] : 0) + B.c[b];
    }
    float plo[3] = {1e30f, 1e30f, 1e30f}, phi[3] = {-1e30f, -1e30f, -1e30f};
    int pc = 0;
    for (int s = 1; s < RC_BVH_NB; ++s) {
        for (int a = 0; a < 3; ++a) { plo[a] = rc_fmin(plo[a], B.lo[s - 1][a]); phi[a] = rc_fmax(phi[a], B.hi[s - 1][a]); }
        pc += B.c[s - 1];
        const int nl = pc, nr = sc[s];
        if (!nl || !nr) continue;
        const float c = (float)nl * rc_bvh_area(plo, phi) + (float)nr * rc_bvh_area(slo[s], shi[s]);
        if (c < best) { best = c; bs = s; bax = axis; }
    }
}

// after the sweeps: 0 leaf, 1 split at bin bs of axis bax, 2 middle split
RC_HD inline int rc_bvh_choice(int count, float best, int bs, float leaf_cost) {
    if (bs > 0 && !(count <= RC_BVH_LEAF && best >= leaf_cost)) return 1;
    if (count <= RC_BVH_LEAF) return 0;
    return 2;
}

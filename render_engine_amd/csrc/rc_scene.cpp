// Host-side scene preparation (see rc_scene.h).
#include "rc_scene.h"
#include "rc_bvh.h"

#include <algorithm>
#include <limits>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include <zlib.h>

namespace rc {

bool g_prep_device = false;

uint64_t mix64(uint64_t x) {
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull; x ^= x >> 27; x *= 0x94d049bb133111ebull; x ^= x >> 31; return x;
}
uint64_t cell_key(int cx, int cy, int cz, int b) {
    return ((((uint64_t)(cx & 0xFFFFF) << 20 | (uint64_t)(cy & 0xFFFFF)) << 20 | (uint64_t)(cz & 0xFFFFF)) << 3 | (uint64_t)b) + 1ull;
}
int normal_bucket(float nx, float ny, float nz) {
    float ax = std::fabs(nx), ay = std::fabs(ny), az = std::fabs(nz);
    if (ax >= ay && ax >= az) return nx > 0.f ? 0 : 1;
    if (ay >= az) return ny > 0.f ? 2 : 3;
    return nz > 0.f ? 4 : 5;
}
void decode_key(uint64_t k, int& cx, int& cy, int& cz, int& b) {
    k -= 1ull;
    b = (int)(k & 7ull); cz = (int)((k >> 3) & 0xFFFFFull); cy = (int)((k >> 23) & 0xFFFFFull); cx = (int)((k >> 43) & 0xFFFFFull);
}
uint64_t spread3(uint64_t x) {   // 21 bits -> every third bit (Morton interleave)
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

void Mesh::add_tri(const float* a, const float* b, const float* c, float albedo, bool dsided, int material) {
    V.insert(V.end(), a, a + 3); V.insert(V.end(), b, b + 3); V.insert(V.end(), c, c + 3);
    alb.push_back(albedo); ds.push_back(dsided ? 1 : 0); mat.push_back(material);
}
void Mesh::add_quad_grid(const float o[3], const float u[3], const float v[3], int nu, int nv, float albedo, bool dsided, int material) {
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nu; ++i) {
            float p00[3], p10[3], p11[3], p01[3];
            for (int a = 0; a < 3; ++a) {
                float fu0 = u[a] * ((float)i / nu), fu1 = u[a] * ((float)(i + 1) / nu);
                float fv0 = v[a] * ((float)j / nv), fv1 = v[a] * ((float)(j + 1) / nv);
                p00[a] = o[a] + fu0 + fv0; p10[a] = o[a] + fu1 + fv0; p11[a] = o[a] + fu1 + fv1; p01[a] = o[a] + fu0 + fv1;
            }
            add_tri(p00, p10, p11, albedo, dsided, material);
            add_tri(p00, p11, p01, albedo, dsided, material);
        }
}

// ------------------------------------------------------------------------------------ BVH
// Level by level (breadth first): the nodes of a level are split in order and each split
// appends its two children, so the node numbering is the device builder's; partitions are
// stable. Triangle boxes use +0 for zero coordinates, so every min / max is order-free.
void build_bvh(const Mesh& M, std::vector<rc_bvh_node>& nodes, std::vector<int>& order, int& depth_out) {
    const int n = M.nt();
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::vector<float> cen(3 * (size_t)n), tlo(3 * (size_t)n), thi(3 * (size_t)n);
    for (int t = 0; t < n; ++t)
        for (int a = 0; a < 3; ++a) {
            const float* v = &M.V[9 * (size_t)t];
            const float x0 = v[a] + 0.f, x1 = v[3 + a] + 0.f, x2 = v[6 + a] + 0.f;
            const float mn = rc_fmin(rc_fmin(x0, x1), x2), mx = rc_fmax(rc_fmax(x0, x1), x2);
            tlo[3 * (size_t)t + a] = mn; thi[3 * (size_t)t + a] = mx;
            cen[3 * (size_t)t + a] = 0.5f * (mn + mx) + 0.f;
        }
    nodes.clear();
    nodes.reserve(2 * (size_t)n + 1);
    nodes.push_back(rc_bvh_node{});
    struct Item { int node, start, count, depth; };
    std::vector<Item> cur{{0, 0, n, 0}}, next;
    std::vector<int> tmp;
    rc_bins B[3];
    float slo[RC_BVH_NB][3], shi[RC_BVH_NB][3];
    int sc[RC_BVH_NB];
    depth_out = 0;
    while (!cur.empty()) {
        next.clear();
        for (const Item& it : cur) {
            depth_out = std::max(depth_out, it.depth);
            float nlo[3] = {1e30f, 1e30f, 1e30f}, nhi[3] = {-1e30f, -1e30f, -1e30f};
            float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
            for (int k = it.start; k < it.start + it.count; ++k) {
                const int t = order[k];
                for (int a = 0; a < 3; ++a) {
                    nlo[a] = rc_fmin(nlo[a], tlo[3 * (size_t)t + a]); nhi[a] = rc_fmax(nhi[a], thi[3 * (size_t)t + a]);
                    clo[a] = rc_fmin(clo[a], cen[3 * (size_t)t + a]); chi[a] = rc_fmax(chi[a], cen[3 * (size_t)t + a]);
                }
            }
            rc_bvh_node& nd = nodes[it.node];
            for (int a = 0; a < 3; ++a) { nd.lo[a] = nlo[a]; nd.hi[a] = nhi[a]; }
            int ax = 0;
            for (int a = 1; a < 3; ++a) if (chi[a] - clo[a] > chi[ax] - clo[ax]) ax = a;
            int choice = 0, bs = -1, bax = -1;
            if (!rc_bvh_leaf_now(it.count, it.depth, chi[ax] - clo[ax])) {
                float best = 1e30f;
                for (int a0 = 0; a0 < 3; ++a0) {
                    if (chi[a0] - clo[a0] < 1e-7f) continue;
                    rc_bins_clear(B[a0]);
                    for (int k = it.start; k < it.start + it.count; ++k) {
                        const int t = order[k], b = rc_bvh_bin(cen[3 * (size_t)t + a0], clo[a0], chi[a0]);
                        B[a0].c[b]++;
                        for (int a = 0; a < 3; ++a) {
                            B[a0].lo[b][a] = rc_fmin(B[a0].lo[b][a], tlo[3 * (size_t)t + a]);
                            B[a0].hi[b][a] = rc_fmax(B[a0].hi[b][a], thi[3 * (size_t)t + a]);
                        }
                    }
                    rc_bins_sweep(B[a0], a0, best, bs, bax, slo, shi, sc);
                }
                choice = rc_bvh_choice(it.count, best, bs, (float)it.count * rc_bvh_area(nlo, nhi));
            }
            if (choice == 0) { nd.lf = it.start; nd.cnt = it.count; continue; }
            int mid = it.start + it.count / 2;
            if (choice == 1) {
                auto pm = std::stable_partition(order.begin() + it.start, order.begin() + it.start + it.count, [&](int t) {
                    return rc_bvh_bin(cen[3 * (size_t)t + bax], clo[bax], chi[bax]) < bs;
                });
                mid = (int)(pm - order.begin());
            }
            const int left = (int)nodes.size();
            nodes.push_back(rc_bvh_node{});
            nodes.push_back(rc_bvh_node{});
            nodes[it.node].lf = left; nodes[it.node].cnt = 0;
            next.push_back({left, it.start, mid - it.start, it.depth + 1});
            next.push_back({left + 1, mid, it.start + it.count - mid, it.depth + 1});
        }
        std::swap(cur, next);
    }
}

float quant_step_floor(const std::vector<rc_bvh_node>& bn, const float lo[3], const float ext[3]) {
    float b = 0.f;
    for (int a = 0; a < 3; ++a) {
        b = std::max(b, std::max(std::fabs(lo[a]), std::fabs(lo[a] + ext[a])));
        if (!bn.empty()) b = std::max(b, std::max(std::fabs(bn[0].lo[a]), std::fabs(bn[0].hi[a])));
    }
    return std::ldexp(1.f, std::max(-60, (int)std::ceil(std::log2(std::max(b, 1e-6f))) - 18));
}

// Slot assignment of the 4-wide collapse by dynamic programming over the binary tree (surface-area
// cost of the wide nodes; the leaves are the binary leaves whatever the choice, so their cost is
// fixed): S[n][k-1] = least cost of representing subtree n by at most k children of its parent,
// either n itself (a leaf, or a wide node: area(n) + the best 4-slot split of its two children) or
// j slots for its left child and k - j for its right one. split[n][k-1] = the j of that split (0:
// n itself). Children follow their parent in the array (the builder's breadth-first order). Against
// opening the largest-area child first: 1.1-1.3 % less wide-node area cost on the streets, the
// 1920x1080 camera pass 0.5 % faster, the same hits.
struct WideDP {
    std::vector<std::array<float, 4>> S;
    std::vector<std::array<uint8_t, 4>> split;
};
static bool wide_dp(const std::vector<rc_bvh_node>& bn, WideDP& dp) {
    const size_t nb = bn.size();
    dp.S.assign(nb, {0.f, 0.f, 0.f, 0.f});
    dp.split.assign(nb, {0, 0, 0, 0});
    for (size_t i = nb; i-- > 0;) {
        const rc_bvh_node& n = bn[i];
        std::array<float, 4>& S = dp.S[i];
        std::array<uint8_t, 4>& sp = dp.split[i];
        if (n.cnt > 0) { S = {0.f, 0.f, 0.f, 0.f}; continue; }   // leaf: its cost is the same in every collapse
        const size_t L = (size_t)n.lf, R = L + 1;
        if (L <= i || R >= nb) return false;
        const std::array<float, 4>&SL = dp.S[L], &SR = dp.S[R];
        float d4 = std::numeric_limits<float>::infinity();
        for (int j = 1; j <= 3; ++j) d4 = std::min(d4, SL[j - 1] + SR[4 - j - 1]);
        S[0] = rc_bvh_area(n.lo, n.hi) + d4; sp[0] = 0;
        for (int k = 2; k <= 4; ++k) {
            S[k - 1] = S[k - 2]; sp[k - 1] = sp[k - 2];
            for (int j = 1; j < k; ++j) {
                const float c = SL[j - 1] + SR[k - j - 1];
                if (c < S[k - 1]) { S[k - 1] = c; sp[k - 1] = (uint8_t)j; }
            }
        }
        // where k slots gain nothing over k - 1, sp repeats the k - 1 choice; wide_collect then hands
        // the right child one slot more, which S (non-increasing in k) prices no higher
    }
    return true;
}
// the binary nodes that take the k slots of subtree n (k >= 1)
static void wide_collect(const std::vector<rc_bvh_node>& bn, const WideDP& dp, int n, int k, std::vector<int>& out) {
    const int j = bn[n].cnt > 0 ? 0 : dp.split[n][k - 1];
    if (j == 0) { out.push_back(n); return; }
    wide_collect(bn, dp, bn[n].lf, j, out);
    wide_collect(bn, dp, bn[n].lf + 1, k - j, out);
}
// the 4-slot split of inner node n's two children
static void wide_children(const std::vector<rc_bvh_node>& bn, const WideDP& dp, int n, std::vector<int>& out) {
    const int L = bn[n].lf, R = L + 1;
    int bj = 1;
    float best = std::numeric_limits<float>::infinity();
    for (int j = 1; j <= 3; ++j) {
        const float c = dp.S[L][j - 1] + dp.S[R][4 - j - 1];
        if (c < best) { best = c; bj = j; }
    }
    wide_collect(bn, dp, L, bj, out);
    wide_collect(bn, dp, R, 4 - bj, out);
}

void build_wide(const std::vector<rc_bvh_node>& bn, std::vector<rc_wnode>& wn, int& depth_out) {
    wn.clear();
    depth_out = 0;
    WideDP dp;
    const bool use_dp = wide_dp(bn, dp);   // else (a child before its parent) the greedy rule below
    // a job is a binary node (first < 0) or a range of more than RC_WLEAF_MAX triangles of the binary leaf
    // bnode (first .. first + count - 1), split over wide children with the leaf's box
    struct Job { int bnode, first, count, wslot, depth; };
    struct Kid { int bnode, first, count; };   // first < 0: binary node bnode
    std::vector<Job> jobs;
    wn.push_back(rc_wnode{});
    jobs.push_back({0, -1, 0, 0, 1});
    while (!jobs.empty()) {
        Job jb = jobs.back(); jobs.pop_back();
        depth_out = std::max(depth_out, jb.depth);
        std::vector<Kid> kids;
        if (jb.first >= 0) {   // a long leaf's range: four near-equal chunks
            const int q = (jb.count + 3) / 4;
            for (int a = jb.first; a < jb.first + jb.count; a += q) kids.push_back({jb.bnode, a, std::min(q, jb.first + jb.count - a)});
        } else if (bn[jb.bnode].cnt > 0) {
            kids.push_back({jb.bnode, -1, 0});
        } else if (use_dp) {
            std::vector<int> ch;
            wide_children(bn, dp, jb.bnode, ch);
            for (int c : ch) kids.push_back({c, -1, 0});
        } else {
            kids.push_back({bn[jb.bnode].lf, -1, 0});
            kids.push_back({bn[jb.bnode].lf + 1, -1, 0});
            while (kids.size() < 4) {   // open the inner child with the largest surface area
                int best = -1; float ba = -1.f;
                for (size_t k = 0; k < kids.size(); ++k) {
                    const rc_bvh_node& c = bn[kids[k].bnode];
                    if (c.cnt > 0) continue;
                    float a = rc_bvh_area(c.lo, c.hi);
                    if (a > ba) { ba = a; best = (int)k; }
                }
                if (best < 0) break;
                int b = kids[best].bnode;
                kids[best] = {bn[b].lf, -1, 0};
                kids.insert(kids.begin() + best + 1, Kid{bn[b].lf + 1, -1, 0});
            }
        }
        rc_wnode w;
        for (int c = 0; c < 4; ++c) {
            w.lox[c] = w.loy[c] = w.loz[c] = std::numeric_limits<float>::infinity();   // empty slot: never entered
            w.hix[c] = w.hiy[c] = w.hiz[c] = -std::numeric_limits<float>::infinity();
            w.child[c] = 0; w.cend[c] = 0;
        }
        std::vector<Job> sub;
        for (size_t k = 0; k < kids.size(); ++k) {
            const rc_bvh_node& c = bn[kids[k].bnode];
            w.lox[k] = c.lo[0]; w.loy[k] = c.lo[1]; w.loz[k] = c.lo[2];
            w.hix[k] = c.hi[0]; w.hiy[k] = c.hi[1]; w.hiz[k] = c.hi[2];
            const bool leaf = kids[k].first >= 0 || c.cnt > 0;
            const int first = kids[k].first >= 0 ? kids[k].first : c.lf, count = kids[k].first >= 0 ? kids[k].count : c.cnt;
            if (leaf && count <= RC_WLEAF_MAX) {   // the traversal's queue entry with the top bit set
                w.child[k] = (int32_t)(0x80000000u | (uint32_t)first << 7 | (uint32_t)count);
                w.cend[k] = first + count;
            } else {   // an inner node, or a longer leaf split over a node of its own
                int slot = (int)wn.size();
                wn.push_back(rc_wnode{});
                w.child[k] = slot; w.cend[k] = 0;
                sub.push_back(leaf ? Job{kids[k].bnode, first, count, slot, jb.depth + 1} : Job{kids[k].bnode, -1, 0, slot, jb.depth + 1});
            }
        }
        wn[jb.wslot] = w;
        for (auto& s : sub) jobs.push_back(s);
    }
}

void build_qnodes(const std::vector<rc_wnode>& wn, float min_step, std::vector<rc_qnode>& qn) {
    qn.assign(wn.size(), rc_qnode{});
    // smallest step: quant_step_floor (a step is then far above the rounding error of either slab
    // test for rays starting inside a few scene radii)
    const int kmin = std::max(-60, (int)std::ceil(std::log2(std::max(min_step, 1e-18f))));
    for (size_t i = 0; i < wn.size(); ++i) {
        const rc_wnode& w = wn[i];
        rc_qnode& q = qn[i];
        const float* lo[3] = {w.lox, w.loy, w.loz};
        const float* hi[3] = {w.hix, w.hiy, w.hiz};
        float org[3] = {0.f, 0.f, 0.f}, step[3] = {1.f, 1.f, 1.f};
        uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
        for (int a = 0; a < 3; ++a) {
            double mn = INFINITY, mx = -INFINITY;
            for (int c = 0; c < 4; ++c)
                if (w.child[c] != 0) { mn = std::min(mn, (double)lo[a][c]); mx = std::max(mx, (double)hi[a][c]); }
            if (!(mn <= mx)) { for (int c = 0; c < 4; ++c) qlo[a] |= 255u << (8 * c); continue; }
            for (int k = kmin;; ++k) {
                const double s = std::ldexp(1.0, k);
                if ((mx - mn) > 250.0 * s) continue;
                float o = (float)(mn - s);
                if ((double)o > mn - s) o = std::nextafter(o, -INFINITY);
                uint32_t l = 0, h = 0;
                bool fits = true;
                for (int c = 0; c < 4; ++c) {
                    if (w.child[c] == 0) { l |= 255u << (8 * c); continue; }
                    const double a0 = std::floor(((double)lo[a][c] - o) / s) - 1.0, a1 = std::ceil(((double)hi[a][c] - o) / s) + 1.0;
                    if (a0 < 0.0 || a1 > 255.0) { fits = false; break; }
                    l |= (uint32_t)a0 << (8 * c); h |= (uint32_t)a1 << (8 * c);
                }
                if (!fits) continue;
                org[a] = o; step[a] = (float)s; qlo[a] = l; qhi[a] = h;
                break;
            }
        }
        q.ox = org[0]; q.oy = org[1]; q.oz = org[2]; q.sx = step[0]; q.sy = step[1]; q.sz = step[2];
        q.qlox = qlo[0]; q.qloy = qlo[1]; q.qloz = qlo[2]; q.qhix = qhi[0]; q.qhiy = qhi[1]; q.qhiz = qhi[2];
        for (int c = 0; c < 4; ++c) q.child[c] = w.child[c];
    }
}

// ------------------------------------------------------------------------------------ cells
void sample_cells(const Mesh& M, float h, const float lo[3], std::vector<uint64_t>& keys_out, std::vector<CellRec>& recs_out) {
    std::unordered_map<uint64_t, CellRec> cells;
    cells.reserve((size_t)M.nt() * 4);
    for (int t = 0; t < M.nt(); ++t) {
        const float* v = &M.V[9 * (size_t)t];
        float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float l = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        if (l < 1e-12f) continue;
        float ar = 0.5f * l, n[3] = {cr[0] / l, cr[1] / l, cr[2] / l};
        int ns = (int)std::min(1e6, std::max(1.0, std::ceil(ar / (h * h) * 4.0)));
        HostRng r((uint64_t)t);
        int sides = M.ds[t] ? 2 : 1;
        for (int k = 0; k < ns; ++k) {
            float a = 1.f / 3.f, b = 1.f / 3.f;
            if (k > 0) { a = r.u(); b = r.u(); if (a + b > 1.f) { a = 1.f - a; b = 1.f - b; } }
            float P[3] = {v[0] + e1[0] * a + e2[0] * b, v[1] + e1[1] * a + e2[1] * b, v[2] + e1[2] * a + e2[2] * b};
            int c[3]; float d2 = 0.f;
            for (int x = 0; x < 3; ++x) {
                c[x] = (int)std::floor((P[x] - lo[x]) / h);
                float cc = lo[x] + ((float)c[x] + 0.5f) * h;
                d2 += (P[x] - cc) * (P[x] - cc);
            }
            for (int sd = 0; sd < sides; ++sd) {
                float nn[3] = {sd ? -n[0] : n[0], sd ? -n[1] : n[1], sd ? -n[2] : n[2]};
                uint64_t key = cell_key(c[0], c[1], c[2], normal_bucket(nn[0], nn[1], nn[2]));
                auto ins = cells.try_emplace(key, CellRec{{0, 0, 0}, 1e30f, {0, 0, 1}, 0.0, 0.0, -1});
                CellRec& rec = ins.first->second;
                if (d2 < rec.d2) {
                    rec.d2 = d2;
                    for (int x = 0; x < 3; ++x) { rec.p[x] = P[x]; rec.n[x] = nn[x]; }
                    rec.mat = (size_t)t < M.mat.size() ? M.mat[t] : -1;
                }
                rec.asum += (double)M.alb[t] * ar / ns;
                rec.wsum += (double)ar / ns;
            }
        }
    }
    std::vector<std::pair<uint64_t, uint64_t>> mk;
    mk.reserve(cells.size());
    for (auto& kv : cells) {
        int cx, cy, cz, b;
        decode_key(kv.first, cx, cy, cz, b);
        mk.push_back({(spread3((uint64_t)cx) | spread3((uint64_t)cy) << 1 | spread3((uint64_t)cz) << 2) * 8 + (uint64_t)b, kv.first});
    }
    std::sort(mk.begin(), mk.end());
    keys_out.resize(mk.size());
    recs_out.resize(mk.size());
    for (size_t i = 0; i < mk.size(); ++i) { keys_out[i] = mk[i].second; recs_out[i] = cells[mk[i].second]; }
}

void build_hash(const std::vector<uint64_t>& keys, int log2, std::vector<uint64_t>& HK, std::vector<int>& HV) {
    size_t HS = (size_t)1 << log2;
    HK.assign(HS, 0ull);
    HV.assign(HS, -1);
    for (size_t i = 0; i < keys.size(); ++i) {
        size_t hh = (size_t)mix64(keys[i]) & (HS - 1);
        while (HK[hh] != 0ull) hh = (hh + 1) & (HS - 1);
        HK[hh] = keys[i];
        HV[hh] = (int)i;
    }
}

// ------------------------------------------------------------------------------------ prepare
static void box_faces(const float c[3], const float e[3], float F[6][12]) {
    // o, u, v, n per face; u x v points along the outward normal n
    const float X = 2 * e[0], Y = 2 * e[1], Z = 2 * e[2];
    const float G[6][12] = {
        {c[0] + e[0], c[1] - e[1], c[2] - e[2], 0, Y, 0, 0, 0, Z, 1, 0, 0},
        {c[0] - e[0], c[1] - e[1], c[2] - e[2], 0, 0, Z, 0, Y, 0, -1, 0, 0},
        {c[0] - e[0], c[1] + e[1], c[2] - e[2], 0, 0, Z, X, 0, 0, 0, 1, 0},
        {c[0] - e[0], c[1] - e[1], c[2] - e[2], X, 0, 0, 0, 0, Z, 0, -1, 0},
        {c[0] - e[0], c[1] - e[1], c[2] + e[2], X, 0, 0, 0, Y, 0, 0, 0, 1},
        {c[0] - e[0], c[1] - e[1], c[2] - e[2], 0, Y, 0, X, 0, 0, 0, 0, -1}};
    memcpy(F, G, sizeof G);
}

static int host_vox(const Scene& S, const float* p) {
    int i = std::min(std::max((int)((p[0] - S.lo[0]) / S.ext[0] * S.vx), 0), S.vx - 1);
    int j = std::min(std::max((int)((p[1] - S.lo[1]) / S.ext[1] * S.vy), 0), S.vy - 1);
    int k = std::min(std::max((int)((p[2] - S.lo[2]) / S.ext[2] * S.vz), 0), S.vz - 1);
    return (k * S.vy + j) * S.vx + i;
}

std::string prepare(const SceneDesc& d, Scene& S) {
    const Mesh& M = d.mesh;
    if (M.nt() <= 0) return "scene has no static triangles";
    if (d.nl <= 0 || d.nl > RC_MAX_LIGHTS) return "light count out of range (1..72)";
    if (d.inten.empty()) return "scene has no states";
    for (auto& v : d.inten) if ((int)v.size() != d.nl) return "state intensity count differs from the light count";
    if (d.vx * d.vy * d.vz % 32 != 0 || d.vx * d.vy * d.vz > 32 * RC_MAX_NW) return "transit grid must hold 32..128 bits in whole words";
    S = Scene();
    S.mesh = d.mesh;
    {   // material table and per-triangle materials
        Mesh& Mm = S.mesh;
        Mm.mat.resize(Mm.nt(), -1);
        S.matalb = d.matalb;
        bool derive = S.matalb.empty();
        for (int t = 0; t < Mm.nt() && !derive; ++t)
            if (Mm.mat[t] < 0 || Mm.mat[t] >= (int)S.matalb.size()) derive = true;
        if (derive) {   // the distinct albedos, sorted, are the materials
            std::vector<float> u(Mm.alb);
            if (!d.objects.empty()) u.push_back(d.dyn_albedo);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            S.matalb = u;
            for (int t = 0; t < Mm.nt(); ++t) Mm.mat[t] = (int)(std::lower_bound(u.begin(), u.end(), Mm.alb[t]) - u.begin());
        }
        S.nmat = (int)S.matalb.size();
    }
    const int dyn_mat = (d.dyn_mat >= 0 && d.dyn_mat < S.nmat)
        ? d.dyn_mat : (int)(std::lower_bound(S.matalb.begin(), S.matalb.end(), d.dyn_albedo) - S.matalb.begin());
    if (!d.objects.empty() && (dyn_mat >= S.nmat || S.matalb[dyn_mat] != d.dyn_albedo)) return "dynamic objects' albedo is not a material";
    S.vx = d.vx; S.vy = d.vy; S.vz = d.vz; S.nw = d.vx * d.vy * d.vz / 32;
    S.nl = d.nl; S.L = d.L; S.ltype = d.ltype; S.cam = d.cam;
    S.NS = (int)d.inten.size();
    S.inten = d.inten;
    // bounds over static vertices and present dynamic boxes, padded (cuda/prep_mesh.cuh:306-310)
    float hi[3];
    for (int a = 0; a < 3; ++a) { S.lo[a] = 1e30f; hi[a] = -1e30f; }
    for (size_t k = 0; k < M.V.size(); ++k) { int a = k % 3; S.lo[a] = std::min(S.lo[a], M.V[k]); hi[a] = std::max(hi[a], M.V[k]); }
    for (auto& st : d.objects)
        for (auto& b : st)
            if (b.present)
                for (int a = 0; a < 3; ++a) { S.lo[a] = std::min(S.lo[a], b.c[a] - b.e[a]); hi[a] = std::max(hi[a], b.c[a] + b.e[a]); }
    float dg = 0.f;
    for (int a = 0; a < 3; ++a) dg += (hi[a] - S.lo[a]) * (hi[a] - S.lo[a]);
    S.diag = std::sqrt(dg);
    for (int a = 0; a < 3; ++a) { float pad = 0.01f * S.diag + 0.05f; S.lo[a] -= pad; hi[a] += pad; S.ext[a] = hi[a] - S.lo[a]; }
    S.sun_len = 2.f * S.diag;

    // 1. BVH + triangles in leaf order (v0, e1, e2) with normal + double-sided flag
    if (g_prep_device) {
        std::string e = build_bvh_device(M, S.BN, S.order, S.bvh_depth);
        if (!e.empty()) return e;
    } else {
        build_bvh(M, S.BN, S.order, S.bvh_depth);
    }
    build_wide(S.BN, S.WN, S.wide_depth);
    if (3 * S.wide_depth + 1 > RC_WSTACK) return "wide BVH too deep for the LDS traversal stack";
    if (S.WN.size() >= (1u << 24)) return "wide BVH has more than 2^24 nodes";
    if (M.nt() >= (1 << 24)) return "more than 2^24 static triangles (the wide nodes' leaf entries hold 24-bit indices)";
    S.nst = M.nt();
    S.ST.resize(9 * (size_t)S.nst); S.STN.resize(4 * (size_t)S.nst); S.Talb.resize(S.nst);
    for (int j = 0; j < S.nst; ++j) {
        const float* v = &M.V[9 * (size_t)S.order[j]];
        float* T = &S.ST[9 * (size_t)j];
        for (int x = 0; x < 3; ++x) { T[x] = v[x]; T[3 + x] = v[3 + x] - v[x]; T[6 + x] = v[6 + x] - v[x]; }
        float cr[3] = {T[4] * T[8] - T[5] * T[7], T[5] * T[6] - T[3] * T[8], T[3] * T[7] - T[4] * T[6]};
        float l = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        l = l > 0 ? l : 1.f;
        int fl = M.ds[S.order[j]] ? 1 : 0;
        S.STN[4 * (size_t)j] = cr[0] / l; S.STN[4 * (size_t)j + 1] = cr[1] / l; S.STN[4 * (size_t)j + 2] = cr[2] / l;
        memcpy(&S.STN[4 * (size_t)j + 3], &fl, 4);
        S.Talb[j] = M.alb[S.order[j]];
    }

    // 2. cache cells (h fixed, or adjusted toward the target count: cuda/prep_mesh.cuh:338-348)
    std::vector<uint64_t> keys; std::vector<CellRec> recs;
    auto cells = [&]() -> std::string {
        if (g_prep_device) return sample_cells_device(S.mesh, S.h, S.lo, keys, recs);
        sample_cells(S.mesh, S.h, S.lo, keys, recs);
        return "";
    };
    if (d.h > 0.f) {
        S.h = d.h;
        std::string e = cells();
        if (!e.empty()) return e;
    } else {
        if (d.target_points <= 0) return "either h or target_points must be set";
        double atot = 0;
        for (int t = 0; t < M.nt(); ++t) {
            const float* v = &M.V[9 * (size_t)t];
            double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
            double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            atot += 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) * (M.ds[t] ? 2 : 1);
        }
        S.h = (float)std::sqrt(atot / d.target_points);
        for (int it = 0; it < 5; ++it) {
            std::string e = cells();
            if (!e.empty()) return e;
            double ratio = (double)keys.size() / d.target_points;
            if (std::fabs(ratio - 1.0) < 0.08 || it == 4) break;
            S.h *= (float)std::min(2.0, std::max(0.5, std::sqrt(ratio)));
        }
    }
    S.npst = (int)keys.size();
    int log2 = 10;
    while (((size_t)1 << log2) < 2 * (size_t)S.npst) ++log2;
    if (d.table_log2 > 0) {
        if (d.table_log2 < log2) return "hash table too small for the cache cells (needs load <= 1/2)";
        log2 = d.table_log2;
    }
    if (g_prep_device) {
        std::string e = build_hash_device(keys, log2, S.HK, S.HV);
        if (!e.empty()) return e;
    } else {
        build_hash(keys, log2, S.HK, S.HV);
    }
    std::unordered_map<uint64_t, int> zone;
    for (int i = 0; i < S.npst; ++i) {
        const CellRec& c = recs[i];
        for (int x = 0; x < 3; ++x) { S.Pst.push_back(c.p[x]); S.Nst.push_back(c.n[x]); }
        S.A.push_back((float)(c.wsum > 0 ? c.asum / c.wsum : 0.5));
        S.mat.push_back(c.mat);
        int cx, cy, cz, b;
        decode_key(keys[i], cx, cy, cz, b);
        auto z = zone.try_emplace(cell_key(cx >> 2, cy >> 2, cz >> 2, b), (int)zone.size());
        S.region.push_back(z.first->second);
    }
    S.NR = (int)zone.size();

    // 3. dynamic boxes -> quads with point grids (cuda/prep_mesh.cuh:215-272)
    const int nobj = d.objects.empty() ? 0 : (int)d.objects[0].size();
    for (auto& st : d.objects) if ((int)st.size() != nobj) return "every state must list the same objects";
    if (!d.objects.empty() && (int)d.objects.size() != S.NS) return "objects must be given for every state";
    S.nbox = 0;
    S.Nq = 6 * nobj;
    int npt = S.npst;
    for (int o = 0; o < nobj; ++o) {
        const DynBox& b0 = d.objects[0][o];
        float F[6][12];
        box_faces(b0.c, b0.e, F);
        for (int q = 0; q < 6; ++q) {
            float lu = std::sqrt(F[q][3] * F[q][3] + F[q][4] * F[q][4] + F[q][5] * F[q][5]);
            float lv = std::sqrt(F[q][6] * F[q][6] + F[q][7] * F[q][7] + F[q][8] * F[q][8]);
            int nu = std::max(2, (int)std::lround(lu / S.h)), nv = std::max(2, (int)std::lround(lv / S.h));
            S.dims.push_back(nu); S.dims.push_back(nv); S.off.push_back(npt);
            int tu = (nu + 3) / 4, tv = (nv + 3) / 4;
            for (int iv = 0; iv < nv; ++iv)
                for (int iu = 0; iu < nu; ++iu) {
                    S.A.push_back(d.dyn_albedo); S.mat.push_back(dyn_mat); S.region.push_back(S.NR + (iv / 4) * tu + iu / 4);
                }
            S.NR += tu * tv;
            npt += nu * nv;
        }
    }
    S.nbox = npt - S.npst;
    S.Np = npt;
    for (int s = 0; s < S.NS; ++s) {
        std::vector<float> P(S.Pst), N(S.Nst), Q, T; std::vector<int> TQ; std::vector<uint8_t> pres; std::vector<float> ob;
        for (int o = 0; o < nobj; ++o) {
            const DynBox& b = d.objects[s][o];
            float c[3] = {b.c[0], b.c[1], b.c[2]};
            if (!b.present) { for (int a = 0; a < 3; ++a) c[a] = S.lo[a] - 1e4f; }   // parked far outside the scene
            pres.push_back(b.present ? 1 : 0);
            for (int a = 0; a < 3; ++a) ob.push_back(c[a] - b.e[a]);
            for (int a = 0; a < 3; ++a) ob.push_back(c[a] + b.e[a]);
            float F[6][12];
            box_faces(c, b.e, F);
            for (int q = 0; q < 6; ++q) {
                int qi = 6 * o + q, nu = S.dims[2 * qi], nv = S.dims[2 * qi + 1];
                Q.insert(Q.end(), F[q], F[q] + 12);
                for (int iv = 0; iv < nv; ++iv)
                    for (int iu = 0; iu < nu; ++iu)
                        for (int x = 0; x < 3; ++x) {
                            P.push_back(F[q][x] + F[q][3 + x] * ((float)iu + 0.5f) / (float)nu + F[q][6 + x] * ((float)iv + 0.5f) / (float)nv);
                            N.push_back(F[q][9 + x]);
                        }
                if (b.present) {
                    const float* oo = F[q]; const float* u = oo + 3; const float* v = oo + 6;
                    float uv[3] = {u[0] + v[0], u[1] + v[1], u[2] + v[2]};
                    T.insert(T.end(), oo, oo + 3); T.insert(T.end(), u, u + 3); T.insert(T.end(), uv, uv + 3);
                    T.insert(T.end(), oo, oo + 3); T.insert(T.end(), uv, uv + 3); T.insert(T.end(), v, v + 3);
                    TQ.push_back(qi); TQ.push_back(qi);
                }
            }
        }
        S.P.push_back(P); S.N.push_back(N); S.Q.push_back(Q); S.T.push_back(T); S.TQ.push_back(TQ);
        S.nt.push_back((int)TQ.size()); S.present.push_back(pres); S.objbox.push_back(ob);
    }

    S.objects = d.objects;
    fine_events(S, d.fine);
    // 4. events per ordered state pair: transit voxels overlapped by every object that moves,
    // appears or disappears (both placements), and per light the voxel of a light turning on
    // (a directional light turning on touches every voxel).
    const int NV = S.vx * S.vy * S.vz;
    for (int a = 0; a < S.NS; ++a)
        for (int b = 0; b < S.NS; ++b) {
            std::vector<char> mark(NV, 0);
            if (a != b)
                for (int o = 0; o < nobj; ++o) {
                    const DynBox &A_ = d.objects[a][o], &B_ = d.objects[b][o];
                    bool same = A_.present == B_.present && (!A_.present || (A_.c[0] == B_.c[0] && A_.c[1] == B_.c[1] && A_.c[2] == B_.c[2]));
                    if (same) continue;
                    for (const DynBox* X : {&A_, &B_}) {
                        if (!X->present) continue;
                        float plo[3] = {X->c[0] - X->e[0], X->c[1] - X->e[1], X->c[2] - X->e[2]};
                        float phi[3] = {X->c[0] + X->e[0], X->c[1] + X->e[1], X->c[2] + X->e[2]};
                        int v0 = host_vox(S, plo), v1 = host_vox(S, phi);
                        int i0 = v0 % S.vx, j0 = (v0 / S.vx) % S.vy, k0 = v0 / (S.vx * S.vy);
                        int i1 = v1 % S.vx, j1 = (v1 / S.vx) % S.vy, k1 = v1 / (S.vx * S.vy);
                        for (int k = k0; k <= k1; ++k) for (int j = j0; j <= j1; ++j) for (int i = i0; i <= i1; ++i) mark[(k * S.vy + j) * S.vx + i] = 1;
                    }
                }
            std::vector<int> occ;
            for (int v = 0; v < NV; ++v) if (mark[v]) occ.push_back(v);
            S.occ.push_back(occ);
            for (int l = 0; l < S.nl; ++l) {
                std::vector<int> em;
                if (a != b && S.inten[a][l] == 0.f && S.inten[b][l] > 0.f) {
                    if (S.ltype[l] == 1) for (int v = 0; v < NV; ++v) em.push_back(v);
                    else em.push_back(host_vox(S, &S.L[3 * l]));
                }
                S.em.push_back(em);
            }
        }
    return "";
}

void fine_events(Scene& S, int g) {
    S.fine = g;
    S.occ_fine.assign((size_t)S.NS * S.NS, {});
    if (g <= 0 || S.objects.empty()) return;
    const int nobj = (int)S.objects[0].size();
    auto cell = [&](float x, int a) { return std::min(std::max((int)((x - S.lo[a]) / S.ext[a] * g), 0), g - 1); };
    for (int a = 0; a < S.NS; ++a)
        for (int b = 0; b < S.NS; ++b) {
            if (a == b) continue;
            std::vector<int64_t> vs;
            for (int o = 0; o < nobj; ++o) {
                const DynBox &A_ = S.objects[a][o], &B_ = S.objects[b][o];
                bool same = A_.present == B_.present && (!A_.present || (A_.c[0] == B_.c[0] && A_.c[1] == B_.c[1] && A_.c[2] == B_.c[2]));
                if (same) continue;
                for (const DynBox* X : {&A_, &B_}) {
                    if (!X->present) continue;
                    int i0 = cell(X->c[0] - X->e[0], 0), i1 = cell(X->c[0] + X->e[0], 0);
                    int j0 = cell(X->c[1] - X->e[1], 1), j1 = cell(X->c[1] + X->e[1], 1);
                    int k0 = cell(X->c[2] - X->e[2], 2), k1 = cell(X->c[2] + X->e[2], 2);
                    for (int k = k0; k <= k1; ++k) for (int j = j0; j <= j1; ++j) for (int i = i0; i <= i1; ++i)
                        vs.push_back(((int64_t)k * g + j) * g + i);
                }
            }
            std::sort(vs.begin(), vs.end());
            vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
            for (int64_t v : vs) S.occ_fine[(size_t)a * S.NS + b].push_back((int)v);
        }
}

// ------------------------------------------------------------------------------------ scene files
namespace {
// The whole file in memory (gzip-compressed files, such as the prototype's .bin.gz scale scenes,
// decompressed on the way; plain files read as they are), then reads bounded by the bytes left: a count
// read from the file can never make a vector larger than what remains (a corrupt count fails the read
// instead of allocating).
struct Reader {
    std::vector<char> buf;
    size_t pos = 0;
    bool ok = true;
    bool open(const char* path) {
        gzFile g = gzopen(path, "rb");
        if (!g) return false;
        char chunk[1 << 16];
        int n;
        while ((n = gzread(g, chunk, sizeof chunk)) > 0) {
            if (buf.size() + (size_t)n > ((size_t)16 << 30)) { gzclose(g); return false; }   // 16 GiB: not a scene file
            buf.insert(buf.end(), chunk, chunk + n);
        }
        const bool good = n == 0;
        gzclose(g);
        return good;
    }
    size_t left() const { return buf.size() - pos; }
    template <class T> void rd(std::vector<T>& v, size_t n) {
        if (!ok || n > left() / sizeof(T)) { ok = false; v.clear(); return; }
        v.resize(n);
        if (n) memcpy(v.data(), buf.data() + pos, n * sizeof(T));
        pos += n * sizeof(T);
    }
    template <class T> T one() {
        T x{};
        if (!ok || left() < sizeof(T)) { ok = false; return x; }
        memcpy(&x, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return x;
    }
};

// a binary BVH read from a file: leaves inside the triangle range, inner nodes' children after the node
// (so no cycle, and the wide collapse terminates) and inside the node range
static std::string check_bvh(const std::vector<rc_bvh_node>& bn, int nst) {
    const long long nn = (long long)bn.size();
    for (long long k = 0; k < nn; ++k) {
        const rc_bvh_node& b = bn[k];
        if (b.cnt > 0) {
            if (b.lf < 0 || (long long)b.lf + b.cnt > nst) return "BVH node " + std::to_string(k) + ": leaf triangles out of range";
        } else if (b.cnt < 0 || b.lf <= k || (long long)b.lf + 1 >= nn) {
            return "BVH node " + std::to_string(k) + ": children out of range or not after the node";
        }
    }
    return "";
}
}  // namespace

// fine 16x8x6 event voxels of a box scene -> transit voxels they overlap (cuda/poc_gpu.cu:720-735)
static std::vector<int> fold_to_transit(const std::vector<int>& fine, int VX, int VY, int VZ) {
    const int FX = 16, FY = 8, FZ = 6;
    const float X1 = 8.f, Y1 = 4.f, Z1 = 3.f;
    std::vector<char> mark((size_t)VX * VY * VZ, 0);
    for (int v : fine) {
        int i = v % FX, j = (v / FX) % FY, k = v / (FX * FY);
        float x0 = i * X1 / FX, x1 = (i + 1) * X1 / FX, y0 = j * Y1 / FY, y1 = (j + 1) * Y1 / FY;
        float z0 = k * Z1 / FZ, z1 = (k + 1) * Z1 / FZ;
        int a0 = std::min((int)(x0 / X1 * VX + 1e-4f), VX - 1), a1 = std::min((int)(x1 / X1 * VX - 1e-4f), VX - 1);
        int b0 = std::min((int)(y0 / Y1 * VY + 1e-4f), VY - 1), b1 = std::min((int)(y1 / Y1 * VY - 1e-4f), VY - 1);
        int c0 = std::min((int)(z0 / Z1 * VZ + 1e-4f), VZ - 1), c1 = std::min((int)(z1 / Z1 * VZ - 1e-4f), VZ - 1);
        for (int a = a0; a <= a1; ++a) for (int b = b0; b <= b1; ++b) for (int c = c0; c <= c1; ++c) mark[((size_t)c * VY + b) * VX + a] = 1;
    }
    std::vector<int> out;
    for (size_t v = 0; v < mark.size(); ++v) if (mark[v]) out.push_back((int)v);
    return out;
}

std::string load_scene_file(const char* path, Scene& S, FileRefs& R) {
    Reader in;
    if (!in.open(path)) return std::string("cannot open ") + path;
    std::vector<int> h;
    in.rd(h, 8);
    if (!in.ok) return "truncated header";
    const int Np = h[0], Nq = h[1], Nt = h[2], NS = h[3], NL = h[4], NR = h[5], has_ref = h[6], fmt = h[7];
    if (fmt != 2 && fmt != 3) return "scene format " + std::to_string(fmt) + " not supported (2 or 3)";
    if (Np <= 0 || Nq < 0 || Nt < 0 || NS <= 0 || NL <= 0 || NL > RC_MAX_LIGHTS || NR <= 0 || NR > Np) return "bad header values";
    if (NS > 64) return "more than 64 states";
    S = Scene();
    S.vx = 8; S.vy = 4; S.vz = 4; S.nw = 4;
    S.Np = Np; S.NS = NS; S.nl = NL; S.NR = NR; S.Nq = Nq;
    S.lo[0] = S.lo[1] = S.lo[2] = 0.f; S.ext[0] = 8.f; S.ext[1] = 4.f; S.ext[2] = 3.f;
    if (fmt == 3) {
        std::vector<float> lo, ext, hc; std::vector<int> m;
        in.rd(lo, 3); in.rd(ext, 3); in.rd(hc, 1); in.rd(m, 4);
        if (!in.ok) return "truncated mesh header";
        for (int a = 0; a < 3; ++a) { S.lo[a] = lo[a]; S.ext[a] = ext[a]; }
        S.h = hc[0];
        S.nst = m[0]; S.npst = m[3];
        if (m[0] < 0 || m[1] < 0 || m[3] < 0 || m[3] > Np) return "bad mesh header counts";
        if (m[0] > 0 && m[1] == 0) return "static triangles without a BVH";
        if (m[2] <= 0 || (m[2] & (m[2] - 1))) return "hash table size must be a power of two";
        in.rd(S.ST, 9 * (size_t)m[0]); in.rd(S.STN, 4 * (size_t)m[0]);
        std::vector<float> bn; in.rd(bn, 8 * (size_t)m[1]);
        S.BN.resize(m[1]);
        if (in.ok && m[1]) memcpy(S.BN.data(), bn.data(), bn.size() * 4);
        in.rd(S.HK, m[2]); in.rd(S.HV, m[2]);
        if (!in.ok) return "truncated mesh data";
        std::string be = check_bvh(S.BN, S.nst);
        if (!be.empty()) return be;
        for (size_t k = 0; k < S.HK.size(); ++k)
            if (S.HK[k] != 0ull && (S.HV[k] < 0 || S.HV[k] >= S.npst)) return "hash value out of the static entries' range";
        if (!S.BN.empty()) build_wide(S.BN, S.WN, S.wide_depth);
        if (3 * S.wide_depth + 1 > RC_WSTACK) return "wide BVH too deep for the LDS traversal stack";
        if (S.WN.size() >= (1u << 24)) return "wide BVH has more than 2^24 nodes";
        if (S.nst >= (1 << 24)) return "more than 2^24 static triangles (the wide nodes' leaf entries hold 24-bit indices)";
        S.order.resize(S.nst); std::iota(S.order.begin(), S.order.end(), 0);
        S.Talb.assign(S.nst, 0.5f);
    }
    float dg = std::sqrt(S.ext[0] * S.ext[0] + S.ext[1] * S.ext[1] + S.ext[2] * S.ext[2]);
    S.diag = dg; S.sun_len = 2.f * dg;
    in.rd(S.L, 3 * (size_t)NL); in.rd(S.A, Np); in.rd(S.region, Np);
    S.ltype.assign(NL, 0);
    if (!in.ok) return "truncated light / albedo / region data";
    for (int i = 0; i < Np; ++i) if (S.region[i] < 0 || S.region[i] >= NR) return "entry region out of range";
    S.matalb = S.A;   // materials: the distinct entry albedos, sorted
    std::sort(S.matalb.begin(), S.matalb.end());
    S.matalb.erase(std::unique(S.matalb.begin(), S.matalb.end()), S.matalb.end());
    S.nmat = (int)S.matalb.size();
    S.mat.resize(Np);
    for (int i = 0; i < Np; ++i) S.mat[i] = (int)(std::lower_bound(S.matalb.begin(), S.matalb.end(), S.A[i]) - S.matalb.begin());
    R = FileRefs(); R.has = has_ref != 0;
    for (int st = 0; st < NS; ++st) {
        std::vector<float> inten, P, N, Q, T; std::vector<int> dims, off, TQ;
        in.rd(inten, NL); in.rd(P, 3 * (size_t)Np); in.rd(N, 3 * (size_t)Np); in.rd(Q, 12 * (size_t)Nq);
        in.rd(dims, 2 * (size_t)Nq); in.rd(off, Nq); in.rd(T, 9 * (size_t)Nt); in.rd(TQ, Nt);
        if (!in.ok) return "truncated state " + std::to_string(st);
        if (st == 0) { S.dims = dims; S.off = off; }
        else if (dims != S.dims || off != S.off) return "quad grids differ between states";
        for (int q = 0; q < Nq; ++q)
            if (dims[2 * q] <= 0 || dims[2 * q + 1] <= 0 || off[q] < 0 || (long long)off[q] + (long long)dims[2 * q] * dims[2 * q + 1] > Np)
                return "quad point grid out of range";
        for (int j = 0; j < Nt; ++j) if (TQ[j] < 0 || TQ[j] >= Nq) return "triangle quad id out of range";
        S.inten.push_back(inten); S.P.push_back(P); S.N.push_back(N); S.Q.push_back(Q); S.T.push_back(T); S.TQ.push_back(TQ);
        S.nt.push_back(Nt);
        S.present.push_back({});
        S.objbox.push_back({});
        if (has_ref) {
            std::vector<float> ref, Pw; std::vector<uint8_t> valid;
            in.rd(ref, Np); in.rd(valid, Np); in.rd(Pw, (size_t)NL * Np);
            if (!in.ok) return "truncated references of state " + std::to_string(st);
            R.ref.push_back(ref); R.valid.push_back(valid); R.Pw.push_back(Pw);
        }
    }
    const int NV = S.vx * S.vy * S.vz;
    for (int e = 0; e < NS * NS; ++e) {
        int no = in.one<int>();
        std::vector<int> o; in.rd(o, no < 0 ? 0 : no);
        if (!in.ok || no < 0) return "truncated event list";
        std::vector<int> oc = fmt == 3 ? o : fold_to_transit(o, S.vx, S.vy, S.vz);
        for (int v : oc) if (v < 0 || v >= NV) return "event voxel out of range";
        S.occ.push_back(oc);
        for (int l = 0; l < NL; ++l) {
            int ne = in.one<int>();
            std::vector<int> m; in.rd(m, ne < 0 ? 0 : ne);
            if (!in.ok || ne < 0) return "truncated emitter list";
            std::vector<int> em = fmt == 3 ? m : fold_to_transit(m, S.vx, S.vy, S.vz);
            for (int v : em) if (v < 0 || v >= NV) return "event voxel out of range";
            S.em.push_back(em);
        }
    }
    S.occ_fine.assign((size_t)NS * NS, {});
    S.fine = 0;
    S.cam = rc_camera{};
    return "";
}

// ------------------------------------------------------------------------------------ generators
static void face_grid(Mesh& M, const float F[12], float cell, float albedo, int material = -1) {
    float lu = std::sqrt(F[3] * F[3] + F[4] * F[4] + F[5] * F[5]), lv = std::sqrt(F[6] * F[6] + F[7] * F[7] + F[8] * F[8]);
    int nu = std::max(1, (int)std::lround(lu / cell)), nv = std::max(1, (int)std::lround(lv / cell));
    M.add_quad_grid(F, F + 3, F + 6, nu, nv, albedo, false, material);
}
static int face_tris(const float F[12], float cell) {
    float lu = std::sqrt(F[3] * F[3] + F[4] * F[4] + F[5] * F[5]), lv = std::sqrt(F[6] * F[6] + F[7] * F[7] + F[8] * F[8]);
    return 2 * std::max(1, (int)std::lround(lu / cell)) * std::max(1, (int)std::lround(lv / cell));
}
// cell size whose face subdivision gives about `target` triangles (bisection on log scale)
static float fit_cell(const std::vector<std::array<float, 12>>& faces, int target) {
    float lo = 1e-3f, hi = 1e3f;
    for (int it = 0; it < 60; ++it) {
        float mid = std::sqrt(lo * hi);
        long long n = 0;
        for (auto& f : faces) n += face_tris(f.data(), mid);
        if (n > target) lo = mid; else hi = mid;
    }
    return hi;
}

static void look_at(rc_camera& c, const float pos[3], const float at[3], float hfov_deg, int W, int H) {
    float f[3] = {at[0] - pos[0], at[1] - pos[1], at[2] - pos[2]};
    float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int a = 0; a < 3; ++a) f[a] /= fl;
    float upw[3] = {0, 0, 1};
    float r[3] = {f[1] * upw[2] - f[2] * upw[1], f[2] * upw[0] - f[0] * upw[2], f[0] * upw[1] - f[1] * upw[0]};
    float rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int a = 0; a < 3; ++a) r[a] /= rl;
    float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    for (int a = 0; a < 3; ++a) { c.pos[a] = pos[a]; c.fwd[a] = f[a]; c.right[a] = r[a]; c.up[a] = u[a]; }
    c.tanx = std::tan(0.5f * hfov_deg * 3.14159265f / 180.f);
    c.tany = c.tanx * (float)H / (float)W;
    c.W = W; c.H = H;
}

SceneDesc gen_box_room(uint64_t seed, int target_tris) {
    SceneDesc d;
    HostRng r(seed);
    const float X = 8.f, Y = 4.f, Z = 3.f;
    std::vector<std::array<float, 12>> faces;
    std::vector<float> falb;
    const float walls[6][12] = {
        {0, 0, 0, X, 0, 0, 0, Y, 0, 0, 0, 1},    // floor
        {0, 0, Z, 0, Y, 0, X, 0, 0, 0, 0, -1},   // ceiling
        {0, 0, 0, 0, 0, Z, X, 0, 0, 0, 1, 0},    // y = 0
        {0, Y, 0, X, 0, 0, 0, 0, Z, 0, -1, 0},   // y = Y
        {0, 0, 0, 0, Y, 0, 0, 0, Z, 1, 0, 0},    // x = 0
        {X, 0, 0, 0, 0, Z, 0, Y, 0, -1, 0, 0}};  // x = X
    for (auto& w : walls) { std::array<float, 12> a; memcpy(a.data(), w, 48); faces.push_back(a); falb.push_back(0.2f + 0.6f * r.u()); }
    const float cubes[2][6] = {{2.5f, 1.5f, 0.5f, 0.5f, 0.5f, 0.5f}, {5.5f, 2.6f, 0.7f, 0.7f, 0.7f, 0.7f}};
    for (auto& cb : cubes) {
        float F[6][12];
        box_faces(cb, cb + 3, F);
        for (int q = 0; q < 5; ++q) {   // the bottom face rests on the floor
            std::array<float, 12> a; memcpy(a.data(), F[q], 48); faces.push_back(a); falb.push_back(0.2f + 0.6f * r.u());
        }
    }
    float cell = fit_cell(faces, target_tris);
    for (size_t k = 0; k < faces.size(); ++k) face_grid(d.mesh, faces[k].data(), cell, falb[k]);
    d.nl = 2;
    for (int l = 0; l < 2; ++l) {
        d.L.push_back(0.5f + 7.f * r.u()); d.L.push_back(0.5f + 3.f * r.u()); d.L.push_back(1.5f + 1.3f * r.u());
        d.ltype.push_back(0);
    }
    d.inten = {{50.f, 50.f}, {50.f, 0.f}};
    d.h = 0.05f;
    d.table_log2 = 18;
    d.fine = 64;
    d.vx = 8; d.vy = 4; d.vz = 4;
    float pos[3] = {0.4f, 0.6f, 2.2f}, at[3] = {6.5f, 2.8f, 0.6f};
    look_at(d.cam, pos, at, 75.f, 256, 256);
    d.cam.spp = 1; d.cam.bounces = 1; d.cam.cache_bounce = 1;
    return d;
}

SceneDesc gen_street(uint64_t seed, int nboxes, int target_tris, int nlamps, bool sun, int target_points, int table_log2,
                     int W, int H, int spp, int bounces) {
    SceneDesc d;
    HostRng r(seed);
    const float pitch = 16.f;
    int cols = (int)std::ceil(std::sqrt(nboxes * 1.25)), rows = (nboxes + cols - 1) / cols;
    std::vector<std::array<float, 12>> faces;
    std::vector<float> falb;
    std::vector<int> fmat;
    // 4096 materials with distinct albedos (stratified in [0.2, 0.8)); every box draws one
    const int nmat = 4096;
    {
        HostRng rm(seed ^ 0x6d6174657269616cull);
        d.matalb.resize(nmat);
        for (int m = 0; m < nmat; ++m) d.matalb[m] = 0.2f + 0.6f * (((float)m + rm.u()) / (float)nmat);
    }
    float gx = cols * pitch, gy = rows * pitch;
    {   // ground
        std::array<float, 12> g = {-pitch, -pitch, 0, gx + 2 * pitch, 0, 0, 0, gy + 2 * pitch, 0, 0, 0, 1};
        faces.push_back(g); falb.push_back(d.matalb[1024]); fmat.push_back(1024);
    }
    for (int b = 0; b < nboxes; ++b) {
        int ci = b % cols, ri = b / cols;
        float e[3] = {0.5f * (0.5f + 9.5f * r.u()), 0.5f * (0.5f + 9.5f * r.u()), 0.5f * (0.5f + 9.5f * r.u())};
        float c[3];
        for (int a = 0; a < 2; ++a) {
            float base = (a == 0 ? ci : ri) * pitch + 3.f;
            float room = 10.f - 2.f * e[a];
            c[a] = base + e[a] + room * r.u();
        }
        c[2] = e[2];
        const int m = std::min((int)(r.u() * (float)nmat), nmat - 1);
        float F[6][12];
        box_faces(c, e, F);
        for (int q = 0; q < 5; ++q) {
            std::array<float, 12> a; memcpy(a.data(), F[q], 48); faces.push_back(a); falb.push_back(d.matalb[m]); fmat.push_back(m);
        }
    }
    float cell = fit_cell(faces, target_tris);
    for (size_t k = 0; k < faces.size(); ++k) face_grid(d.mesh, faces[k].data(), cell, falb[k], fmat[k]);
    d.dyn_mat = 3413;
    d.dyn_albedo = d.matalb[d.dyn_mat];
    // lamps at street crossings (x = i*pitch, y = j*pitch), lamp 0 at the crossing nearest the middle
    int cx0 = cols / 2, cy0 = rows / 2;
    std::vector<std::pair<int, int>> used{{cx0, cy0}};
    while ((int)used.size() < nlamps) {
        int i = (int)(r.u() * (cols + 1)), j = (int)(r.u() * (rows + 1));
        if (std::find(used.begin(), used.end(), std::make_pair(i, j)) == used.end()) used.push_back({i, j});
    }
    for (auto& ij : used) { d.L.push_back(ij.first * pitch); d.L.push_back(ij.second * pitch); d.L.push_back(6.f); d.ltype.push_back(0); }
    if (sun) {
        float s[3] = {0.45f, 0.3f, 0.84f};
        float sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int a = 0; a < 3; ++a) d.L.push_back(s[a] / sl);
        d.ltype.push_back(1);
    }
    d.nl = (int)d.ltype.size();
    std::vector<float> base(d.nl, 60.f);
    if (sun) base[d.nl - 1] = 1.0f;
    std::vector<float> night = base, lamp0 = base, occl = base;
    if (sun) night[d.nl - 1] = 0.f;
    lamp0[0] = 0.f;
    d.inten = {base, night, lamp0, occl};
    // occluder: a 12 m x 10 m wall (0.4 m thick) standing on the street next to lamp 0, present in state 3
    DynBox wall{{cx0 * pitch + 4.f, cy0 * pitch, 5.f}, {0.2f, 6.f, 5.f}, false};
    d.objects.resize(4, std::vector<DynBox>{wall});
    d.objects[3][0].present = true;
    d.h = 0.f;
    d.target_points = target_points;
    d.fine = target_points > 800000 ? 256 : 128;
    d.table_log2 = table_log2;
    d.vx = 8; d.vy = 4; d.vz = 4;
    float at[3] = {cx0 * pitch + 2.f, cy0 * pitch + 1.f, 0.5f}, pos[3] = {cx0 * pitch - 22.f, cy0 * pitch - 18.f, 14.f};
    look_at(d.cam, pos, at, 70.f, W, H);
    d.cam.spp = spp; d.cam.bounces = bounces; d.cam.cache_bounce = 1;
    return d;
}

}  // namespace rc

// ------------------------------------------------------------------------------------ scene files (writer, MSH1)
namespace rc {
namespace {
struct Writer {
    FILE* f = nullptr;
    bool ok = true;
    template <class T> void wr(const std::vector<T>& v) { if (ok && !v.empty() && fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) ok = false; }
    template <class T> void one(const T& x) { if (ok && fwrite(&x, sizeof(T), 1, f) != 1) ok = false; }
    ~Writer() { if (f) fclose(f); }
};
}  // namespace

std::string save_scene_file(const Scene& S, const FileRefs& R, const char* path) {
    for (int l = 0; l < S.nl; ++l) if (S.ltype[l] != 0) return "the scene file holds point lights only";
    for (int st = 1; st < S.NS; ++st) if (S.nt[st] != S.nt[0]) return "dynamic triangle counts differ between states";
    if (S.BN.empty() || S.HK.empty()) return "scene has no BVH / hash table";
    Writer w;
    w.f = fopen(path, "wb");
    if (!w.f) return std::string("cannot create ") + path;
    const int Nt = S.NS ? S.nt[0] : 0;
    w.wr(std::vector<int>{S.Np, S.Nq, Nt, S.NS, S.nl, S.NR, R.has ? 1 : 0, 3});
    w.wr(std::vector<float>{S.lo[0], S.lo[1], S.lo[2]});
    w.wr(std::vector<float>{S.ext[0], S.ext[1], S.ext[2]});
    w.one(S.h);
    w.wr(std::vector<int>{S.nst, (int)S.BN.size(), (int)S.HK.size(), S.npst});
    w.wr(S.ST); w.wr(S.STN); w.wr(S.BN); w.wr(S.HK); w.wr(S.HV);
    w.wr(S.L); w.wr(S.A); w.wr(S.region);
    for (int st = 0; st < S.NS; ++st) {
        w.wr(S.inten[st]); w.wr(S.P[st]); w.wr(S.N[st]); w.wr(S.Q[st]); w.wr(S.dims); w.wr(S.off); w.wr(S.T[st]); w.wr(S.TQ[st]);
        if (R.has) { w.wr(R.ref[st]); w.wr(R.valid[st]); w.wr(R.Pw[st]); }
    }
    for (int e = 0; e < S.NS * S.NS; ++e) {
        w.one((int)S.occ[e].size()); w.wr(S.occ[e]);
        for (int l = 0; l < S.nl; ++l) { const auto& m = S.em[(size_t)e * S.nl + l]; w.one((int)m.size()); w.wr(m); }
    }
    return w.ok ? "" : std::string("write error on ") + path;
}

std::string read_msh(const char* path, Mesh& M) {
    Reader in;
    if (!in.open(path)) return std::string("cannot open ") + path;
    std::vector<char> magic;
    in.rd(magic, 4);
    const int nt = in.one<int>();
    if (!in.ok || memcmp(magic.data(), "MSH1", 4) != 0) return "not an MSH1 mesh";
    if (nt <= 0 || nt > (1 << 27)) return "bad triangle count";
    std::vector<float> V, alb; std::vector<uint8_t> ds;
    in.rd(V, 9 * (size_t)nt); in.rd(alb, nt); in.rd(ds, nt);
    if (!in.ok) return "truncated mesh";
    M = Mesh();
    for (int t = 0; t < nt; ++t) M.add_tri(&V[9 * (size_t)t], &V[9 * (size_t)t + 3], &V[9 * (size_t)t + 6], alb[t], ds[t] != 0);
    return "";
}

SceneDesc desc_from_mesh(const Mesh& M, int target_points, float* cube_m) {
    SceneDesc d;
    d.mesh = M;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (size_t k = 0; k < M.V.size(); ++k) { lo[k % 3] = std::min(lo[k % 3], M.V[k]); hi[k % 3] = std::max(hi[k % 3], M.V[k]); }
    float ext[3], diag = 0.f;
    for (int a = 0; a < 3; ++a) { ext[a] = std::max(hi[a] - lo[a], 1e-3f); diag += ext[a] * ext[a]; }
    diag = std::sqrt(diag);
    // surface samples (triangle centroids, at most 200k) in a uniform grid for nearest-distance queries
    const int G = 48;
    std::vector<std::vector<int>> grid((size_t)G * G * G);
    std::vector<float> pts;
    const int nt = M.nt(), stride = std::max(1, nt / 200000);
    auto cell = [&](const float* p, int a) { return std::min(G - 1, std::max(0, (int)((p[a] - lo[a]) / ext[a] * G))); };
    for (int t = 0; t < nt; t += stride) {
        float c[3];
        for (int a = 0; a < 3; ++a) c[a] = (M.V[9 * (size_t)t + a] + M.V[9 * (size_t)t + 3 + a] + M.V[9 * (size_t)t + 6 + a]) / 3.f;
        grid[((size_t)cell(c, 2) * G + cell(c, 1)) * G + cell(c, 0)].push_back((int)(pts.size() / 3));
        pts.insert(pts.end(), c, c + 3);
    }
    auto openness = [&](const float* p) {   // distance to the nearest sample, searched ring by ring
        int c0 = cell(p, 0), c1 = cell(p, 1), c2 = cell(p, 2);
        float best = 1e30f;
        for (int r = 0; r < G; ++r) {
            for (int z = std::max(0, c2 - r); z <= std::min(G - 1, c2 + r); ++z)
                for (int y = std::max(0, c1 - r); y <= std::min(G - 1, c1 + r); ++y)
                    for (int x = std::max(0, c0 - r); x <= std::min(G - 1, c0 + r); ++x) {
                        if (std::max(std::abs(x - c0), std::max(std::abs(y - c1), std::abs(z - c2))) != r) continue;
                        for (int i : grid[((size_t)z * G + y) * G + x]) {
                            float dx = pts[3 * i] - p[0], dy = pts[3 * i + 1] - p[1], dz = pts[3 * i + 2] - p[2];
                            best = std::min(best, std::sqrt(dx * dx + dy * dy + dz * dz));
                        }
                    }
            float reach = r * std::min(ext[0], std::min(ext[1], ext[2])) / G;
            if (best < reach) break;
        }
        return best;
    };
    struct Cand { float p[3]; float open; };
    std::vector<Cand> cands;
    const float hz[3] = {0.15f, 0.35f, 0.6f};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 12; ++j)
            for (int i = 0; i < 12; ++i) {
                Cand c;
                c.p[0] = lo[0] + ext[0] * (i + 0.5f) / 12.f; c.p[1] = lo[1] + ext[1] * (j + 0.5f) / 12.f; c.p[2] = lo[2] + ext[2] * hz[k];
                c.open = openness(c.p);
                cands.push_back(c);
            }
    std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.open > b.open; });
    const size_t top = std::max<size_t>(3, cands.size() * 3 / 10);   // the most open 30 %
    auto dist = [](const float* a, const float* b) { float s = 0; for (int x = 0; x < 3; ++x) s += (a[x] - b[x]) * (a[x] - b[x]); return std::sqrt(s); };
    std::vector<int> lamps{0};
    while (lamps.size() < 3) {   // farthest from the lamps placed, among the most open
        int bi = -1; float bd = -1.f;
        for (size_t c = 0; c < top; ++c) {
            float dmin = 1e30f;
            for (int l : lamps) dmin = std::min(dmin, dist(cands[c].p, cands[l].p));
            if (dmin > bd) { bd = dmin; bi = (int)c; }
        }
        lamps.push_back(bi);
    }
    const float b = std::min(0.1f * diag, std::max(0.02f * diag, cands[top - 1].open));   // cube side
    int box[2] = {-1, -1};
    for (int k = 0; k < 2; ++k) {   // open points away from the lamps, the first near lamp 0, the second near lamp 1
        float bd = 1e30f;
        for (size_t c = 0; c < top; ++c) {
            bool free = (int)c != box[0];
            for (int l : lamps) if (dist(cands[c].p, cands[l].p) < 1.2f * b) free = false;
            if (!free) continue;
            float dl = dist(cands[c].p, cands[lamps[k]].p);
            if (dl < bd) { bd = dl; box[k] = (int)c; }
        }
        if (box[k] < 0) box[k] = lamps[k];
    }
    d.nl = 3;
    for (int l : lamps) { for (int x = 0; x < 3; ++x) d.L.push_back(cands[l].p[x]); d.ltype.push_back(0); }
    const float I[5][3] = {{1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 0}};
    const int BP[5] = {0, 0, 1, 0, 1};
    for (int s = 0; s < 5; ++s) {
        d.inten.push_back(std::vector<float>(I[s], I[s] + 3));
        const float* c = cands[box[BP[s]]].p;
        d.objects.push_back({DynBox{{c[0], c[1], c[2]}, {0.5f * b, 0.5f * b, 0.5f * b}, true}});
    }
    d.h = 0.f;
    d.target_points = target_points;
    d.vx = 8; d.vy = 4; d.vz = 4;
    d.fine = 64;
    if (cube_m) *cube_m = b;
    return d;
}
}  // namespace rc

// Host-side scene model and preparation: procedural generators, binned-SAH binary BVH,
// 4-wide collapse, cache-cell sampling, 64-bit-key hash table, dynamic quads and events.
// Mirrors the prototype's scene model (cuda/poc_gpu.cu:737-801, cuda/prep_mesh.cuh:215-531):
// static triangles in a BVH, cache points found through a cell hash grid, per-state light
// intensities and dynamic quads (objects that move / appear), per state-pair event voxels.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include "rc_types.h"

namespace rc {

// SplitMix64-based host RNG of the scene preparation (cuda/prep_mesh.cuh:32-36).
uint64_t mix64(uint64_t x);
struct HostRng {
    uint64_t s;
    explicit HostRng(uint64_t x) : s(mix64(x * 0x9E3779B97F4A7C15ull + 1)) {}
    float u() { s = mix64(s + 0x9E3779B97F4A7C15ull); return (float)(s >> 40) * (1.0f / 16777216.0f); }
};

uint64_t cell_key(int cx, int cy, int cz, int b);
int normal_bucket(float nx, float ny, float nz);
void decode_key(uint64_t k, int& cx, int& cy, int& cz, int& b);
uint64_t spread3(uint64_t x);

struct Mesh {                       // input mesh: triangles, per-triangle albedo, material and double-sided flag
    std::vector<float> V;           // 9 per triangle
    std::vector<float> alb;
    std::vector<int> mat;           // material index per triangle (-1: derived from the albedo)
    std::vector<uint8_t> ds;
    int nt() const { return (int)alb.size(); }
    void add_tri(const float* a, const float* b, const float* c, float albedo, bool dsided, int material = -1);
    void add_quad_grid(const float o[3], const float u[3], const float v[3], int nu, int nv, float albedo, bool dsided, int material = -1);
};

struct DynBox { float c[3]; float e[3]; bool present; };   // axis-aligned box object per state

struct SceneDesc {
    Mesh mesh;
    int nl = 0;
    std::vector<float> L;           // 3 per light
    std::vector<int> ltype;         // 0 point, 1 directional
    std::vector<std::vector<float>> inten;         // per state, nl each
    std::vector<std::vector<DynBox>> objects;      // per state, same object count in every state
    float h = 0.f;                  // cache cell size; 0 -> adjusted toward target_points
    int target_points = 0;
    int table_log2 = 0;             // hash table size 2^k (0 -> smallest pow2 >= max(1024, 2 npst))
    int vx = 8, vy = 4, vz = 4;     // transit grid (cuda/build.sh:12)
    rc_camera cam{};
    float dyn_albedo = 0.7f;
    std::vector<float> matalb;      // material albedo table (empty: the distinct albedos, sorted)
    int dyn_mat = -1;               // material of the dynamic objects' points (-1: by dyn_albedo)
    int fine = 0;                   // fine voxel grid fine^3 for per-light volumes (0: none)
};

struct Scene {
    Mesh mesh;                              // input mesh (input triangle order)
    // static geometry (BVH leaf order)
    int nst = 0;
    std::vector<float> ST, STN, Talb;
    std::vector<int> order;                 // BVH position -> input triangle
    std::vector<rc_bvh_node> BN;
    std::vector<rc_wnode> WN;
    int bvh_depth = 0, wide_depth = 0;
    float lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1}, diag = 0.f, sun_len = 0.f;
    // cache entries
    float h = 0.f;
    int npst = 0, Np = 0, NR = 0;
    std::vector<uint64_t> HK; std::vector<int> HV;
    std::vector<float> Pst, Nst, A;
    std::vector<int> region;
    // materials: per-entry index into the albedo table (the material of the cell's
    // representative sample; dynamic points: the objects' material)
    int nmat = 0;
    std::vector<float> matalb;
    std::vector<int> mat;
    // lights
    int nl = 0;
    std::vector<float> L; std::vector<int> ltype;
    // dynamic quads
    int Nq = 0, nbox = 0;
    std::vector<int> dims, off;
    // states
    int NS = 0;
    std::vector<std::vector<float>> inten, P, N, Q, T;
    std::vector<std::vector<int>> TQ;
    std::vector<int> nt;
    std::vector<std::vector<uint8_t>> present;       // per state, per object
    // events, per ordered state pair (a, b): occluder voxels, per-light emitter voxels
    std::vector<std::vector<int>> occ, em;
    int vx = 8, vy = 4, vz = 4, nw = 4;
    rc_camera cam{};
    std::vector<std::vector<float>> objbox;          // per state, per object: lo3 hi3 (for fine voxel events)
    std::vector<std::vector<DynBox>> objects;        // per state, per object
    int fine = 0;
    std::vector<std::vector<int>> occ_fine;          // per ordered state pair: fine voxels of moving objects
};

// Fine-grid voxels (g^3 over the bounds) overlapped by every object that moves, appears or
// disappears between states a and b (both placements), for every ordered pair.
void fine_events(Scene& S, int g);

// Binned-SAH binary BVH over the mesh (split rule in rc_bvh.h: 64 bins on each of the three
// axes, leaves of <= 2 triangles or when SAH says so, a middle split for equal centroids,
// depth <= 56; the reference bins 16 along the longest axis, cuda/prep_mesh.cuh:42-114), built
// breadth first with stable partitions. Hits do not depend on the tree's shape: the
// closest-hit rule orders ties by triangle index.
void build_bvh(const Mesh& M, std::vector<rc_bvh_node>& nodes, std::vector<int>& order, int& depth);
// The same tree built on the GPU (rc_prep.hip), bit for bit; "" on success, else the error.
std::string build_bvh_device(const Mesh& M, std::vector<rc_bvh_node>& nodes, std::vector<int>& order, int& depth);
// Collapse of the binary BVH into 4-wide nodes (children: up to 4 descendants chosen by least
// surface-area cost over the binary tree, leaves keep their triangle ranges; a leaf of more than RC_WLEAF_MAX triangles becomes a node of its own whose
// children are near-equal chunks of its range, recursively, each with the leaf's box).
void build_wide(const std::vector<rc_bvh_node>& bn, std::vector<rc_wnode>& wn, int& depth);
// the smallest quantization step of a scene's quantized nodes: 2^-18 of its largest absolute coordinate
// (the BVH root box and the scene bounds), rounded up to a power of two
float quant_step_floor(const std::vector<rc_bvh_node>& bn, const float lo[3], const float ext[3]);
// quantized copy of the wide nodes (rc_qnode); min_step: the smallest quantization step
void build_qnodes(const std::vector<rc_wnode>& wn, float min_step, std::vector<rc_qnode>& qn);

struct CellRec { float p[3]; float d2; float n[3]; double asum, wsum; int mat; };
// Cache-cell sampling on the triangles (cuda/prep_mesh.cuh:129-157): returns cells sorted by
// (morton*8 + bucket, key) with their representative point, normal and area-weighted albedo.
void sample_cells(const Mesh& M, float h, const float lo[3], std::vector<uint64_t>& keys, std::vector<CellRec>& recs);
// The same cells computed on the GPU (rc_prep.hip), bit for bit; "" on success, else the error.
std::string sample_cells_device(const Mesh& M, float h, const float lo[3], std::vector<uint64_t>& keys, std::vector<CellRec>& recs);
// the cells' hash table built on the GPU (64-bit CAS inserts, ballot dedup): the host table's keys and
// lookups, slots possibly elsewhere within their probe runs
std::string build_hash_device(const std::vector<uint64_t>& keys, int log2, std::vector<uint64_t>& HK, std::vector<int>& HV);
// Where prepare() builds the BVH and samples the cache cells: false (default) host, true GPU
// (set by rc_scene_prep_device).
extern bool g_prep_device;
// Table of size 2^log2 (linear probing in the sorted-key order, cuda/prep_mesh.cuh:356-366).
void build_hash(const std::vector<uint64_t>& keys, int log2, std::vector<uint64_t>& HK, std::vector<int>& HV);

// Full preparation of a scene description (bounds, BVH, cells, table, zones, quads, events).
std::string prepare(const SceneDesc& d, Scene& S);

// Reader of the prototype's scene files (cuda/poc_gpu.cu:751-801): format 2 (box scenes: all
// geometry dynamic, bounds (0,0,0)-(8,4,3), event voxels on a 16x8x6 grid folded onto the transit
// grid) and format 3 (mesh scenes: static triangles + binary BVH + cell hash table). Converged
// references (ref, valid, per-light Pw) are returned when the file holds them.
struct FileRefs { bool has = false; std::vector<std::vector<float>> ref, Pw; std::vector<std::vector<uint8_t>> valid; };
std::string load_scene_file(const char* path, Scene& S, FileRefs& R);
// Format-3 scene file (the layout load_scene_file reads; cuda/prep_mesh.cuh:242-297): point lights only.
std::string save_scene_file(const Scene& S, const FileRefs& R, const char* path);
// MSH1 mesh file: "MSH1", int nt, float V[9 nt], float albedo[nt], uint8 double-sided[nt] (cuda/prep_mesh.cuh:20-30).
std::string read_msh(const char* path, Mesh& M);
// Bench scene around a mesh (the prep mode's scene, cuda/prep_mesh.cuh:303-453): 3 point lamps and a cube of
// side cube_m at open points, 5 states (lamp intensities {1,1,0} {0,1,0} {1,1,0} {1,1,1} {0,1,0}, cube positions
// 0 0 1 0 1), cache cells adjusted toward target_points. Open points: candidates on a grid at 15-60 % of the
// height (z), scored by the distance to the nearest surface sample.
SceneDesc desc_from_mesh(const Mesh& M, int target_points, float* cube_m);

// Procedural generators.
SceneDesc gen_box_room(uint64_t seed, int target_tris);
SceneDesc gen_street(uint64_t seed, int nboxes, int target_tris, int nlamps, bool sun, int target_points, int table_log2,
                     int W, int H, int spp, int bounces);

}  // namespace rc

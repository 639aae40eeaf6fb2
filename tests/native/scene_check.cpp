// Host-only check of the scene preparation and the CPU oracle, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (Makefile beside this file). It prepares the procedural box room and a
// small street, writes and reads back the room's scene file, runs the oracle's state build, frames of
// the plain / targeted / flush + compact-snapshot methods over an event, the camera pass and the adjoint
// on the room, and then reads every scene file named on the command line (corrupt files must fail
// cleanly). Exit status 0 when every step behaves; the sanitizers abort on the first fault.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../oracle/rc_oracle.h"
#include "../../render_engine_amd/csrc/rc_scene.h"

namespace rc {   // the device preparation is not part of the host-only build
std::string build_bvh_device(const Mesh&, std::vector<rc_bvh_node>&, std::vector<int>&, int&) { return "no GPU in this build"; }
std::string sample_cells_device(const Mesh&, float, const float[3], std::vector<uint64_t>&, std::vector<CellRec>&) { return "no GPU in this build"; }
std::string build_hash_device(const std::vector<uint64_t>&, int, std::vector<uint64_t>&, std::vector<int>&) { return "no GPU in this build"; }
}  // namespace rc

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { ++fails; fprintf(stderr, "check failed: " __VA_ARGS__); fprintf(stderr, "\n"); } } while (0)

struct OracleView {   // one orc_scene per state over a prepared scene, with the oracle's own state build
    std::vector<orc_scene> st;
    std::vector<std::vector<float>> E;
    std::vector<std::vector<uint32_t>> SV;
    std::vector<std::vector<uint64_t>> LM;
    explicit OracleView(const rc::Scene& S) {
        st.resize(S.NS); E.resize(S.NS); SV.resize(S.NS); LM.resize(S.NS);
        for (int k = 0; k < S.NS; ++k) {
            orc_scene& o = st[k];
            memset(&o, 0, sizeof o);
            o.ST = S.ST.data(); o.STN = S.STN.data(); o.BN = (const orc_node*)S.BN.data(); o.WN = (const orc_wnode*)S.WN.data();
            o.HK = S.HK.data(); o.HV = S.HV.data(); o.hmask = (uint32_t)S.HK.size() - 1; o.hcell = S.h;
            o.L = S.L.data(); o.ltype = S.ltype.data(); o.Alb = S.A.data(); o.region = S.region.data();
            for (int a = 0; a < 3; ++a) { o.lo[a] = S.lo[a]; o.ext[a] = S.ext[a]; }
            o.sun_len = S.sun_len;
            o.nst = S.nst; o.nl = S.nl; o.Np = S.Np; o.NR = S.NR; o.vx = S.vx; o.vy = S.vy; o.vz = S.vz; o.nw = S.nw;
            o.P = S.P[k].data(); o.N = S.N[k].data(); o.Q = S.Q[k].data(); o.dims = S.dims.data(); o.off = S.off.data();
            o.T = S.T[k].data(); o.TQ = S.TQ[k].data(); o.nt = S.nt[k]; o.inten = S.inten[k].data();
            E[k].assign((size_t)S.Np * S.nl, 0.f); SV[k].assign((size_t)S.Np * S.nw, 0u); LM[k].assign(S.Np, 0ull);
            o.Edir = E[k].data(); o.SV = SV[k].data(); o.LM = LM[k].data();
            orc_k_build(&o, 1, E[k].data(), SV[k].data(), LM[k].data());
        }
    }
};

static void flat(const std::vector<std::vector<int>>& lists, std::vector<int32_t>& data, std::vector<int64_t>& off) {
    data.clear(); off.assign(1, 0);
    for (const auto& l : lists) { data.insert(data.end(), l.begin(), l.end()); off.push_back((int64_t)data.size()); }
    data.push_back(0);
}

static void run_oracle(const rc::Scene& S) {
    OracleView V(S);
    std::vector<int32_t> occ, em, occf; std::vector<int64_t> oo, eo, fo;
    flat(S.occ, occ, oo); flat(S.em, em, eo); flat(S.occ_fine, occf, fo);
    orc_method ms[3];
    memset(ms, 0, sizeof ms);
    for (auto& m : ms) { m.nmax = 32.f; m.budget = 1.f; m.track_every = 1; m.nmax_resp = 8.f; m.ddgi_h = 0.97f; m.ddgi_thr = 0.25f; }
    ms[1].deps = 1; ms[1].targeted = 1;
    ms[2].flush = 1; ms[2].snaps = 1; ms[2].compact = 1; ms[2].eventboost = 30;
    for (const orc_method& m : ms) {
        orc_runner* r = orc_runner_create(V.st.data(), S.NS, &m, 1, occ.data(), oo.data(), em.data(), eo.data(), nullptr, nullptr, nullptr,
                                          occf.data(), fo.data());
        CHECK(r != nullptr, "runner");
        orc_runner_begin(r, 0, 5);
        int inv = 0;
        for (int f = 0; f < 26; ++f) orc_runner_frame(r, f < 21 ? 0 : (f < 24 ? 1 : 0), 0, &inv);
        const orc_cache* c = orc_runner_cache(r);
        double sum = 0;
        for (int i = 0; i < S.Np; ++i) sum += c->R[i];
        CHECK(sum > 0, "resolved radiance");
        orc_runner_destroy(r);
    }
    orc_camera cam;
    memcpy(&cam, &S.cam, sizeof cam);
    cam.W = 48; cam.H = 32; cam.spp = 2; cam.bounces = 3;
    std::vector<float> img(cam.W * cam.H), R(S.Np, 0.5f);
    long long rays[2] = {0, 0};
    orc_k_query(&V.st[0], 1, R.data(), &cam, 3, 0, 0, img.data(), rays, 0, cam.W * cam.H);
    orc_k_query(&V.st[0], 1, nullptr, &cam, 3, 0, 0, img.data(), rays, 0, cam.W * cam.H);
    CHECK(rays[0] > 0, "camera rays");
    const int K = 2, B = 2;
    std::vector<float> G((size_t)S.Np * S.nl), Rr(S.Np), LB(S.Np);
    std::vector<uint64_t> SG(S.Np), SR(S.Np);
    std::vector<double> G64((size_t)S.Np * S.nl), R64(S.Np), g(S.Np, 1.0), dI(S.nl), dalb(S.nmat), alb64(S.matalb.begin(), S.matalb.end());
    std::vector<double> I64(S.inten[0].begin(), S.inten[0].end());
    std::vector<int32_t> hits((size_t)B * S.Np * K);
    orc_adj_unit(&V.st[0], 1, G.data(), SG.data(), G64.data());
    orc_adj_hits(&V.st[0], 1, K, B, 9, hits.data());
    orc_adj_forward(S.Np, S.nl, K, B, hits.data(), S.mat.data(), S.matalb.data(), S.inten[0].data(), G.data(), SG.data(), Rr.data(), SR.data(),
                    LB.data());
    orc_adj_grad64(S.Np, S.nl, S.nmat, K, B, hits.data(), S.mat.data(), alb64.data(), I64.data(), G64.data(), g.data(), R64.data(), dI.data(),
                   dalb.data());
}

// quantized wide nodes: every child box decoded in exact (double) arithmetic holds its fp32 box with at
// least one step of slack on each side, steps are powers of two at or above the floor, empty slots and
// children are as in the fp32 nodes
static void check_qnodes(const rc::Scene& S) {
    std::vector<rc_qnode> qn;
    const float floor_step = rc::quant_step_floor(S.BN, S.lo, S.ext);
    rc::build_qnodes(S.WN, floor_step, qn);
    CHECK(qn.size() == S.WN.size(), "quantized node count");
    long long boxes = 0;
    for (size_t i = 0; i < qn.size(); ++i) {
        const rc_wnode& w = S.WN[i];
        const rc_qnode& q = qn[i];
        const float o[3] = {q.ox, q.oy, q.oz}, st[3] = {q.sx, q.sy, q.sz};
        const uint32_t ql[3] = {q.qlox, q.qloy, q.qloz}, qh[3] = {q.qhix, q.qhiy, q.qhiz};
        const float* lo[3] = {w.lox, w.loy, w.loz};
        const float* hi[3] = {w.hix, w.hiy, w.hiz};
        for (int a = 0; a < 3; ++a) {
            int e2 = 0;
            CHECK(std::frexp((double)st[a], &e2) == 0.5 && st[a] >= floor_step, "node %zu axis %d: step %g not a power of two >= %g", i,
                  a, st[a], floor_step);
        }
        for (int c = 0; c < 4; ++c) {
            CHECK(q.child[c] == w.child[c], "node %zu child %d", i, c);
            for (int a = 0; a < 3; ++a) {
                const uint32_t l = (ql[a] >> (8 * c)) & 255u, h = (qh[a] >> (8 * c)) & 255u;
                if (w.child[c] == 0) { CHECK(l == 255u && h == 0u, "node %zu empty slot %d", i, c); continue; }
                const double dl = (double)o[a] + (double)l * st[a], dh = (double)o[a] + (double)h * st[a];
                CHECK(dl <= (double)lo[a][c] - st[a] && dh >= (double)hi[a][c] + st[a], "node %zu child %d axis %d: [%g, %g] vs [%g, %g]",
                      i, c, a, dl, dh, lo[a][c], hi[a][c]);
                ++boxes;
            }
        }
    }
    CHECK(boxes > 0, "no quantized boxes checked");
}

int main(int argc, char** argv) {
    rc::Scene room, street;
    std::string e = rc::prepare(rc::gen_box_room(1, 600), room);
    CHECK(e.empty(), "box room: %s", e.c_str());
    e = rc::prepare(rc::gen_street(3, 40, 3000, 4, true, 4000, 0, 64, 48, 2, 3), street);
    CHECK(e.empty(), "street: %s", e.c_str());
    const char* path = "_build/room_check.bin";
    rc::FileRefs none;
    e = rc::save_scene_file(room, none, path);
    CHECK(e.empty(), "save: %s", e.c_str());
    rc::Scene back; rc::FileRefs refs;
    e = rc::load_scene_file(path, back, refs);
    CHECK(e.empty() && back.Np == room.Np && back.nst == room.nst && back.BN.size() == room.BN.size(), "reload: %s", e.c_str());
    run_oracle(room);
    run_oracle(street);
    check_qnodes(room);
    check_qnodes(street);
    for (int k = 1; k < argc; ++k) {   // scene files given by the caller: read or refused, never a fault
        rc::Scene S; rc::FileRefs R;
        e = rc::load_scene_file(argv[k], S, R);
        printf("%s: %s\n", argv[k], e.empty() ? "loaded" : e.c_str());
    }
    printf("scene_check: %d failed checks\n", fails);
    return fails ? 1 : 0;
}

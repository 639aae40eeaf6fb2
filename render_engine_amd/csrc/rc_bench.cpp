// rc_bench: the prototype's command line (cuda/poc_gpu.cu:1286-1572) on top of librc, writing the
// same JSON-lines records ("scene", "timing", "scen", "vram", "cpu") to stdout.
//
//   rc_bench timing <scene>                                    per-phase GPU ms (:1344-1367)
//   rc_bench scenf  <scene> <nseeds> <scenarios.txt> [budget]  error curves per scenario and method (:1370-1436)
//   rc_bench res    <scene>                                    memory and host cost per frame (:1321-1343)
//
// <scene> is a prototype scene file (format 2 / 3, uncompressed), "room" (procedural box room) or
// "street" (procedural street, 960x540 camera). Errors go to stderr with a non-zero exit code.
#include <sys/resource.h>
#include <time.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rc.h"

namespace {

struct Named { std::string name; rc_method m; };

rc_method base(int mode, float nmax, float budget, bool deps = false, bool verify = false, bool oracle = false, int track = 1) {
    rc_method m;
    memset(&m, 0, sizeof m);
    m.mode = mode; m.nmax = nmax; m.budget = budget; m.deps = deps; m.verify = verify; m.oracle = oracle; m.track_every = track;
    m.nmax_resp = 8.f; m.ddgi_h = 0.97f; m.ddgi_thr = 0.25f; m.ddgi_clamp = 1;
    return m;
}

std::string fmt2(const char* f, float b) { char buf[128]; snprintf(buf, sizeof buf, f, b); return buf; }

// cuda/poc_gpu.cu:1348-1358
std::vector<Named> timing_methods() {
    std::vector<Named> M;
    M.push_back({"Classique 1.0x", base(0, 16.f, 1.f)});
    M.push_back({"Classique 2.0x", base(0, 16.f, 2.f)});
    { rc_method r = base(0, 32.f, 1.f); r.resp_mask = 1; r.rescale = 1; M.push_back({"SHaRC responsive 1.0x", r}); }
    { rc_method r = base(0, 64.f, 1.f); r.ddgi = 1; r.rescale = 1; M.push_back({"DDGI 1.0x", r}); }
    M.push_back({"v3 sans suivi", base(1, 64.f, 1.f, false, true)});
    M.push_back({"v3 complet", base(1, 64.f, 1.f, true, true, false, 1)});
    M.push_back({"v3 suivi 1 frame sur 2", base(1, 64.f, 1.f, true, true, false, 2)});
    M.push_back({"v3 suivi 1 frame sur 4", base(1, 64.f, 1.f, true, true, false, 4)});
    { rc_method r = base(1, 64.f, 1.f, true, true, false, 4); r.memory = 1; r.wake_verify = 1; r.ramp = 1; M.push_back({"v4-A2C suivi 1/4", r}); }
    return M;
}

// cuda/poc_gpu.cu:1404-1418
std::vector<Named> scen_methods(uint64_t dyn, float bS, float bA) {
    std::vector<Named> M;
    auto sharc = [&](bool mem, bool snaps) {
        rc_method r = base(0, 64.f, bS); r.resp_mask = dyn; r.nmax_resp = 8.f; r.rescale = 1; r.memory = mem; r.snaps = snaps; return r;
    };
    auto accum = [&](float N, bool mem, bool snaps) { rc_method r = base(0, N, bA); r.rescale = 1; r.memory = mem; r.snaps = snaps; return r; };
    M.push_back({fmt2("SHaRC N=64 resp=8, %.2fx", bS), sharc(false, false)});
    M.push_back({fmt2("SHaRC + memoire naive, %.2fx", bS), sharc(true, false)});
    M.push_back({fmt2("Accumulation+parlampe N=32, %.2fx", bA), accum(32.f, false, false)});
    M.push_back({fmt2("Accum+parlampe + memoire naive N=32, %.2fx", bA), accum(32.f, true, false)});
    M.push_back({fmt2("Accum+parlampe + memoire naive N=64, %.2fx", bA), accum(64.f, true, false)});
    M.push_back({fmt2("SHaRC + memoire + instantanes, %.2fx", bS), sharc(true, true)});
    M.push_back({fmt2("Accum + memoire + instantanes N=32, %.2fx", bA), accum(32.f, true, true)});
    M.push_back({fmt2("Accum + memoire + instantanes N=64, %.2fx", bA), accum(64.f, true, true)});
    M.push_back({"Oracle", base(1, 64.f, 1.f, false, false, true)});
    M.push_back({"v3", base(1, 64.f, 1.f, true, true, false, 4)});
    { rc_method r = base(1, 64.f, 1.f, true, true, false, 4); r.memory = 1; r.wake_verify = 1; r.ramp = 1; M.push_back({"v4-A2C memoire verifiee+rampe", r}); }
    { rc_method r = base(1, 64.f, 1.f, true, true, false, 4); r.memory = 1; r.wake_verify = 1; r.ramp = 1; r.snaps = 1; r.snap_verify = 1;
      M.push_back({"v5-A instantanes verifies", r}); }
    { rc_method r = base(1, 64.f, 1.f, true, true, false, 4); r.memory = 1; r.wake_verify = 1; r.ramp = 1; r.snaps = 1;
      M.push_back({"v5-A instantanes", r}); }
    return M;
}

int die(const char* what) {
    fprintf(stderr, "rc_bench: %s: %s\n", what, rc_last_error());
    return 1;
}

double cpu_seconds() {
    struct rusage u;
    getrusage(RUSAGE_SELF, &u);
    return u.ru_utime.tv_sec + u.ru_utime.tv_usec * 1e-6 + u.ru_stime.tv_sec + u.ru_stime.tv_usec * 1e-6;
}
double wall_ms() { timespec t; clock_gettime(CLOCK_MONOTONIC, &t); return t.tv_sec * 1e3 + t.tv_nsec * 1e-6; }
double ram_mb() { struct rusage u; getrusage(RUSAGE_SELF, &u); return u.ru_maxrss / 1024.0; }

rc_status open_scene(const char* arg, rc_scene** s) {
    if (!strcmp(arg, "room")) return rc_scene_box_room(1, 2000, s);
    if (!strcmp(arg, "street")) return rc_scene_street(1, 2000, 200000, 32, 1, 530000, 22, 960, 540, 4, 3, s);
    return rc_scene_load(arg, s);
}

// run() over a state sequence: per-frame errors (want_err), phase times accumulated in T[5]
rc_status run(rc_cache* c, const std::vector<int>& seq, uint32_t seed, bool want_err, std::vector<double>* errs, double* T, int* frames,
              int* events) {
    rc_status st = rc_cache_begin(c, seq[0], seed);
    if (st != RC_OK) return st;
    for (int s : seq) {
        rc_frame_stats fs;
        st = rc_cache_frame(c, s, want_err, T != nullptr, &fs);
        if (st != RC_OK) return st;
        if (errs) errs->push_back(fs.err);
        if (T) {
            T[0] += fs.ms_swap; T[1] += fs.ms_trace; T[2] += fs.ms_decide; T[3] += fs.ms_update;
            if (fs.event) { T[4] += fs.ms_event; ++*events; }
            ++*frames;
        }
    }
    return rc_cache_sync(c);
}

// prep: mesh (or the procedural room) -> bench scene with converged references, written as a format-3 file;
// one "prep" record (cuda/prep_mesh.cuh:303-531, record :522-528)
int prep_mode(int argc, char** argv) {
    if (argc < 4) { fprintf(stderr, "usage: rc_bench prep <mesh.msh|room> <out.bin> [points]\n"); return 1; }
    const double t0 = wall_ms();
    const int target = argc > 4 ? atoi(argv[4]) : 300000;
    rc_scene* S = nullptr;
    float cube = 0.f;
    if (!strcmp(argv[2], "room")) { if (rc_scene_box_room(1, 2000, &S) != RC_OK) return die("scene"); }
    else if (rc_scene_mesh(argv[2], target, &S, &cube) != RC_OK) return die("mesh");
    if (rc_scene_upload(S, 2) != RC_OK) return die("upload");
    rc_scene_info info;
    rc_scene_info_get(S, &info);
    std::vector<float> noise(info.NS);
    std::vector<int> iters(info.NS);
    if (rc_scene_reference(S, 32, 16, 64, 512, 0.03f, noise.data(), iters.data()) != RC_OK) return die("reference");
    double vfrac = 0;
    for (int s = 0; s < info.NS; ++s) {
        std::vector<uint8_t> v(info.Np);
        if (rc_scene_array(S, RC_A_VALID, s, v.data(), v.size()) != RC_OK) return die("valid");
        for (uint8_t x : v) vfrac += x;
    }
    vfrac /= (double)info.Np * info.NS;
    if (rc_scene_save(S, argv[3]) != RC_OK) return die("save");
    std::vector<float> L(3 * (size_t)info.nl);
    rc_scene_array(S, RC_A_L, 0, L.data(), L.size() * 4);
    printf("{\"type\":\"prep\",\"triangles\":%d,\"np\":%d,\"cellule_m\":%.4f,\"zones\":%d,\"diag_m\":%.2f,\"cube_m\":%.2f,\"bvh_noeuds\":%d,"
           "\"points_valides\":%.4f,\"bruit_ref\":[", info.nst, info.Np, info.h, info.NR, info.diag, cube, info.nnodes, vfrac);
    for (int k = 0; k < info.NS; ++k) printf("%s%.5f", k ? "," : "", noise[k]);
    printf("],\"iterations_ref\":[");
    for (int k = 0; k < info.NS; ++k) printf("%s%d", k ? "," : "", iters[k]);
    printf("],\"lampes\":[");
    for (int l = 0; l < info.nl; ++l) printf("%s[%.3f,%.3f,%.3f]", l ? "," : "", L[3 * l], L[3 * l + 1], L[3 * l + 2]);
    printf("],\"secondes\":%.1f}\n", (wall_ms() - t0) / 1e3);
    rc_scene_destroy(S);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: rc_bench {timing|scenf|res} <scene.bin|room|street> [nseeds scenarios.txt [budget]]\n"
                        "       rc_bench prep <mesh.msh|room> <out.bin> [points]\n");
        return 1;
    }
    std::string mode = argv[1];
    if (mode == "prep") return prep_mode(argc, argv);
    rc_scene* S = nullptr;
    if (open_scene(argv[2], &S) != RC_OK) return die("scene");
    if (rc_scene_upload(S, 2) != RC_OK) return die("upload");
    rc_scene_info info;
    rc_scene_info_get(S, &info);
    printf("{\"type\":\"scene\",\"np\":%d,\"nt\":%d,\"gpu\":\"MI355X\",\"mask_bits\":%d,\"triangles_statiques\":%d,\"lampes\":%d}\n", info.Np,
           info.nst, 32 * info.nw, info.nst, info.nl);
    fflush(stdout);
    std::vector<int> seq;
    for (int s = 0; s < info.NS; ++s) for (int f = 0; f < 60; ++f) seq.push_back(s);
    if (mode == "timing") {
        for (auto& M : timing_methods()) {
            rc_cache* c = nullptr;
            if (rc_cache_create(S, &M.m, &c) != RC_OK) return die("cache");
            double warm[5] = {0}, T[5] = {0};
            int fw = 0, ew = 0, fr = 0, ev = 0;
            if (run(c, seq, 7, false, nullptr, warm, &fw, &ew) != RC_OK) return die("run");
            for (int rep = 0; rep < 3; ++rep)
                if (run(c, seq, 11 + rep, false, nullptr, T, &fr, &ev) != RC_OK) return die("run");
            printf("{\"type\":\"timing\",\"name\":\"%s\",\"trace\":%.5f,\"decide\":%.5f,\"update\":%.5f,\"swap\":%.5f,\"event\":%.5f}\n",
                   M.name.c_str(), T[1] / fr, T[2] / fr, T[3] / fr, T[0] / fr, ev ? T[4] / ev : 0.0);
            fflush(stdout);
            rc_cache_destroy(c);
        }
    } else if (mode == "res") {
        std::vector<Named> M;
        M.push_back({"Classique", base(0, 16.f, 1.f)});
        { rc_method r = base(0, 64.f, 1.f); r.resp_mask = 1; r.rescale = 1; M.push_back({"SHaRC responsive", r}); }
        M.push_back({"v3", base(1, 64.f, 1.f, true, true, false, 4)});
        { rc_method r = base(1, 64.f, 1.f, true, true, false, 4); r.memory = 1; r.wake_verify = 1; r.ramp = 1; M.push_back({"v4-A2C", r}); }
        printf("{\"type\":\"vram\",\"ram_pic_mb\":%.1f}\n", ram_mb());
        for (auto& m : M) {
            rc_cache* c = nullptr;
            if (rc_cache_create(S, &m.m, &c) != RC_OK) return die("cache");
            if (run(c, seq, 5, false, nullptr, nullptr, nullptr, nullptr) != RC_OK) return die("run");
            double c0 = cpu_seconds(), t0 = wall_ms();
            const int reps = 3;
            for (int k = 0; k < reps; ++k)
                if (run(c, seq, 9 + k, false, nullptr, nullptr, nullptr, nullptr) != RC_OK) return die("run");
            double c1 = cpu_seconds(), t1 = wall_ms();
            int frames = reps * (int)seq.size();
            printf("{\"type\":\"cpu\",\"name\":\"%s\",\"cpu_ms_par_frame\":%.4f,\"mur_ms_par_frame\":%.4f}\n", m.name.c_str(),
                   (c1 - c0) * 1000.0 / frames, (t1 - t0) / frames);
            fflush(stdout);
            rc_cache_destroy(c);
        }
    } else if (mode == "scenf") {
        if (argc < 5) { fprintf(stderr, "usage: rc_bench scenf <scene> <nseeds> <scenarios.txt> [budget]\n"); return 1; }
        if (!info.has_ref) {   // procedural scenes: converged references first (cuda/prep_mesh.cuh:454-520)
            std::vector<float> noise(info.NS);
            std::vector<int> iters(info.NS);
            if (rc_scene_reference(S, 32, 16, 64, 512, 0.03f, noise.data(), iters.data()) != RC_OK) return die("reference");
            rc_scene_info_get(S, &info);
        }
        int nseeds = atoi(argv[3]);
        FILE* sf = fopen(argv[4], "r");
        if (!sf) { fprintf(stderr, "rc_bench: scenario file %s\n", argv[4]); return 1; }
        std::vector<std::pair<std::string, std::vector<std::pair<int, int>>>> SC;
        char line[1024];
        while (fgets(line, sizeof line, sf)) {
            std::string L(line);
            while (!L.empty() && (L.back() == '\n' || L.back() == '\r')) L.pop_back();
            if (L.empty() || L[0] == '#') continue;
            size_t sc = L.find(';');
            if (sc == std::string::npos) continue;
            std::vector<std::pair<int, int>> parts;
            std::string rest = L.substr(sc + 1);
            size_t pos = 0;
            while (pos < rest.size()) {
                size_t comma = rest.find(',', pos);
                if (comma == std::string::npos) comma = rest.size();
                std::string tok = rest.substr(pos, comma - pos);
                size_t colon = tok.find(':');
                int st = atoi(tok.substr(0, colon).c_str()), nf = atoi(tok.substr(colon + 1).c_str());
                if (st < 0 || st >= info.NS || nf <= 0) { fprintf(stderr, "rc_bench: bad scenario token '%s'\n", tok.c_str()); return 1; }
                parts.push_back({st, nf});
                pos = comma + 1;
            }
            SC.push_back({L.substr(0, sc), parts});
        }
        fclose(sf);
        // lights whose intensity changes between states are "responsive" for SHaRC (:1395-1396)
        uint64_t dyn = 0;
        std::vector<std::vector<float>> inten(info.NS, std::vector<float>(info.nl));
        for (int s = 0; s < info.NS; ++s) rc_scene_array(S, RC_A_INTEN, s, inten[s].data(), inten[s].size() * 4);
        for (int s = 1; s < info.NS; ++s) for (int l = 0; l < info.nl && l < 64; ++l) if (inten[s][l] != inten[0][l]) dyn |= 1ull << l;
        float bS = 1.57f, bA = 1.53f;
        if (argc > 5) bS = bA = (float)atof(argv[5]);
        for (auto& sc : SC) {
            std::vector<int> q;
            for (auto& pr : sc.second) for (int k = 0; k < pr.second; ++k) q.push_back(pr.first);
            int last = (int)q.size() - sc.second.back().second;
            for (auto& M : scen_methods(dyn, bS, bA)) {
                rc_cache* c = nullptr;
                if (rc_cache_create(S, &M.m, &c) != RC_OK) return die("cache");
                double pre = 0, post = 0;
                std::vector<double> curve(q.size(), 0.0);
                for (int sd = 0; sd < nseeds; ++sd) {
                    std::vector<double> er;
                    if (run(c, q, 3000 + sd, true, &er, nullptr, nullptr, nullptr) != RC_OK) return die("run");
                    for (int f = std::max(0, last - 20); f < last; ++f) pre += er[f] / 20.0 / nseeds;
                    for (int f = last; f < std::min((int)q.size(), last + 30); ++f) post += er[f] / 30.0 / nseeds;
                    for (size_t f = 0; f < q.size(); ++f) curve[f] += er[f] / nseeds;
                }
                printf("{\"type\":\"scen\",\"scen\":\"%s\",\"name\":\"%s\",\"pre\":%.5f,\"post\":%.5f,\"last\":%d,\"curve\":[", sc.first.c_str(),
                       M.name.c_str(), pre, post, last);
                for (size_t f = 0; f < curve.size(); ++f) printf("%s%.5f", f ? "," : "", curve[f]);
                printf("]}\n");
                fflush(stdout);
                rc_cache_destroy(c);
            }
        }
    } else {
        fprintf(stderr, "rc_bench: unknown mode %s\n", mode.c_str());
        return 1;
    }
    rc_scene_destroy(S);
    return 0;
}

"""Benchmark of the radiance-cache frame on one MI355X.

Workload (default): procedural street of 2,000 boxes (~1M triangles), 64 point lamps + sun, ~1M cache
points in a 2^23-entry hash grid, 256^3 light volumes, camera 1920x1080 at 4 spp. One step = one
frame of the plain-cache method: entry pass (k_trace2: one cosine ray per cache entry, cache read at
the hit) + resolve (k_update EMA) + camera query pass (k_query: primary ray, one-light NEE shadow ray,
one bounce ray ending on the cached radiance; without a cache a path takes up to 3 bounces).

Prints one JSON line: value = Mrays/s over all rays traced in the timed frames (entry rays + query
extension rays + query shadow rays), with inputs resident in HBM. Extra fields: cache Mpts/s,
frames/s, per-kernel times, the roofline of the dominant kernel (VALU issue for the camera pass, from
rocprofv3 SQ counters of this workload, with its HBM traffic beside it), the CPU baseline (the C
oracle on the host cores, bounded sample), the relative RMS of the FP32 results against their FP64
twins, and the event timeline (light-off, night, occluder: each held 90 frames, then reverted).

Other modes: --config 4 (adjoint), --config 5 (learned cache), --events (the event kernels' workload,
for profiling), --pmc-summary (rocprofv3 counter CSVs -> the JSON summary read by the roofline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # small box room (CPU-runnable case): 256x256, 1 spp, 1 bounce, 2^18 grid
    1: dict(kind="room", seed=1, tris=2000),
    # street at 960x540, 4 spp, 3 bounces, ~530k points, 2^22 grid
    2: dict(kind="street", seed=1, nboxes=2000, tris=200_000, nlamps=32, sun=True, points=530_000, table_log2=22, W=960, H=540,
            spp=4, bounces=3),
    # street at 1920x1080, 4 spp, 3 bounces, ~1M points, 2^23 grid, 64 lamps
    3: dict(kind="street", seed=1, nboxes=2000, tris=1_000_000, nlamps=64, sun=True, points=1_000_000, table_log2=23, W=1920,
            H=1080, spp=4, bounces=3),
    # differentiation: the 960x540 street at 2 spp, 63 lamps + sun = 64 emission scalars, 4,096 material albedos;
    # adjoint of all cached radiance, 3 bounces, 4 rays per entry and bounce
    4: dict(kind="street", seed=1, nboxes=2000, tris=200_000, nlamps=63, sun=True, points=530_000, table_log2=22, W=960, H=540,
            spp=2, bounces=3, adjoint=dict(K=4, B=3, seed=17)),
    # learned cache: config 2's street; hash-encoded MLP (16 levels x 2 features, 2^19 table, 5 x 64 ReLU), BF16 MFMA,
    # Adam lr 1e-2, 2^16 self-training samples per frame, queried at 1920x1080
    5: dict(kind="street", seed=1, nboxes=2000, tris=200_000, nlamps=32, sun=True, points=530_000, table_log2=22, W=960, H=540,
            spp=4, bounces=3, mlp=dict(log2_table=19, base_res=16, max_res=2048, batch=1 << 16, lr=1e-2, W=1920, H=1080)),
}


def make_scene(rea, cfg):
    if cfg["kind"] == "room":
        return rea.Scene.box_room(cfg["seed"], cfg["tris"])
    return rea.Scene.street(cfg["seed"], cfg["nboxes"], cfg["tris"], cfg["nlamps"], cfg["sun"], cfg["points"], cfg["table_log2"],
                            cfg["W"], cfg["H"], cfg["spp"], cfg["bounces"])


def update_bytes_per_entry(nl, tracked=False):
    """k_update algorithmic bytes per entry: read kk, n, albedo, Lc[nl], m[nl], Edir[nl];
    write Lc2[nl], n, R"""
    return 4 + 4 + 4 + 16 * nl + 4 + 4 + (8 if tracked else 0)


def oracle_scene(scene):
    """the CPU oracle's view of the scene, with the GPU-built (bit-identical) state arrays"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    osc = po.OracleScene(scene, build=False)
    for st in range(scene.info.NS):
        osc.Edir[st][:] = scene.array("EDIR", st)
        osc.SV[st][:] = scene.array("SV", st)
        osc.LM[st][:] = scene.array("LM", st)
        if scene.info.fine > 0:
            osc.VOL[st][:] = scene.array("VOL", st)
    return osc


def query_bytes(osc, R, cam, seed, frame, log):
    """Algorithmic bytes of one query launch: per ray 32 B per binary-BVH node visited and 36 B
    per triangle tested, 8 B per hash probe, 4 B per resolved-radiance read, plus the image;
    counted by the CPU oracle on the same paths over 1/8 of the rows (8 bands) and scaled."""
    import ctypes as C

    import numpy as np
    import pyoracle as po

    W, H = cam.W, cam.H
    band = max(1, H // 64)
    img = np.zeros(W * H, np.float32)
    po.lib.orc_count_enable(1)
    sampled = 0
    for b in range(8):
        r0 = b * H // 8
        r1 = min(H, r0 + band)
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(R), C.byref(cam), seed, frame, 0, po.ptr(img), None, r0 * W, r1 * W)
        sampled += (r1 - r0) * W
    cnt = np.zeros(4, np.int64)
    po.lib.orc_count_get(po.ptr(cnt))
    po.lib.orc_count_enable(0)
    scale = W * H / sampled
    per = dict(nodes=cnt[0] * scale, tris=cnt[1] * scale, probes=cnt[2] * scale, resolved=cnt[3] * scale)
    total = 32 * per["nodes"] + 36 * per["tris"] + 8 * per["probes"] + 4 * per["resolved"] + 4 * W * H
    log(f"query work per launch (oracle count, 1/{scale:.0f} of rows): {json.dumps({k: round(v) for k, v in per.items()})}")
    return float(total), per


def cpu_baseline(osc, scene, seconds, log):
    """The C oracle on the host cores: the same frame (entry pass + update + query) on the same
    scene in state 0, timed per frame."""
    import ctypes as C

    import numpy as np
    import pyoracle as po
    import render_engine_amd as rea

    m = rea.methods.plain()
    run = po.OracleRunner(osc, m)
    run.begin(0, 1234)
    cam = scene.camera()
    img = np.zeros(cam.W * cam.H, np.float32)
    frames, rays, t_tot = 0, 0, 0.0
    while frames < 1 or (t_tot < seconds and frames < 50):
        t0 = time.perf_counter()
        run.frame(0)
        t1 = time.perf_counter()
        R = run.get("R")   # a copy, outside the timed parts
        er = (C.c_longlong * 2)()
        t2 = time.perf_counter()
        po.lib.orc_k_query(osc.state(0), 1, po.ptr(R), C.byref(cam), 99, frames, 0, po.ptr(img), er, 0, cam.W * cam.H)
        t3 = time.perf_counter()
        rays += int(run.get("KK").sum()) + er[0] + er[1]
        t_tot += (t1 - t0) + (t3 - t2)
        frames += 1
    log(f"cpu baseline: {frames} frames in {t_tot:.1f} s on {po.lib.orc_num_threads()} threads")
    return dict(value=rays / t_tot / 1e6, unit="Mrays/s", cores=int(po.lib.orc_num_threads()), kind="port",
                sample=f"{frames} full plain-cache frames (entry pass + update + {cam.W}x{cam.H}x{cam.spp} query) of the "
                       f"same scene, state 0", fps=frames / t_tot)


def mlp_cpu_baseline(mlp, cfg, P, N, D, seconds, log):
    """The learned cache's oracle (numpy, one thread) on the host: encoding + network inference of
    the same queries with the same table and weights, timed per batch."""
    import numpy as np
    import mlp_oracle as MO
    import render_engine_amd.mlp as M
    from threadpoolctl import threadpool_limits

    table, W = mlp.get(M.MLP_TABLE), mlp.get(M.MLP_W)
    lo, ext = np.array(list(cfg.lo), np.float32), np.array(list(cfg.ext), np.float32)
    n, done, t_tot = len(P), 0, 0.0
    with threadpool_limits(limits=1):
        while done == 0 or (t_tot < seconds and done < 8 * n):
            t0 = time.perf_counter()
            X, _ = MO.encode(table, P, N, D, lo, ext, cfg.log2_table, cfg.base_res, cfg.max_res)
            MO.forward(W, X)
            t_tot += time.perf_counter() - t0
            done += n
    log(f"mlp cpu baseline: {done} queries in {t_tot:.1f} s on 1 thread")
    return dict(value=done / t_tot / 1e6, unit="Mqueries/s (encoding + network inference only, no path tracing)", cores=1, kind="port",
                sample=f"{done} queries ({n} cache-entry positions and normals) through the numpy oracle, one thread")


def reference_images(rea, scene, cam, log, ref_spp_rounds=16):
    """converged per-state references (two-chain Jacobi solve), then per state the camera image at
    ref_spp_rounds x spp through the converged cache"""
    import numpy as np
    import torch

    t0 = time.time()
    noise, iters = scene.reference(K=32, warm=8, block=32, max_iter=256, target=0.03)
    log(f"converged references: self-noise {[round(x, 4) for x in noise]}, iterations {iters}, {time.time() - t0:.1f} s")
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    refs = {}
    for st in range(scene.info.NS):
        c = rea.Cache(scene, rea.methods.plain())
        c.begin(st, 1)
        acc = np.zeros(cam.W * cam.H)
        for k in range(ref_spp_rounds):
            c.render(img, cam, seed=100000 + k)
            acc += img.cpu().numpy()
        refs[st] = (acc / ref_spp_rounds).reshape(cam.H, cam.W)
    return refs


def event_list(scene):
    """the street's events (config 2: sun to 0, lamp 0 off, 12 x 10 m occluder inserted), each from the
    base state; the box room's light-1-off event. (name, base state, event state, region kind)"""
    if scene.info.NS >= 4:
        return [("night (sun off)", 0, 1, "light"), ("lamp 0 off", 0, 2, "light"), ("occluder inserted", 0, 3, "panel")]
    return [("light 1 off", 0, 1, "light")]


def quality(rea, scene, cam, log, pre=20, hold=90, after=30, ref_spp_rounds=16, methods=None):
    """Event timeline per method against converged references (cuda/prep_mesh.cuh:454-520 solve, then
    the same camera pass at 16 x spp through the converged cache): `pre` frames in the base state, the
    event state held `hold` frames, then reverted for `after` frames (snapshot restore and relighting).
    Per frame the region-luminance bias sum(est) / sum(ref) - 1 over the affected region: for a light
    event the pixels where the light carried > 50 % of the converged radiance, for a geometry event the
    pixels whose converged radiance changes by > 25 % (9 x 9 blur at 960 px, windows/sondes.py:30-47);
    frames until |bias| < 5 % after the event and after the revert."""
    import numpy as np
    import torch

    from render_engine_amd import scenarios as sc

    refs = reference_images(rea, scene, cam, log, ref_spp_rounds)
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    if methods is None:
        methods = (("plain", rea.methods.plain()), ("flush_snapshots", rea.methods.flush_snapshots()),
                   ("targeted", rea.methods.targeted()), ("targeted_grad", rea.methods.targeted_grad()))
    out = {}
    for name, a, b, kind in event_list(scene):
        region = sc.light_off_region(refs[a], refs[b]) if kind == "light" else sc.panel_region(refs[a], refs[b])
        ev = {"region": "light > 50 % of the converged radiance" if kind == "light" else "converged radiance changes > 25 %",
              "region_pixels": int(region.sum()), "timeline": f"{pre} frames base, {hold} event, {after} reverted", "methods": {}}
        if region.sum() == 0:
            out[name] = ev
            continue
        reg = torch.from_numpy(region.ravel()).cuda()
        ref_sum = {st: float(refs[st][region].sum()) for st in (a, b)}
        for mname, m in methods:
            c = rea.Cache(scene, m)
            c.begin(a, 777)
            seq = [a] * pre + [b] * hold + [a] * after
            biases, errs, inval, ev_ms = [], [], {}, {}
            for f, st in enumerate(seq):
                tr = "event" if f == pre else ("revert" if f == pre + hold else None)
                fs = c.frame(st, want_err=pre <= f < pre + 5, timed=tr is not None)
                if tr:
                    ev_ms[tr] = round(float(fs.ms_event), 4)   # GPU time of the event block
                    inval[tr] = int(fs.invalidated)
                if pre <= f < pre + 5:
                    errs.append(round(float(fs.err), 4))
                if f >= pre:
                    c.render(img, cam, seed=5000 + f)
                    biases.append(float(img[reg].double().sum()) / ref_sum[st] - 1.0)
            post, rev = biases[:hold], biases[hold:]
            ev["methods"][mname] = {
                "frames_to_5pct": sc.frames_to_below(post, 0.05),
                "bias_first_10": [round(x, 4) for x in post[:10]],
                "bias_mean_30": round(float(np.mean(np.abs(post[:30]))), 4),
                "bias_last_10": round(float(np.mean(np.abs(post[-10:]))), 4),
                "revert_frames_to_5pct": sc.frames_to_below(rev, 0.05),
                "revert_bias_first_10": [round(x, 4) for x in rev[:10]],
                "entry_relrms_first_5": errs,
                "entries_invalidated": inval,
                "event_block_ms": ev_ms,
            }
        out[name] = ev
        log(f"{name}: {json.dumps(ev)}")
    return out


def fp64_results(rea, osc, scene, cam, log, frames=3, seed=1234):
    """Relative RMS of the FP32 results against their FP64 twins (the oracle replaying the same rays and
    hits with the radiometry in double), after `frames` plain-cache frames from the same start: cache
    channels (every entry and light), resolved radiance (every entry), camera image (8 row bands of 8
    rows). The FP32 cache arrays are first checked equal to the oracle's bit for bit."""
    import numpy as np
    import torch
    import pyoracle as po

    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    cache.begin(0, seed)
    orun = po.OracleRunner(osc, m)
    orun.begin(0, seed)
    twin = po.FP64Twin(osc, orun, 0)
    for _ in range(frames):
        cache.frame(0)
        twin.frame()
    lc = cache.get("LC")
    same = bool(np.array_equal(lc.view(np.uint32), orun.get("LC").view(np.uint32)))
    W, H = cam.W, cam.H
    img = torch.zeros(W * H, dtype=torch.float32, device="cuda")
    cache.render(img, cam, seed=seed + 1)
    got = img.cpu().numpy()
    a, b = [], []
    for k in range(8):
        r0 = k * H // 8
        a.append(got[r0 * W:(r0 + 8) * W])
        b.append(twin.image64(cam, seed + 1, frames, r0 * W, (r0 + 8) * W)[r0 * W:(r0 + 8) * W])
    out = {"cache_channels": po.rel_rms(lc, twin.L64), "resolved_radiance": po.rel_rms(cache.get("R"), twin.R64()),
           "camera_image": po.rel_rms(np.concatenate(a), np.concatenate(b)), "tolerance": 1e-4,
           "fp32_bit_exact_with_oracle": same,
           "sample": f"{frames} plain-cache frames; channels and resolved radiance of all {scene.info.Np} entries; image rows "
                     f"in 8 bands of 8"}
    log(f"fp64 twins: {json.dumps(out)}")
    return out


def pmc_profile(tag):
    """the committed rocprofv3 counter summary of this workload (profiles/pmc_<tag>.json), or None"""
    path = os.path.join(ROOT, "profiles", f"pmc_{tag}.json")
    return json.load(open(path)) if os.path.exists(path) else None


def kernel_counters(prof, prefix):
    if not prof:
        return None
    for kname, v in prof["kernels"].items():
        if kname.split("<")[0] == prefix:
            return v
    return None


def pmc_traffic(tag, kname, grid):
    """HBM bytes per launch (copy-calibrated FETCH_SIZE + WRITE_SIZE) of kernel kname's dispatches of grid
    threads in the committed counter summary of this workload, or None"""
    prof = pmc_profile(tag)
    e = ((prof or {}).get("kernels_by_grid", {}).get(kname) or {}).get(str(grid))
    return e.get("traffic_bytes") if e else None


# G wave-instructions/s: the published FP32 vector rate (157.3 TFLOP/s = 256 CUs x 128 FMA lanes per cycle x
# 2.4 GHz) is two wave64 VALU instructions per CU per cycle; render_engine_amd/valu_probe (independent FP32
# multiply / add / max chains) issues 1125 G/s of them on the part (profiles/valu_issue_ceiling.md)
VALU_PEAK = 256 * 2 * 2.4
VALU_BASIS = ("peak = 256 CUs x 2 VALU wave-instructions per cycle x 2.4 GHz (the published FP32 vector rate, 128 FMA "
              "lanes per CU and cycle; the VALU probe reaches 0.92 of it)")


def bench_events(args, rea, scene, log):
    """The event kernels' workload (for rocprofv3): each event method runs 21 frames in the base state
    (a snapshot slot), then every event and its revert with 2 frames each, the event blocks timed with
    HIP events; then the DDA bit-volume marking of 2^20 random segments into a fine volume."""
    import numpy as np
    import torch

    info = scene.info
    ms = rea.methods
    tl = dict(ms.timing_list())
    methods = (("flush_snapshots", ms.flush_snapshots()), ("flush_snapshots_compact", ms.flush_snapshots_compact()),
               ("targeted", ms.targeted()), ("targeted_grad", ms.targeted_grad()), ("verify (v3 complet)", tl["v3 complet"]))
    seq = [0] * 21
    for _, a, b, _k in event_list(scene):
        seq += [b] * 2 + [a] * 2
    res = {}
    for name, m in methods:
        c = rea.Cache(scene, m)
        c.begin(0, 9)
        times = []
        for f, st in enumerate(seq):
            ev = f > 0 and seq[f - 1] != st
            fs = c.frame(st, timed=ev)
            if ev:
                times.append({"from": seq[f - 1], "to": st, "ms": round(float(fs.ms_event), 4), "invalidated": int(fs.invalidated)})
        c.sync()
        res[name] = times
        log(f"events {name}: {json.dumps(times)}")
    g = info.fine if info.fine > 0 else 128
    n = 1 << 20
    rng = np.random.default_rng(3)
    lo, ext = np.array(list(info.lo)), np.array(list(info.ext))
    A = torch.from_numpy((lo + ext * rng.random((n, 3))).astype(np.float32)).cuda()
    B = torch.from_numpy((lo + ext * rng.random((n, 3))).astype(np.float32)).cuda()
    wpv = (g ** 3 + 31) // 32
    vol = torch.zeros(wpv, dtype=torch.int32, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dda = []
    for _ in range(3):
        vol.zero_()
        e0.record()
        rea.dda_volume(scene, (g, g, g), A, B, None, vol, wpv)
        e1.record()
        e1.synchronize()
        dda.append(e0.elapsed_time(e1))
    out = {"metric": "event block GPU ms per method and transition", "workload": f"street {info.nst} tris, {info.Np} entries, "
           f"{info.nl} lights", "events": res, "dda_volume": {"segments": n, "grid": g, "ms": round(float(np.median(dda)), 4)}}
    print(json.dumps(out))


def summarize_pmc(dirs, tag, workload):
    """rocprofv3 --pmc CSV outputs -> profiles/pmc_<tag>.json: per kernel the mean per dispatch of every
    counter, FETCH_SIZE / WRITE_SIZE scaled by what the same counter reads for a 256 MiB device copy
    (the directories named cal_FETCH_SIZE / cal_WRITE_SIZE), and the HBM bytes per dispatch"""
    import csv
    import glob

    import numpy as np

    per = {}
    cal = {}
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows = list(csv.DictReader(open(f)))
            for r in rows:
                k = r.get("Kernel_Name", "")
                ctr, val = r.get("Counter_Name"), float(r.get("Counter_Value", 0) or 0)
                key = (os.path.basename(d.rstrip("/")), k, ctr, r.get("Dispatch_Id"), r.get("Grid_Size"))
                per.setdefault(key, 0.0)
                per[key] += val   # counters of one dispatch may come in several rows (per dimension)
    kernels, by_grid = {}, {}
    for (d, k, ctr, disp, grid), v in per.items():
        name = k.split("(")[0].replace("void ", "").strip()
        if d.startswith("cal_"):
            cal.setdefault(ctr, []).append(v)
            continue
        kernels.setdefault(name, {}).setdefault(ctr, []).append(v)
        by_grid.setdefault(name, {}).setdefault(str(int(float(grid or 0))), {}).setdefault(ctr, []).append(v)
    copy_kib = 256 * 1024
    factor = {c: copy_kib / max(v) for c, v in cal.items()}   # the copy kernel: the largest dispatch of the calibration run
    out = {"workload": workload, "copy_bytes": copy_kib * 1024.0, "fetch_factor": factor.get("FETCH_SIZE"),
           "write_factor": factor.get("WRITE_SIZE"), "kernels": {}, "kernels_by_grid": {}}

    def entry(ctrs):
        e = {c: float(np.mean(v)) for c, v in ctrs.items()}
        e["dispatches"] = max(len(v) for v in ctrs.values())
        if "FETCH_SIZE" in e and "WRITE_SIZE" in e and factor.get("FETCH_SIZE") and factor.get("WRITE_SIZE"):
            e["fetch_bytes"] = e["FETCH_SIZE"] * 1024 * factor["FETCH_SIZE"]
            e["write_bytes"] = e["WRITE_SIZE"] * 1024 * factor["WRITE_SIZE"]
            e["traffic_bytes"] = e["fetch_bytes"] + e["write_bytes"]
        return e
    for name, ctrs in kernels.items():
        out["kernels"][name] = entry(ctrs)
        if len(by_grid.get(name, {})) > 1:   # a kernel launched at several sizes: per dispatch size too
            out["kernels_by_grid"][name] = {g: entry(c) for g, c in by_grid[name].items()}
    path = os.path.join(ROOT, "profiles", f"pmc_{tag}.json")
    json.dump(out, open(path, "w"), indent=1, sort_keys=True)
    print(path)


def adjoint_cpu(scene, adj, a, seconds, rays_step, log):
    """CPU leg of the adjoint workload. (1) Algorithmic bytes of one k_adj_fwd launch, counted by
    the oracle on the same rays (every bounce, 1/8 of the entries in 8 bands, scaled): per entry
    36 B, per ray 32 B per binary-BVH node visited + 36 B per triangle tested + 8 B per hash
    probe, per hit 16 B (28 B when the previous bounce's L and S are read); the mean over the
    bounces. (2) The oracle's step on the host cores: every bounce's rays traced once (the hits
    kept for the reverse pass), the FP32 forward and the FP64 reverse pass, with the unit direct
    light copied from the GPU (computed once per handle, outside the step on both sides)."""
    import numpy as np

    from render_engine_amd import adjoint as ad

    osc = oracle_scene(scene)   # puts the oracle on the path
    import pyoracle as po
    info = scene.info
    Np, nl, K, B, seed = info.Np, info.nl, a["K"], a["B"], a["seed"]
    st = osc.state(0)
    band = max(1, Np // 64)
    per_launch, work = [], []
    for b in range(B):
        po.lib.orc_count_enable(1)
        sampled, nhit = 0, 0
        for j in range(8):
            i0 = j * Np // 8
            i1 = min(Np, i0 + band)
            h = np.empty((i1 - i0) * K, np.int32)
            po.lib.orc_adj_hits_rows(st, 1, K, b, seed, i0, i1, po.ptr(h))
            nhit += int((h >= 0).sum())
            sampled += i1 - i0
        cnt = np.zeros(4, np.int64)
        po.lib.orc_count_get(po.ptr(cnt))
        po.lib.orc_count_enable(0)
        sc = Np / sampled
        per_launch.append(36.0 * Np + (32 * cnt[0] + 36 * cnt[1] + 8 * cnt[2]) * sc + nhit * sc * (16 if b == 0 else 28))
        work.append({"nodes": round(cnt[0] * sc), "tris": round(cnt[1] * sc), "probes": round(cnt[2] * sc), "hits": round(nhit * sc)})
    fb = float(np.mean(per_launch))
    log(f"adjoint forward work per bounce (oracle count, 1/{Np / sampled:.0f} of entries): {json.dumps(work)}")
    # the oracle's step
    mat = np.ascontiguousarray(scene.array("MAT"), np.int32)
    G = adj.get(ad.ADJ_G)
    SG = adj.get(ad.ADJ_SG)
    G64 = G.astype(np.float64)
    I32 = np.ascontiguousarray(scene.array("INTEN", 0), np.float32)
    alb32 = np.ascontiguousarray(scene.array("MATALB"), np.float32)
    I64, alb64 = I32.astype(np.float64), alb32.astype(np.float64)
    g64 = np.ones(Np)
    hits = np.zeros(B * Np * K, np.int32)
    R, SR = np.zeros(Np, np.float32), np.zeros(Np, np.uint64)
    R64, dI, dalb = np.zeros(Np), np.zeros(nl), np.zeros(info.nmat)
    steps, t_tot = 0, 0.0
    while steps < 1 or (t_tot < seconds and steps < 20):
        t0 = time.perf_counter()
        po.lib.orc_adj_hits(st, 1, K, B, seed, po.ptr(hits))
        po.lib.orc_adj_forward(Np, nl, K, B, po.ptr(hits), po.ptr(mat), po.ptr(alb32), po.ptr(I32), po.ptr(G), po.ptr(SG),
                               po.ptr(R), po.ptr(SR), None)
        po.lib.orc_adj_grad64(Np, nl, info.nmat, K, B, po.ptr(hits), po.ptr(mat), po.ptr(alb64), po.ptr(I64), po.ptr(G64),
                              po.ptr(g64), po.ptr(R64), po.ptr(dI), po.ptr(dalb))
        t_tot += time.perf_counter() - t0
        steps += 1
    log(f"cpu baseline (adjoint): {steps} steps in {t_tot:.1f} s on {po.lib.orc_num_threads()} threads")
    cpu = dict(value=rays_step * steps / t_tot / 1e6, unit="Mrays/s", cores=int(po.lib.orc_num_threads()), kind="port",
               sample=f"{steps} full forward + reverse steps of the same workload (the oracle traces each of the "
                      f"{Np * K * B} rays once and keeps the hits; value = the step's ray count over the step time)",
               steps_per_s=steps / t_tot)
    return fb, work, cpu


def bench_adjoint(args, rea, scene, cfg, log):
    """Config 4: one step = forward + backward of the cached-radiance adjoint (B bounces, K rays
    per entry and bounce; the backward re-traces them) for J = sum of all cached radiance, giving
    the gradient w.r.t. every light's emission and every material's albedo. Reports rays/s, the
    central-difference check (h = 1e-3) of the largest components, the gradient-support masks,
    and the light-off / occluder events with targeted invalidation from recorded masks vs from
    the gradient support."""
    import numpy as np
    import torch

    import render_engine_amd.adjoint as ad

    info = scene.info
    a = cfg["adjoint"]
    t0 = time.time()
    adj = ad.Adjoint(scene, state=0, K=a["K"], bounces=a["B"], seed=a["seed"])
    t_unit = time.time() - t0
    I = torch.tensor(scene.array("INTEN", 0), dtype=torch.float32, device="cuda")
    alb = torch.tensor(scene.array("MATALB"), dtype=torch.float32, device="cuda")
    g = torch.ones(info.Np, dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        adj.forward(I, alb)
        adj.backward(g)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd, bwd = [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record()
        R = adj.forward(I, alb)
        ev[1].record()
        gI, galb = adj.backward(g)
        ev[2].record()
        ev[2].synchronize()
        fwd.append(ev[0].elapsed_time(ev[1]))
        bwd.append(ev[1].elapsed_time(ev[2]))
    elapsed = time.perf_counter() - t_start
    rays_step = 2 * info.Np * a["K"] * a["B"]   # forward rays + the backward's re-traced rays (every entry's adjoint is nonzero)
    gI_h, galb_h = gI.cpu().numpy(), galb.cpu().numpy()
    h = 1e-3

    def J(Iv, Av):
        return float(adj.forward(Iv, Av).double().sum())

    fd = {}
    for l in np.argsort(-np.abs(gI_h))[:2]:
        e = torch.zeros_like(I); e[l] = h
        d = (J(I + e, alb) - J(I - e, alb)) / (2 * h)
        fd[f"emission[{l}]"] = {"adjoint": float(gI_h[l]), "central_difference": d, "rel_err": float(abs(d - gI_h[l]) / abs(gI_h[l]))}
    for m in np.argsort(-np.abs(galb_h))[:2]:
        e = torch.zeros_like(alb); e[m] = h
        d = (J(I, alb + e) - J(I, alb - e)) / (2 * h)
        fd[f"albedo[{m}]"] = {"adjoint": float(galb_h[m]), "central_difference": d, "rel_err": float(abs(d - galb_h[m]) / abs(galb_h[m]))}
    adj.forward(I, alb)
    sr = adj.support()
    bits = np.unpackbits(sr.view(np.uint8)).reshape(-1, 64).sum(1)
    support = {"mean_lights_per_entry": round(float(bits.mean()), 2),
               "entries_depending_on_lamp0": int(((sr & np.uint64(1)) != 0).sum()),
               "entries_depending_on_sun": int(((sr >> np.uint64(min(info.nl - 1, 63))) & np.uint64(1)).sum())}
    log(f"adjoint: fwd {np.mean(fwd):.3f} ms, bwd {np.mean(bwd):.3f} ms, FD {json.dumps(fd)}")
    # per-launch time of the traced bounces (k_adj_fwd, k_adj_bwd): HIP events on the stream they run on
    adj.set_timed(True)
    tf, tb = [], []
    for _ in range(max(3, args.steps // 4)):
        adj.forward(I, alb)
        adj.backward(g)
        f_ms, b_ms = adj.trace_ms()
        tf.append(f_ms / a["B"]); tb.append(b_ms / a["B"])
    adj.set_timed(False)
    k_fwd, k_bwd = float(np.mean(tf)), float(np.mean(tb))
    cpu, fb = None, None
    if not args.no_cpu:
        try:
            fb, fwork, cpu = adjoint_cpu(scene, adj, a, args.cpu_seconds, rays_step, log)
        except Exception as e:  # the baseline is reported, not required
            log(f"cpu leg failed: {e!r}")
    # roofline of the forward bounce: VALU issue (rocprofv3 SQ_INSTS_VALU per launch of this workload over the
    # in-run launch time), the counter HBM traffic and the oracle-counted touched bytes beside it
    prof = pmc_profile(f"street{scene.camera().W}x{scene.camera().H}_adjoint")
    kc = kernel_counters(prof, "k_adj_fwd")
    valu = kc.get("SQ_INSTS_VALU") if kc else None
    traffic = kc.get("traffic_bytes") if kc else None
    ach = valu / (k_fwd * 1e-3) / 1e9 if valu else None
    roof = {"bound": "valu", "achieved": ach, "peak": VALU_PEAK, "unit": "G wave-instructions/s",
            "frac": ach / VALU_PEAK if ach else None, "traffic": traffic, "kernel": "k_adj_fwd",
            "valu_per_launch": valu, "hbm_GB_s": traffic / k_fwd / 1e6 if traffic else None,
            "hbm_frac": traffic / k_fwd / 1e6 / 8000.0 if traffic else None,
            "touched_bytes": fb, "touched_frac_of_hbm_peak": fb / k_fwd / 1e6 / 8000.0 if fb else None,
            "counters": f"profiles/pmc_street{scene.camera().W}x{scene.camera().H}_adjoint.json" if prof else None,
            "basis": "SQ_INSTS_VALU per launch (rocprofv3, this workload) / in-run launch time; " + VALU_BASIS + ". traffic = "
                     "FETCH_SIZE + WRITE_SIZE per launch (copy-calibrated); "
                     "touched bytes = per entry 36 B (P, N, L and S out) + per ray 32 B/binary node + 36 B/triangle + 8 B/hash "
                     "probe + per hit 16 B (A, E, SG; 28 B with the previous bounce's L, S), oracle-counted, mean over the bounces",
            "work": fwork if fb else None}
    qual = None
    if not args.no_quality:
        try:
            qual = quality(rea, scene, scene.camera(), log, methods=(("plain", rea.methods.plain()), ("targeted", rea.methods.targeted()),
                                                                    ("targeted_grad", rea.methods.targeted_grad())))
        except Exception as e:
            log(f"quality phase failed: {e!r}")
    ms = float(np.mean(fwd) + np.mean(bwd))
    out = {
        "metric": "Mrays/s + cache Mpts/s; frames to <5% error after light-off, 1 GPU",
        "value": round(rays_step / (ms * 1e3), 2),
        "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": f"adjoint of all cached radiance: street {info.nst} tris, {info.Np} entries, {info.nl} emission "
                               f"scalars, {info.nmat} material albedos, {a['B']} bounces x {a['K']} rays per entry",
                   "config_index": args.config},
        "kernels": {"forward_ms": float(np.mean(fwd)), "backward_ms": float(np.mean(bwd)), "unit_direct_light_s": round(t_unit, 3),
                    "rays_per_step": rays_step,
                    "k_adj_fwd": {"ms": k_fwd, "rays": info.Np * a["K"], "mrays_s": info.Np * a["K"] / k_fwd / 1e3,
                                  "alg_bytes": fb, "GB_s": (fb / k_fwd / 1e6) if fb else None},
                    "k_adj_bwd": {"ms": k_bwd, "rays": info.Np * a["K"], "mrays_s": info.Np * a["K"] / k_bwd / 1e3}},
        "gradient_check": fd,
        "gradient_support": support,
        "roofline": roof,
        "cpu_baseline": cpu,
        "light_off": qual,
    }
    print(json.dumps(out))


def bench_mlp(args, rea, scene, cfg, log):
    """Config 5: one step = one frame of the plain grid cache + one online training step of the
    hash-encoded MLP cache on 2^16 self-training samples + the camera pass at 1920x1080 ending on
    the MLP (batched BF16 MFMA inference at the cache bounce). Reports frame throughput, inference
    TFLOP/s against the BF16 MFMA peak, the error of the MLP judged against the grid cache, and the
    light-off events with the grid cache as judge and fallback: a grid cache with targeted
    invalidation runs beside the MLP, and the frame's image comes from it while the MLP's error
    judged against it is above 25 %."""
    import numpy as np
    import torch

    import render_engine_amd.mlp as M
    from render_engine_amd import scenarios as sc

    info = scene.info
    mc = cfg["mlp"]
    mlp = M.MLPCache(M.config(list(info.lo), list(info.ext), log2_table=mc["log2_table"], base_res=mc["base_res"],
                              max_res=mc["max_res"], max_batch=mc["batch"], lr=mc["lr"]))
    cache = rea.Cache(scene, rea.methods.plain())
    cache.set_render_mode(rea.RENDER_STAGED if args.render == "staged" else rea.RENDER_FUSED)
    qcam = scene.camera()
    qcam.W, qcam.H, qcam.spp, qcam.cache_bounce = mc["W"], mc["H"], 1, 1
    qimg = torch.zeros(qcam.W * qcam.H, dtype=torch.float32, device="cuda")
    B = mc["batch"]
    # warm-up: the grid cache and the MLP train together from a cold MLP
    cache.begin(0, 11)
    judges = []
    for f in range(args.warmup + 200):
        cache.frame(0)
        _, j = mlp.train_scene(cache, n=B, seed=7 + f)
        judges.append(j)
    log(f"judge after warm-up: {judges[-1]:.3f} (first {judges[0]:.3f})")
    # timed frames
    tr, rd, fr, inf = [], [], [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for f in range(args.steps):
        t0 = time.perf_counter()
        cache.frame(0)
        cache.sync()
        t1 = time.perf_counter()
        _, j = mlp.train_scene(cache, n=B, seed=1000 + f)
        t2 = time.perf_counter()
        ms = mlp.render(cache, qimg, qcam, seed=2000 + f, timed=True)
        fr.append((t1 - t0) * 1e3); tr.append((t2 - t1) * 1e3); rd.append(ms); inf.append(mlp.last_infer_ms())
        judges.append(j)
    elapsed = time.perf_counter() - t_start
    # inference throughput of the MFMA network on one full batch (positions of random cache entries)
    P = torch.from_numpy(np.concatenate([scene.array("P", 0).reshape(-1, 3)[:B], np.zeros((B, 1), np.float32)], 1)).cuda()
    Nn = torch.from_numpy(np.concatenate([scene.array("N", 0).reshape(-1, 3)[:B], np.zeros((B, 1), np.float32)], 1)).cuda()
    y = torch.empty(B, dtype=torch.float32, device="cuda")
    for _ in range(3):
        mlp.infer(P, Nn, Nn, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        mlp.infer(P, Nn, Nn, y)
    e1.record()
    e1.synchronize()
    ms_inf = e0.elapsed_time(e1) / reps
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        cpu = mlp_cpu_baseline(mlp, mlp.cfg, P.cpu().numpy(), Nn.cpu().numpy(), Nn.cpu().numpy(), min(args.cpu_seconds, 8.0), log)
    flop_per_query = 2 * (M.HIDDEN * 64 * 64 + 64 * 16)
    tflops = B * flop_per_query / (ms_inf * 1e-3) / 1e12
    nq = qcam.W * qcam.H * qcam.spp
    ms_inf_r = float(np.mean(inf))
    tflops_r = nq * flop_per_query / (ms_inf_r * 1e-3) / 1e12
    log(f"frame {np.mean(fr):.3f} ms, train {np.mean(tr):.3f} ms, render {np.mean(rd):.3f} ms (its inference {ms_inf_r:.4f} ms, "
        f"{tflops_r:.1f} TFLOP/s), infer {B} {ms_inf:.4f} ms ({tflops:.1f} TFLOP/s), judge {np.mean(judges[-args.steps:]):.3f}")
    qual = None
    if not args.no_quality:
        try:
            cam = scene.camera()
            refs = reference_images(rea, scene, cam, log)
            img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
            qual = {}
            for name, (a, b) in {"lamp 0 off": (0, 2), "night (sun off)": (0, 1)}.items():
                region = sc.light_off_region(refs[a], refs[b])
                c = rea.Cache(scene, rea.methods.targeted())
                c.set_render_mode(rea.RENDER_STAGED if args.render == "staged" else rea.RENDER_FUSED)
                c.begin(a, 777)
                ml = M.MLPCache(M.config(list(info.lo), list(info.ext), log2_table=mc["log2_table"], base_res=mc["base_res"],
                                         max_res=mc["max_res"], max_batch=mc["batch"], lr=mc["lr"]))
                bm, bg, bo, bq, jj, fb = [], [], [], [], [], 0
                for f in range(200 + 60):
                    st = a if f < 200 else b
                    c.frame(st)
                    _, j = ml.train_scene(c, n=B, seed=5000 + f)
                    if f >= 200:
                        ml.render(c, img, cam, seed=9000 + f)
                        em = img.cpu().numpy().reshape(cam.H, cam.W)
                        ml.set_grid_fallback(0.5)   # the grid judges every query: off by > 50 % -> the grid's value
                        ml.render(c, img, cam, seed=9000 + f)
                        ml.set_grid_fallback(0.0)
                        eq = img.cpu().numpy().reshape(cam.H, cam.W)
                        c.render(img, cam, seed=9000 + f)
                        eg = img.cpu().numpy().reshape(cam.H, cam.W)
                        bm.append(sc.region_bias(em, refs[b], region))
                        bq.append(sc.region_bias(eq, refs[b], region))
                        bg.append(sc.region_bias(eg, refs[b], region))
                        use_grid = j > 0.25
                        fb += int(use_grid)
                        bo.append(bg[-1] if use_grid else bm[-1])
                        jj.append(round(j, 4))
                qual[name] = {"region_pixels": int(region.sum()),
                              "mlp": {"frames_to_5pct": sc.frames_to_below(bm, 0.05), "bias_first_10": [round(x, 4) for x in bm[:10]]},
                              "grid_targeted": {"frames_to_5pct": sc.frames_to_below(bg, 0.05)},
                              "mlp_with_grid_fallback": {"frames_to_5pct": sc.frames_to_below(bo, 0.05), "fallback_frames": fb,
                                                         "judge": "global relative RMS of the training samples > 0.25: whole frame"},
                              "mlp_with_per_query_grid_judge": {"frames_to_5pct": sc.frames_to_below(bq, 0.05),
                                                                "bias_first_10": [round(x, 4) for x in bq[:10]],
                                                                "judge": "per query: prediction off the grid's value by > 50 %"},
                              "judge_first_10": jj[:10]}
                log(f"{name}: {json.dumps(qual[name])}")
        except Exception as e:
            log(f"quality phase failed: {e!r}")
    step_ms = float(np.mean(fr) + np.mean(tr) + np.mean(rd))
    out = {
        "metric": "Mrays/s + cache Mpts/s; frames to <5% error after light-off, 1 GPU",
        "value": round(qcam.W * qcam.H / (np.mean(rd) * 1e3), 2),
        "unit": "Mqueries/s (1920x1080 camera pass ending on the learned cache)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": f"hash-encoded MLP cache on the street ({info.Np} entries): 16 levels x 2 features, 2^{mc['log2_table']} "
                               f"table, 5 x 64 ReLU, {B} training samples per frame, queried at {qcam.W}x{qcam.H}",
                   "config_index": args.config},
        "kernels": {"grid_frame_ms": float(np.mean(fr)), "mlp_train_ms": float(np.mean(tr)), "mlp_render_1080p_ms": float(np.mean(rd)),
                    "mlp_render_inference_ms": ms_inf_r, "mlp_infer_65536_ms": ms_inf, "mlp_infer_65536_tflops": tflops, "step_ms": step_ms},
        "roofline": {"bound": "mfma", "achieved": tflops_r, "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops_r / 2500.0,
                     "traffic": pmc_traffic(f"street{cfg['W']}x{cfg['H']}_mlp", "k_mlp_fwd<false>", nq),
                     "kernel": f"k_mlp_fwd (the batched inference of the {qcam.W}x{qcam.H} pass's {nq} recorded queries)",
                     "basis": f"{flop_per_query} FLOP per query (5 x 64x64 + 64x16 multiply-adds) / the inference's time inside the "
                              f"render (HIP events on its stream, rc_mlp_last_infer_ms); traffic = FETCH_SIZE + WRITE_SIZE per launch "
                              f"of that dispatch size (copy-calibrated)"},
        "judge_relrms": {"after_warmup": round(float(judges[args.warmup + 199]), 4), "timed_mean": round(float(np.mean(judges[-args.steps:])), 4)},
        "cpu_baseline": cpu,
        "gpu_inference_over_cpu": round(B / (ms_inf * 1e-3) / 1e6 / cpu["value"], 1) if cpu else None,
        "light_off": qual,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=3, choices=sorted(CONFIGS))
    ap.add_argument("--trav", type=int, default=2, choices=[0, 1, 2, 3],
                    help="0 binary (prototype order), 1 binary, 2 4-wide fp32 nodes, 3 camera pass on 4-wide quantized nodes")
    ap.add_argument("--render", choices=["staged", "fused"], default="fused",
                    help="camera-pass schedule (same image): staged ray lists or one thread per path")
    ap.add_argument("--stack-levels", type=int, default=0,
                    help="LDS levels of the camera pass's traversal stacks (0: the library's choice)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--events", action="store_true", help="the event kernels' workload (for profiling) instead of frames")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 --pmc CSV directories of this config's frame workload into profiles/pmc_<tag>.json")
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the path would shard the camera "
              f"pass by screen tiles and the cache entries by Morton ranges; this measures one GPU's share)", file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    tag = (f"{cfg['kind']}{cfg['W']}x{cfg['H']}" if cfg["kind"] == "street" else "room") + \
        ("_adjoint" if "adjoint" in cfg else "") + ("_mlp" if "mlp" in cfg else "")
    if args.pmc_summary:
        summarize_pmc(args.pmc_summary, tag + ("_events" if args.events else ""), f"bench.py --config {args.config}"
                      + (" --events" if args.events else ""))
        return

    def log(msg):
        if not args.quiet:
            print(f"[bench] {msg}", file=sys.stderr, flush=True)

    import numpy as np
    import torch

    import render_engine_amd as rea

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    rea.prep_device(True)   # BVH build and cache cells on the GPU (the host path's arrays, bit for bit)
    t0 = time.time()
    scene = make_scene(rea, cfg)
    prep_s = time.time() - t0
    info = scene.info
    log(f"scene: {info.nst} triangles, {info.nnodes} binary / {info.nwnodes} wide nodes, {info.Np} cache entries, "
        f"table {info.hsize}, {info.nl} lights, prep {prep_s:.2f} s")
    t0 = time.time()
    scene.upload(trav=args.trav)
    log(f"upload + per-state build {time.time() - t0:.1f} s")
    if args.events:
        bench_events(args, rea, scene, log)
        return
    if "adjoint" in cfg:
        bench_adjoint(args, rea, scene, cfg, log)
        return
    if "mlp" in cfg:
        bench_mlp(args, rea, scene, cfg, log)
        return
    m = rea.methods.plain()
    cache = rea.Cache(scene, m)
    cache.set_render_mode(rea.RENDER_STAGED if args.render == "staged" else rea.RENDER_FUSED)
    if args.stack_levels:
        cache.set_stack_levels(args.stack_levels)
    cam = scene.camera()
    img = torch.zeros(cam.W * cam.H, dtype=torch.float32, device="cuda")
    seed = 4242

    def one_frame(f):
        cache.frame(0)
        cache.render(img, cam, seed=seed + f)

    # 1) ray counts of exactly the frames that are timed below (same seeds -> same paths)
    cache.begin(0, seed)
    for f in range(args.warmup):
        one_frame(f)
    n_entry, n_ext, n_sh = 0, 0, 0
    for f in range(args.steps):
        cache.frame(0)
        n_entry += int(cache.get("KK").sum())
        e, s_ = cache.render(img, cam, seed=seed + args.warmup + f, count_rays=True)
        n_ext += e
        n_sh += s_
    # 2) timed region: the same frames again from the same start
    cache.begin(0, seed)
    for f in range(args.warmup):
        one_frame(f)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for f in range(args.steps):
        one_frame(args.warmup + f)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    total_rays = n_entry + n_ext + n_sh
    mrays = total_rays / elapsed / 1e6
    # 3) per-kernel GPU time with HIP events on the stream each kernel runs on (same frames again)
    cache.begin(0, seed)
    for f in range(args.warmup):
        one_frame(f)
    k_trace, k_update, k_query = [], [], []
    for f in range(args.steps):
        fs = cache.frame(0, timed=True)
        ms_q = cache.render(img, cam, seed=seed + args.warmup + f, timed=True)
        k_trace.append(fs.ms_trace)
        k_update.append(fs.ms_update)
        k_query.append(ms_q)
    t_tr, t_up, t_q = (float(np.mean(x)) for x in (k_trace, k_update, k_query))
    upd_bytes = update_bytes_per_entry(info.nl) * info.Np
    kernels = {
        "k_trace2": {"ms": t_tr, "rays": n_entry / args.steps, "mrays_s": n_entry / args.steps / t_tr / 1e3},
        "k_update": {"ms": t_up, "entries": info.Np, "mpts_s": info.Np / t_up / 1e3, "alg_bytes": upd_bytes, "GB_s": upd_bytes / t_up / 1e6,
                     "hbm_frac": upd_bytes / t_up / 1e6 / 8000.0},
        "k_query": {"ms": t_q, "rays": (n_ext + n_sh) / args.steps, "mrays_s": (n_ext + n_sh) / args.steps / t_q / 1e3},
    }
    log(f"kernels: {json.dumps(kernels)}")
    # 4) the steady-state frame cost of each method (no event in the timed frames): what dependency
    # tracking and snapshots cost every frame, beside the plain cache timed above
    method_frames = {}
    for mname, mk in (("plain", rea.methods.plain), ("flush_snapshots", rea.methods.flush_snapshots),
                      ("targeted", rea.methods.targeted), ("targeted_grad", rea.methods.targeted_grad)):
        mc = rea.Cache(scene, mk())
        mc.begin(0, seed)
        for f in range(args.warmup):
            mc.frame(0)
            mc.render(img, cam, seed=seed + f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nfr = max(5, args.steps // 2)
        for f in range(nfr):
            mc.frame(0)
            mc.render(img, cam, seed=seed + args.warmup + f)
        torch.cuda.synchronize()
        ms_frame = (time.perf_counter() - t0) * 1e3 / nfr
        tr, up = [], []
        for f in range(nfr):
            fs = mc.frame(0, timed=True)
            tr.append(fs.ms_trace); up.append(fs.ms_update)
        method_frames[mname] = {"ms_per_frame": round(ms_frame, 4), "entry_pass_ms": round(float(np.mean(tr)), 4),
                                "update_ms": round(float(np.mean(up)), 4)}
        del mc
    log(f"method frames: {json.dumps(method_frames)}")
    # CPU leg (the C oracle on the host cores): the algorithmic bytes of the query pass counted on the
    # same paths (nodes / triangles / probes / resolved reads), the FP64 twins, the baseline timing
    cpu, qb, fp64 = None, None, None
    if not args.no_cpu:
        try:
            osc = oracle_scene(scene)
            qb, qwork = query_bytes(osc, cache.get("R"), cam, seed + args.warmup + args.steps - 1, args.warmup + args.steps, log)
            kernels["k_query"].update({"touched_bytes": qb, "touched_GB_s": qb / t_q / 1e6, "work": qwork})
            fp64 = fp64_results(rea, osc, scene, cam, log)
            cpu = cpu_baseline(osc, scene, args.cpu_seconds, log)
        except Exception as e:  # the baseline is reported, not required
            log(f"cpu leg failed: {e!r}")
    # roofline of the camera pass: VALU issue (rocprofv3 SQ_INSTS_VALU of this workload per launch over the
    # launch time measured above), its counter HBM traffic and the oracle-counted touched bytes beside it
    prof = pmc_profile(tag)
    kq = kernel_counters(prof, "k_query")
    valu = kq.get("SQ_INSTS_VALU") if kq else None
    traffic = kq.get("traffic_bytes") if kq else None
    ach = valu / (t_q * 1e-3) / 1e9 if valu else None
    roof = {"bound": "valu", "achieved": ach, "peak": VALU_PEAK, "unit": "G wave-instructions/s",
            "frac": ach / VALU_PEAK if ach else None, "traffic": traffic, "kernel": "k_query",
            "valu_per_launch": valu, "hbm_GB_s": traffic / t_q / 1e6 if traffic else None,
            "hbm_frac": traffic / t_q / 1e6 / 8000.0 if traffic else None,
            "touched_bytes": qb, "touched_frac_of_hbm_peak": qb / t_q / 1e6 / 8000.0 if qb else None,
            "counters": f"profiles/pmc_{tag}.json" if prof else None,
            "basis": "SQ_INSTS_VALU per launch (rocprofv3, this workload) / in-run launch time; " + VALU_BASIS + ". traffic = "
                     "FETCH_SIZE + WRITE_SIZE per launch (copy-calibrated); "
                     "touched bytes = 32 B/binary node + 36 B/triangle + 8 B/hash probe + 4 B/resolved read, oracle-counted"}
    qual = None
    if not args.no_quality:
        try:
            qual = quality(rea, scene, cam, log)
        except Exception as e:
            log(f"quality phase failed: {e!r}")
    cb = cam.cache_bounce
    out = {
        "metric": "Mrays/s + cache Mpts/s; frames to <5% error after light-off, 1 GPU",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": (f"street {cfg.get('nboxes')} boxes / {info.nst} tris, {info.nl} lights, {info.Np} cache entries, "
                                f"table 2^{int(np.log2(info.hsize))}, {info.fine}^3 light volumes, {cam.W}x{cam.H} x {cam.spp} spp, "
                                f"plain cache; camera paths end on the cache at bounce {cb} (up to {cam.bounces} bounces "
                                f"without a cache)")
                   if cfg["kind"] == "street" else f"box room {info.nst} tris, {info.Np} entries, {cam.W}x{cam.H}",
                   "config_index": args.config, "traversal": ["binary-ref", "binary", "wide4", "wide4-quantized"][args.trav],
                   "render": args.render if args.trav == 2 else "fused",
                   "stack_levels_lds": args.stack_levels or "auto"},
        "cache_mpts_s": round(info.Np * args.steps / elapsed / 1e6, 2),
        "frames_per_s": round(args.steps / elapsed, 2),
        "scene_prep_s": round(prep_s, 3),
        "rays_per_frame": {"entry": n_entry / args.steps, "query_ext": n_ext / args.steps, "query_shadow": n_sh / args.steps},
        "kernels": kernels,
        "method_frames": method_frames,
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64_relrms": fp64,
        "light_off": qual,
    }
    if cpu:
        out["gpu_over_cpu_fps"] = round((args.steps / elapsed) / cpu["fps"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

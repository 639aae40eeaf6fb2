"""Method presets: the prototype's method lists (cuda/poc_gpu.cu:1348-1358 timing,
:1404-1418 scenario runs) and the four cache methods of the SDK bench suite
(windows/run.sh:98-103,128): no cache, plain, flush + snapshots, targeted invalidation.
Field defaults follow the Method struct (cuda/poc_gpu.cu:803-805)."""
from ._ffi import Method


def make(mode=0, nmax=16.0, budget=1.0, deps=False, verify=False, oracle=False, track_every=1, rescale=False, resp_mask=0,
         nmax_resp=8.0, ddgi=False, ddgi_h=0.97, ddgi_thr=0.25, ddgi_clamp=True, memory=False, ramp=False,
         wake_verify=False, snaps=False, snap_verify=False, flush=False, targeted=False, compact=False, grad_masks=0,
         eventboost=0):
    return Method(mode, nmax, budget, int(deps), int(verify), int(oracle), track_every, int(rescale), resp_mask, nmax_resp,
                  int(ddgi), ddgi_h, ddgi_thr, int(ddgi_clamp), int(memory), int(ramp), int(wake_verify), int(snaps),
                  int(snap_verify), int(flush), int(targeted), int(compact), int(grad_masks), int(eventboost))


# ---- the four methods of the north-star interface
def plain(nmax=32.0, budget=1.0):
    """plain cache: capped-history accumulation, no event handling (SHaRC without extras)"""
    return make(mode=0, nmax=nmax, budget=budget)


def flush_snapshots(nmax=32.0, budget=1.0, eventboost=30):
    """whole-cache flush on a new state, snapshot restore on a revisited one (rtxgi_bench.patch:323-366), and the
    post-event history-cap ramp over `eventboost` frames (windows/run.sh:100: snapshots 1, eventboost 30, eventclear 1)"""
    return make(mode=0, nmax=nmax, budget=budget, flush=True, snaps=True, eventboost=eventboost)


def flush_snapshots_compact(nmax=32.0, budget=1.0, eventboost=30):
    """flush on a new state, restore from 24-byte records of the entries with samples on a revisited
    one (rtxgi_bench.patch:1289-1345, compactsnap 1), with the post-event ramp (windows/run.sh:101)"""
    return make(mode=0, nmax=nmax, budget=budget, flush=True, snaps=True, compact=True, eventboost=eventboost)


def targeted(nmax=32.0, budget=1.0, track_every=1):
    """entries whose recorded light / voxel dependencies meet the event restart (rtxgi_bench.patch:1028-1050)"""
    return make(mode=0, nmax=nmax, budget=budget, deps=True, targeted=True, track_every=track_every)


def targeted_grad(nmax=32.0, budget=1.0, rays=4):
    """targeted invalidation whose per-entry light dependencies are the nonzero support of the
    cached-radiance gradient d R_i / d I_l (adjoint forward per state, `rays` rays per entry and
    bounce, 3 bounces) instead of masks recorded by the entry pass"""
    return make(mode=0, nmax=nmax, budget=budget, deps=True, targeted=True, grad_masks=rays)


# ---- the prototype's timing list (cuda/poc_gpu.cu:1348-1358)
def timing_list():
    return [
        ("Classique 1.0x", make(0, 16.0, 1.0)),
        ("Classique 2.0x", make(0, 16.0, 2.0)),
        ("SHaRC responsive 1.0x", make(0, 32.0, 1.0, resp_mask=1, rescale=True)),
        ("DDGI 1.0x", make(0, 64.0, 1.0, ddgi=True, rescale=True)),
        ("v3 sans suivi", make(1, 64.0, 1.0, verify=True)),
        ("v3 complet", make(1, 64.0, 1.0, deps=True, verify=True, track_every=1)),
        ("v3 suivi 1 frame sur 2", make(1, 64.0, 1.0, deps=True, verify=True, track_every=2)),
        ("v3 suivi 1 frame sur 4", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4)),
        ("v4-A2C suivi 1/4", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4, memory=True, wake_verify=True, ramp=True)),
    ]


# ---- the prototype's scenario-run list (cuda/poc_gpu.cu:1404-1418); dyn = lights that change
def scenario_list(dyn, bS=1.57, bA=1.53):
    M = [
        (f"SHaRC N=64 resp=8, {bS:.2f}x", make(0, 64.0, bS, resp_mask=dyn, nmax_resp=8.0, rescale=True)),
        (f"SHaRC + memoire naive, {bS:.2f}x", make(0, 64.0, bS, resp_mask=dyn, nmax_resp=8.0, rescale=True, memory=True)),
        (f"Accumulation+parlampe N=32, {bA:.2f}x", make(0, 32.0, bA, rescale=True)),
        (f"Accum+parlampe + memoire naive N=32, {bA:.2f}x", make(0, 32.0, bA, rescale=True, memory=True)),
        (f"Accum+parlampe + memoire naive N=64, {bA:.2f}x", make(0, 64.0, bA, rescale=True, memory=True)),
        (f"SHaRC + memoire + instantanes, {bS:.2f}x", make(0, 64.0, bS, resp_mask=dyn, nmax_resp=8.0, rescale=True, memory=True, snaps=True)),
        (f"Accum + memoire + instantanes N=32, {bA:.2f}x", make(0, 32.0, bA, rescale=True, memory=True, snaps=True)),
        (f"Accum + memoire + instantanes N=64, {bA:.2f}x", make(0, 64.0, bA, rescale=True, memory=True, snaps=True)),
        ("Oracle", make(1, 64.0, 1.0, oracle=True)),
        ("v3", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4)),
        ("v4-A2C memoire verifiee+rampe", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4, memory=True, wake_verify=True, ramp=True)),
        ("v5-A instantanes verifies", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4, memory=True, wake_verify=True, ramp=True,
                                           snaps=True, snap_verify=True)),
        ("v5-A instantanes", make(1, 64.0, 1.0, deps=True, verify=True, track_every=4, memory=True, wake_verify=True, ramp=True, snaps=True)),
    ]
    return M

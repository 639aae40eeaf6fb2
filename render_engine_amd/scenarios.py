"""Event interface and light-off metric.

* Scenario lines ``name;state:frames,state:frames,...`` (cuda/poc_gpu.cu:1378-1394, e.g.
  cuda/scenes/scenarios_main.txt) expand to a per-frame state sequence; an event fires when
  the state changes between frames (:856).
* Affected region of a light-off event: pixels where the switched-off light carried more
  than half of the converged radiance, after a k x k box blur with k = 9 at 960 px width
  scaled with the width (windows/sondes.py:30-47). Region bias = sum(est)/sum(ref) - 1
  over the region (rtxgi_bench.patch:1258-1272).
* Affected region of a geometry event (an object inserted or moved): pixels whose blurred converged
  radiance changes by more than 25 % relative to the larger of the two states (the moving panel's
  probe, windows/sondes.py:41,44).
"""
import numpy as np


def parse_scenarios(text):
    out = []
    for line in text.splitlines():
        line = line.rstrip("\r\n")
        if not line or line.startswith("#"):
            continue
        name, rest = line.split(";", 1)
        parts = []
        for tok in rest.split(","):
            st, nf = tok.split(":")
            parts.append((int(st), int(nf)))
        out.append((name, parts))
    return out


def expand(parts):
    seq = []
    for st, nf in parts:
        seq += [st] * nf
    return seq


def box_blur(a, k):
    """k x k moving mean with edge padding (integral image)"""
    r = k // 2
    p = np.pad(np.asarray(a, np.float64), ((r + 1, r), (r + 1, r)), mode="edge")
    s = p.cumsum(0).cumsum(1)
    return (s[k:, k:] - s[:-k, k:] - s[k:, :-k] + s[:-k, :-k]) / (k * k)


def blur_kernel(W):
    return max(9, int(round(9 * W / 960)) | 1)


def light_off_region(ref_on, ref_off, thr=0.5):
    """pixels where (on - off) / on > thr after the blur"""
    H, W = ref_on.shape
    k = blur_kernel(W)
    b_on, b_off = box_blur(ref_on, k), box_blur(ref_off, k)
    part = np.clip(b_on - b_off, 0, None) / np.maximum(b_on, 1e-6)
    return part > thr


def panel_region(ref_a, ref_b, thr=0.25):
    """pixels where |a - b| / max(a, b) > thr after the blur (the moving panel's probe)"""
    H, W = ref_a.shape
    k = blur_kernel(W)
    a, b = box_blur(ref_a, k), box_blur(ref_b, k)
    return np.abs(a - b) / np.maximum(np.maximum(a, b), 1e-6) > thr


def region_bias(est, ref, region):
    sr = float(np.asarray(ref, np.float64)[region].sum())
    if not region.any() or sr <= 0:
        return 0.0
    return float(np.asarray(est, np.float64)[region].sum()) / sr - 1.0


def frames_to_below(errors, thr=0.05):
    """frames after the event until |error| < thr (0 = already below at the event frame)"""
    for k, e in enumerate(errors):
        if abs(e) < thr:
            return k
    return None

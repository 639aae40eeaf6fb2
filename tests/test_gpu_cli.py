"""rc_bench (the prototype's CLI modes over librc) on the GPU: JSON-lines records of the same
types as cuda/poc_gpu.cu:1317-1436, and scenario curves equal to the Python runner's."""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import render_engine_amd as rea

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "render_engine_amd", "rc_bench")


def _run(args, timeout=240):
    p = subprocess.run([BIN] + args, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr
    return [json.loads(l) for l in p.stdout.splitlines() if l.strip()]


def test_timing_records(torch_gpu):
    recs = _run(["timing", "room"])
    assert recs[0]["type"] == "scene" and recs[0]["np"] > 50000
    names = [r["name"] for r in recs if r["type"] == "timing"]
    assert names[0] == "Classique 1.0x" and names[-1] == "v4-A2C suivi 1/4" and len(names) == 9
    for r in recs[1:]:
        assert r["trace"] > 0 and r["update"] > 0 and r["event"] >= 0


def test_scenf_matches_python_runner(torch_gpu):
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write("# short\nI1 off;0:12,1:10\n")
        path = f.name
    try:
        recs = _run(["scenf", "room", "1", path])
    finally:
        os.unlink(path)
    scen = [r for r in recs if r["type"] == "scen"]
    assert len(scen) == 13 and scen[0]["last"] == 12
    # the same method, seed and scene from Python: identical errors (double sums, 1e-9)
    s = rea.Scene.box_room(1, 2000).upload()
    s.reference(K=32, warm=16, block=64, max_iter=512, target=0.03)
    inten = [s.array("INTEN", k) for k in range(s.info.NS)]
    dyn = sum(1 << l for l in range(s.info.nl) if any(inten[k][l] != inten[0][l] for k in range(1, s.info.NS)))
    name, m = rea.methods.scenario_list(dyn)[0]
    assert scen[0]["name"] == name
    errs = rea.Cache(s, m).run([0] * 12 + [1] * 10, 3000)
    assert np.allclose(scen[0]["curve"], np.round(errs, 5), atol=2e-5)


def test_prep_record_and_file(torch_gpu, tmp_path):
    """rc_bench prep: the prep record (cuda/prep_mesh.cuh:522-528 fields) and a format-3 file with the
    converged references that rc_scene_load reads back"""
    out = str(tmp_path / "room_prep.bin")
    recs = _run(["prep", "room", out])
    r = recs[-1]
    assert r["type"] == "prep"
    for k in ("triangles", "np", "cellule_m", "zones", "diag_m", "cube_m", "bvh_noeuds", "points_valides", "bruit_ref",
              "iterations_ref", "lampes", "secondes"):
        assert k in r, k
    s = rea.Scene.load(out)
    i = s.info
    assert i.has_ref == 1 and i.Np == r["np"] and len(r["bruit_ref"]) == i.NS and len(r["lampes"]) == i.nl
    valid = np.mean([s.array("VALID", k).mean() for k in range(i.NS)])
    assert abs(valid - r["points_valides"]) < 1e-3
    assert all(0 < x < 0.1 for x in r["bruit_ref"])

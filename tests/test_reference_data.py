"""The runner and the radiance model against the reference's own data (CPU, via the oracle,
which the GPU path matches bit for bit):

* scene_main.bin, scene_holdout.bin (the scene never used for tuning) and scene_stress.bin (8 lights,
  6 states) (format 2, read by rc_scene_load) hold converged per-point references from the
  prototype's reference solver: the oracle's Jacobi solve on the same scene must agree within the
  references' self-noise;
* reference/cuda_rtx4060/scen_{main,holdout,stress}.jsonl hold the published error summaries of every
  method on every scenario of scenarios_main.txt / scenarios_stress.txt
  (tests/golden/scen_*_prepost.json): the runner must reproduce them.

Needs the reference checkout (only the scene files are read; skipped where they are absent)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
import render_engine_amd as rea
from render_engine_amd import scenarios as sc

REF = "/root/reference/cuda/scenes"
MAIN = os.path.join(REF, "scene_main.bin")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scen_{}_prepost.json")
SCENARIOS = {"main": "scenarios_main.txt", "holdout": "scenarios_main.txt", "stress": "scenarios_stress.txt"}
pytestmark = pytest.mark.skipif(not os.path.exists(MAIN), reason="reference scene files not present")
_SCENES = {}


def _scene(name):
    if name not in _SCENES:
        s = rea.Scene.load(os.path.join(REF, f"scene_{name}.bin"))
        _SCENES[name] = (s, po.OracleScene(s))
    return _SCENES[name]


@pytest.fixture(scope="module")
def main_scene():
    return _scene("main")


def test_scene_file_header(main_scene):
    s, _ = main_scene
    i = s.info
    assert (i.Np, i.Nq, i.NS, i.nl, i.NR, i.has_ref) == (10728, 29, 5, 3, 714, 1)
    assert i.nst == 0 and len(s.array("TQ", 0)) == 58
    assert list(s.array("INTEN", 1)) == [0.0, 3.0, 5.0]
    # emitter voxel of light 0 turning back on (1 -> 0) folded onto the 8x4x4 transit grid
    assert len(s.array("EM", (1 * i.NS + 0) * i.nl + 0)) >= 1


def test_holdout_and_stress_headers():
    h, _ = _scene("holdout")
    st, _ = _scene("stress")
    assert (h.info.Np, h.info.NS, h.info.nl, h.info.has_ref) == (10728, 5, 3, 1)
    assert (st.info.Np, st.info.NS, st.info.nl, st.info.has_ref) == (11268, 6, 8, 1)
    assert len(st.array("TQ", 0)) == 88


@pytest.mark.parametrize("name,np_", [("43k", None), ("530k", 539987), ("1M", 1082802)])
def test_scale_scenes_load(name, np_):
    """the prototype's scale scenes (format 2, the 530k and 1M ones gzip-compressed; cuda/run_cuda.sh:131),
    whose entry counts the reference's timing records state (reference/cuda_rtx4060/timing_*.jsonl)"""
    path = os.path.join(REF, f"scene_scale_{name}.bin" + ("" if name == "43k" else ".gz"))
    s = rea.Scene.load(path)
    i = s.info
    assert (i.NS, i.nl, i.nst, len(s.array("TQ", 0))) == (5, 3, 0, 58)
    if np_:
        assert i.Np == np_
    P = s.array("P", 0).reshape(-1, 3)
    assert np.all(np.isfinite(P)) and P.min() >= -0.5 and P.max() <= 8.5


@pytest.mark.parametrize("name,state", [("main", 0), ("main", 3), ("holdout", 1), ("stress", 0), ("stress", 5)])
def test_converged_solve_matches_stored_references(name, state):
    s, osc = _scene(name)
    i = s.info
    ref, valid = s.array("REF", state), s.array("VALID", state).astype(bool)
    Pw = s.array("PW", state).reshape(-1, i.nl)
    K, warm, T = 32, 16, 48
    Lin = np.zeros(i.Np * i.nl, np.float32)
    Lout, acc, cl = np.zeros_like(Lin), np.zeros_like(Lin), np.zeros(i.Np, np.float32)
    for it in range(warm + T):
        po.lib.orc_k_ref(osc.state(state), 1, po.ptr(Lin), po.ptr(Lout), po.ptr(acc), po.ptr(cl), (7919 * it + 17) & 0xFFFFFFFF, K,
                         int(it >= warm), 0)
        Lin, Lout = Lout, Lin
    est = (acc / T).reshape(-1, i.nl)
    e = est.sum(1)
    rel = np.sqrt(np.mean((e[valid] - ref[valid]) ** 2)) / ref[valid].mean()
    assert rel < 0.05, rel                                   # stored refs carry ~3 % self-noise
    assert abs(e[valid].sum() / ref[valid].sum() - 1) < 0.01
    on = Pw[valid].sum(0) > 0                                  # lights switched off in this state carry nothing
    per_light = est[valid].sum(0)[on] / Pw[valid].sum(0)[on]
    assert np.all(np.abs(per_light - 1) < 0.015), per_light
    assert np.all(est[valid][:, ~on] == 0)


@pytest.mark.parametrize("name,scen", [("main", "R1 allumee-eteinte-rallumee"), ("main", "I1 lampe eteinte"),
                                       ("main", "I3 lampe allumee 1re fois"), ("main", "I4 boite"),
                                       ("holdout", "R2 allumee-eteinte-porte-rallumee"), ("holdout", "I2 porte"),
                                       ("stress", "G1 mur effondre"), ("stress", "L2 quatre lampes eteintes puis rallumees"),
                                       ("stress", "L3 trois lampes allumees 1re fois"),
                                       ("stress", "M1 extinction, mur effondre dans le noir, rallumage")])
def test_runner_reproduces_published_curves(name, scen):
    s, osc = _scene(name)
    i = s.info
    gold = json.load(open(GOLD.format(name)))[scen]
    parts = dict(sc.parse_scenarios(open(os.path.join(REF, SCENARIOS[name])).read()))[scen]
    seq = sc.expand(parts)
    last = len(seq) - parts[-1][1]
    NS = i.NS
    ref = [s.array("REF", k) for k in range(NS)]
    Pw = [s.array("PW", k) for k in range(NS)]
    val = [s.array("VALID", k) for k in range(NS)]
    inten = [s.array("INTEN", k) for k in range(NS)]
    dyn = sum(1 << l for l in range(i.nl) if any(inten[k][l] != inten[0][l] for k in range(1, NS)))
    checked = 0
    for name, m in rea.methods.scenario_list(dyn):
        if name not in gold:
            continue
        curve = np.zeros(len(seq))
        for sd in range(2):
            run = po.OracleRunner(osc, m, Pw=Pw, ref=ref, valid=val)
            run.begin(seq[0], 3000 + sd)
            curve += np.array([run.frame(st, True)[0] for st in seq]) / 2
        pre, post = curve[last - 20:last].mean(), curve[last:last + 30].mean()
        g = gold[name]
        assert abs(pre - g["pre"]) <= 0.05 * g["pre"] + 1e-3, (name, pre, g["pre"])
        assert abs(post - g["post"]) <= 0.05 * g["post"] + 1e-3, (name, post, g["post"])
        checked += 1
    assert checked >= 12

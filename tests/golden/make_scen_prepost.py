"""Extract the published per-scenario, per-method error summaries (pre = mean relative RMS over
the 20 frames before the last event, post = over the 30 frames after it; 10 seeds, RTX 4060) from
the reference's data files reference/cuda_rtx4060/scen_{main,holdout,stress}.jsonl into
scen_<name>_prepost.json. Reads the data files only."""
import json
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/reference/cuda_rtx4060"
for name in ("main", "holdout", "stress"):
    out = {}
    for line in open(os.path.join(SRC, f"scen_{name}.jsonl")):
        d = json.loads(line)
        if d.get("type") == "scen":
            out.setdefault(d["scen"], {})[d["name"]] = {"pre": d["pre"], "post": d["post"], "last": d["last"]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"scen_{name}_prepost.json")
    json.dump(out, open(dst, "w"), indent=1, ensure_ascii=False)
    print(f"{sum(len(v) for v in out.values())} entries -> {dst}")

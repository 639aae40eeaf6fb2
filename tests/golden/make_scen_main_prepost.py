"""Extract the published per-scenario, per-method error summaries (pre = mean relative RMS over
the 20 frames before the last event, post = over the 30 frames after it; 10 seeds, RTX 4060) from
the reference's data file reference/cuda_rtx4060/scen_main.jsonl into scen_main_prepost.json.
Reads the data file only."""
import json
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/reference/cuda_rtx4060/scen_main.jsonl"
out = {}
for line in open(SRC):
    d = json.loads(line)
    if d.get("type") == "scen":
        out.setdefault(d["scen"], {})[d["name"]] = {"pre": d["pre"], "post": d["post"], "last": d["last"]}
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scen_main_prepost.json")
json.dump(out, open(dst, "w"), indent=1, ensure_ascii=False)
print(f"{sum(len(v) for v in out.values())} entries -> {dst}")

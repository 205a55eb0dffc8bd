"""Benchmark: zone-updates per second of the CTU Sedov step on one MI355X.

A step is one Amr::coarseTimeStep (computeNewDt with the estdt reduction,
Castro::advance, post_timestep) over every box of the level, inputs resident
in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "mini-castro_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--mode", default="plain", choices=["plain", "fused"])
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import minicastro as mc
    from minicastro import configs

    cfg = configs.config(a.config)
    mc.load_table()
    mode = mc.MODE_FUSED if a.mode == "fused" else mc.MODE_PLAIN
    L = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mode)
    L.init_data(**cfg)
    L.run(a.warmup)
    L.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.run(a.steps)
    L.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    zones = int(cfg["n_cell"][0]) * int(cfg["n_cell"][1]) * int(cfg["n_cell"][2])
    ms = (t1 - t0) * 1e3 / a.steps
    out = {"metric": "zones/s at 256^3 Sedov FP64, batched vs per-box (1e-12 rel); adjoint/fwd time",
           "value": zones * a.steps / (t1 - t0), "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": cfg["name"], "n_cell": list(cfg["n_cell"]), "boxes": len(cfg["boxes"]),
                      "mode": a.mode}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

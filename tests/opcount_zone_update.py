"""FP64 work per zone-update of the reference algorithm, counted by the
op-count build of the oracle (oracle/build/liboracle_cnt.so).

One zone-update is one valid cell advanced through Castro::estTimeStep,
Castro::advance and Castro::post_timestep (Source/Castro.cpp:162-216,
Source/Castro_advance.cpp:11-100), the unit of the figure of merit of
Source/main.cpp:148-160.  The count follows the reference's per-box kernel
order, so it includes the work the reference repeats on ghost cells
(ctoprim and the traces over the grown boxes, clean_state of Sborder).

The plain oracle advances the workload `--skip` steps (multi-threaded); the
op-count build then repeats the next `--count` steps from the same state with
the same dt and must reproduce the plain oracle's state bit for bit.

    python tests/opcount_zone_update.py --config 3 --skip 2 --count 1 > summary.txt

prints a summary whose last lines are machine-readable
(`flop_per_zone_update <workload> <value>`).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "mini-castro_amd")]

import orc  # noqa: E402
from helpers import oracle_level, table_array  # noqa: E402
from minicastro import configs  # noqa: E402


def count_steps(cfg, skip, count, nthreads=8, log=None):
    """Per counted step: (dt, zones, {op: count}, identical_to_plain_oracle)."""
    D = oracle_level(cfg, nthreads=nthreads)
    for s in range(skip):
        D.coarse_step()
        if log:
            log("plain oracle: step %d done" % (s + 1))
    orc.cnt_eos_init_arrays(table_array())
    C = orc.CntLevel(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"])
    zones = int(np.prod(cfg["n_cell"]))
    out = []
    for s in range(count):
        S = D.state()
        dt = D.coarse_step()
        C.set_state(S)
        C.opcount_reset()
        C.est_time_step()
        C.advance(dt)
        C.post_timestep()
        ops = C.opcount()
        same = bool(np.array_equal(C.state(), D.state()))
        out.append((dt, zones, ops, same))
        if log:
            log("counted step %d: %s" % (skip + s + 1, ops))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--skip", type=int, default=2)
    ap.add_argument("--count", type=int, default=1)
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    cfg = configs.config(a.config)

    def log(msg):
        print("# [%s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)

    res = count_steps(cfg, a.skip, a.count, a.threads, log)
    tot = {k: sum(r[2][k] for r in res) for k in orc.OPCOUNT_FIELDS}
    zu = sum(r[1] for r in res)
    flop = tot["add"] + tot["mul"] + tot["div"] + tot["sqrt"]
    print("# FP64 operations per zone-update of the reference algorithm (per-box kernel order)")
    print("# workload %s, %d boxes; plain oracle advanced %d steps, then %d step(s) counted"
          % (cfg["name"], len(cfg["boxes"]), a.skip, a.count))
    print("# command (repo root, CPU): python tests/opcount_zone_update.py --config %d --skip %d --count %d"
          % (a.config, a.skip, a.count))
    print("# counted: + - (add), * (mul), / (div), sqrt = 1 flop each (no FMA: -ffp-contract=off);")
    print("# log/log10 are listed apart and not in the flop total")
    for dt, z, ops, same in res:
        print("# step dt %.9e  zones %d  identical to the plain oracle: %s" % (dt, z, same))
    for k in orc.OPCOUNT_FIELDS:
        print("# %-5s per zone-update %12.1f" % (k, tot[k] / zu))
    print("flop_per_zone_update %s %.1f" % (cfg["name"], flop / zu))
    print("log_per_zone_update %s %.1f" % (cfg["name"], tot["log"] / zu))
    if not all(r[3] for r in res):
        sys.exit("op-count build diverged from the plain oracle")


if __name__ == "__main__":
    main()

"""Benchmark: zone-updates per second of the CTU Sedov step on one MI355X.

A step is one Amr::coarseTimeStep of mini-Castro (time-step control with the
estdt reduction, Castro::advance, post_timestep) over every box of the
level, with the state resident in HBM.  Default workload: 256^3 Sedov tiled
by 64 seeded uneven boxes (sides in {16,...,96}), batched fused path.  The
line also carries the per-box (reference kernel order) rate on the same
workload, the roofline of the dominant stage measured with HIP events on the
stream it runs on, and the CPU oracle timed on a bounded sample.

Prints one JSON line.
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

METRIC = "zones/s at 256^3 Sedov FP64, batched vs per-box (1e-12 rel); adjoint/fwd time"
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
FP64_PEAK = 78.6e12  # MI355X FP64 vector, FLOP/s

# Algorithmic HBM bytes per launch of each fused stage: every distinct array
# element the launch reads counted once, every element it writes counted
# once (doubles).  A launch covers all directions (trace, flux1, final) or
# all six (t, d) pairs (trans1) of its stage, so an input that several
# directions or pairs read -- the primitives Q, the traced t-states shared
# by the two pairs of a face direction, the first-stage fluxes shared by the
# two pairs of a flux direction -- counts once.  n: cells per side; cell
# arrays hold n^3 elements, ring arrays (n+2)^3, face arrays (n+1)(n+2)^2
# per direction (the [-1, n] ring in the transverse directions), the
# transverse-stage fluxes (n+1) n (n+2) per pair and the final fluxes
# (n+1) n^2 per direction.
def _counts(n):
    n0, n1, n2 = n
    cell = n0 * n1 * n2
    ring = (n0 + 2) * (n1 + 2) * (n2 + 2)
    face = (n0 + 1) * (n1 + 2) * (n2 + 2)      # one direction of the face arrays (cube: all equal)
    tface = (n0 + 1) * n1 * (n2 + 2)           # one (t, d) pair of transverse-stage faces
    fface = (n0 + 1) * n1 * n2                 # one direction of final faces
    return cell, ring, face, tface, fface


def _stage_bytes(stage, n):
    cell, ring, face, tface, fface = _counts(n)
    NS, NF, NQ, NN = 19, 20, 21, 22   # state, flux, primitive, final-flux components
    doubles = {
        "prim": (20 + 20 + 20 + NQ) * cell,                        # U in; clean(U) in place, U1, Q out
        "flat": 4 * cell + 2 * ring,                               # p, u, v, w in; flatn, div out
        "trace": NQ * cell + ring + 3 * 2 * NS * ring,             # Q, flatn in; minus/plus states out
        "flux1": 3 * (2 * NS + NF) * face + 2 * cell,              # states in, F^d out; c, gamc in
        "trans1": 3 * 2 * NS * face + 3 * NF * face + cell + 6 * NF * tface,  # t-states, F^d, gamc in; F^{t|d} out
        "final": (3 * 2 * NS * face + 6 * NF * tface + ring + 20 * cell + 2 * cell
                  + 3 * NN * fface),                               # states, F^{t|d}, div, U1, c/gamc in; final flux out
        "update": 3 * NN * fface + 20 * cell + 20 * cell,          # final fluxes, U in; U out
    }[stage]
    return 8 * doubles


STAGES_WITH_BYTES = ("prim", "flat", "trace", "flux1", "trans1", "final", "update")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain
    copy), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fused_stages_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if len(w) == 4 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel and w[2] == workload:
                return float(w[3])
    return None


def reference_flop_per_zone_update(workload):
    """FP64 operations per zone-update of the reference algorithm, counted by
    the op-count build of the CPU restatement (tests/opcount_zone_update.py),
    from the committed summary, or None."""
    path = os.path.join(ROOT, "profiles", "fp64_work_per_zone_update_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if len(w) == 3 and w[0] == "flop_per_zone_update" and w[1] == workload:
                return float(w[2])
    return None


# Unit bytes of a zone-update: the 20 conserved components read and written
# once (Source/Castro_advance.cpp:11-100 reads S_old and writes S_new).
B_ALG_PER_ZONE_UPDATE = 2 * 20 * 8


def measured_fp64_ops(kernel, workload):
    """FP64 operations issued per launch of `kernel` on `workload` from the
    committed rocprofv3 SQ_INSTS_VALU_*_F64 summary, or None."""
    path = os.path.join(ROOT, "profiles", "fp64_fused_stages_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if len(w) == 4 and w[0] == "fp64_ops_per_launch" and w[1] == kernel and w[2] == workload:
                return float(w[3])
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cfg_full, n=64, steps=2):
    """The oracle (CPU restatement, OpenMP over (1024,16,16) tiles) on an n^3
    single-box sample of the same problem, in zone-updates/s: the timing build
    (-O3, AVX2/FMA) on up to 16 host threads and, for one more step from the
    same state, on one thread; and the parity build (no FMA contraction, the
    one the tests compare against) on the same threads."""
    import numpy as np
    import orc
    from helpers import table_array
    from minicastro import configs
    threads = max(1, min(16, os.cpu_count() or 1))
    cfg = dict(cfg_full)
    cfg.update(n_cell=(n, n, n), boxes=configs.grid_boxes(n, n))
    tab = np.ascontiguousarray(table_array(), dtype=np.float64)

    def level(lib, nthreads):
        lib.orc_eos_init_arrays(orc.ptr(tab))
        L = orc.Level(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], nthreads,
                      _library=lib)
        L.init_data(cfg["p_ambient"], cfg["dens_ambient"], cfg["exp_energy"], cfg["r_init"], cfg["nsub"])
        return L

    def timed(L, k):
        t0 = time.perf_counter()
        for _ in range(k):
            L.coarse_step()
        return n ** 3 * k / (time.perf_counter() - t0)

    rate = {}
    for name, lib in (("fast", orc.fast_lib()), ("parity", orc.lib())):
        L = level(lib, threads)
        L.coarse_step()  # first step (the initial-temperature Newton solves are longer); untimed
        rate[name] = timed(L, steps)
        if name == "fast":
            one = orc.Level(cfg["n_cell"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], 1,
                            _library=lib)
            one.set_state(L.state())
            rate["fast_1core"] = timed(one, 1)
            del one
        del L
    return {"value": rate["fast"], "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "value_1core": rate["fast_1core"], "value_parity_build": rate["parity"],
            "build": orc.FAST_FLAGS, "parity_build": "-O2 -ffp-contract=off",
            "host": {"cpu": _cpu_model(), "os_cpu_count": os.cpu_count()},
            "sample": "%d^3 single-box Sedov (same problem constants), %d steps after one untimed step, "
                      "oracle OpenMP over (1024,16,16) tiles; value_1core: one more step on one thread"
                      % (n, steps)}


def ensemble_bench(steps=50, n_inst=24):
    """24 independent Sedov instances (sides in {32,48,64,96}, seeded
    energies, radii, densities), every instance with its own dt, advanced
    together with one launch per stage (hipGraph).  Aggregate zone-updates/s,
    and every instance's dt sequence end point and state compared bit for bit
    with the same instance advanced alone by the per-box path."""
    import numpy as np
    import minicastro as mc
    from minicastro import configs
    insts = configs.ensemble(n_inst, seed=7)
    E = mc.Ensemble(insts)
    E.init_data()
    tuned = E.autotune(reps=2)
    E.run(1, use_graph=True)      # graph capture
    E.init_data()
    E.run(1, use_graph=True)
    E.init_data()
    t0 = time.perf_counter()
    E.run(steps, use_graph=True)
    t1 = time.perf_counter()
    zones = E.zones()
    dt_equal = state_equal = 0
    for m, d in enumerate(insts):
        L = mc.Level(d["n_cell"], d["boxes"], d["prob_lo"], d["prob_hi"], d["bc"], mc.MODE_PLAIN)
        L.init_data(**d)
        L.run(steps)
        dt, tm, st = E.clock(m)
        dt_equal += int(dt == L.dt and tm == L.time and st == L.steps == steps)
        state_equal += int(np.array_equal(E.state(m), L.state()))
        L.close()
    E.close()
    return {"workload": "%d Sedov instances, sides %s" % (n_inst, sorted({d["n_cell"][0] for d in insts})),
            "cells": zones, "steps": steps, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "zone_updates_per_s": zones * steps / (t1 - t0),
            "dt_bitwise_equal_sequential": "%d/%d" % (dt_equal, n_inst),
            "state_bitwise_equal_sequential": "%d/%d" % (state_equal, n_inst),
            "tiling": {k: list(v[0]) for k, v in tuned.items()}}


def adjoint_bench(reps=3):
    """Discrete adjoint on the 64^3 Sedov adjoint workload, 10 steps: time of
    the checkpointing forward pass and of the adjoint pass (the reversed
    stages, reading each step's intermediates as the forward pass kept them),
    the same with every step recomputed from its checkpoint instead, and the
    central-difference check of two smooth random directions (eps 1e-6 and
    1e-7)."""
    import numpy as np
    import torch
    from minicastro import configs
    from adjoint_harness import SedovAdjointProblem, smooth_direction
    cfg = configs.adjoint_problem(n=64, steps=10)
    P = SedovAdjointProblem(cfg)
    J, g_rho, g_E = P.gradient()
    U = P.U0(P.rho0, torch.tensor(P.E, dtype=torch.float64, device="cuda")).contiguous()
    out, gin = torch.empty_like(U), torch.empty_like(U)
    seed = torch.zeros_like(U)
    seed[0] = 2.0 * P.mask
    def timed():
        fwd, adj = [], []
        for r in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            P.adj.forward(U.data_ptr(), out.data_ptr(), cfg["steps"], P.dts)
            t1 = time.perf_counter()
            P.adj.backward(seed.data_ptr(), gin.data_ptr())
            t2 = time.perf_counter()
            if r:
                fwd.append(t1 - t0)
                adj.append(t2 - t1)
        return float(np.mean(fwd)) * 1e3, float(np.mean(adj)) * 1e3

    kept = P.adj.keeps_intermediates
    f, a_ = timed()
    rec = None
    if kept:
        P.adj.keep_intermediates(False)
        fr, ar = timed()
        P.adj.keep_intermediates(True)
        rec = {"forward_ms": fr, "adjoint_ms": ar, "adjoint_over_forward": ar / fr}
    rng = np.random.default_rng(7)
    errs = []
    for q in range(2):
        v = P.rho0 * torch.from_numpy(smooth_direction(rng, 64)).to("cuda")
        vE = P.E * rng.standard_normal()
        ad = float((g_rho * v).sum()) + g_E * vE
        e = []
        for eps in (1e-6, 1e-7):
            fd = (P.J(P.rho0 + eps * v, P.E + eps * vE) - P.J(P.rho0 - eps * v, P.E - eps * vE)) / (2 * eps)
            e.append(abs(fd - ad) / abs(ad))
        errs.append(e)
    return {"workload": cfg["name"], "steps": cfg["steps"], "forward_ms": f, "adjoint_ms": a_,
            "adjoint_over_forward": a_ / f, "intermediates": "kept" if kept else "recomputed",
            "recompute_mode": rec, "J": J, "dJ_dE": g_E,
            "fd_rel_err_smooth_dirs": {"eps_1e-6": [e[0] for e in errs], "eps_1e-7": [e[1] for e in errs]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3, help="per-box comparison steps (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble run")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="only time the adjoint workload and print its JSON (for profiling)")
    ap.add_argument("--tiling-json", default=None,
                    help="take the stage tiling from config.tiling of an earlier bench.py JSON line "
                         "instead of running the autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU; this build does not run across GPUs (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    if a.copy_calibration:
        nbytes = a.copy_calibration << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 5}))
        return
    import minicastro as mc
    from minicastro import configs

    if a.only_adjoint:
        mc.load_table()
        print(json.dumps(adjoint_bench()))
        return
    cfg = configs.config(a.config)
    n = cfg["n_cell"]
    zones = int(n[0]) * int(n[1]) * int(n[2])
    mc.load_table()

    # batched fused path
    L = mc.Level(n, cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    L.init_data(**cfg)
    if a.tiling_json:
        with open(a.tiling_json) as f:
            prev = json.loads([ln for ln in f if ln.startswith("{")][-1])
        for st, b in prev["config"]["tiling"].items():
            L.set_tiling(L.STAGES.index(st), *b)
        tuned = None
    else:
        tuned = L.autotune(reps=2)   # restores the state it times on
    tiling = {L.STAGES[st]: list(L.get_tiling(st)) for st in range(1, 8)}
    L.run(a.warmup, use_graph=True)
    L.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stages = L.run_timed(a.steps)
    L.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = zones * a.steps / (t1 - t0)
    # graph replay rate (dt on the device, no host work between steps)
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    L.run(a.steps, use_graph=True)
    L.sync()
    g1 = time.perf_counter()
    graph_ms = (g1 - g0) * 1e3 / a.steps
    fused_steps_done = L.steps
    L.close()

    dom = max((k for k in stages if k in STAGES_WITH_BYTES), key=lambda k: stages[k][0])
    tot_ms, nl = stages[dom]
    avg_s = tot_ms / nl / 1e3
    traffic_alg = _stage_bytes(dom, n)
    achieved = traffic_alg / avg_s
    traffic = measured_traffic("mc::K_" + dom, cfg["name"])
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": "mc::K_" + dom,
                "avg_launch_ms": tot_ms / nl, "algorithmic_bytes_per_launch": traffic_alg,
                "traffic_GB_per_s": (traffic / avg_s / 1e9) if traffic else None}

    # FP64 pipe load of the EOS-bound stages (clean_state + ctoprim, update + estdt)
    fp64 = {}
    for st in ("prim", "update"):
        ops = measured_fp64_ops("mc::K_" + st, cfg["name"])
        if ops and stages[st][1]:
            rate = ops / (stages[st][0] / stages[st][1] / 1e3)
            fp64[st] = {"bound": "fp64 valu", "achieved": rate / 1e12, "peak": FP64_PEAK / 1e12,
                        "unit": "TFLOP/s", "frac": rate / FP64_PEAK, "ops_per_launch": ops}
    roofline["fp64_stages"] = fp64 or None

    # whole step: work of the reference algorithm per zone-update at the
    # measured rate, and the FP64 operations the fused stages issue
    # (rocprof counters, summed over the stages of one step)
    w_flop = reference_flop_per_zone_update(cfg["name"])
    hw = [measured_fp64_ops("mc::K_" + st, cfg["name"]) for st in STAGES_WITH_BYTES]
    hw_step = sum(hw) if all(hw) else None
    roofline["step"] = {
        "unit_work": "zone-update",
        "reference_flop_per_zone_update": w_flop,
        "fp64_achieved_TFLOPs": (value * w_flop / 1e12) if w_flop else None,
        "fp64_frac": (value * w_flop / FP64_PEAK) if w_flop else None,
        "issued_fp64_ops_per_zone_update": (hw_step / zones) if hw_step else None,
        "issued_fp64_TFLOPs": (hw_step / (ms_step / 1e3) / 1e12) if hw_step else None,
        "issued_fp64_frac": (hw_step / (ms_step / 1e3) / FP64_PEAK) if hw_step else None,
        "bytes_per_zone_update": B_ALG_PER_ZONE_UPDATE,
        "hbm_achieved_GBs": value * B_ALG_PER_ZONE_UPDATE / 1e9,
        "hbm_frac": value * B_ALG_PER_ZONE_UPDATE / HBM_PEAK,
    }

    # per-box path (the reference's kernel order) on the same workload
    plain = None
    if a.plain_steps > 0:
        P = mc.Level(n, cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
        P.init_data(**cfg)
        P.run(1)
        P.sync()
        p0 = time.perf_counter()
        P.run(a.plain_steps)
        P.sync()
        p1 = time.perf_counter()
        plain = (p1 - p0) * 1e3 / a.plain_steps
        P.close()

    adjoint = None if a.no_adjoint else adjoint_bench()
    ensemble = None if a.no_ensemble else ensemble_bench()
    cpu = None if a.no_cpu_baseline else cpu_baseline(cfg)
    out = {"metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": cfg["name"], "n_cell": list(n), "boxes": int(len(cfg["boxes"])),
                      "path": "batched fused, one launch per stage over all boxes",
                      "graph_replay_ms_per_step": graph_ms, "steps_advanced": fused_steps_done,
                      "fom_zones_per_usec": value / 1e6,
                      "tiling": tiling, "tiling_source": "autotuner" if tuned else a.tiling_json,
                      "per_box_ms_per_step": plain,
                      "batched_vs_per_box": (plain / ms_step) if plain else None,
                      "stage_ms_per_step": {k: v[0] / a.steps for k, v in stages.items() if v[1]}},
           "roofline": roofline, "cpu_baseline": cpu, "adjoint": adjoint,
           "ensemble": ensemble}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

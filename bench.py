"""Benchmark: zone-updates per second of the CTU Sedov step on one MI355X.

A step is one Amr::coarseTimeStep of mini-Castro (time-step control with the
estdt reduction, Castro::advance, post_timestep) over every box of the
level, with the state resident in HBM.  Default workload: 256^3 Sedov tiled
by 64 seeded uneven boxes (sides in {16,...,96}), batched fused path.  The
line also carries: the whole-step FP64 roofline against the reference
algorithm's work per zone-update and the measured HBM traffic per step; the
dominant stage measured with HIP events on the stream it runs on; the
per-box (reference kernel order) rate and the rate of the drop-in batched
entry mc_advance_batch on the same workload; the discrete adjoint; the
ensemble; and the CPU oracle timed on one step of the same workload.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

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
        # trace + flux1 in one launch: the states it stores are read back by
        # the same launch, so they count once (written)
        "tflux": NQ * cell + ring + 3 * 2 * NS * ring + 3 * NF * face,  # Q, flatn in; states, F^d out
        "trans1": 3 * 2 * NS * face + 3 * NF * face + cell + 6 * NF * tface,  # t-states, F^d, gamc in; F^{t|d} out
        "trans1c": 3 * 2 * NS * face + 3 * NF * face + cell + 6 * NF * tface,  # the same, by cells
        "final": (3 * 2 * NS * face + 6 * NF * tface + ring + 20 * cell + 2 * cell
                  + 3 * NN * fface),                               # states, F^{t|d}, div, U1, c/gamc in; final flux out
        "finalc": (3 * 2 * NS * face + 6 * NF * tface + ring + 20 * cell + 2 * cell
                   + 3 * NN * fface),                              # the same, by cells
        "update": 3 * NN * fface + 20 * cell + 20 * cell,          # final fluxes, U in; U out
    }[stage]
    return 8 * doubles


STAGES_WITH_BYTES = ("prim", "flat", "trace", "flux1", "tflux", "trans1", "trans1c", "final", "finalc", "update")


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


def log(msg):
    """Progress line on stderr (the JSON line alone goes to stdout)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cfg, S, dt_prev, one_core=False):
    """The oracle (CPU restatement, OpenMP over the reference's (1024,16,16)
    tiles) on the benchmark workload itself: one coarse step of every box of
    the level (estTimeStep, computeNewDt from the previous step's dt,
    Castro::advance, post_timestep), starting from the state S the GPU run
    reached, timed with the timing build (-O3, AVX2/FMA) on up to 16 host
    threads, in zone-updates/s.
    one_core: the same step once more on one thread (about three minutes at
    256^3)."""
    import orc
    from helpers import table_array
    threads = max(1, min(16, os.cpu_count() or 1))
    n = cfg["n_cell"]
    zones = int(n[0]) * int(n[1]) * int(n[2])
    lib = orc.fast_lib()
    lib.orc_eos_init_arrays(orc.ptr(np.ascontiguousarray(table_array(), dtype=np.float64)))

    def timed_step(nthreads):
        log("cpu baseline: one step on %d thread(s)" % nthreads)
        L = orc.Level(n, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], cfg["boxes"], nthreads, _library=lib)
        L.set_state(S)
        t0 = time.perf_counter()
        dt = min(L.est_time_step(), 1.1 * dt_prev)
        L.advance(dt)
        L.post_timestep()
        t1 = time.perf_counter()
        del L
        return zones / (t1 - t0)

    out = {"value": timed_step(threads), "unit": "zone-updates/s", "cores": threads, "kind": "port",
           "build": orc.FAST_FLAGS, "host": {"cpu": _cpu_model(), "os_cpu_count": os.cpu_count()},
           "sample": "%s itself (%d^3, %d uneven boxes): one coarse step (estTimeStep, computeNewDt, "
                     "advance, post_timestep) from the state the GPU run reached; oracle OpenMP over "
                     "(1024,16,16) tiles" % (cfg["name"], n[0], len(cfg["boxes"]))}
    if one_core:
        out["value_1core"] = timed_step(1)
    else:
        rec = recorded_cpu_1core(cfg["name"])
        if rec:
            out["value_1core_recorded"] = rec
    return out


def recorded_cpu_1core(workload):
    """One-thread rate of the same step from the committed summary of a
    `bench.py --cpu-1core` run on the GPU box, or None."""
    path = os.path.join(ROOT, "profiles", "cpu_baseline_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            w = ln.split()
            if len(w) == 3 and w[0] == "cpu_1core_zone_updates_per_s" and w[1] == workload:
                return {"value": float(w[2]), "source": os.path.relpath(path, ROOT)}
    return None


def batch_bench(cfg, steps=10, warmup=2):
    """mc_advance_batch, the drop-in batched entry over caller-owned FABs, on
    the benchmark workload: per call, gather every box's S_old FAB into the
    engine, one fused step, scatter into the S_new FABs and return
    estTimeStep (a host synchronisation per call, as the reference host
    reads dt every step).  The FABs are the per-box state of a plain-mode
    level (S_old) and one device tensor per box (S_new), swapped every call.
    dt follows computeNewDt from the returned estimates."""
    import torch
    import minicastro as mc
    n = cfg["n_cell"]
    zones = int(n[0]) * int(n[1]) * int(n[2])
    P = mc.Level(n, cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
    P.init_data(**cfg)
    dt = P.initial_dt()
    boxes = [tuple(int(v) for v in b) for b in cfg["boxes"]]
    a = [int(C.cast(P.L.mc_level_state_ptr(P.h, ib), C.c_void_p).value) for ib in range(len(boxes))]
    bt = [torch.empty(20 * (b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1), dtype=torch.float64,
                      device="cuda") for b in boxes]
    b_ = [t.data_ptr() for t in bt]

    def work(src, dst):
        return [mc.work_item(b[:3], b[3:], src[ib], b[:3], b[3:], dst[ib], b[:3], b[3:])
                for ib, b in enumerate(boxes)]
    fwd, bwd = work(a, b_), work(b_, a)
    k = 0
    for _ in range(warmup):
        est = mc.advance_batch(n, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], fwd if k % 2 == 0 else bwd, dt)
        dt = min(est, 1.1 * dt)
        k += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        est = mc.advance_batch(n, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], fwd if k % 2 == 0 else bwd, dt)
        dt = min(est, 1.1 * dt)
        k += 1
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mc.batch.release()
    P.close()
    del bt
    return {"entry": "mc_advance_batch", "steps": steps, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "zone_updates_per_s": zones * steps / (t1 - t0),
            "per_call": "gather S_old FABs + one fused step + scatter S_new FABs + estTimeStep read back"}


def large_domain_bench(n=512, steps=3):
    """A domain whose whole-level face arrays would not fit one GPU (n^3
    Sedov, one instance, outflow): the engine runs the face stages in
    z-slabs chosen by free memory.  Steps are hipGraph replays."""
    import ctypes as C
    import torch
    import minicastro as mc
    from minicastro import configs
    torch.cuda.empty_cache()
    d = dict(configs.UNIT_SEDOV)
    d.update(name="sedov%d_1box" % n, n_cell=(n, n, n), boxes=configs.grid_boxes(n, n))
    E = mc.Ensemble([d])
    slab = E.slab
    E.init_data()
    E.run(1, use_graph=True)      # capture; the first step's EOS solves start from the initial guess
    t0 = time.perf_counter()
    E.run(steps, use_graph=True)  # returns after the stream has drained
    t1 = time.perf_counter()
    dt, tm, st = E.clock(0)
    E.close()
    nc = (C.c_int * 3)(n, n, n)
    used = mc.lib().mc_fused_bytes(1, nc, slab)
    whole = mc.lib().mc_fused_bytes(1, nc, 0)
    return {"workload": d["name"], "cells": n ** 3, "steps": steps, "slab_planes": slab,
            "ms_per_step": (t1 - t0) * 1e3 / steps, "zone_updates_per_s": n ** 3 * steps / (t1 - t0),
            "engine_GB": used / 1e9, "engine_bytes_per_cell": used / n ** 3,
            "whole_level_engine_GB": whole / 1e9, "steps_advanced": st}


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


def adjoint_bench(reps=3, ndir=16, n_tl=2):
    """Discrete adjoint on the 64^3 Sedov adjoint workload, 10 steps: time of
    the checkpointing forward pass and of the adjoint pass (the reversed
    stages, reading each step's intermediates as the forward pass kept
    them), the same with every step recomputed from its checkpoint instead;
    central differences at eps 1e-6 along `ndir` per-cell white-noise
    directions (the 16 random directions of the workload) and `ndir` smooth
    directions (random low-order cosine modes), each with a random E
    component; and the relative difference from the CPU tangent-linear of
    the oracle along `n_tl` of them (smooth and white noise alternating)."""
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
        P.gradient()  # the kept intermediates of the unperturbed forward pass again

    def fd_err(v, vE, eps=1e-6):
        ad = float((g_rho * v).sum()) + g_E * vE
        fd = (P.J(P.rho0 + eps * v, P.E + eps * vE) - P.J(P.rho0 - eps * v, P.E - eps * vE)) / (2 * eps)
        return abs(fd - ad) / abs(ad)

    rng = np.random.default_rng(7)
    n = cfg["n_cell"][0]
    dirs = {"white_noise": [], "smooth": []}
    for q in range(ndir):
        dirs["white_noise"].append((cfg["dens_ambient"] * rng.standard_normal((n, n, n)), P.E * rng.standard_normal()))
        dirs["smooth"].append((P.rho0.cpu().numpy() * smooth_direction(rng, n), P.E * rng.standard_normal()))
    fd = {}
    for kind, lst in dirs.items():
        e = [fd_err(torch.from_numpy(v).to("cuda"), vE) for v, vE in lst]
        fd[kind] = {"n": len(e), "below_1e-6": int(sum(x < 1e-6 for x in e)), "max": max(e),
                    "median": float(np.median(e)), "rel_err": e}
    fd["E_only"] = fd_err(torch.zeros_like(P.rho0), P.E)
    tl = None
    if n_tl > 0:
        from tl_problem import OracleSedovProblem
        O = OracleSedovProblem(cfg)
        threads = max(1, min(16, os.cpu_count() or 1))
        rel = []
        for q in range(n_tl):
            v, vE = dirs["smooth" if q % 2 == 0 else "white_noise"][q // 2]
            ad = float((g_rho * torch.from_numpy(v).to("cuda")).sum()) + g_E * vE
            _, t, _ = O.tangent(np.ascontiguousarray(v.transpose(2, 1, 0)), vE, P.dts, nthreads=threads)
            rel.append(abs(ad - t) / abs(t))
        tl = {"directions": n_tl, "max_rel_diff": max(rel), "rel_diff": rel,
              "what": "CPU tangent-linear of the oracle (forward-mode dual numbers through the restated "
                      "reference arithmetic), smooth and white-noise directions alternating"}
    return {"workload": cfg["name"], "steps": cfg["steps"], "forward_ms": f, "adjoint_ms": a_,
            "adjoint_over_forward": a_ / f, "intermediates": "kept" if kept else "recomputed",
            "recompute_mode": rec, "J": J, "dJ_dE": g_E,
            "fd_eps": 1e-6, "fd_central_differences": fd, "vs_tangent_linear": tl,
            "deviations_from_workload": {
                "r_init": cfg["r_init"], "exp_energy": cfg["exp_energy"], "init_shrink": None,
                "why": "the unit-cube Sedov blast (r_init 0.01, E 1e12) with the initial 0.01 dt shrink "
                       "does not reach the objective's shell r in [0.2, 0.3] in 10 steps; dt_0 = "
                       "cfl * estdt(U0) and the blast starts at r = 0.15 with E = 1e14"}}


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
    ap.add_argument("--no-batch", action="store_true", help="skip the mc_advance_batch timing")
    ap.add_argument("--no-large", action="store_true", help="skip the 512^3 z-slab run")
    ap.add_argument("--cpu-1core", action="store_true",
                    help="also time the CPU baseline step on one thread (about three minutes at 256^3)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="only time the adjoint workload and print its JSON (for profiling)")
    ap.add_argument("--stage-variant", type=int, default=-1,
                    help="kernel variant flags of the fused step instead of the autotuner's choice (0: tflux, "
                         "trans1c, final; 1: trace + flux1 instead of tflux; 2: trans1 instead of trans1c; "
                         "4: finalc instead of final)")
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
            L.set_tiling(L.STAGES.index(st), *b[:3])
            if len(b) > 3:
                L.set_brick_order(L.STAGES.index(st), b[3])
        L.set_stage_variant(prev["config"].get("stage_variant", 0))
        tuned = None
    else:
        tuned = L.autotune(reps=4)   # restores the state it times on; picks the kernel variant
        log("autotuned")
    if a.stage_variant >= 0:
        L.set_stage_variant(a.stage_variant)
    variant = L.stage_variant
    # brick (bx, by, bz) and brick order of every stage
    tiling = {L.STAGES[st]: list(L.get_tiling(st)) + [L.brick_order(st)] for st in range(1, len(L.STAGES) - 1)}
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
    log("fused path timed")
    cpu_state = None if a.no_cpu_baseline else (L.state(), L.dt)
    L.close()

    # Whole step against the reference algorithm's work per zone-update
    # (op-count build of the CPU restatement, committed summary) and the
    # summed measured HBM traffic of the stages of one step (rocprof
    # counters, committed summary) against the zone-update's 320 B.
    w_flop = reference_flop_per_zone_update(cfg["name"])
    ran = [st for st in STAGES_WITH_BYTES if stages[st][1]]   # the stages this step runs
    tr = [measured_traffic("mc::K_" + st, cfg["name"]) for st in ran]
    step_traffic = sum(tr) if all(tr) else None
    hw = [measured_fp64_ops("mc::K_" + st, cfg["name"]) for st in ran]
    hw_step = sum(hw) if all(hw) else None
    roofline = {
        "bound": "fp64", "unit": "TFLOP/s", "peak": FP64_PEAK / 1e12,
        "achieved": (value * w_flop / 1e12) if w_flop else None,
        "frac": (value * w_flop / FP64_PEAK) if w_flop else None,
        "traffic": step_traffic,
        "scope": "whole coarse step (%d fused stage launches + dt control); achieved = value x "
                 "flop_per_zone_update (reference algorithm, counted by the op-count oracle build)" % len(ran),
        "flop_per_zone_update": w_flop,
        "traffic_scope": "HBM bytes per step, summed over the %d stages (rocprof FETCH_SIZE + WRITE_SIZE, "
                         "calibrated on a plain copy)" % len(ran),
        "traffic_per_zone_update_bytes": (step_traffic / zones) if step_traffic else None,
        "traffic_over_320B_per_zone_update": (step_traffic / zones / B_ALG_PER_ZONE_UPDATE) if step_traffic else None,
        "hbm": {"bytes_per_zone_update": B_ALG_PER_ZONE_UPDATE,
                "achieved_GBs": value * B_ALG_PER_ZONE_UPDATE / 1e9,
                "frac": value * B_ALG_PER_ZONE_UPDATE / HBM_PEAK},
        "issued_fp64": {"ops_per_zone_update": (hw_step / zones) if hw_step else None,
                        "TFLOPs": (hw_step / (ms_step / 1e3) / 1e12) if hw_step else None,
                        "frac": (hw_step / (ms_step / 1e3) / FP64_PEAK) if hw_step else None},
    }

    # the dominant stage (HIP events on the level's stream) against its own
    # intermediate arrays' bytes -- not the zone-update unit
    dom = max(ran, key=lambda k: stages[k][0])
    tot_ms, nl = stages[dom]
    avg_s = tot_ms / nl / 1e3
    traffic_alg = _stage_bytes(dom, n)
    achieved = traffic_alg / avg_s
    traffic = measured_traffic("mc::K_" + dom, cfg["name"])
    roofline["dominant_stage"] = {
        "kernel": "mc::K_" + dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
        "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic, "avg_launch_ms": tot_ms / nl,
        "bytes": "the stage's own input and output arrays (every distinct element once per launch)",
        "algorithmic_bytes_per_launch": traffic_alg,
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

    log("per-box path done")
    large = None if a.no_large else large_domain_bench()
    log("large domain done")
    batch = None if a.no_batch else batch_bench(cfg)
    log("mc_advance_batch done")
    adjoint = None if a.no_adjoint else adjoint_bench()
    log("adjoint done")
    ensemble = None if a.no_ensemble else ensemble_bench()
    log("ensemble done")
    cpu = None if cpu_state is None else cpu_baseline(cfg, cpu_state[0], cpu_state[1], one_core=a.cpu_1core)
    log("cpu baseline done")
    out = {"metric": METRIC, "value": value, "unit": "zone-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": cfg["name"], "n_cell": list(n), "boxes": int(len(cfg["boxes"])),
                      "path": "batched fused, one launch per stage over all boxes",
                      "graph_replay_ms_per_step": graph_ms, "steps_advanced": fused_steps_done,
                      "fom_zones_per_usec": value / 1e6,
                      "tiling": tiling, "tiling_source": "autotuner" if tuned else a.tiling_json,
                      "stage_variant": variant,
                      "stage_variant_source": ("--stage-variant" if a.stage_variant >= 0 else
                                               "autotuner" if tuned else a.tiling_json),
                      "stages": [st for st in STAGES_WITH_BYTES if stages[st][1]],
                      "per_box_ms_per_step": plain,
                      "batched_vs_per_box": (plain / ms_step) if plain else None,
                      "batch_entry": batch,
                      "engine_bytes_per_cell": {
                          "whole_level": mc.lib().mc_fused_bytes(1, (C.c_int * 3)(*n), 0) / zones,
                          "slabs_of_32_planes": mc.lib().mc_fused_bytes(1, (C.c_int * 3)(*n), 32) / zones,
                          "state": 20 * 8},
                      "large_domain": large,
                      "stage_ms_per_step": {k: v[0] / a.steps for k, v in stages.items() if v[1]}},
           "roofline": roofline, "cpu_baseline": cpu, "adjoint": adjoint,
           "ensemble": ensemble}
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Batched fused path: bit-identical to the per-box path and to the oracle.

The fused engine advances every box with one launch per stage and keeps dt
on the device; its state after every step must equal the oracle's (and the
plain per-box GPU path's) bit for bit, for periodic and outflow boundaries,
uniform and uneven box lists, eager launches and hipGraph replays.
"""
import numpy as np
import pytest

import minicastro as mc
from minicastro import configs
from helpers import oracle_level, rel_err, small_cfg, table_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())


def gpu_level(cfg, boxes=None, mode=mc.MODE_FUSED):
    L = mc.Level(cfg["n_cell"], cfg["boxes"] if boxes is None else boxes, cfg["prob_lo"], cfg["prob_hi"],
                 cfg["bc"], mode)
    L.init_data(**cfg)
    return L


@pytest.mark.parametrize("bc", [0, 1])
def test_fused_matches_oracle(bc):
    cfg = small_cfg(n=16, bc=bc, r_init=0.12)
    O = oracle_level(cfg, nthreads=8)
    G = gpu_level(cfg, boxes=configs.grid_boxes(16, 8))
    for s in range(4):
        dto, dtg = O.coarse_step(), G.coarse_step()
        assert dtg == dto, (s, dtg, dto)
        np.testing.assert_array_equal(G.state(), O.state(), err_msg="step %d" % s)


def test_fused_config1_bitwise():
    cfg = configs.config(1)
    O = oracle_level(cfg, nthreads=8)
    G = gpu_level(cfg)
    for s in range(cfg["steps"]):
        dto, dtg = O.coarse_step(), G.coarse_step()
        assert dtg == dto, (s, dtg, dto)
    np.testing.assert_array_equal(G.state(), O.state())


def test_fused_uneven_equals_plain_and_graph_replay():
    n = 48
    cfg = small_cfg(n=n, bc=1, r_init=0.08)
    boxes = configs.uneven_boxes(n, seed=3, parts=2, sizes=tuple(range(16, 33, 8)))
    P = gpu_level(cfg, boxes=boxes, mode=mc.MODE_PLAIN)
    F = gpu_level(cfg, boxes=boxes, mode=mc.MODE_FUSED)
    for s in range(2):
        assert P.coarse_step() == F.coarse_step()
    F.run(3, use_graph=True)      # graph replays with device-resident dt
    for s in range(3):
        P.coarse_step()
    assert F.steps == P.steps == 5
    assert F.dt == P.dt and F.time == P.time
    np.testing.assert_array_equal(F.state(), P.state())


def test_tiling_does_not_change_results():
    """Other bricks for every stage, with the default kernels (tflux,
    trans1c, final) and the stages tflux and trans1c replace (trace +
    flux1, trans1)."""
    cfg = small_cfg(n=24, bc=0, r_init=0.1)
    A = gpu_level(cfg)
    B = gpu_level(cfg)
    C = gpu_level(cfg)
    C.set_stage_variant(C.SPLIT_TRACE | C.FACE_TRANS1)
    for st, t in ((2, (16, 4, 4)), (3, (64, 1, 4)), (4, (8, 8, 4)), (5, (32, 2, 4)), (6, (4, 4, 16)),
                  (1, (256, 1, 1)), (7, (8, 8, 2)), (8, (8, 16, 32)), (9, (32, 4, 128)), (10, (16, 4, 32))):
        B.set_tiling(st, *t)
        C.set_tiling(st, *t)
        B.set_brick_order(st, (2, 3, 4, 8)[st % 4])
    assert B.brick_order(6) == 4 and C.brick_order(6) == 0
    A.run(3, use_graph=False)
    B.run(3, use_graph=True)
    C.run(3, use_graph=True)
    np.testing.assert_array_equal(A.state(), B.state())
    np.testing.assert_array_equal(A.state(), C.state())
    with pytest.raises(RuntimeError):
        B.set_tiling(8, 16, 16, 3)      # X must divide W*L
    with pytest.raises(RuntimeError):
        B.set_tiling(8, 64, 1, 64)      # L >= 2
    with pytest.raises(RuntimeError):
        B.set_brick_order(6, 1)


def test_autotune_keeps_state_and_results():
    """The autotuner leaves state and clock untouched, picks valid bricks, and
    the tuned level advances bit-identically to an untuned one."""
    cfg = small_cfg(n=32, bc=1, r_init=0.1)
    boxes = configs.uneven_boxes(32, seed=5, parts=2, sizes=(8, 16, 24))
    A = gpu_level(cfg, boxes=boxes)
    B = gpu_level(cfg, boxes=boxes)
    A.run(1, use_graph=False)
    B.run(1, use_graph=False)
    s0 = A.state()
    clock0 = (A.steps, A.dt, A.time)
    tuned = A.autotune(reps=1)
    assert set(tuned) == {"prim", "flat", "trace", "flux1", "tflux", "trans1", "trans1c", "final", "finalc",
                          "update"}
    for st, ((bx, by, bz), ms) in tuned.items():
        nt = bx * by if st in ("tflux", "trans1c", "finalc") else bx * by * bz
        assert nt % 64 == 0 and nt <= 256 and ms > 0.0, (st, bx, by, bz, ms)
        assert A.get_tiling(A.STAGES.index(st)) == (bx, by, bz)
    # the faster kernel of each alternative
    ms = {st: v[1] for st, v in tuned.items()}
    want = ((A.SPLIT_TRACE if ms["trace"] + ms["flux1"] < ms["tflux"] else 0)
            | (A.FACE_TRANS1 if ms["trans1"] < ms["trans1c"] else 0)
            | (A.CELL_FINAL if ms["finalc"] < ms["final"] else 0))
    assert A.stage_variant == want
    np.testing.assert_array_equal(A.state(), s0)
    assert (A.steps, A.dt, A.time) == clock0
    A.run(3, use_graph=True)
    B.run(3, use_graph=True)
    assert (A.steps, A.dt, A.time) == (B.steps, B.dt, B.time)
    np.testing.assert_array_equal(A.state(), B.state())


def test_fused_odd_domain_matches_oracle():
    """A 15^3 domain in ragged 8/7-cell boxes: its cell count (3375) and
    extended count (17^3) are odd, so the engine's even-padded pair arrays
    carry a padding element; the state must still equal the oracle's bit for
    bit, outflow and periodic."""
    for bc in (1, 0):
        cfg = small_cfg(n=15, bc=bc, r_init=0.14)
        O = oracle_level(cfg, nthreads=8)
        G = gpu_level(cfg, boxes=configs.grid_boxes(15, 8))
        for s in range(3):
            dto, dtg = O.coarse_step(), G.coarse_step()
            assert dtg == dto, (bc, s, dtg, dto)
        np.testing.assert_array_equal(G.state(), O.state(), err_msg="bc %d" % bc)


ENSEMBLE_SETS = {
    "even": ((16, 2e12, 0.12, 0.7, 1), (24, 5e12, 0.09, 1.6, 0), (20, 1e12, 0.15, 1.1, 1)),
    # odd sides: every instance's cell and extended counts are odd, so each
    # instance offset in the pair arrays includes the padding of the ones
    # before it
    "odd": ((15, 2e12, 0.12, 0.7, 1), (13, 5e12, 0.09, 1.6, 0), (17, 1e12, 0.15, 1.1, 1), (11, 3e12, 0.2, 0.9, 0)),
}


@pytest.mark.parametrize("sizes", sorted(ENSEMBLE_SETS))
def test_ensemble_matches_sequential_runs(sizes):
    """Instances of different sizes and constants advanced together, each with
    its own dt, equal the same instances advanced one at a time by the per-box
    path: every state bit and every instance's dt."""
    insts = []
    for side, E, r, rho, bc in ENSEMBLE_SETS[sizes]:
        d = small_cfg(n=side, bc=bc, r_init=r, exp_energy=E, dens_ambient=rho)
        insts.append(d)
    ens = mc.Ensemble(insts)
    ens.init_data()
    ens.run(2, use_graph=False)
    ens.run(3, use_graph=True)
    for m, d in enumerate(insts):
        L = gpu_level(d, mode=mc.MODE_PLAIN)
        for _ in range(5):
            L.coarse_step()
        dt, t, steps = ens.clock(m)
        assert steps == 5 and dt == L.dt and t == L.time, (m, dt, L.dt)
        np.testing.assert_array_equal(ens.state(m), L.state(), err_msg="instance %d" % m)


def test_config4_ensemble_24_instances_50_steps():
    """Config 4: the 24 seeded Sedov instances (sides 32-96, energies, radii
    and densities drawn per instance), 50 steps each with its own dt, all in
    one launch per stage.  Every instance's clock and state equal the same
    instance advanced alone by the per-box path, bit for bit; every 32^3 and
    48^3 instance (eight of them) also equals the oracle advanced alone with
    its own problem constants, bit for bit over all 50 steps."""
    import os
    insts = configs.ensemble(24, seed=7)
    ens = mc.Ensemble(insts)
    ens.init_data()
    ens.run(50, use_graph=True)
    threads = max(1, min(16, os.cpu_count() or 1))
    checked = 0
    for m, d in enumerate(insts):
        L = gpu_level(d, mode=mc.MODE_PLAIN)
        L.run(50)
        dt, t, steps = ens.clock(m)
        assert steps == L.steps == 50 and dt == L.dt and t == L.time, (m, dt, L.dt)
        S = ens.state(m)
        np.testing.assert_array_equal(S, L.state(), err_msg="instance %d" % m)
        L.close()
        if d["n_cell"][0] in (32, 48):
            O = oracle_level(d, nthreads=threads)
            for _ in range(50):
                O.coarse_step()
            assert (O.steps, O.dt, O.time) == (steps, dt, t), (m, O.dt, dt)
            np.testing.assert_array_equal(S, O.state(), err_msg="instance %d vs oracle" % m)
            checked += 1
            print("instance %d (%d^3, E %.3e, r %.4f, rho %.3f): oracle == ensemble over 50 steps" %
                  (m, d["n_cell"][0], d["exp_energy"], d["r_init"], d["dens_ambient"]), flush=True)
    assert checked == 8
    ens.close()


def test_fused_bit_patterns_equal_oracle():
    """Bit patterns, not just values: the batched path's state (signed zeros
    included) equals the oracle's after several steps, outflow and periodic."""
    for bc in (1, 0):
        cfg = small_cfg(n=16, bc=bc, r_init=0.12)
        O = oracle_level(cfg, nthreads=8)
        G = gpu_level(cfg, boxes=configs.grid_boxes(16, 8))
        for _ in range(4):
            O.coarse_step()
            G.coarse_step()
        a, b = G.state(), O.state()
        diff = a.view(np.uint64) != b.view(np.uint64)
        assert not diff.any(), (bc, int(diff.sum()), a[diff][:4], b[diff][:4])

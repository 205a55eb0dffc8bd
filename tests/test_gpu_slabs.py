"""The fused engine in z-slabs: the face stages and the update run slab
after slab, the face arrays holding one slab and its ring, and the results
equal the oracle's (and the whole-level engine's) bit for bit: thin slabs,
a last slab of one plane, periodic and outflow boundaries, uneven boxes,
hipGraph replay, ensembles whose instances are shallower than a slab, the
autotuner, and the benchmark workload itself."""
import ctypes as C

import numpy as np
import pytest

import minicastro as mc
from minicastro import configs
from helpers import oracle_level, small_cfg, table_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())


def fused(cfg, boxes, slab):
    L = mc.Level(cfg["n_cell"], boxes, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    L.set_fused_slab(slab)
    L.init_data(**cfg)
    assert L.fused_slab == slab
    return L


@pytest.mark.parametrize("bc", [1, 0])
@pytest.mark.parametrize("slab", [5, 1])
def test_slabs_match_oracle(bc, slab):
    """16^3 in uneven boxes: slabs of 5 planes (5, 5, 5, 1) or of 1 plane;
    three eager steps and two graph replays against the oracle."""
    cfg = small_cfg(n=16, bc=bc, r_init=0.14)
    boxes = configs.uneven_boxes(16, seed=4, parts=2, sizes=(4, 8, 12))
    O = oracle_level(cfg, nthreads=8)
    G = fused(cfg, boxes, slab)
    for s in range(3):
        dto, dtg = O.coarse_step(), G.coarse_step()
        assert dtg == dto, (s, dtg, dto)
        np.testing.assert_array_equal(G.state(), O.state(), err_msg="step %d" % s)
    G.run(2, use_graph=True)
    O.coarse_step()
    O.coarse_step()
    assert (G.dt, G.time) == (O.dt, O.time)
    np.testing.assert_array_equal(G.state(), O.state())


@pytest.mark.parametrize("flags", [0, 4, 3])
def test_ensemble_slabs_match_sequential_runs(flags):
    """Instances deeper and shallower than the slab (sides 16, 13, 9, 20 with
    slabs of 6 planes), with the default kernels, the final stage by cells,
    and the split trace/flux with face-centred trans1: every instance equals
    its per-box run."""
    insts = []
    for side, E, r, rho, bc in ((16, 2e12, 0.12, 0.7, 1), (13, 5e12, 0.09, 1.6, 0), (9, 1e12, 0.2, 1.1, 1),
                                (20, 3e12, 0.1, 0.9, 0)):
        insts.append(small_cfg(n=side, bc=bc, r_init=r, exp_energy=E, dens_ambient=rho))
    ens = mc.Ensemble(insts, slab=6)
    assert ens.slab == 6
    ens.set_stage_variant(flags)
    assert ens.stage_variant == flags
    ens.init_data()
    ens.run(2, use_graph=False)
    ens.run(2, use_graph=True)
    for m, d in enumerate(insts):
        L = mc.Level(d["n_cell"], d["boxes"], d["prob_lo"], d["prob_hi"], d["bc"], mc.MODE_PLAIN)
        L.init_data(**d)
        for _ in range(4):
            L.coarse_step()
        dt, t, steps = ens.clock(m)
        assert steps == 4 and dt == L.dt and t == L.time, (m, dt, L.dt)
        np.testing.assert_array_equal(ens.state(m), L.state(), err_msg="instance %d" % m)
        L.close()
    ens.close()


def test_autotune_and_timed_steps_in_slabs():
    """The autotuner (timed on the first slab) leaves state and clock alone;
    timed steps report every stage; results equal the whole-level engine."""
    cfg = small_cfg(n=24, bc=1, r_init=0.1)
    boxes = configs.grid_boxes(24, 12)
    A = fused(cfg, boxes, 7)
    B = fused(cfg, boxes, 0)
    A.run(1, use_graph=False)
    B.run(1, use_graph=False)
    s0, c0 = A.state(), (A.steps, A.dt, A.time)
    A.autotune(reps=1)
    np.testing.assert_array_equal(A.state(), s0)
    assert (A.steps, A.dt, A.time) == c0
    st = A.run_timed(2)
    v = A.stage_variant  # the kernels the autotuner chose
    ran = ["prim", "flat", "update", "trans1" if v & A.FACE_TRANS1 else "trans1c",
           "finalc" if v & A.CELL_FINAL else "final"] + (["trace", "flux1"] if v & A.SPLIT_TRACE else ["tflux"])
    for k in A.STAGES[1:-1]:
        want = (2 if k == "prim" else 2 * 4) if k in ran else 0
        assert st[k][1] == want and (st[k][0] > 0) == (want > 0), (k, st[k], v)
    B.run(2, use_graph=False)
    assert (A.steps, A.dt, A.time) == (B.steps, B.dt, B.time)
    np.testing.assert_array_equal(A.state(), B.state())


def test_slab_choice_and_switching_keep_the_state():
    cfg = small_cfg(n=16, bc=1, r_init=0.12)
    L = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    L.init_data(**cfg)
    assert L.fused_slab == 0          # a small level keeps whole-level arrays
    L.run(2, use_graph=False)
    L.set_fused_slab(3)               # rebuilt with the state and clock kept
    assert L.fused_slab == 3
    R = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_PLAIN)
    R.init_data(**cfg)
    for _ in range(2):
        R.coarse_step()
    L.run(2, use_graph=True)
    for _ in range(2):
        R.coarse_step()
    assert (L.steps, L.dt, L.time) == (R.steps, R.dt, R.time)
    np.testing.assert_array_equal(L.state(), R.state())
    with pytest.raises(RuntimeError):
        L.set_fused_slab(-2)
    nc = (C.c_int * 3)(512, 512, 512)
    assert mc.lib().mc_fused_bytes(1, nc, 64) < mc.lib().mc_fused_bytes(1, nc, 0) / 4


def test_config3_slabs_equal_whole_level():
    """The benchmark workload (256^3, 64 uneven boxes) in slabs of 32 planes
    equals the whole-level engine bit for bit over three steps."""
    cfg = configs.config(3)
    A = fused(cfg, cfg["boxes"], 0)
    A.run(3, use_graph=True)
    SA, ca = A.state(), (A.steps, A.dt, A.time)
    A.close()
    B = fused(cfg, cfg["boxes"], 32)
    B.run(3, use_graph=True)
    assert (B.steps, B.dt, B.time) == ca
    np.testing.assert_array_equal(B.state(), SA)
    B.close()

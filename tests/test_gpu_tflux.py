"""The stages whose bricks overlap their neighbour along a direction by one
cell and pass data between threads through LDS: K_tflux (the PPM traces and
the first-stage fluxes in one launch, the flux of a face reading the minus
state the layer below traced), K_trans1c and K_finalc (trans1 and trans2 by
cells, each cell correcting both its states with the transverse fluxes on
its own faces).  Every brick shape the autotuner may pick (one brick
spanning the whole direction, two-cell bricks, 64 to 256 threads) gives the
oracle's state bit for bit, and so does every combination of these kernels
with the stages they replace."""
import numpy as np
import pytest

import minicastro as mc
from minicastro import configs
from helpers import oracle_level, small_cfg, table_bin

pytestmark = pytest.mark.gpu

# (W, L, X): y bricks (W, L, 1), z bricks (W, 1, L), x bricks (X, W*L/X, 1)
TILINGS = [(16, 16, 64), (32, 8, 64), (8, 32, 64), (16, 16, 128), (64, 4, 64), (16, 8, 64), (8, 8, 64),
           (16, 16, 256), (32, 4, 64), (64, 2, 2)]


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())


def tflux_tile_fits(t):
    """K_tflux stages its bricks' primitives with two cells either side
    along the brick's direction: (L + 4) W cells for the y and z bricks,
    (X + 4) W L / X for the x bricks, at most 448."""
    w, l, x = t
    return (l + 4) * w <= 448 and (x + 4) * (w * l // x) <= 448


@pytest.mark.parametrize("t", TILINGS)
def test_overlap_bricks_match_oracle(t):
    """20^3 in uneven boxes (22 ring cells per direction: one brick of L = 32,
    or 7, 11 or 21 overlapping bricks), outflow and periodic, two eager steps
    and a graph replay; the three overlapping stages take bricks of different
    shapes and brick orders (the final stage by cells on every other
    tiling)."""
    bc = TILINGS.index(t) % 2
    cfg = small_cfg(n=20, bc=bc, r_init=0.13)
    boxes = configs.uneven_boxes(20, seed=7, parts=2, sizes=(4, 8, 12))
    O = oracle_level(cfg, nthreads=8)
    G = mc.Level(cfg["n_cell"], boxes, cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    if tflux_tile_fits(t):
        G.set_tiling(G.STAGES.index("tflux"), *t)
    else:  # K_tflux's LDS tile holds at most 448 cells a component
        with pytest.raises(RuntimeError):
            G.set_tiling(G.STAGES.index("tflux"), *t)
        G.set_tiling(G.STAGES.index("tflux"), 8, 8, 2)
    G.set_tiling(G.STAGES.index("trans1c"), *TILINGS[(TILINGS.index(t) + 3) % len(TILINGS)])
    G.set_tiling(G.STAGES.index("finalc"), *TILINGS[(TILINGS.index(t) + 6) % len(TILINGS)])
    if TILINGS.index(t) % 2:
        G.set_stage_variant(G.CELL_FINAL)
    for st in ("tflux", "trans1c", "finalc"):  # raster, or groups of brick planes in z
        G.set_brick_order(G.STAGES.index(st), (0, 2, 3, 4, 8)[(TILINGS.index(t) + len(st)) % 5])
    G.init_data(**cfg)
    for s in range(2):
        dto, dtg = O.coarse_step(), G.coarse_step()
        assert dtg == dto, (s, dtg, dto)
        np.testing.assert_array_equal(G.state(), O.state(), err_msg="step %d" % s)
    G.run(1, use_graph=True)
    O.coarse_step()
    assert (G.dt, G.time) == (O.dt, O.time)
    np.testing.assert_array_equal(G.state(), O.state())
    G.close()


@pytest.mark.parametrize("flags", [4, 3])
def test_variants_in_slabs_match_oracle(flags):
    """16^3 in z-slabs of 5 planes (5, 5, 5, 1) with the final stage by cells
    and with the stages the overlapping kernels replace."""
    cfg = small_cfg(n=16, bc=flags % 2, r_init=0.14)
    O = oracle_level(cfg, nthreads=8)
    G = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    G.set_fused_slab(5)
    G.set_stage_variant(flags)
    G.init_data(**cfg)
    for s in range(2):
        dto, dtg = O.coarse_step(), G.coarse_step()
        assert dtg == dto, (s, dtg, dto)
        np.testing.assert_array_equal(G.state(), O.state(), err_msg="step %d" % s)
    G.close()


def test_variants_agree_on_the_benchmark_workload():
    """256^3 in 64 uneven boxes: two steps with every combination of the
    overlapping kernels and the stages they replace give the same state; the
    timed run reports the stages each variant runs."""
    cfg = configs.config(3)
    ref = None
    runs = {0: ("tflux", "trans1c", "final"), 1: ("trace", "flux1", "trans1c", "final"),
            2: ("tflux", "trans1", "final"), 4: ("tflux", "trans1c", "finalc"),
            3: ("trace", "flux1", "trans1", "final")}
    for flags, ran in runs.items():
        L = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
        L.set_stage_variant(flags)
        L.init_data(**cfg)
        st = L.run_timed(2)
        for k in ("trace", "flux1", "tflux", "trans1", "trans1c", "final", "finalc"):
            assert st[k][1] == (2 if k in ran else 0), (flags, k, st[k])
        got = (L.state(), (L.steps, L.dt, L.time))
        L.close()
        if ref is None:
            ref = got
            continue
        assert got[1] == ref[1], flags
        np.testing.assert_array_equal(got[0], ref[0], err_msg="variant %d" % flags)


def test_bad_variant_refused():
    cfg = small_cfg(n=8, bc=1, r_init=0.2)
    L = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    with pytest.raises(RuntimeError):
        L.set_stage_variant(8)
    with pytest.raises(RuntimeError):
        L.set_tiling(L.STAGES.index("trans1c"), 16, 16, 3)
    L.close()


@pytest.mark.parametrize("flags,slab", [(0, 0), (4, 3), (3, 0)])
def test_tiny_instances_match_sequential_runs(flags, slab):
    """An ensemble of instances smaller than every brick (sides 4, 5, 7; one
    brick spans a whole direction, most threads of a brick are idle), with
    and without z-slabs and with each kernel set: every instance equals its
    per-box run over three steps."""
    insts = []
    for side, bc, r in ((4, 1, 0.3), (5, 0, 0.25), (7, 1, 0.2), (5, 1, 0.35)):
        insts.append(small_cfg(n=side, bc=bc, r_init=r))
    ens = mc.Ensemble(insts, slab=slab)
    ens.set_stage_variant(flags)
    ens.init_data()
    ens.run(3, use_graph=True)
    for m, d in enumerate(insts):
        L = mc.Level(d["n_cell"], d["boxes"], d["prob_lo"], d["prob_hi"], d["bc"], mc.MODE_PLAIN)
        L.init_data(**d)
        for _ in range(3):
            L.coarse_step()
        dt, t, steps = ens.clock(m)
        assert steps == 3 and dt == L.dt and t == L.time, (m, dt, L.dt)
        np.testing.assert_array_equal(ens.state(m), L.state(), err_msg="instance %d" % m)
        L.close()
    ens.close()

"""Whole-step parity of the GPU level driver against the oracle driver.

The GPU driver replays Castro::advance / post_timestep / computeNewDt with
the drop-in kernels; every step's state and dt must equal the oracle's bit
for bit (the bar is written as max-norm relative error == 0 for the state
and exact equality for dt).  Box-decomposition invariance is checked on the
GPU against a single-box oracle run.
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


def gpu_level(cfg, boxes=None, mode=mc.MODE_PLAIN):
    L = mc.Level(cfg["n_cell"], cfg["boxes"] if boxes is None else boxes, cfg["prob_lo"], cfg["prob_hi"],
                 cfg["bc"], mode)
    L.init_data(**cfg)
    return L


def run_pair(cfg, steps, boxes=None, mode=mc.MODE_PLAIN, tol=0.0, nthreads=8):
    O = oracle_level(cfg, nthreads=nthreads)
    G = gpu_level(cfg, boxes, mode)
    np.testing.assert_array_equal(G.state(), O.state())
    for s in range(steps):
        dto = O.coarse_step()
        dtg = G.coarse_step()
        assert dtg == dto, "step %d dt %r != %r" % (s, dtg, dto)
        err = rel_err(G.state(), O.state())
        assert err.max() <= tol, "step %d max rel err %g (comp %d)" % (s, err.max(), err.argmax())
    return G, O


def test_config1_sedov32_outflow_bitwise():
    cfg = configs.config(1)
    run_pair(cfg, cfg["steps"])


def test_config1_sedov32_periodic_bitwise():
    """Config 1's periodic variant: 32^3, one box, 10 steps."""
    cfg = dict(configs.config(1), bc=0)
    run_pair(cfg, cfg["steps"])


def test_periodic_8boxes_bitwise():
    cfg = small_cfg(n=16, bc=0, r_init=0.1)
    run_pair(cfg, 4, boxes=configs.grid_boxes(16, 8))


def test_uneven_boxes_bitwise():
    cfg = small_cfg(n=24, bc=1, r_init=0.12)
    boxes = np.array([[0, 0, 0, 7, 15, 23], [8, 0, 0, 23, 15, 23], [0, 16, 0, 15, 23, 7],
                      [16, 16, 0, 23, 23, 7], [0, 16, 8, 23, 23, 23]], dtype=np.int32)
    run_pair(cfg, 3, boxes=boxes)


def test_ragged_thin_boxes_bitwise():
    """Boxes one to three cells thick: every stencil of the step reaches
    across several boxes, so the ghost fill and the box algebra are
    exercised at their limits.  Per-box path == oracle, batched == per-box."""
    cfg = small_cfg(n=12, bc=1, r_init=0.2)
    cx, cy, cz = [0, 1, 3, 12], [0, 2, 5, 12], [0, 3, 4, 12]
    boxes = np.array([[cx[a], cy[b], cz[c], cx[a + 1] - 1, cy[b + 1] - 1, cz[c + 1] - 1]
                      for c in range(3) for b in range(3) for a in range(3)], dtype=np.int32)
    G, _ = run_pair(cfg, 3, boxes=boxes)
    F = gpu_level(cfg, boxes=boxes, mode=mc.MODE_FUSED)
    F.run(3, use_graph=False)
    assert (F.dt, F.time) == (G.dt, G.time)
    np.testing.assert_array_equal(F.state(), G.state())


def test_config2_sedov128_8x64_20_steps():
    """Config 2: 128^3 in eight 64^3 boxes, 20 steps.  The per-box GPU path
    equals the oracle bit for bit at every one of the 20 steps (state and
    dt), and the batched path equals both at the end, dt included."""
    import os
    cfg = configs.config(2)
    threads = max(1, min(16, os.cpu_count() or 1))
    G, O = run_pair(cfg, cfg["steps"], nthreads=threads)
    F = gpu_level(cfg, mode=mc.MODE_FUSED)
    F.run(2, use_graph=False)
    F.run(cfg["steps"] - 2, use_graph=True)
    assert (F.steps, F.dt, F.time) == (G.steps, G.dt, G.time) == (cfg["steps"], O.dt, O.time)
    np.testing.assert_array_equal(F.state(), O.state())


def test_config3_sedov256_uneven64():
    """Config 3, the benchmark workload: 256^3 in the 64 seeded uneven boxes.
    The per-box GPU path equals the oracle bit for bit over the first step
    (about a minute of oracle time on 16 host threads), and the batched path
    equals the per-box path bit for bit, dt included, over the first 12 steps
    (the benchmark times steps 3-12)."""
    import os
    cfg = configs.config(3)
    G, O = run_pair(cfg, 1, nthreads=max(1, min(16, os.cpu_count() or 1)))
    del O
    F = gpu_level(cfg, mode=mc.MODE_FUSED)
    F.run(1, use_graph=False)
    np.testing.assert_array_equal(F.state(), G.state())
    for _ in range(11):
        G.coarse_step()
    F.run(11, use_graph=True)
    assert (F.steps, F.dt, F.time) == (G.steps, G.dt, G.time)
    np.testing.assert_array_equal(F.state(), G.state())


def test_periodic_conservation_on_gpu():
    """Periodic boundaries: the domain sums of mass and total energy are
    conserved to round-off by the batched path (the flux divergence
    telescopes); momentum too, the Sedov state being symmetric."""
    cfg = small_cfg(n=32, bc=0, r_init=0.15)
    F = gpu_level(cfg, boxes=configs.grid_boxes(32, 16), mode=mc.MODE_FUSED)
    S0 = F.state()
    F.run(10, use_graph=True)
    S1 = F.state()
    for c in (0, 4):   # URHO, UEDEN
        a, b = S0[..., c].sum(), S1[..., c].sum()
        assert abs(b - a) <= 1e-12 * abs(a), (c, a, b)
    scale = np.abs(S1[..., 1:4]).sum()
    assert abs(S1[..., 1:4].sum()) <= 1e-12 * scale


def test_multispecies_state_bitwise():
    """Three species (He, C, O) whose mass fractions vary across the blast:
    every species stream of the path (tracing, both transverse stages, the
    flux normalisation, the species sums of the EOS inputs) carries distinct
    non-trivial data.  Per-box and batched paths equal the oracle bit for bit
    at every step, dt included."""
    import os
    cfg = small_cfg(n=24, bc=1, r_init=0.15)
    threads = max(1, min(16, os.cpu_count() or 1))
    O = oracle_level(cfg, nthreads=threads)
    S = O.state()
    n = cfg["n_cell"][0]
    x = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    xc = 0.3 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    xo = 0.1 + 0.08 * Z
    rho = S[..., 0]
    S[..., 7] = rho * (1.0 - xc - xo)
    S[..., 8] = rho * xc
    S[..., 9] = rho * xo
    O.set_state(S)
    G = gpu_level(cfg, boxes=configs.grid_boxes(n, 12))
    F = gpu_level(cfg, boxes=configs.grid_boxes(n, 12), mode=mc.MODE_FUSED)
    G.set_state(S)
    F.set_state(S)
    for s in range(4):
        dto = O.coarse_step()
        assert G.coarse_step() == dto and F.coarse_step() == dto, s
        So = O.state()
        np.testing.assert_array_equal(G.state(), So, err_msg="per-box, step %d" % s)
        np.testing.assert_array_equal(F.state(), So, err_msg="batched, step %d" % s)
    # the composition is still non-trivial in all three species
    for c in (7, 8, 9):
        assert np.ptp(So[..., c]) > 0.05, c

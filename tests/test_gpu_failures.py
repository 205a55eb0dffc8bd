"""Failure detection on the GPU path.

* EOS Newton solves that end unconverged pass silently in the reference
  (Source/eos.F90:193,550); the library counts them.  The condition is forced
  from the host with non-finite energies (the oracle shows such a solve runs
  all 100 iterations), and the results stay those of the oracle.
* A healthy Sedov run leaves the counter at zero.
* Castro::initData rejects cells that are not cubes (Source/Castro.cpp:95-100).
"""
import ctypes as C

import numpy as np
import pytest

import orc
import minicastro as mc
from helpers import oracle_table, small_cfg, table_bin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    oracle_table()


def eval_re(rho, T, e, abar, zbar):
    import torch
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    P = lambda t: C.cast(t.data_ptr(), C.POINTER(C.c_double))
    n = len(rho)
    out = torch.zeros(10 * n, dtype=torch.float64, device="cuda")
    args = [dev(rho), dev(T), dev(e), dev(abar), dev(zbar)]
    assert mc.lib().mc_eos_eval_batch(n, 2, *[P(a) for a in args], P(out)) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(n, 10)


def test_newton_failures_counted():
    n = 256
    rng = np.random.default_rng(5)
    rho = 10.0 ** rng.uniform(-3, 3, n)
    T = 10.0 ** rng.uniform(4, 8, n)
    abar, zbar = np.full(n, 4.0), np.full(n, 2.0)
    e = np.array([orc.eos(orc.EOS_RT, r, t, 4.0, 2.0).e for r, t in zip(rho, T)])
    mc.newton_failures(reset=True)
    eval_re(rho, 1.3 * T, e, abar, zbar)
    assert mc.newton_failures() == 0
    bad = np.array([3, 17, 100, 101, 255])
    e2 = e.copy()
    e2[bad[:3]] = np.nan
    e2[bad[3:]] = np.inf
    o = eval_re(rho, 1.3 * T, e2, abar, zbar)
    assert mc.newton_failures(reset=True) == len(bad)
    assert mc.newton_failures() == 0
    for i in bad:       # the solve's result is still the oracle's (100 iterations)
        s = orc.eos(orc.EOS_RE, rho[i], 1.3 * T[i], 4.0, 2.0, e=e2[i])
        assert s.niter == 100
        assert o[i, 0] == s.T or (np.isnan(o[i, 0]) and np.isnan(s.T))


def test_sedov_run_has_no_newton_failures():
    cfg = small_cfg(n=24, r_init=0.15)
    mc.newton_failures(reset=True)
    for mode in (mc.MODE_PLAIN, mc.MODE_FUSED):
        L = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mode)
        L.init_data(**cfg)
        L.run(4, use_graph=False)
        L.sync()
        L.close()
    assert mc.newton_failures() == 0


def test_non_cubic_cells_rejected():
    cfg = small_cfg(n=16)
    L = mc.Level(cfg["n_cell"], cfg["boxes"], (0.0, 0.0, 0.0), (1.0, 1.0, 2.0), cfg["bc"], mc.MODE_PLAIN)
    with pytest.raises(RuntimeError, match="dx != dy"):
        L.init_data(**cfg)
    L.close()
    E = mc.Ensemble([dict(n_cell=(16, 16, 16), prob_lo=(0, 0, 0), prob_hi=(1, 2, 1), bc=1,
                          p_ambient=1e13, dens_ambient=1.0, exp_energy=1e12, r_init=0.1)])
    with pytest.raises(RuntimeError, match="dx != dy"):
        E.init_data()
    E.close()


def test_oversized_instance_rejected_before_allocation():
    """The fused engine addresses an instance's cells with 32-bit 8-byte lane
    offsets and its face pairs with 32-bit 16-byte offsets: an instance of
    2^29 cells or more (820^3) is refused, and so is one whose whole-instance
    face arrays would exceed 2^28 extended cells when z-slabs are turned off
    (700^3: 702^3 > 2^28), both with an error and before any device memory
    is taken.  In z-slabs a 700^3 instance is within the limits."""
    big = dict(prob_lo=(0, 0, 0), prob_hi=(1, 1, 1), bc=1, p_ambient=1e13, dens_ambient=1.0, exp_energy=1e12,
               r_init=0.1)
    with pytest.raises(RuntimeError, match="2\\^29"):
        mc.Ensemble([dict(big, n_cell=(820, 820, 820))])
    with pytest.raises(RuntimeError, match="2\\^28"):
        mc.Ensemble([dict(big, n_cell=(700, 700, 700))], slab=0)
    nc = (C.c_int * 3)(700, 700, 700)
    assert mc.lib().mc_fused_bytes(1, nc, 32) < 0.3 * mc.lib().mc_fused_bytes(1, nc, 0)


def test_runs_refuse_without_table_and_recapture_after_reload():
    """After eos_finalize every fused entry point reports an error instead of
    launching with no table; after a new load a level that had captured its
    step in a hipGraph captures it again (the table may have moved) and
    still equals a level that never lost its table."""
    cfg = small_cfg(n=16, bc=1, r_init=0.12)
    A = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    B = mc.Level(cfg["n_cell"], cfg["boxes"], cfg["prob_lo"], cfg["prob_hi"], cfg["bc"], mc.MODE_FUSED)
    A.init_data(**cfg)
    B.init_data(**cfg)
    A.run(2, use_graph=True)
    B.run(2, use_graph=True)
    L = mc.lib()
    try:
        L.eos_finalize()
        for what, call in (("run", lambda: L.mc_level_run(A.h, 1, -1.0, 1)),
                           ("run eager", lambda: L.mc_level_run(A.h, 1, -1.0, 0)),
                           ("run_stage", lambda: L.mc_level_run_stage(A.h, 1)),
                           ("autotune", lambda: L.mc_level_autotune(A.h, 1, None, None))):
            assert call() != 0, what
            assert b"EOS table" in L.mc_last_error(), what
    finally:
        mc.load_table(table_bin())
    assert A.steps == 2
    A.run(3, use_graph=True)
    B.run(3, use_graph=True)
    assert A.dt == B.dt and A.time == B.time
    np.testing.assert_array_equal(A.state(), B.state())
    A.close()
    B.close()

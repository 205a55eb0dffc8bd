"""AMR coupling on the GPU against the oracle, bit for bit: the flux register
operations (CrseInit, FineAdd, ClearInternalBorders, Reflux) and the
average-down on random data, outflow and periodic; then a two-level Sedov
step with the register fed by the per-box path's store_flux output
(Source/Castro_advance.cpp:82-88), Castro::reflux and Castro::avgDown
(Source/Castro.cpp:451-505), whose composite mass and energy are conserved."""
import numpy as np
import pytest

import orc
import minicastro as mc
from minicastro import amr, configs
from helpers import Dev, NVAR, oracle_table, small_cfg, table_bin

pytestmark = pytest.mark.gpu

RATIO = (2, 2, 2)


@pytest.fixture(scope="module", autouse=True)
def gpu_table():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    oracle_table()


def _pair(host, lo, hi, nc=NVAR):
    """The same FAB for the oracle (host) and the GPU (device)."""
    return orc.Fab(lo, hi, nc, host), Dev(lo, hi, nc, host)


@pytest.mark.parametrize("bc", [orc.BC_OUTFLOW, orc.BC_PERIODIC])
def test_register_operations_match_oracle(bc):
    rng = np.random.default_rng(11 + bc)
    nc = 8
    fine = np.array([[0, 0, 2, 7, 7, 9], [8, 0, 2, 15, 3, 9], [4, 8, 0, 11, 15, 15]], dtype=np.int32)
    O = orc.FluxRegister(fine, RATIO, (nc,) * 3, bc)
    G = amr.FluxRegister(fine, RATIO, (nc,) * 3, bc)
    O.set_val(0.0)
    G.set_val(0.0)
    # CrseInit from two coarse boxes' flux FABs per direction (they share a face)
    cboxes = [((0, 0, 0), (3, 7, 7)), ((4, 0, 0), (7, 7, 7))]
    for d in range(3):
        for lo, hi in cboxes:
            fhi = list(hi)
            fhi[d] += 1
            h = rng.standard_normal(tuple(b - a + 1 for a, b in zip(lo, fhi)) + (NVAR,))
            o, g = _pair(h, lo, fhi)
            O.crse_init(o, d, -1.0)
            G.crse_init(g.t.data_ptr(), lo, fhi, d, -1.0)
    np.testing.assert_array_equal(G.data(), O.data())
    # FineAdd of two substeps for every fine box and direction
    for _ in range(2):
        for j, fb in enumerate(fine):
            for d in range(3):
                hi = list(fb[3:])
                hi[d] += 1
                h = rng.standard_normal(tuple(b - a + 1 for a, b in zip(fb[:3], hi)) + (NVAR,))
                o, g = _pair(h, fb[:3], hi)
                O.fine_add(j, o, d, 1.0)
                G.fine_add(j, g.t.data_ptr(), fb[:3], hi, d, 1.0)
    np.testing.assert_array_equal(G.data(), O.data())
    O.clear_internal_borders()
    G.clear_internal_borders()
    np.testing.assert_array_equal(G.data(), O.data())
    assert (O.data() == 0).any() and (O.data() != 0).any()
    # Reflux into a coarse state held as two boxes with ghost cells
    for lo, hi in cboxes:
        glo, ghi = tuple(v - 1 for v in lo), tuple(v + 1 for v in hi)
        hs = rng.random(tuple(b - a + 1 for a, b in zip(glo, ghi)) + (NVAR,)) + 1.0
        hv = np.full(tuple(b - a + 1 for a, b in zip(glo, ghi)) + (1,), 0.125)
        os_, gs = _pair(hs, glo, ghi)
        ov, gv = _pair(hv, glo, ghi, 1)
        O.reflux(os_, lo, hi, ov, 1.0)
        G.reflux(gs.t.data_ptr(), glo, ghi, lo, hi, gv.t.data_ptr(), glo, ghi, 1.0)
        got = gs.host()
        np.testing.assert_array_equal(got, os_.a)
        assert (got != hs).any()
    # average-down of a fine box into a coarse FAB
    fb = fine[2]
    hf = rng.random(tuple(fb[3:] - fb[:3] + 1) + (NVAR,))
    hc = rng.random((nc, nc, nc, NVAR))
    of, gf = _pair(hf, fb[:3], fb[3:])
    oc, gc = _pair(hc, (0, 0, 0), (nc - 1,) * 3)
    clo, chi = fb[:3] // 2, (fb[3:] + 1) // 2 - 1
    orc.average_down(of, oc, clo, chi, RATIO)
    amr.average_down(gf.t.data_ptr(), fb[:3], fb[3:], gc.t.data_ptr(), (0, 0, 0), (nc - 1,) * 3, clo, chi, RATIO)
    np.testing.assert_array_equal(gc.host(), oc.a)


def test_bad_register_arguments_are_refused():
    L = mc.lib()
    with pytest.raises(RuntimeError):
        amr.FluxRegister([[1, 0, 0, 8, 7, 7]], RATIO, (8, 8, 8), 1)    # not coarsenable
    with pytest.raises(RuntimeError):
        amr.FluxRegister([[0, 0, 0, 17, 7, 7]], RATIO, (8, 8, 8), 1)   # outside the refined domain
    R = amr.FluxRegister([[0, 0, 0, 7, 7, 7]], RATIO, (8, 8, 8), 1)
    small = Dev((0, 0, 0), (3, 7, 7), NVAR, np.zeros((4, 8, 8, NVAR)))
    with pytest.raises(RuntimeError):
        R.fine_add(0, small.t.data_ptr(), (0, 0, 0), (3, 7, 7), 0, 1.0)   # FAB lacks the box's faces
    assert b"FineAdd" in L.mc_last_error()


def test_two_level_sedov_matches_oracle_and_conserves():
    """Coarse 16^3 and fine 32^3 levels of one Sedov problem; the fine boxes
    covering fine cells [8, 23] x [8, 23] x [8, 15] are the refined region,
    which the blast crosses.  One coarse step, two fine half steps, then
    Castro::post_timestep's reflux, avgDown and clean_state: both levels'
    states equal the oracle's bit for bit, and the coarse level, holding the
    fine average under the region, keeps its total mass and energy."""
    from test_oracle_amr import two_level
    nc = 16
    Lc, Lf, R, region, totals = two_level(nc=nc, nthreads=8)
    cfg_c = small_cfg(n=nc, bc=1, r_init=0.15)
    cfg_c["boxes"] = configs.grid_boxes(nc, 8)
    cfg_f = small_cfg(n=2 * nc, bc=1, r_init=0.15)
    cfg_f["boxes"] = configs.grid_boxes(2 * nc, 8)
    Gc = mc.Level(cfg_c["n_cell"], cfg_c["boxes"], cfg_c["prob_lo"], cfg_c["prob_hi"], 1, mc.MODE_PLAIN)
    Gf = mc.Level(cfg_f["n_cell"], cfg_f["boxes"], cfg_f["prob_lo"], cfg_f["prob_hi"], 1, mc.MODE_PLAIN)
    Gc.init_data(**cfg_c)
    Gf.init_data(**cfg_f)
    Rg = amr.FluxRegister(cfg_f["boxes"][region], RATIO, (nc,) * 3, 1)
    amr.set_flux_registers(Gc, crse_side=Rg)
    amr.set_flux_registers(Gf, fine_side=Rg)
    amr.average_down_levels(Gf, Gc, RATIO, region)
    S0 = Gc.state()
    dt = Gc.initial_dt()
    Gc.advance(dt)
    for _ in range(2):
        Gf.advance(dt / 2)
        Gf.post_timestep()
    amr.reflux(Gc, Rg)
    amr.average_down_levels(Gf, Gc, RATIO, region)
    Gc.post_timestep()
    np.testing.assert_array_equal(Gf.state(), Lf.state())
    S1 = Gc.state()
    np.testing.assert_array_equal(S1, Lc.state())
    assert (Rg.data() == 0).all()
    for c in (0, 4):
        assert abs(S1[..., c].sum() - S0[..., c].sum()) <= 2e-14 * S0[..., c].sum(), c


@pytest.mark.parametrize("crse_mode,fine_mode,fine_slab,steps", [("plain", "fused", 0, 1), ("fused", "plain", 0, 1),
                                                                  ("fused", "fused", 0, 1), ("fused", "fused", 0, 2),
                                                                  ("plain", "fused", 5, 1)])
def test_two_level_sedov_with_fused_levels(crse_mode, fine_mode, fine_slab, steps):
    """The same two-level step with either level, or both, on the fused
    engine: the registers are fed from the engine's final fluxes, a fused
    coarse level stops its step after the advance's clean_state and takes
    post_timestep's clean after the reflux and average-down, and both levels
    still equal the oracle bit for bit, also over a second coarse step that
    starts from the refluxed, averaged and cleaned state.  A register is not
    fed from z-slabs."""
    from test_oracle_amr import two_level
    nc = 16
    Lc, Lf, R, region, totals = two_level(nc=nc, nthreads=8, steps=steps)
    mode = {"plain": mc.MODE_PLAIN, "fused": mc.MODE_FUSED}
    cfg_c = small_cfg(n=nc, bc=1, r_init=0.15)
    cfg_c["boxes"] = configs.grid_boxes(nc, 8)
    cfg_f = small_cfg(n=2 * nc, bc=1, r_init=0.15)
    cfg_f["boxes"] = configs.grid_boxes(2 * nc, 8)
    Gc = mc.Level(cfg_c["n_cell"], cfg_c["boxes"], cfg_c["prob_lo"], cfg_c["prob_hi"], 1, mode[crse_mode])
    Gf = mc.Level(cfg_f["n_cell"], cfg_f["boxes"], cfg_f["prob_lo"], cfg_f["prob_hi"], 1, mode[fine_mode])
    if fine_slab:
        Gf.set_fused_slab(fine_slab)
    Gc.init_data(**cfg_c)
    Gf.init_data(**cfg_f)
    Rg = amr.FluxRegister(cfg_f["boxes"][region], RATIO, (nc,) * 3, 1)
    amr.set_flux_registers(Gc, crse_side=Rg)
    amr.set_flux_registers(Gf, fine_side=Rg)
    amr.average_down_levels(Gf, Gc, RATIO, region)
    dt = Gc.initial_dt()
    if fine_slab:
        Gc.advance(dt)
        with pytest.raises(RuntimeError):
            Gf.advance(dt / 2)
        return
    for _ in range(steps):
        Gc.advance(dt)
        for _ in range(2):
            Gf.advance(dt / 2)
            Gf.post_timestep()
        amr.reflux(Gc, Rg)
        amr.average_down_levels(Gf, Gc, RATIO, region)
        Gc.post_timestep()
    np.testing.assert_array_equal(Gf.state(), Lf.state())
    np.testing.assert_array_equal(Gc.state(), Lc.state())
    assert (Rg.data() == 0).all()
    # a level with registers steps through advance / post_timestep
    for G in (Gc, Gf):
        with pytest.raises(RuntimeError):
            G.run(1)

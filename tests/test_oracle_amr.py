"""Pins of the oracle's AMR coupling (oracle/orc_amr.c): the flux register of
a fine level (Source/Castro_advance.cpp:82-88, Castro::reflux at
Source/Castro.cpp:451-486) and Castro::avgDown (Source/Castro.cpp:488-505).

What the mathematics fixes:
* average-down of a linear field gives the coarse cell-centre value, and it
  conserves the volume integral;
* FineAdd of a uniform fine flux adds ratio^2 times it per coarse face;
  CrseInit copies mult * flux onto the register faces a FAB holds and no
  others;
* ClearInternalBorders zeroes exactly the faces two fine boxes share (through
  the periodic wrap too);
* conservation of the composite solution: a coarse step, two fine substeps,
  reflux and average-down leave the total of every conserved quantity over
  the composite grid unchanged, with arbitrary fluxes on a periodic domain
  and with the real Sedov step's fluxes on an outflow domain.
"""
import numpy as np
import pytest

import orc
from minicastro import configs
from helpers import NVAR, oracle_table, small_cfg

RATIO = (2, 2, 2)


@pytest.fixture(scope="module", autouse=True)
def table():
    oracle_table()


def test_average_down_linear_exact_and_conservative():
    rng = np.random.default_rng(1)
    flo, fhi = (4, 2, 6), (11, 9, 13)   # coarse cells (2..5, 1..4, 3..6)
    x = [np.arange(flo[d], fhi[d] + 1) + 0.5 for d in range(3)]   # fine cell centres (fine units)
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    lin = 1.5 + 0.25 * X - 0.5 * Y + 0.125 * Z
    F = orc.Fab(flo, fhi, 2)
    F.a[..., 0] = lin
    F.a[..., 1] = rng.random(lin.shape)
    clo, chi = (2, 1, 3), (5, 4, 6)
    Cf = orc.Fab(clo, chi, 2)
    orc.average_down(F, Cf, clo, chi, RATIO)
    xc = [(np.arange(clo[d], chi[d] + 1) + 0.5) * 2 for d in range(3)]  # coarse centres in fine units
    Xc, Yc, Zc = np.meshgrid(*xc, indexing="ij")
    np.testing.assert_allclose(Cf.a[..., 0], 1.5 + 0.25 * Xc - 0.5 * Yc + 0.125 * Zc, rtol=1e-15)
    assert abs(Cf.a[..., 1].sum() * 8 - F.a[..., 1].sum()) <= 1e-13 * F.a[..., 1].sum()


def test_fine_add_and_crse_init():
    fine = [[4, 4, 4, 11, 7, 9]]            # coarse (2..5, 2..3, 2..4)
    R = orc.FluxRegister(fine, RATIO, (8, 8, 8), orc.BC_OUTFLOW)
    R.set_val(0.0)
    for d in range(3):
        lo, hi = [4, 4, 4], [11, 7, 9]
        hi[d] += 1
        f = orc.Fab(lo, hi, NVAR)
        f.a[...] = (d + 1) * 0.5 + np.arange(NVAR)
        R.fine_add(0, f, d, 1.0)
        R.fine_add(0, f, d, 1.0)   # a second substep adds again
    for (b, d, side), (lo, hi, a) in R.face_arrays().items():
        np.testing.assert_array_equal(a, 2 * 4 * ((d + 1) * 0.5 + np.arange(NVAR)) * np.ones(a.shape))
    # CrseInit from a coarse flux FAB covering only part of the low x face
    g = orc.Fab((2, 2, 2), (2, 3, 3), NVAR)
    g.a[...] = 3.0
    R.crse_init(g, 0, -1.0)
    lo, hi, a = R.face_arrays()[(0, 0, 0)]
    np.testing.assert_array_equal(a[:, :, :2], -3.0)       # k = 2, 3 held by the FAB: copied
    np.testing.assert_array_equal(a[:, :, 2], 2 * 4 * (0.5 + np.arange(NVAR)) + np.zeros((1, 2, NVAR)))
    lo, hi, a = R.face_arrays()[(0, 0, 1)]                 # high x face (i = 6): not in the FAB
    assert (a != -3.0).all()


def test_clear_internal_borders_shared_faces_only():
    # two fine boxes meeting at fine x = 8 (coarse face 4), and a periodic
    # domain where the second box touches the high end and the first the low end
    fine = [[0, 0, 0, 7, 7, 7], [8, 0, 0, 15, 3, 7]]
    for bc in (orc.BC_OUTFLOW, orc.BC_PERIODIC):
        # a box spanning the whole periodic z range meets itself
        Rz = orc.FluxRegister([[0, 0, 0, 3, 3, 15]], RATIO, (8, 8, 8), bc)
        Rz.set_val(1.0)
        Rz.clear_internal_borders()
        Fz = Rz.face_arrays()
        assert (Fz[(0, 2, 0)][2] == (0 if bc == orc.BC_PERIODIC else 1)).all()
        assert (Fz[(0, 2, 1)][2] == (0 if bc == orc.BC_PERIODIC else 1)).all()
        assert (Fz[(0, 0, 0)][2] == 1).all()
        R = orc.FluxRegister(fine, RATIO, (8, 8, 8), bc)
        R.set_val(1.0)
        R.clear_internal_borders()
        F = R.face_arrays()
        # box 0 high x face (coarse i = 4, j 0..3, k 0..3): shared with box 1's
        # low x face only where box 1 exists (j 0..1)
        a = F[(0, 0, 1)][2]
        assert (a[:, :2] == 0).all() and (a[:, 2:] == 1).all()
        b = F[(1, 0, 0)][2]
        assert (b == 0).all()
        # domain-boundary x faces: box 0 low (i = 0), box 1 high (i = 8):
        # shared through the periodic wrap on the overlap j 0..1
        a0, a1 = F[(0, 0, 0)][2], F[(1, 0, 1)][2]
        if bc == orc.BC_PERIODIC:
            assert (a0[:, :2] == 0).all() and (a0[:, 2:] == 1).all() and (a1 == 0).all()
        else:
            assert (a0 == 1).all() and (a1 == 1).all()
        # faces in y and z touch no other box (the boxes span half the
        # domain in z, so even the periodic wrap finds no partner there)
        for key in ((0, 1, 0), (1, 1, 1), (0, 2, 0), (1, 2, 1)):
            assert (F[key][2] == 1).all()


def _divergence_update(S, flux, lo, vol):
    """S -= (flux(i+1) - flux(i)) / vol per direction on a FAB of cells from lo."""
    for d in range(3):
        f = flux[d]
        sl_hi = [slice(None)] * 3
        sl_lo = [slice(None)] * 3
        sl_hi[d] = slice(1, None)
        sl_lo[d] = slice(0, -1)
        S -= (f[tuple(sl_hi)] - f[tuple(sl_lo)]) / vol


@pytest.mark.parametrize("bc", [orc.BC_PERIODIC, orc.BC_OUTFLOW])
def test_reflux_and_average_down_conserve_the_composite(bc):
    """Arbitrary coarse and fine face fluxes (already dt * area weighted): a
    conservative coarse update, two conservative fine substeps, CrseInit /
    FineAdd, ClearInternalBorders, Reflux and average-down leave the total of
    every component over the composite grid unchanged.  With outflow
    boundaries the fluxes through the domain boundary are zeroed so the total
    is closed."""
    rng = np.random.default_rng(2 + bc)
    nc = 8
    Vc, Vf = 1.0, 1.0 / 8
    fine = np.array([[4, 4, 2, 9, 11, 9], [10, 4, 2, 13, 7, 9]], dtype=np.int32)   # meet at fine x = 10
    S = rng.random((nc, nc, nc, NVAR)) + 1.0
    Sf = {}
    for b, fb in enumerate(fine):
        shp = tuple(fb[3:] - fb[:3] + 1)
        Sf[b] = rng.random(shp + (NVAR,)) + 1.0
    # average-down first (post_init), so the composite total is the coarse total
    Cfab = orc.Fab((0, 0, 0), (nc - 1,) * 3, NVAR, S)
    for b, fb in enumerate(fine):
        F = orc.Fab(fb[:3], fb[3:], NVAR, Sf[b])
        clo, chi = fb[:3] // 2, (fb[3:] + 1) // 2 - 1
        orc.average_down(F, Cfab, clo, chi, RATIO)
    total0 = Cfab.a.sum(axis=(0, 1, 2)) * Vc

    def face_field(n, d):
        shp = [n, n, n]
        shp[d] += 1
        g = rng.standard_normal(tuple(shp) + (NVAR,))
        if bc == orc.BC_PERIODIC:
            sl_a, sl_b = [slice(None)] * 3, [slice(None)] * 3
            sl_a[d], sl_b[d] = -1, 0
            g[tuple(sl_a)] = g[tuple(sl_b)]
        else:
            sl_a, sl_b = [slice(None)] * 3, [slice(None)] * 3
            sl_a[d], sl_b[d] = -1, 0
            g[tuple(sl_a)] = 0.0
            g[tuple(sl_b)] = 0.0
        return g
    G = [face_field(nc, d) for d in range(3)]
    gf = [[face_field(2 * nc, d) for d in range(3)] for _ in range(2)]   # two fine substeps

    R = orc.FluxRegister(fine, RATIO, (nc,) * 3, bc)
    R.set_val(0.0)
    # coarse step over one box holding the whole domain
    Sc = Cfab.a.copy()
    _divergence_update(Sc, G, (0, 0, 0), Vc)
    for d in range(3):
        hi = [nc - 1] * 3
        hi[d] = nc
        R.crse_init(orc.Fab((0, 0, 0), hi, NVAR, G[d]), d, -1.0)
    # fine substeps
    for s in range(2):
        for b, fb in enumerate(fine):
            sl = tuple(slice(fb[d], fb[3 + d] + 1) for d in range(3))
            fl = []
            for d in range(3):
                hi = list(fb[3:])
                hi[d] += 1
                slf = list(sl)
                slf[d] = slice(fb[d], fb[3 + d] + 2)
                a = np.asfortranarray(gf[s][d][tuple(slf)])
                fl.append(a)
                R.fine_add(b, orc.Fab(fb[:3], hi, NVAR, a), d, 1.0)
            _divergence_update(Sf[b], fl, fb[:3], Vf)
    Cfab2 = orc.Fab((0, 0, 0), (nc - 1,) * 3, NVAR, Sc)
    vol = orc.Fab((0, 0, 0), (nc - 1,) * 3, 1, np.full((nc,) * 3, Vc))
    R.clear_internal_borders()
    R.reflux(Cfab2, (0, 0, 0), (nc - 1,) * 3, vol, 1.0)
    for b, fb in enumerate(fine):
        F = orc.Fab(fb[:3], fb[3:], NVAR, Sf[b])
        orc.average_down(F, Cfab2, fb[:3] // 2, (fb[3:] + 1) // 2 - 1, RATIO)
    total1 = Cfab2.a.sum(axis=(0, 1, 2)) * Vc
    np.testing.assert_allclose(total1, total0, rtol=1e-13)
    # without the reflux the composite is not conserved
    Cfab3 = orc.Fab((0, 0, 0), (nc - 1,) * 3, NVAR, Sc)
    for b, fb in enumerate(fine):
        orc.average_down(orc.Fab(fb[:3], fb[3:], NVAR, Sf[b]), Cfab3, fb[:3] // 2, (fb[3:] + 1) // 2 - 1, RATIO)
    assert np.abs(Cfab3.a.sum(axis=(0, 1, 2)) * Vc - total0).max() > 1e-3


def two_level(nc=16, nthreads=4, steps=1, reflux=True):
    """A coarse level (nc^3, boxes of 8) and a fine level (2 nc^3, boxes of 8)
    of the same Sedov problem; the fine level's boxes covering fine cells
    [8, 23] x [8, 23] x [8, 15] form the refined region.  One coarse step of
    dt and two fine steps of dt/2 feed the region's flux register; then
    Castro::post_timestep's reflux, avgDown and clean_state."""
    cfg_c = small_cfg(n=nc, bc=1, r_init=0.15)
    cfg_c["boxes"] = configs.grid_boxes(nc, 8)
    cfg_f = small_cfg(n=2 * nc, bc=1, r_init=0.15)
    cfg_f["boxes"] = configs.grid_boxes(2 * nc, 8)
    Lc = orc.Level(cfg_c["n_cell"], cfg_c["prob_lo"], cfg_c["prob_hi"], 1, cfg_c["boxes"], nthreads)
    Lf = orc.Level(cfg_f["n_cell"], cfg_f["prob_lo"], cfg_f["prob_hi"], 1, cfg_f["boxes"], nthreads)
    for L, c in ((Lc, cfg_c), (Lf, cfg_f)):
        L.init_data(c["p_ambient"], c["dens_ambient"], c["exp_energy"], c["r_init"], c["nsub"])
    region = [b for b, bx in enumerate(cfg_f["boxes"])
              if bx[0] >= 8 and bx[3] <= 23 and bx[1] >= 8 and bx[4] <= 23 and bx[2] >= 8 and bx[5] <= 15]
    R = orc.FluxRegister(cfg_f["boxes"][region], RATIO, (nc,) * 3, orc.BC_OUTFLOW)
    Lc.set_flux_registers(crse_side=R)
    Lf.set_flux_registers(fine_side=R)
    Lc.average_down_from(Lf, RATIO, region)       # Castro::post_init
    totals = [_composite_totals(Lc)]
    dt = Lc.initial_dt()
    for _ in range(steps):
        Lc.advance(dt)
        for _ in range(2):
            Lf.advance(dt / 2)
            Lf.post_timestep()
        if reflux:
            Lc.reflux(R)                           # Castro::post_timestep, Source/Castro.cpp:354-362
        Lc.average_down_from(Lf, RATIO, region)
        Lc.post_timestep()
        totals.append(_composite_totals(Lc))
    return Lc, Lf, R, region, totals


def _composite_totals(Lc):
    S = Lc.state()
    return S[..., 0].sum(), S[..., 4].sum()


def test_two_level_sedov_conserves_composite_mass_and_energy():
    """The real step's fluxes (store_flux output) through the register: the
    coarse level, with the region's coarse cells replaced by the fine
    average, keeps its total mass and total energy to round-off (nothing
    crosses the outflow boundary in one step of the centred blast), and
    the reflux changed the coarse cells next to the region."""
    Lc, Lf, R, region, totals = two_level()
    (m0, e0), (m1, e1) = totals
    assert abs(m1 - m0) <= 1e-14 * m0 and abs(e1 - e0) <= 2e-14 * e0, (m0, m1, e0, e1)
    assert (R.data() == 0).all()    # reflux clears the register
    # the correction matters: without the reflux the composite loses mass
    _, _, _, _, t2 = two_level(reflux=False)
    assert abs(t2[1][0] - t2[0][0]) > 1e-9 * t2[0][0], t2

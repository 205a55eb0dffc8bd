"""Known-answer pins of the oracle's characteristic tracing, transverse
updates, source terms and species fluxes (oracle/orc_hydro.c restating
Source/ppm.F90:309-743, Source/trans.F90:14-505,
Source/Castro_hydro_nd.F90:248-297,344-441), and of the whole CTU step by
its order of accuracy.

Every expected value here comes from the mathematics of the equations the
reference discretises, not from its code:

* PPM tracing of a linear simple wave: a uniform state plus an
  eps-amplitude linear profile in one characteristic family (u-c, u, u+c,
  the internal-energy entropy mode, transverse velocities, a passive
  species) arrives at the face it travels to as the profile shifted by
  lambda dt/2, with no leakage into the other families (to O(eps^2)).
* Transverse updates of a uniform state by the fluxes of a uniform
  expansion (Godunov velocity linear across the zone): the density follows
  the mass flux, velocities are unchanged, the internal energy obeys the
  first law, d(rho e) = -(rho e + p) div u dt, the pressure the adiabatic
  law dp/p = gamma d rho/rho with the zone's gamma, and the composition is
  unchanged.
* fill_hydro_source of linear fluxes and a linear Godunov velocity:
  minus the flux gradient, and for rho e also minus p div u at the cell
  centre.
* normalize_species_fluxes: the species fluxes sum to the mass flux and
  keep their ratios.
* The whole step on periodic acoustic and entropy waves along the (1,1,1)
  diagonal converges at second order to the exact linear solution, whose
  sound speed comes from the EOS alone.
* A two-species composition profile is advected with the flow.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import orc
from helpers import NVAR, QVAR, NQAUX, NGDNV, NSPEC, oracle_table

QRHO, QU, QV, QW, QGAME, QPRES, QREINT, QTEMP, QFS = range(9)
QGAMC, QC = 0, 1
GDRHO, GDU, GDV, GDW, GDPRES, GDGAME = range(6)
URHO, UMX, UMY, UMZ, UEDEN, UEINT, UTEMP, UFS = range(8)


@pytest.fixture(scope="module", autouse=True)
def table():
    oracle_table()


def _axes(idir):
    """(normal, transverse, second transverse) velocity components of the
    trace in direction idir (Source/ppm.F90:350-366)."""
    return {1: (QU, QV, QW), 2: (QV, QW, QU), 3: (QW, QU, QV)}[idir]


# ------------------------------------------------------------ PPM tracing
def _trace_linear_wave(idir, mach, fams, eps=1e-6, n=8):
    """Trace a uniform state (rho 1, p 1, gamma 1.4, normal velocity
    mach * c) plus eps times a linear profile in the families `fams`, in
    direction idir.  Returns (qm, qp, expected(face x, side), families'
    speeds)."""
    rho0, p0, gam = 1.0, 1.0, 1.4
    c0 = math.sqrt(gam * p0 / rho0)
    rhoe0 = p0 / (gam - 1.0)
    h0 = (p0 + rhoe0) / rho0
    u0 = mach * c0
    vt0, vtt0 = 0.2, -0.1
    QUN, QUT, QUTT = _axes(idir)
    dx = 0.1
    dt = 0.4 * dx / (abs(u0) + c0)
    # right eigenvectors of the primitive system (rho, u_n, p, rho e, u_t,
    # u_tt, X) and their speeds
    R = {"u-c": ({QRHO: 1.0, QUN: -c0 / rho0, QPRES: c0 * c0, QREINT: h0}, u0 - c0),
         "u+c": ({QRHO: 1.0, QUN: c0 / rho0, QPRES: c0 * c0, QREINT: h0}, u0 + c0),
         "rho": ({QRHO: 1.0}, u0),
         "rhoe": ({QREINT: 1.0}, u0),
         "ut": ({QUT: 1.0}, u0),
         "utt": ({QUTT: 1.0}, u0),
         "X": ({QFS: -1.0, QFS + 1: 1.0}, u0)}
    base = {QRHO: rho0, QUN: u0, QUT: vt0, QUTT: vtt0, QPRES: p0, QREINT: rhoe0, QGAME: p0 / rhoe0 + 1.0,
            QTEMP: 1.0, QFS: 0.7, QFS + 1: 0.3}
    G = 4
    d = idir - 1
    qlo, qhi = [-G] * 3, [G] * 3
    qhi[d] = n - 1 + G
    q = orc.Fab(qlo, qhi, QVAR)
    qa = orc.Fab(qlo, qhi, NQAUX)
    # cell centres along idir (cell averages of linear data are the centre values)
    shape = q.a.shape[:3]
    idx = np.arange(qlo[d], qhi[d] + 1)
    xc = (idx + 0.5) * dx
    ell = lambda x: x / dx - 3.3
    bshape = [1, 1, 1]
    bshape[d] = len(idx)
    prof = np.broadcast_to(ell(xc).reshape(bshape), shape)
    for v, b in base.items():
        q.a[..., v] = b
    for f in fams:
        vec, _ = R[f]
        for v, r in vec.items():
            q.a[..., v] += eps * r * prof
    qa.a[..., QGAMC] = gam
    qa.a[..., QC] = c0
    lo, hi = [0, 0, 0], [0, 0, 0]
    hi[d] = n - 1
    flo, fhi = list(lo), list(hi)
    fhi[d] = n
    qm = orc.Fab(flo, fhi, QVAR)
    qp = orc.Fab(flo, fhi, QVAR)
    orc.call("trace_ppm", orc.iv(*lo), orc.iv(*hi), orc.iv(*lo), orc.iv(*hi), idir, q, qa, qm, qp,
             orc.iv(*lo), orc.iv(*hi), orc.dv(dx, dx, dx), dt)

    def expected(x_face):
        out = {v: np.full_like(x_face, b) for v, b in base.items()}
        for f in fams:
            vec, lam = R[f]
            for v, r in vec.items():
                out[v] = out[v] + eps * r * ell(x_face - lam * dt / 2.0)
        return out
    return qm, qp, expected, {f: R[f][1] for f in R}, dx


def _check_trace_side(qstate, sl, x_face, expected, eps):
    exp = expected(x_face)
    for v, e in exp.items():
        if v in (QGAME, QTEMP):
            continue  # not written by the trace (Source/ppm.F90:419-421)
        got = qstate.a[sl + (v,)].ravel()
        tol = 1e-4 * eps * 4.0 + 1e-14 * abs(e).max()
        np.testing.assert_allclose(got, e, rtol=0, atol=tol, err_msg="component %d" % v)


@pytest.mark.parametrize("idir", [1, 2, 3])
@pytest.mark.parametrize("mach", [-1.5, -0.3, 0.3, 1.5])
def test_trace_linear_simple_waves(idir, mach):
    """Each family arrives at the face it travels to shifted by lambda dt/2,
    alone and superposed with the others that travel the same way: the plus
    state (left face of a zone) takes the families with lambda < 0, the
    minus state (right face) those with lambda > 0."""
    eps = 1e-6
    _, _, _, speeds, _ = _trace_linear_wave(idir, mach, [])
    n = 8
    d = idir - 1
    for side in ("plus", "minus"):
        moving = [f for f, lam in speeds.items() if (lam < 0 if side == "plus" else lam > 0)]
        if "u-c" not in moving and side == "plus" or "u+c" not in moving and side == "minus":
            # the reference state (the u-c / u+c wave) itself moves away from
            # the face: nothing is traced to it (edge values)
            continue
        for fams in [[f] for f in moving] + [moving]:
            qm, qp, expected, _, dx = _trace_linear_wave(idir, mach, fams, eps=eps, n=n)
            sl = [slice(None)] * 3
            if side == "plus":      # zone i's state at face i
                sl[d] = slice(0, n)
                x_face = np.arange(0, n) * dx
                _check_trace_side(qp, tuple(sl), x_face, expected, eps)
            else:                   # zone i's state at face i+1
                sl[d] = slice(1, n + 1)
                x_face = np.arange(1, n + 1) * dx
                _check_trace_side(qm, tuple(sl), x_face, expected, eps)


def test_trace_supersonic_sees_every_family():
    """In supersonic flow toward a face every family is traced to it: the
    face state is the full linear solution (all seven families at once)."""
    for mach, side in ((-1.5, "plus"), (1.5, "minus")):
        fams = ["u-c", "u+c", "rho", "rhoe", "ut", "utt", "X"]
        qm, qp, expected, _, dx = _trace_linear_wave(1, mach, fams)
        if side == "plus":
            _check_trace_side(qp, (slice(0, 8), 0, 0), np.arange(0, 8) * dx, expected, 1e-6)
        else:
            _check_trace_side(qm, (slice(1, 9), 0, 0), np.arange(1, 9) * dx, expected, 1e-6)


# ------------------------------------------------------- transverse updates
def _expansion_fluxes(glo, ghi, ax, a, dx, rho0, vel, p0, E0, X, gam):
    """Godunov states and fluxes of a uniform expansion along axis ax: on
    the faces m = 0, 1 of zone 0 the Godunov velocity is a (m - 1/2) dx, the
    Godunov pressure p0; fluxes carry the zone's density, transverse
    velocities, total energy plus pressure and composition."""
    f = orc.Fab(glo, ghi, NVAR)
    g = orc.Fab(glo, ghi, NGDNV)
    qa = orc.Fab(glo, ghi, NQAUX)
    idx = np.arange(glo[ax], ghi[ax] + 1)
    ug = a * (idx - 0.5) * dx
    bshape = [1, 1, 1]
    bshape[ax] = len(idx)
    ug = np.broadcast_to(ug.reshape(bshape), f.a.shape[:3])
    g.a[..., GDPRES] = p0
    g.a[..., GDU + ax] = ug
    for t in range(3):
        g.a[..., GDU + t] = ug if t == ax else vel[t]
    f.a[..., URHO] = rho0 * ug
    for t in range(3):
        f.a[..., UMX + t] = rho0 * ug * (ug if t == ax else vel[t])
    f.a[..., UEDEN] = ug * (E0 + p0)
    for s in range(NSPEC):
        f.a[..., UFS + s] = rho0 * ug * X[s]
    qa.a[..., QGAMC] = gam
    return f, g, qa


def _uniform_state(lo, hi, rho0, vel, p0, rhoe0, X):
    q = orc.Fab(lo, hi, QVAR)
    q.a[..., QRHO] = rho0
    q.a[..., QU], q.a[..., QV], q.a[..., QW] = vel
    q.a[..., QPRES] = p0
    q.a[..., QREINT] = rhoe0
    q.a[..., QGAME] = p0 / rhoe0 + 1.0
    for s in range(NSPEC):
        q.a[..., QFS + s] = X[s]
    return q


def _check_expanded(o, theta, gam, rho0, vel, p0, rhoe0, X):
    s = o.a[0, 0, 0]
    rho = s[QRHO]
    np.testing.assert_allclose(rho, rho0 * (1.0 - theta), rtol=1e-14)
    # adiabatic: dp / p = gamma d rho / rho, with the zone's gamma
    np.testing.assert_allclose(s[QPRES] / p0 - 1.0, gam * (rho / rho0 - 1.0), rtol=1e-11)
    # first law: d(rho e) = -(rho e + p) div u dt
    np.testing.assert_allclose(s[QREINT], rhoe0 - theta * (rhoe0 + p0), rtol=1e-13)
    np.testing.assert_allclose([s[QU], s[QV], s[QW]], vel, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(s[QFS:QFS + NSPEC], X, rtol=1e-13, atol=1e-15)


EXPANSIONS = [(1.2, 0.7), (1.4, -0.4), (5.0 / 3.0, 0.25)]


@pytest.mark.parametrize("idir1,idir2", [(1, 2), (1, 3), (2, 1), (2, 3), (3, 1), (3, 2)])
def test_trans1_uniform_expansion_closed_form(idir1, idir2):
    """trans1 corrects the idir2-face states by the idir1 fluxes."""
    rho0, p0, rhoe0 = 1.3, 0.9, 2.1
    X = np.zeros(NSPEC)
    X[0], X[1] = 0.6, 0.4
    ax = idir1 - 1
    for gam, a in EXPANSIONS:
        vel = [0.3, -0.2, 0.15]
        vel[ax] = 0.0
        E0 = rhoe0 + 0.5 * rho0 * sum(v * v for v in vel)
        lo = hi = (0, 0, 0)
        q2 = _uniform_state(lo, hi, rho0, vel, p0, rhoe0, X)
        dx, cdtdx = 1.0, 0.12
        f1, q1, qa = _expansion_fluxes((-1, -1, -1), (1, 1, 1), ax, a, dx, rho0, vel, p0, E0, X, gam)
        outm, outp = orc.Fab(lo, hi, QVAR), orc.Fab(lo, hi, QVAR)
        orc.call("trans1", orc.iv(*lo), orc.iv(*hi), idir1, idir2, q2, outm, q2, outp, qa, f1, q1, cdtdx)
        theta = cdtdx * a * dx
        for o in (outm, outp):
            _check_expanded(o, theta, gam, rho0, vel, p0, rhoe0, X)


@pytest.mark.parametrize("idir1,idir2,idir3", [(1, 2, 3), (2, 1, 3), (3, 1, 2)])
def test_trans2_uniform_expansion_closed_form(idir1, idir2, idir3):
    """trans2 corrects the idir1-face states by the idir2 and idir3 fluxes
    (the two expansions add)."""
    rho0, p0, rhoe0 = 0.8, 1.7, 3.2
    X = np.zeros(NSPEC)
    X[0], X[2] = 0.25, 0.75
    for gam, a in EXPANSIONS:
        a2, a3 = a, -0.5 * a + 0.1
        vel = [0.0, 0.0, 0.0]
        vel[idir1 - 1] = 0.35
        E0 = rhoe0 + 0.5 * rho0 * sum(v * v for v in vel)
        lo = hi = (0, 0, 0)
        q = _uniform_state(lo, hi, rho0, vel, p0, rhoe0, X)
        dx, h2, h3 = 1.0, 0.07, 0.05
        f2, g2, qa = _expansion_fluxes((-1, -1, -1), (1, 1, 1), idir2 - 1, a2, dx, rho0, vel, p0, E0, X, gam)
        f3, g3, _ = _expansion_fluxes((-1, -1, -1), (1, 1, 1), idir3 - 1, a3, dx, rho0, vel, p0, E0, X, gam)
        outm, outp = orc.Fab(lo, hi, QVAR), orc.Fab(lo, hi, QVAR)
        orc.call("trans2", orc.iv(*lo), orc.iv(*hi), idir1, idir2, idir3, q, outm, q, outp, qa, f2, f3, g2, g3,
                 0.3, h2, h3)
        theta = h2 * a2 * dx + h3 * a3 * dx
        for o in (outm, outp):
            _check_expanded(o, theta, gam, rho0, vel, p0, rhoe0, X)


# ------------------------------------------------------------- source terms
def test_fill_hydro_source_closed_form():
    """Linear fluxes F_d(n) = s_dn x_d + b_dn give src_n = -sum_d s_dn; a
    Godunov velocity u_d = w_d x_d with a linear Godunov pressure adds
    -p div u at the cell centre to rho e (Source/Castro_hydro_nd.F90:425-434).
    Cells are not cubes, so every direction's dx is used."""
    rng = np.random.default_rng(4)
    lo, hi = (0, 0, 0), (3, 2, 4)
    dx = np.array([0.1, 0.25, 0.4])
    slopes = rng.uniform(-2, 2, (3, NVAR))
    offs = rng.uniform(-1, 1, (3, NVAR))
    w = np.array([0.7, -1.3, 0.4])
    P0, P1 = 2.0, np.array([0.5, -0.3, 0.8])
    fl, qe, ar = [], [], []
    for d in range(3):
        flo, fhi = list(lo), list(hi)
        fhi[d] += 1
        F = orc.Fab(flo, fhi, NVAR)
        Q = orc.Fab(flo, fhi, NGDNV)
        A = orc.Fab(flo, fhi, 1)
        shape = F.a.shape[:3]
        xf = np.arange(flo[d], fhi[d] + 1) * dx[d]   # face coordinates along d
        bsh = [1, 1, 1]
        bsh[d] = len(xf)
        xf = np.broadcast_to(xf.reshape(bsh), shape)
        for nn in range(NVAR):
            F.a[..., nn] = slopes[d, nn] * xf + offs[d, nn]
        Q.a[..., GDU + d] = w[d] * xf
        Q.a[..., GDPRES] = P0 + P1[d] * xf
        A.a[...] = np.prod(dx) / dx[d]
        fl.append(F)
        qe.append(Q)
        ar.append(A)
    vol = orc.Fab(lo, hi, 1)
    vol.a[...] = np.prod(dx)
    src = orc.Fab(lo, hi, NVAR)
    U = orc.Fab(lo, hi, NVAR)
    q = orc.Fab(lo, hi, QVAR)
    orc.call("fill_hydro_source", orc.iv(*lo), orc.iv(*hi), U, q, src, fl[0], fl[1], fl[2], qe[0], qe[1], qe[2],
             ar[0], ar[1], ar[2], vol, orc.dv(*dx), 0.01)
    shape = src.a.shape[:3]
    xc = [np.broadcast_to(((np.arange(lo[d], hi[d] + 1) + 0.5) * dx[d]).reshape([-1 if e == d else 1
                                                                              for e in range(3)]), shape)
          for d in range(3)]
    pdivu = sum((P0 + P1[d] * xc[d]) * w[d] for d in range(3))
    for nn in range(NVAR):
        exp = -(slopes[0, nn] + slopes[1, nn] + slopes[2, nn]) + np.zeros(shape)
        if nn == UEINT:
            exp = exp - pdivu
        np.testing.assert_allclose(src.a[..., nn], exp, rtol=1e-11, atol=1e-11, err_msg="component %d" % nn)


def test_normalize_species_fluxes_sum_to_mass_flux():
    """After normalize_species_fluxes the species fluxes sum to the mass
    flux and keep their ratios; a zero species sum leaves them unchanged."""
    rng = np.random.default_rng(8)
    lo, hi = (0, 0, 0), (3, 2, 1)
    f = orc.Fab(lo, hi, NVAR)
    f.a[...] = rng.uniform(0.1, 2.0, f.a.shape)
    f.a[..., URHO] = rng.uniform(-3.0, 3.0, f.a.shape[:3])
    f.a[0, 0, 0, UFS:] = 0.0
    before = f.a.copy()
    orc.call("normalize_species_fluxes", orc.iv(*lo), orc.iv(*hi), f)
    sp = f.a[..., UFS:UFS + NSPEC]
    mask = np.ones(f.a.shape[:3], bool)
    mask[0, 0, 0] = False
    np.testing.assert_allclose(sp.sum(axis=-1)[mask], f.a[..., URHO][mask], rtol=1e-13)
    ratio = sp[mask] / before[..., UFS:UFS + NSPEC][mask]
    np.testing.assert_allclose(ratio, ratio[:, :1] * np.ones((1, NSPEC)), rtol=1e-14)
    np.testing.assert_array_equal(f.a[0, 0, 0], before[0, 0, 0])
    np.testing.assert_array_equal(f.a[..., :UFS], before[..., :UFS])


# ----------------------------------------------------- order of accuracy
def _quirk_sound_speed(rho0, rhoe0):
    """Sound speed of the gas the step evolves, from the EOS alone.  After
    the first clean_state the species slots hold mass fractions, so the EOS
    is called with abar = 4 rho, zbar = 2 (Source/Castro_nd.F90:152-154,
    285-286): p is a function P(rho e, rho).  Along an adiabat
    d(rho e) = h d rho with h = (rho e + p) / rho, so c^2 = dP/d rho there,
    taken by a centred difference."""
    def p_of(rho, rhoe):
        st = orc.eos(orc.EOS_RE, rho, 1e5, 4.0 * rho, 2.0, e=rhoe / rho)
        return st.p
    p0 = p_of(rho0, rhoe0)
    h = (rhoe0 + p0) / rho0
    dr = 1e-4 * rho0
    c2 = (p_of(rho0 + dr, rhoe0 + h * dr) - p_of(rho0 - dr, rhoe0 - h * dr)) / (2 * dr)
    return math.sqrt(c2), p0


def _diagonal_wave_error(n, kind, eps=1e-5, nthreads=None):
    """L1 error of the density after the wave has moved a fixed distance
    along (1,1,1), relative to eps rho0, on an n^3 periodic unit cube."""
    rho0, T0 = 1.0, 1.6e5
    st = orc.eos(orc.EOS_RT, rho0, T0, 4.0 * rho0, 2.0)
    rhoe0 = rho0 * st.e
    c0, p0 = _quirk_sound_speed(rho0, rhoe0)
    nhat = np.ones(3) / math.sqrt(3.0)
    kvec = 2.0 * math.pi * np.ones(3)
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    damp = np.prod([math.sin(k * dx / 2) / (k * dx / 2) for k in kvec])   # cell averages of a sine

    def s(t, speed):
        return damp * np.sin(kvec[0] * (X - speed * t * nhat[0]) + kvec[1] * (Y - speed * t * nhat[1]) +
                             kvec[2] * (Z - speed * t * nhat[2]))
    S = np.zeros((n, n, n, NVAR), order="F")
    if kind == "acoustic":
        u0, speed = 0.0, c0
        h = (rhoe0 + p0) / rho0
        ph = s(0.0, speed)
        rho = rho0 * (1.0 + eps * ph)
        un = eps * c0 * ph
        rhoe = rhoe0 + h * (rho - rho0)
    else:
        u0 = c0
        speed = u0
        rho = rho0 * (1.0 + eps * s(0.0, speed))
        un = u0 + 0.0 * rho
        rhoe = rhoe0 + 0.0 * rho
    S[..., URHO] = rho
    for d in range(3):
        S[..., UMX + d] = rho * un * nhat[d]
    S[..., UEINT] = rhoe
    S[..., UEDEN] = rhoe + 0.5 * rho * un * un
    S[..., UTEMP] = T0
    S[..., UFS] = rho
    # a fixed distance: the wave crosses a sixth of its wavelength 1/sqrt(3)
    T = (1.0 / math.sqrt(3.0)) / 6.0 / speed
    lam_max = c0 + abs(u0) * nhat[0]
    nsteps = int(math.ceil(T / (0.4 * dx / lam_max)))
    dt = T / nsteps
    nthreads = nthreads or max(1, min(8, os.cpu_count() or 1))
    L = orc.Level((n, n, n), (0, 0, 0), (1, 1, 1), orc.BC_PERIODIC, [[0, 0, 0, n - 1, n - 1, n - 1]], nthreads)
    L.set_state(S)
    for _ in range(nsteps):
        L.advance(dt)
        L.post_timestep()
    got = L.state()[..., URHO]
    exact = rho0 * (1.0 + eps * s(T, speed))
    return float(np.abs(got - exact).mean() / (eps * rho0))


def test_diagonal_entropy_wave_converges_at_second_order():
    """A density wave at uniform pressure and velocity along (1,1,1) is an
    exact solution of the equations the step discretises; the error falls
    by about 4 per doubling of the resolution (12^3 -> 24^3 -> 48^3).
    Tracing, both transverse stages and the flux divergence are all active
    along the diagonal, and an O(1) slip in any of them costs the second
    order."""
    kind = "entropy"
    e = [_diagonal_wave_error(n, kind) for n in (12, 24, 48)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    print(kind, "errors", e, "ratios", r1, r2)
    assert e[2] < 5e-3, e
    assert r1 > 3.0 and r2 > 3.4, (e, r1, r2)


def test_two_species_composition_is_advected():
    """A He / C composition profile X_C = 0.5 + 0.3 sin(2 pi x) at uniform
    density 1, uniform pressure (each cell's internal energy set by the EOS
    for its own composition) and uniform velocity is a contact wave: after
    the flow has crossed a quarter of the domain the profile is the shifted
    one, the mass fractions sum to one, density and pressure stay uniform,
    and the error falls at second order with resolution."""
    aion = np.array([4.0, 12.0])
    zion = np.array([2.0, 6.0])
    errs = []
    for n in (16, 32):
        rho0, T0 = 1.0, 1.6e5
        p0 = orc.eos(orc.EOS_RT, rho0, T0, 4.0, 2.0).p
        dx = 1.0 / n
        x = (np.arange(n) + 0.5) * dx
        damp = math.sin(math.pi * dx) / (math.pi * dx)
        xc = 0.5 + 0.3 * damp * np.sin(2 * math.pi * x)
        e = np.zeros(n)
        cmax = 0.0
        for i in range(n):
            X = np.array([1.0 - xc[i], xc[i]])
            abar = 1.0 / (X / aion).sum()
            zbar = abar * (X * zion / aion).sum()
            st = orc.eos(orc.EOS_RP, rho0, T0, abar, zbar, p=p0)
            e[i] = st.e
            cmax = max(cmax, st.cs)
        u0 = 0.5 * cmax
        S = np.zeros((n, 4, 4, NVAR), order="F")
        S[..., URHO] = rho0
        S[..., UMX] = rho0 * u0
        S[..., UEINT] = (rho0 * e)[:, None, None]
        S[..., UEDEN] = (rho0 * e + 0.5 * rho0 * u0 * u0)[:, None, None]
        S[..., UTEMP] = T0
        S[..., UFS] = rho0 * (1.0 - xc)[:, None, None]
        S[..., UFS + 1] = rho0 * xc[:, None, None]
        T = 0.25 / u0
        nsteps = int(math.ceil(T / (0.4 * dx / (u0 + cmax))))
        dt = T / nsteps
        L = orc.Level((n, 4, 4), (0, 0, 0), (1, 4.0 / n, 4.0 / n), orc.BC_PERIODIC, [[0, 0, 0, n - 1, 3, 3]], 4)
        L.set_state(S)
        for _ in range(nsteps):
            L.advance(dt)
            L.post_timestep()
        R = L.state()
        Xs = R[..., UFS:UFS + NSPEC]
        np.testing.assert_allclose(Xs.sum(axis=-1), 1.0, rtol=1e-14)
        exact = 0.5 + 0.3 * damp * np.sin(2 * math.pi * (x - u0 * T))
        got = R[:, 0, 0, UFS + 1]
        assert np.ptp(R[..., UFS + 1], axis=(1, 2)).max() == 0.0   # stays one-dimensional
        errs.append(np.abs(got - exact).mean() / 0.3)
        np.testing.assert_allclose(R[..., URHO], rho0, rtol=1e-4)
        np.testing.assert_allclose(R[..., UMX] / R[..., URHO], u0, rtol=1e-4)
    print("composition errors", errs)
    assert errs[1] < 0.01 and errs[0] / errs[1] > 3.2, errs

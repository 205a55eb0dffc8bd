"""Pins for the oracle EOS (oracle/orc_eos.c, restating Source/eos.F90).

* closed form: with the generated table (ideal non-degenerate electrons) the
  EOS must reproduce textbook ideal gas + radiation thermodynamics written out
  independently here (pressure, energy, c_v, dp/dT, dp/drho, Gamma_1);
* exactness: the biquintic / bicubic Hermite interpolation reproduces
  polynomials of degree <= 5 / <= 3 in each variable, so a table filled from
  such a polynomial must be interpolated to round-off at any point of a cell
  (every weight, every one of the 36 + 16 terms and every derivative order is
  exercised);
* Newton: (rho, e) and (rho, p) solves invert the (rho, T) call;
* text table == binary cache.
"""
import numpy as np
import pytest

import orc
from helpers import oracle_table, table_array, table_bin
import minicastro as mc

AVO = 6.0221417930e23
KERG = 1.380650424e-16
CLIGHT = 2.99792458e10
ARAD = 4.0 * 5.67051e-5 / CLIGHT


def ideal(rho, T, abar, zbar):
    """Ions + non-degenerate electrons + radiation, from first principles."""
    ye = zbar / abar
    n_i = rho * AVO / abar          # ions per cm^3
    n_e = ye * rho * AVO            # electrons per cm^3
    p = (n_i + n_e) * KERG * T + ARAD * T ** 4 / 3.0
    e = 1.5 * (n_i + n_e) * KERG * T / rho + ARAD * T ** 4 / rho
    cv = 1.5 * (n_i + n_e) * KERG / rho + 4.0 * ARAD * T ** 3 / rho
    dpdT = (n_i + n_e) * KERG + 4.0 * ARAD * T ** 3 / 3.0
    dpdr = (n_i + n_e) * KERG * T / rho
    chit = T / p * dpdT
    chid = rho / p * dpdr
    gam1 = chit ** 2 * p / (rho * T * cv) + chid
    return dict(p=p, e=e, cv=cv, dpdT=dpdT, dpdr=dpdr, gam1=gam1, cs=np.sqrt(gam1 * p / rho))


@pytest.fixture(scope="module", autouse=True)
def table():
    oracle_table()


@pytest.mark.parametrize("abar_scale", [1.0, "rho"])
def test_ideal_gas_closed_form(abar_scale):
    rng = np.random.default_rng(1)
    worst = {}
    for _ in range(300):
        rho = 10.0 ** rng.uniform(-4, 5)
        T = 10.0 ** rng.uniform(3.5, 9.5)
        abar = 4.0 * (rho if abar_scale == "rho" else 1.0)
        zbar = 2.0
        if not (1e-10 < zbar / abar * rho < 1e13):
            continue
        st = orc.eos(orc.EOS_RT, rho, T, abar, zbar)
        ref = ideal(rho, T, abar, zbar)
        for k in ("p", "e", "dpdT", "dpdr", "gam1", "cs"):
            got = {"p": st.p, "e": st.e, "dpdT": st.dpdT, "dpdr": st.dpdr, "gam1": st.gam1, "cs": st.cs}[k]
            worst[k] = max(worst.get(k, 0.0), abs(got / ref[k] - 1.0))
        worst["cv"] = max(worst.get("cv", 0.0), abs(st.cv / ref["cv"] - 1.0))
    for k, v in worst.items():
        assert v < 2e-6, (k, v)


def _poly_table(iat, jat, cd, ct, P, Q):
    """Table whose f-columns are the polynomial P(x, y) and dpdf-columns Q(x, y),
    x = (d - d0)/sd, y = (T - t0)/st around node (iat, jat) (1-based)."""
    t, d = orc.eos_grid()
    tab = table_array().copy()
    d0, t0 = d[iat - 1], t[jat - 1]
    sd, st = d[iat] - d0, t[jat] - t0
    cols_f = ["f", "fd", "ft", "fdd", "ftt", "fdt", "fddt", "fdtt", "fddtt"]
    order = {"f": (0, 0), "fd": (1, 0), "ft": (0, 1), "fdd": (2, 0), "ftt": (0, 2), "fdt": (1, 1), "fddt": (2, 1),
             "fdtt": (1, 2), "fddtt": (2, 2)}
    cols_q = {"dpdf": (0, 0), "dpdfd": (1, 0), "dpdft": (0, 1), "dpdfdt": (1, 1)}
    idx = {"f": 0, "fd": 1, "ft": 2, "fdd": 3, "ftt": 4, "fdt": 5, "fddt": 6, "fdtt": 7, "fddtt": 8,
           "dpdf": 9, "dpdfd": 10, "dpdft": 11, "dpdfdt": 12}
    from numpy.polynomial import polynomial as Pn
    for jj in (jat, jat + 1):
        for ii in (iat, iat + 1):
            x = (d[ii - 1] - d0) / sd
            y = (t[jj - 1] - t0) / st
            for name in cols_f:
                a, b = order[name]
                c = Pn.polyder(Pn.polyder(P, m=a, axis=0), m=b, axis=1) if (a or b) else P
                tab[idx[name], jj - 1, ii - 1] = Pn.polyval2d(x, y, c) / (sd ** a * st ** b)
            for name, (a, b) in cols_q.items():
                c = Pn.polyder(Pn.polyder(Q, m=a, axis=0), m=b, axis=1) if (a or b) else Q
                tab[idx[name], jj - 1, ii - 1] = Pn.polyval2d(x, y, c) / (sd ** a * st ** b)
    return tab, (d0, sd, t0, st)


def test_hermite_reproduces_polynomials():
    """Biquintic (f) and bicubic (dpdf) interpolation are exact on polynomials."""
    from numpy.polynomial import polynomial as Pn
    rng = np.random.default_rng(2)
    iat, jat = 234, 12   # density row used by ye*rho = 0.5, a low-temperature column
    P = rng.uniform(-1, 1, (6, 6)) * 1e13
    Q = rng.uniform(0.5, 1, (4, 4)) * 1e9
    tab, (d0, sd, t0, st) = _poly_table(iat, jat, None, None, P, Q)
    orc.eos_init_arrays(tab)
    try:
        for _ in range(40):
            x, y = rng.uniform(0.02, 0.98, 2)
            din = d0 + x * sd
            T = t0 + y * st
            abar, zbar = 4.0e6, 2.0e6          # ye = 0.5; tiny ion pressure
            rho = din / 0.5
            s = orc.eos(orc.EOS_RT, rho, T, abar, zbar)
            assert (s.iat, s.jat) == (iat, jat)
            ye = 0.5
            # remove the closed-form radiation and ion parts
            prad = ARAD * T ** 4 / 3.0
            pion = KERG * AVO / abar * rho * T
            erad, eion = 3.0 * prad / rho, 1.5 * pion / rho
            f = Pn.polyval2d(x, y, P)
            fd = Pn.polyval2d(x, y, Pn.polyder(P, axis=0)) / sd
            ft = Pn.polyval2d(x, y, Pn.polyder(P, axis=1)) / st
            ftt = Pn.polyval2d(x, y, Pn.polyder(P, m=2, axis=1)) / st ** 2
            fdt = Pn.polyval2d(x, y, Pn.polyder(Pn.polyder(P, axis=0), axis=1)) / (sd * st)
            q = Pn.polyval2d(x, y, Q)
            scale_p = din ** 2 * abs(fd) + 1e-300
            assert abs((s.p - prad - pion) - din ** 2 * fd) < 1e-9 * scale_p
            assert abs((s.e - erad - eion) - ye * (f - T * ft)) < 1e-9 * ye * (abs(f) + abs(T * ft))
            dedT_e = s.dedT - 4.0 * ARAD * T ** 3 / rho - 1.5 * pion / rho / T
            assert abs(dedT_e - (-ye * T * ftt)) < 1e-8 * abs(ye * T * ftt)
            dpdT_e = s.dpdT - 4.0 * prad / T - pion / T
            assert abs(dpdT_e - din ** 2 * fdt) < 1e-8 * abs(din ** 2 * fdt)
            dpdr_e = s.dpdr - KERG * AVO / abar * T
            assert abs(dpdr_e - max(ye * q, 0.0)) < 1e-9 * abs(ye * q)
    finally:
        orc.eos_init_arrays(table_array())


def test_newton_inverts_rt():
    rng = np.random.default_rng(3)
    for _ in range(200):
        rho = 10.0 ** rng.uniform(-3, 4)
        T = 10.0 ** rng.uniform(3.5, 9.0)
        abar, zbar = 4.0 * rho, 2.0
        e = orc.eos(orc.EOS_RT, rho, T, abar, zbar).e
        p = orc.eos(orc.EOS_RT, rho, T, abar, zbar).p
        g = T * 10.0 ** rng.uniform(-2, 2)
        s = orc.eos(orc.EOS_RE, rho, g, abar, zbar, e=e)
        assert abs(s.T / T - 1) < 1e-7 and s.e == e and s.niter < 100
        s = orc.eos(orc.EOS_RP, rho, g, abar, zbar, p=p)
        assert abs(s.T / T - 1) < 1e-7 and s.p == p


def test_newton_floor_at_mintemp():
    """e below e(T = 1e3 K) drives the solve onto the mintemp floor (eos.F90:543)."""
    rho, abar, zbar = 1.0, 4.0, 2.0
    e_floor = orc.eos(orc.EOS_RT, rho, 1.0e3, abar, zbar).e
    s = orc.eos(orc.EOS_RE, rho, 1.0e5, abar, zbar, e=0.5 * e_floor)
    assert s.T == 1.0e3


def test_hash_index_brackets_temperature():
    t, d = orc.eos_grid()
    rng = np.random.default_rng(4)
    for T in 10.0 ** rng.uniform(3.0, 12.9, 500):
        s = orc.eos(orc.EOS_RT, 1.0, T, 4.0, 2.0)
        j = s.jat
        assert t[j - 1] * (1 - 1e-14) <= T < t[j] * (1 + 1e-14)
    # ye*rho = 0.5 for the cleaned single-species state sits in row 234
    assert orc.eos(orc.EOS_RT, 0.25, 1e6, 1.0, 2.0).iat == 234


def test_text_table_equals_binary():
    import os
    txt = mc.table_path("dat")
    if not os.path.exists(txt):
        pytest.skip("text table not generated")
    pts = [(1.0, 1e5), (3.0, 2e7), (0.1, 4e9), (1e3, 1e4)]
    ref = [orc.eos(orc.EOS_RT, r, T, 4.0, 2.0) for r, T in pts]
    orc.eos_init(txt)
    try:
        got = [orc.eos(orc.EOS_RT, r, T, 4.0, 2.0) for r, T in pts]
    finally:
        orc.eos_init_arrays(table_array())
    for a, b in zip(got, ref):
        assert (a.p, a.e, a.cs, a.gam1) == (b.p, b.e, b.cs, b.gam1)


def test_generated_table_closed_form_columns():
    """Spot-check generated columns against the ideal electron gas free energy."""
    tab = table_array()
    t, d = orc.eos_grid()
    A = KERG * AVO
    for (i, j) in ((0, 0), (233, 40), (540, 200), (100, 150)):
        T, dd = t[j], d[i]
        np.testing.assert_allclose(tab[1, j, i], A * T / dd, rtol=1e-13)     # fd
        np.testing.assert_allclose(tab[4, j, i], -1.5 * A / T, rtol=1e-13)   # ftt
        np.testing.assert_allclose(tab[9, j, i], A * T, rtol=1e-13)          # dpdf
        # f - T ft = 1.5 A T: internal energy per gram / ye
        np.testing.assert_allclose(tab[0, j, i] - T * tab[2, j, i], 1.5 * A * T, rtol=1e-9)

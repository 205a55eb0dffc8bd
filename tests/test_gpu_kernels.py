"""Per-kernel parity of the HIP drop-in entry points against the oracle.

Each kernel is called through the C ABI (libmccastro.so, the symbols of
include/mc_castro.h) on device copies of the same inputs the oracle gets;
inputs of later stages are the oracle's outputs of earlier stages, so every
kernel is checked in isolation.  The bar is bit-identical output (both sides
evaluate the same expressions without FMA contraction).
"""
import ctypes as C

import numpy as np
import pytest

import orc
import minicastro as mc
from helpers import (NVAR, QVAR, NQAUX, NGDNV, NSPEC, Dev, gpu_call, oracle_level, oracle_table, small_cfg,
                     table_bin)

pytestmark = pytest.mark.gpu

QRHO, QU, QV, QW, QGAME, QPRES, QREINT, QTEMP, QFS = range(9)
# components a traced / transverse-corrected state defines (QGAME, QTEMP are never written)
QDEF = [QRHO, QU, QV, QW, QPRES, QREINT] + list(range(QFS, QFS + NSPEC))


def grow(lo, hi, g):
    return tuple(l - g for l in lo), tuple(h + g for h in hi)


def grow3(lo, hi, g):
    return tuple(l - a for l, a in zip(lo, g)), tuple(h + a for h, a in zip(hi, g))


def nodes(lo, hi, d):
    hi = list(hi)
    hi[d] += 1
    return tuple(lo), tuple(hi)


def sub(dom, lo, hi):
    """Extract [lo, hi] from a domain array with outflow (clamp) extension."""
    n = dom.shape[:3]
    idx = [np.clip(np.arange(lo[d], hi[d] + 1), 0, n[d] - 1) for d in range(3)]
    return np.asfortranarray(dom[np.ix_(idx[0], idx[1], idx[2])])


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    return make_setup()


def make_setup():
    oracle_table()
    cfg = small_cfg(n=16, r_init=0.15)
    L = oracle_level(cfg)
    for _ in range(4):
        L.coarse_step()
    S = L.state()
    rng = np.random.default_rng(11)
    # perturb so every limiter / Riemann branch is visited
    S = S.copy()
    S[..., 0] *= 1.0 + 0.2 * rng.random(S.shape[:3])
    S[..., 1:4] += 0.3 * np.sqrt(S[..., 5:6] / S[..., 0:1]) * S[..., 0:1] * rng.standard_normal(S.shape[:3] + (3,))
    S[..., 5] *= 1.0 + 0.3 * rng.random(S.shape[:3])
    S[..., 4] = S[..., 5] + 0.5 * (S[..., 1] ** 2 + S[..., 2] ** 2 + S[..., 3] ** 2) / S[..., 0]
    S[..., 7] = S[..., 0]
    S[..., 6] *= 1.0 + 0.1 * rng.random(S.shape[:3])
    bx = ((3, 4, 2), (12, 10, 11))
    dt = 0.3 * L.dt * 20
    dx = (1.0 / 16,) * 3
    return dict(S=S, bx=bx, dt=dt, dx=dx, cfg=cfg)


def test_clean_state_kernels(setup):
    lo, hi = grow(*setup["bx"], 4)
    U = sub(setup["S"], lo, hi)
    # plant cells that take the rare branches: negative E, tiny rho, small e
    U[0, 0, 0, 4] = -abs(U[0, 0, 0, 4])
    U[1, 0, 0, 4] = -abs(U[1, 0, 0, 4]); U[1, 0, 0, 5] = -1.0
    U[2, 0, 0, 0] = 1e-14
    rho3 = U[3, 0, 0, 0]
    U[3, 0, 0, 1] = 1e3 * rho3
    U[3, 0, 0, 4] = 0.5 * U[3, 0, 0, 1] ** 2 / rho3 * (1.0 + 1e-7)
    U[3, 0, 0, 5] = -1.0
    for name in ("enforce_minimum_density", "normalize_species", "reset_internal_e", "compute_temp"):
        o = orc.Fab(lo, hi, NVAR, U)
        g = Dev(lo, hi, NVAR, U)
        orc.call(name, orc.iv(*lo), orc.iv(*hi), o)
        gpu_call(name, orc.iv(*lo), orc.iv(*hi), g)
        np.testing.assert_array_equal(g.host(), o.a, err_msg=name)
        U = o.a.copy()


def test_ctoprim_divu(setup):
    lo, hi = grow(*setup["bx"], 4)
    U = sub(setup["S"], lo, hi)
    o_u = orc.Fab(lo, hi, NVAR, U)
    o_q, o_qa = orc.Fab(lo, hi, QVAR), orc.Fab(lo, hi, NQAUX)
    orc.call("ctoprim", orc.iv(*lo), orc.iv(*hi), o_u, o_q, o_qa)
    g_u, g_q, g_qa = Dev(lo, hi, NVAR, U), Dev(lo, hi, QVAR), Dev(lo, hi, NQAUX)
    gpu_call("ctoprim", orc.iv(*lo), orc.iv(*hi), g_u, g_q, g_qa)
    np.testing.assert_array_equal(g_q.host(), o_q.a)
    np.testing.assert_array_equal(g_qa.host(), o_qa.a)
    olo, ohi = grow(*setup["bx"], 1)
    o_d, g_d = orc.Fab(olo, ohi, 1), Dev(olo, ohi, 1)
    dx = orc.dv(*setup["dx"])
    orc.call("divu", orc.iv(*olo), orc.iv(*ohi), o_q, dx, o_d)
    gpu_call("divu", orc.iv(*olo), orc.iv(*ohi), Dev(lo, hi, QVAR, o_q.a), dx, g_d)
    np.testing.assert_array_equal(g_d.host(), o_d.a)


def _prim(setup):
    lo, hi = grow(*setup["bx"], 4)
    U = sub(setup["S"], lo, hi)
    o_q, o_qa = orc.Fab(lo, hi, QVAR), orc.Fab(lo, hi, NQAUX)
    orc.call("ctoprim", orc.iv(*lo), orc.iv(*hi), orc.Fab(lo, hi, NVAR, U), o_q, o_qa)
    return (lo, hi), o_q, o_qa


def _tbx(bx):
    lo, hi = bx
    e = [nodes(lo, hi, d) for d in range(3)]
    g = {(0, 0): (0, 1, 1), (0, 1): (0, 0, 1), (0, 2): (0, 1, 0), (1, 0): (0, 0, 1), (1, 1): (1, 0, 1),
         (1, 2): (1, 0, 0), (2, 0): (0, 1, 0), (2, 1): (1, 0, 0), (2, 2): (1, 1, 0)}
    return {k: grow3(*e[k[0]], v) for k, v in g.items()}


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_trace_ppm(setup, idir):
    qb, o_q, o_qa = _prim(setup)
    bx = setup["bx"]
    olo, ohi = grow(*bx, 1)
    tb = _tbx(bx)[(idir - 1, idir - 1)]
    o_qm, o_qp = orc.Fab(*tb, QVAR), orc.Fab(*tb, QVAR)
    args = (orc.iv(*olo), orc.iv(*ohi), orc.iv(*bx[0]), orc.iv(*bx[1]), idir)
    tail = (orc.iv(0, 0, 0), orc.iv(15, 15, 15), orc.dv(*setup["dx"]), setup["dt"])
    orc.call("trace_ppm", *args, o_q, o_qa, o_qm, o_qp, *tail)
    g_qm, g_qp = Dev(*tb, QVAR), Dev(*tb, QVAR)
    gpu_call("trace_ppm", *args, Dev(*qb, QVAR, o_q.a), Dev(*qb, NQAUX, o_qa.a), g_qm, g_qp, *tail)
    np.testing.assert_array_equal(g_qm.host()[..., QDEF], o_qm.a[..., QDEF])
    np.testing.assert_array_equal(g_qp.host()[..., QDEF], o_qp.a[..., QDEF])


def _traced(setup):
    qb, o_q, o_qa = _prim(setup)
    bx = setup["bx"]
    olo, ohi = grow(*bx, 1)
    T = _tbx(bx)
    qm, qp = {}, {}
    for d in range(3):
        tb = T[(d, d)]
        qm[d], qp[d] = orc.Fab(*tb, QVAR), orc.Fab(*tb, QVAR)
        orc.call("trace_ppm", orc.iv(*olo), orc.iv(*ohi), orc.iv(*bx[0]), orc.iv(*bx[1]), d + 1, o_q, o_qa,
                 qm[d], qp[d], orc.iv(0, 0, 0), orc.iv(15, 15, 15), orc.dv(*setup["dx"]), setup["dt"])
    return qb, o_q, o_qa, T, qm, qp


@pytest.mark.parametrize("idir", [1, 2, 3])
def test_compute_flux(setup, idir):
    qb, o_q, o_qa, T, qm, qp = _traced(setup)
    d = idir - 1
    tb = T[(d, d)]
    olo, ohi = grow(*setup["bx"], 1)
    o_f, o_qi, o_qg = orc.Fab(olo, ohi, NVAR), orc.Fab(olo, ohi, QVAR), orc.Fab(olo, ohi, NGDNV)
    orc.call("compute_flux", orc.iv(*tb[0]), orc.iv(*tb[1]), qm[d], qp[d], o_f, o_qi, o_qg, o_qa, idir)
    g_f, g_qi, g_qg = Dev(olo, ohi, NVAR), Dev(olo, ohi, QVAR), Dev(olo, ohi, NGDNV)
    gpu_call("compute_flux", orc.iv(*tb[0]), orc.iv(*tb[1]), Dev(*tb, QVAR, qm[d].a), Dev(*tb, QVAR, qp[d].a),
             g_f, g_qi, g_qg, Dev(*qb, NQAUX, o_qa.a), idir)
    sl = tuple(slice(a - l, b - l + 1) for a, b, l in zip(tb[0], tb[1], olo))
    np.testing.assert_array_equal(g_f.host()[sl], o_f.a[sl])
    np.testing.assert_array_equal(g_qg.host()[sl], o_qg.a[sl])
    np.testing.assert_array_equal(g_qi.host()[sl][..., QDEF + [QGAME]], o_qi.a[sl][..., QDEF + [QGAME]])


def _trans_inputs(setup):
    """Oracle intermediates through trans1 for every direction."""
    qb, o_q, o_qa, T, qm, qp = _traced(setup)
    olo, ohi = grow(*setup["bx"], 1)
    dt, dx = setup["dt"], setup["dx"]
    cdtdx = [dt / dx[d] / 3.0 for d in range(3)]
    f, qg = {}, {}
    qmt, qpt = {}, {}
    for d in range(3):
        tb = T[(d, d)]
        f[d], qg[d] = orc.Fab(olo, ohi, NVAR), orc.Fab(olo, ohi, NGDNV)
        orc.call("compute_flux", orc.iv(*tb[0]), orc.iv(*tb[1]), qm[d], qp[d], f[d], orc.Fab(olo, ohi, QVAR),
                 qg[d], o_qa, d + 1)
    return qb, o_q, o_qa, T, qm, qp, f, qg, cdtdx


@pytest.mark.parametrize("d,t", [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)])
def test_trans1(setup, d, t):
    qb, o_q, o_qa, T, qm, qp, f, qg, cdtdx = _trans_inputs(setup)
    olo, ohi = grow(*setup["bx"], 1)
    tb = T[(t, d)]
    o_mo, o_po = orc.Fab(*tb, QVAR), orc.Fab(*tb, QVAR)
    orc.call("trans1", orc.iv(*tb[0]), orc.iv(*tb[1]), d + 1, t + 1, qm[t], o_mo, qp[t], o_po, o_qa, f[d], qg[d],
             cdtdx[d])
    g_mo, g_po = Dev(*tb, QVAR), Dev(*tb, QVAR)
    tt = T[(t, t)]
    gpu_call("trans1", orc.iv(*tb[0]), orc.iv(*tb[1]), d + 1, t + 1, Dev(*tt, QVAR, qm[t].a), g_mo,
             Dev(*tt, QVAR, qp[t].a), g_po, Dev(*qb, NQAUX, o_qa.a), Dev(olo, ohi, NVAR, f[d].a),
             Dev(olo, ohi, NGDNV, qg[d].a), cdtdx[d])
    np.testing.assert_array_equal(g_mo.host()[..., QDEF], o_mo.a[..., QDEF])
    np.testing.assert_array_equal(g_po.host()[..., QDEF], o_po.a[..., QDEF])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_trans2_final_flux_av_source(setup, d):
    qb, o_q, o_qa, T, qm, qp, f, qg, cdtdx = _trans_inputs(setup)
    bx = setup["bx"]
    olo, ohi = grow(*bx, 1)
    dt, dx = setup["dt"], setup["dx"]
    hdtdx = [0.5 * dt / dx[e] for e in range(3)]
    t1, t2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
    # trans1 for the two pairs feeding this direction's transverse fluxes
    qmt, qpt = {}, {}
    for (tt, dd) in ((t1, t2), (t2, t1)):
        tb = T[(tt, dd)]
        qmt[tt], qpt[tt] = orc.Fab(*tb, QVAR), orc.Fab(*tb, QVAR)
        orc.call("trans1", orc.iv(*tb[0]), orc.iv(*tb[1]), dd + 1, tt + 1, qm[tt], qmt[tt], qp[tt], qpt[tt], o_qa,
                 f[dd], qg[dd], cdtdx[dd])
    f1, f2 = orc.Fab(olo, ohi, NVAR), orc.Fab(olo, ohi, NVAR)
    g1, g2 = orc.Fab(olo, ohi, NGDNV), orc.Fab(olo, ohi, NGDNV)
    orc.call("compute_flux", orc.iv(*T[(t1, t2)][0]), orc.iv(*T[(t1, t2)][1]), qmt[t1], qpt[t1], f1,
             orc.Fab(olo, ohi, QVAR), g1, o_qa, t1 + 1)
    orc.call("compute_flux", orc.iv(*T[(t2, t1)][0]), orc.iv(*T[(t2, t1)][1]), qmt[t2], qpt[t2], f2,
             orc.Fab(olo, ohi, QVAR), g2, o_qa, t2 + 1)
    elo, ehi = nodes(*bx, d)
    o_ql, o_qr = orc.Fab(olo, ohi, QVAR), orc.Fab(olo, ohi, QVAR)
    targs = (orc.iv(*elo), orc.iv(*ehi), d + 1, t1 + 1, t2 + 1)
    orc.call("trans2", *targs, qm[d], o_ql, qp[d], o_qr, o_qa, f1, f2, g1, g2, hdtdx[d], hdtdx[t1], hdtdx[t2])
    g_ql, g_qr = Dev(olo, ohi, QVAR), Dev(olo, ohi, QVAR)
    tb = T[(d, d)]
    gpu_call("trans2", *targs, Dev(*tb, QVAR, qm[d].a), g_ql, Dev(*tb, QVAR, qp[d].a), g_qr,
             Dev(*qb, NQAUX, o_qa.a), Dev(olo, ohi, NVAR, f1.a), Dev(olo, ohi, NVAR, f2.a),
             Dev(olo, ohi, NGDNV, g1.a), Dev(olo, ohi, NGDNV, g2.a), hdtdx[d], hdtdx[t1], hdtdx[t2])
    sl = tuple(slice(a - l, b - l + 1) for a, b, l in zip(elo, ehi, olo))
    np.testing.assert_array_equal(g_ql.host()[sl][..., QDEF], o_ql.a[sl][..., QDEF])
    np.testing.assert_array_equal(g_qr.host()[sl][..., QDEF], o_qr.a[sl][..., QDEF])

    # final flux, artificial viscosity, species normalisation, store_flux
    gelo, gehi = grow(elo, ehi, 1)
    o_f, o_qe = orc.Fab(gelo, gehi, NVAR), orc.Fab(gelo, gehi, NGDNV)
    orc.call("compute_flux", orc.iv(*elo), orc.iv(*ehi), o_ql, o_qr, o_f, orc.Fab(olo, ohi, QVAR), o_qe, o_qa, d + 1)
    o_div = orc.Fab(olo, ohi, 1)
    orc.call("divu", orc.iv(*olo), orc.iv(*ohi), o_q, orc.dv(*dx), o_div)
    slo, shi = grow(*bx, 4)
    U = sub(setup["S"], slo, shi)
    g_f = Dev(gelo, gehi, NVAR, o_f.a)
    orc.call("apply_av", orc.iv(*elo), orc.iv(*ehi), d + 1, orc.dv(*dx), o_div, orc.Fab(slo, shi, NVAR, U), o_f)
    gpu_call("apply_av", orc.iv(*elo), orc.iv(*ehi), d + 1, orc.dv(*dx), Dev(olo, ohi, 1, o_div.a),
             Dev(slo, shi, NVAR, U), g_f)
    sle = tuple(slice(a - l, b - l + 1) for a, b, l in zip(elo, ehi, gelo))
    np.testing.assert_array_equal(g_f.host()[sle], o_f.a[sle])
    g_f = Dev(gelo, gehi, NVAR, o_f.a)
    orc.call("normalize_species_fluxes", orc.iv(*elo), orc.iv(*ehi), o_f)
    gpu_call("normalize_species_fluxes", orc.iv(*elo), orc.iv(*ehi), g_f)
    np.testing.assert_array_equal(g_f.host()[sle], o_f.a[sle])
    alo, ahi = grow(*nodes(*bx, d), 4)
    area = np.full(tuple(h - l + 1 for l, h in zip(alo, ahi)), dx[1] * dx[2])
    o_fo, g_fo = orc.Fab(elo, ehi, NVAR), Dev(elo, ehi, NVAR)
    orc.call("store_flux", orc.iv(*elo), orc.iv(*ehi), o_fo, o_f, orc.Fab(alo, ahi, 1, area), dt)
    gpu_call("store_flux", orc.iv(*elo), orc.iv(*ehi), g_fo, Dev(gelo, gehi, NVAR, o_f.a), Dev(alo, ahi, 1, area),
             dt)
    np.testing.assert_array_equal(g_fo.host(), o_fo.a)


def test_fill_hydro_source(setup):
    bx = setup["bx"]
    lo, hi = bx
    rng = np.random.default_rng(5)
    dx = (0.0625, 0.0625, 0.0625)
    fl, qe, ar = [], [], []
    for d in range(3):
        e = grow(*nodes(lo, hi, d), 1)
        shp = tuple(h - l + 1 for l, h in zip(*e))
        fl.append((e, rng.standard_normal(shp + (NVAR,))))
        qe.append((e, rng.random(shp + (NGDNV,)) + 0.5))
        a = grow(*nodes(lo, hi, d), 4)
        ar.append((a, np.full(tuple(h - l + 1 for l, h in zip(*a)), dx[0] * dx[1])))
    vlo, vhi = grow(lo, hi, 4)
    vol = np.full(tuple(h - l + 1 for l, h in zip(vlo, vhi)), dx[0] * dx[1] * dx[2])
    dummy = orc.Fab(lo, hi, NVAR)
    o_s = orc.Fab(lo, hi, NVAR)
    orc.call("fill_hydro_source", orc.iv(*lo), orc.iv(*hi), dummy, dummy, o_s,
             *[orc.Fab(*e, NVAR, v) for e, v in fl], *[orc.Fab(*e, NGDNV, v) for e, v in qe],
             *[orc.Fab(*a, 1, v) for a, v in ar], orc.Fab(vlo, vhi, 1, vol), orc.dv(*dx), 1.0)
    g_s = Dev(lo, hi, NVAR, np.zeros(o_s.a.shape))
    gd = Dev(lo, hi, NVAR, np.zeros(o_s.a.shape))
    gpu_call("fill_hydro_source", orc.iv(*lo), orc.iv(*hi), gd, gd, g_s,
             *[Dev(*e, NVAR, v) for e, v in fl], *[Dev(*e, NGDNV, v) for e, v in qe],
             *[Dev(*a, 1, v) for a, v in ar], Dev(vlo, vhi, 1, vol), orc.dv(*dx), 1.0)
    np.testing.assert_array_equal(g_s.host(), o_s.a)


def test_estdt_bitwise(setup):
    import torch
    lo, hi = setup["bx"]
    U = sub(setup["S"], lo, hi)
    dto = C.c_double(1e300)
    orc.call("estdt", orc.iv(*lo), orc.iv(*hi), orc.Fab(lo, hi, NVAR, U), orc.dv(*setup["dx"]), C.byref(dto))
    # host pointer
    dth = C.c_double(1e300)
    gpu_call("estdt", orc.iv(*lo), orc.iv(*hi), Dev(lo, hi, NVAR, U), orc.dv(*setup["dx"]), C.byref(dth))
    assert dth.value == dto.value
    # device pointer
    dd = torch.full((1,), 1e300, dtype=torch.float64, device="cuda")
    gpu_call("estdt", orc.iv(*lo), orc.iv(*hi), Dev(lo, hi, NVAR, U), orc.dv(*setup["dx"]),
             C.cast(dd.data_ptr(), C.POINTER(C.c_double)))
    assert dd.item() == dto.value


def test_eos_batch_bitwise():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    oracle_table()
    rng = np.random.default_rng(3)
    n = 4000
    rho = 10.0 ** rng.uniform(-6, 6, n)
    T = 10.0 ** rng.uniform(3.2, 12.5, n)
    abar = rng.choice([4.0, 12.0, 40.0], n) * np.where(rng.random(n) < 0.5, 1.0, rho)
    zbar = abar * rng.uniform(0.3, 0.5, n)
    ref = [orc.eos(orc.EOS_RT, *a) for a in zip(rho, T, abar, zbar)]
    e = np.array([r.e for r in ref])
    Tg = T * 10.0 ** rng.uniform(-1.5, 1.5, n)
    ref_re = [orc.eos(orc.EOS_RE, r_, t_, a_, z_, e=e_) for r_, t_, a_, z_, e_ in zip(rho, Tg, abar, zbar, e)]
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    P = lambda t: C.cast(t.data_ptr(), C.POINTER(C.c_double))
    out = torch.zeros(10 * n, dtype=torch.float64, device="cuda")
    for mode, Tin, ev in ((1, T, e), (2, Tg, e)):
        args = [dev(rho), dev(Tin), dev(ev), dev(abar), dev(zbar)]
        assert mc.lib().mc_eos_eval_batch(n, mode, *[P(a) for a in args], P(out)) == 0
        torch.cuda.synchronize()
        o = out.cpu().numpy().reshape(n, 10)
        rr = ref if mode == 1 else ref_re
        np.testing.assert_array_equal(o[:, 0], [r.T for r in rr])
        np.testing.assert_array_equal(o[:, 1], [r.p for r in rr])
        np.testing.assert_array_equal(o[:, 3], [r.cs for r in rr])
        np.testing.assert_array_equal(o[:, 4], [r.gam1 for r in rr])
        np.testing.assert_array_equal(o[:, 5], [r.dpde for r in rr])
        np.testing.assert_array_equal(o[:, 6], [r.dpdr_e for r in rr])
        np.testing.assert_array_equal(o[:, 8], [r.iat for r in rr])
        np.testing.assert_array_equal(o[:, 9], [r.jat for r in rr])
        if mode == 1:
            np.testing.assert_array_equal(o[:, 2], [r.e for r in rr])
            np.testing.assert_array_equal(o[:, 7], [r.dedT for r in rr])


def test_eos_hash_on_grid_lines():
    """The table index hash takes log10 (ocml on the GPU, glibc in the
    oracle).  Inputs on every grid line of T and of ye*rho, and one ulp to
    either side, are where the two could pick different table cells: the
    indices (and the results) must still agree bit for bit."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    mc.load_table(table_bin())
    oracle_table()
    t, d = orc.eos_grid()
    near = lambda g: np.concatenate([np.nextafter(g, 0.0), g, np.nextafter(g, np.inf)])
    T_pts = near(t[:-1])
    din_pts = near(d[:-1])
    rho = np.concatenate([np.full(T_pts.size, 2.0), 2.0 * din_pts])   # ye = 1/2: ye*rho = din exactly
    T = np.concatenate([T_pts, np.full(din_pts.size, 1.0e7)])
    n = rho.size
    abar, zbar = np.full(n, 4.0), np.full(n, 2.0)
    ref = [orc.eos(orc.EOS_RT, r_, t_, 4.0, 2.0) for r_, t_ in zip(rho, T)]
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
    P = lambda x: C.cast(x.data_ptr(), C.POINTER(C.c_double))
    out = torch.zeros(10 * n, dtype=torch.float64, device="cuda")
    args = [dev(rho), dev(T), dev(np.zeros(n)), dev(abar), dev(zbar)]
    assert mc.lib().mc_eos_eval_batch(n, 1, *[P(a) for a in args], P(out)) == 0
    torch.cuda.synchronize()
    o = out.cpu().numpy().reshape(n, 10)
    np.testing.assert_array_equal(o[:, 8], [r.iat for r in ref])
    np.testing.assert_array_equal(o[:, 9], [r.jat for r in ref])
    np.testing.assert_array_equal(o[:, 1], [r.p for r in ref])
    np.testing.assert_array_equal(o[:, 2], [r.e for r in ref])

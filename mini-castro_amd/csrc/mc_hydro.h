// Per-cell / per-face CTU hydro arithmetic as device functions.
//
// Each function evaluates the same expressions, in the same order, as the
// reference routine it names, so that without FMA contraction the results
// are bit-identical to the reference arithmetic.  Data movement (which
// arrays, which layout, what is staged in LDS) is left to the kernels.
#pragma once

#include "mc_common.h"

namespace mc {

// Source/ppm.F90:132-236; s[0..4] = s(-2..2)
__device__ __forceinline__ void ppm_reconstruct(const double *s5, double flatn, double &smo, double &spo)
{
    const double *s = s5 + 2;
    double dsl, dsr, dsc, dsvl_l, dsvl_r, sm, sp;
    dsl = 2.0 * (s[-1] - s[-2]);
    dsr = 2.0 * (s[0] - s[-1]);
    if (dsl * dsr > 0.0) {
        dsc = HALF * (s[0] - s[-2]);
        dsvl_l = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_l = 0.0;
    dsl = 2.0 * (s[0] - s[-1]);
    dsr = 2.0 * (s[1] - s[0]);
    if (dsl * dsr > 0.0) {
        dsc = HALF * (s[1] - s[-1]);
        dsvl_r = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_r = 0.0;
    sm = HALF * (s[0] + s[-1]) - SIXTH * (dsvl_r - dsvl_l);
    sm = fmax(sm, fmin(s[0], s[-1]));
    sm = fmin(sm, fmax(s[0], s[-1]));

    // the slope at cell 0 is the same van Leer slope as dsvl_r above
    dsvl_l = dsvl_r;
    dsl = 2.0 * (s[1] - s[0]);
    dsr = 2.0 * (s[2] - s[1]);
    if (dsl * dsr > 0.0) {
        dsc = HALF * (s[2] - s[0]);
        dsvl_r = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_r = 0.0;
    sp = HALF * (s[1] + s[0]) - SIXTH * (dsvl_r - dsvl_l);
    sp = fmax(sp, fmin(s[1], s[0]));
    sp = fmin(sp, fmax(s[1], s[0]));

    sm = flatn * sm + (1.0 - flatn) * s[0];
    sp = flatn * sp + (1.0 - flatn) * s[0];

    if ((sp - s[0]) * (s[0] - sm) <= 0.0) {
        sp = s[0];
        sm = s[0];
    } else if (fabs(sp - s[0]) >= 2.0 * fabs(sm - s[0])) {
        sp = 3.0 * s[0] - 2.0 * sm;
    } else if (fabs(sm - s[0]) >= 2.0 * fabs(sp - s[0])) {
        sm = 3.0 * s[0] - 2.0 * sp;
    }
    smo = sm;
    spo = sp;
}

// Source/ppm.F90:240-305 for one wave speed.
__device__ __forceinline__ void ppm_int_wave(double sm, double sp, double s6, double speed, double dtdx, double &Ip,
                                             double &Im)
{
    const double sigma = fabs(speed) * dtdx;
    if (speed <= 0.0) {
        Ip = sp;
        Im = sm + HALF * sigma * (sp - sm + (1.0 - TWO3RD * sigma) * s6);
    } else {
        Ip = sp - HALF * sigma * (sp - sm - (1.0 - TWO3RD * sigma) * s6);
        Im = sm;
    }
}

__device__ __forceinline__ double ppm_s6(double sm, double sp, double sc) { return 6.0 * sc - 3.0 * (sm + sp); }

// Source/ppm.F90:11-128 for one direction; q5p[m] = p(i+m-3), q5u likewise
// for the normal velocity (m = 0..6 covers offsets -3..+3).
__device__ __forceinline__ double flatten_dir(const double *p7, const double *u7)
{
    const double shktst = 0.33, zcut1 = 0.75, zcut2 = 0.85;
    const double dzcut = 1.0 / (zcut2 - zcut1);
    const double *p = p7 + 3, *u = u7 + 3;
    double pr = p[1], pl = p[-1];
    double dp2 = p[2] - p[-2];
    double du = u[1] - u[-1];
    double dp = pr - pl;
    const int ishft = (dp > 0.0) ? 1 : -1;
    double denom = fmax(small_pres, fabs(dp2));
    double denominv = 1.0 / denom;
    double zeta = fabs(dp) * denominv;
    const double z = fmin(1.0, fmax(0.0, dzcut * (zeta - zcut1)));
    double tst = (du <= 0.0) ? 1.0 : 0.0;
    double tmp = fmin(pr, pl);
    const double chi = (fabs(dp) > shktst * tmp) ? tst : 0.0;

    pr = p[1 - ishft];
    pl = p[-1 - ishft];
    dp2 = p[2 - ishft] - p[-2 - ishft];
    du = u[1 - ishft] - u[-1 - ishft];
    dp = pr - pl;
    denom = fmax(small_pres, fabs(dp2));
    denominv = 1.0 / denom;
    zeta = fabs(dp) * denominv;
    const double z2 = fmin(1.0, fmax(0.0, dzcut * (zeta - zcut1)));
    tst = (du <= 0.0) ? 1.0 : 0.0;
    tmp = fmin(pr, pl);
    const double chi2 = (fabs(dp) > shktst * tmp) ? tst : 0.0;
    return 1.0 - fmax(chi2 * z2, chi * z);
}

// Characteristic projection of Source/ppm.F90:529-625 (plus state, face i,
// PLUS=true; reference state Im(1,:)) and :631-737 (minus state, face i+1,
// PLUS=false; reference state Ip(3,:)).  I[w][v]: v = 0 rho, 1 un, 2 p,
// 3 rhoe, w = wave (u-c, u, u+c).  gam_ref = Im_gc(1) or Ip_gc(3).
// Outputs rho, un, rhoe, p of the interface state.
template <bool PLUS>
__device__ __forceinline__ void ppm_project(const double I[3][4], double gam_ref, double un, double cc, double &o_rho,
                                            double &o_un, double &o_rhoe, double &o_p)
{
    const int w = PLUS ? 0 : 2;
    double rho_ref = I[w][0], un_ref = I[w][1], p_ref = I[w][2], rhoe_g_ref = I[w][3];
    rho_ref = fmax(rho_ref, small_dens);
    const double rho_ref_inv = 1.0 / rho_ref;
    p_ref = fmax(p_ref, small_pres);
    const double csq_ref = gam_ref * p_ref * rho_ref_inv;
    const double cc_ref = sqrt(csq_ref);
    const double cc_ref_inv = 1.0 / cc_ref;
    const double h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv;
    const double dum = un_ref - I[0][1];
    const double dptotm = p_ref - I[0][2];
    const double drho = rho_ref - I[1][0];
    const double dptot = p_ref - I[1][2];
    const double drhoe_g = rhoe_g_ref - I[1][3];
    const double dup = un_ref - I[2][1];
    const double dptotp = p_ref - I[2][2];
    double alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
    double alphap = HALF * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
    double alpha0r = drho - dptot / csq_ref;
    double alpha0e_g = drhoe_g - dptot * h_g_ref / csq_ref;
    if (PLUS) {
        alpham = (un - cc > 0.0) ? 0.0 : -alpham;
        alphap = (un + cc > 0.0) ? 0.0 : -alphap;
        alpha0r = (un > 0.0) ? 0.0 : -alpha0r;
        alpha0e_g = (un > 0.0) ? 0.0 : -alpha0e_g;
    } else {
        alpham = (un - cc > 0.0) ? -alpham : 0.0;
        alphap = (un + cc > 0.0) ? -alphap : 0.0;
        alpha0r = (un > 0.0) ? -alpha0r : 0.0;
        alpha0e_g = (un > 0.0) ? -alpha0e_g : 0.0;
    }
    o_rho = fmax(small_dens, rho_ref + alphap + alpham + alpha0r);
    o_un = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
    o_rhoe = rhoe_g_ref + (alphap + alpham) * h_g_ref + alpha0e_g;
    o_p = fmax(small_pres, p_ref + (alphap + alpham) * csq_ref);
}

// Interface state without species, velocities permuted to the face direction
// (un normal, ut / utt the two transverse components in the reference's order).
struct FaceH {
    double rho, un, ut, utt, p, rhoe;
};

// Riemann solution without species (Source/riemann.F90:114-318).  Species are
// passive: their interface value is fpl*XL + fmr*XR (riemann.F90:285-288) and
// their flux f_rho times that (:321-325).
struct RiemannH {
    double f_rho, f_mn, f_mt, f_mtt, f_E, f_e;
    double g_rho, g_un, g_ut, g_utt, g_p, g_game, g_rhoe;
    double fpl, fmr;
};

__device__ __forceinline__ void riemann_hydro(const FaceH &L, const FaceH &R, double cL, double cR, double gamcl,
                                              double gamcr, RiemannH &o)
{
    const double rl = fmax(L.rho, small_dens);
    const double ul = L.un, v1l = L.ut, v2l = L.utt;
    const double pl = fmax(L.p, small_pres);
    const double rel = L.rhoe;
    const double rr = fmax(R.rho, small_dens);
    const double ur = R.un, v1r = R.ut, v2r = R.utt;
    const double pr = fmax(R.p, small_pres);
    const double rer = R.rhoe;

    const double csmall = fmax(small, small * fmax(cR, cL));
    const double cavg = HALF * (cR + cL);
    const double wsmall = small_dens * csmall;
    const double wl = fmax(wsmall, sqrt(fabs(gamcl * pl * rl)));
    const double wr = fmax(wsmall, sqrt(fabs(gamcr * pr * rr)));
    const double wwinv = 1.0 / (wl + wr);
    double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) * wwinv;
    double ustar = ((wl * ul + wr * ur) + (pl - pr)) * wwinv;
    pstar = fmax(pstar, small_pres);
    if (fabs(ustar) < smallu * HALF * (fabs(ul) + fabs(ur))) ustar = 0.0;

    double sgnm = fsign1(ustar);
    if (ustar == 0.0) sgnm = 0.0;
    const double fpl = HALF * (1.0 + sgnm);
    const double fmr = HALF * (1.0 - sgnm);
    double ro = fpl * rl + fmr * rr;
    const double uo = fpl * ul + fmr * ur;
    const double po = fpl * pl + fmr * pr;
    const double reo = fpl * rel + fmr * rer;
    const double gamco = fpl * gamcl + fmr * gamcr;
    ro = fmax(small_dens, ro);
    const double roinv = 1.0 / ro;
    double co = sqrt(fabs(gamco * po * roinv));
    co = fmax(csmall, co);
    const double co2inv = 1.0 / (co * co);
    o.g_ut = fpl * v1l + fmr * v1r;
    o.g_utt = fpl * v2l + fmr * v2r;

    const double drho = (pstar - po) * co2inv;
    double rstar = ro + drho;
    rstar = fmax(small_dens, rstar);
    const double rstarinv = 1.0 / rstar;
    const double entho = (reo + po) * roinv * co2inv;
    const double estar = reo + (pstar - po) * entho;
    double cstar = sqrt(fabs(gamco * pstar * rstarinv));
    cstar = fmax(cstar, csmall);

    sgnm = fsign1(ustar);
    double spout = co - sgnm * uo;
    double spin = cstar - sgnm * ustar;
    const double ushock = HALF * (spin + spout);
    if (pstar - po > 0.0) {
        spin = ushock;
        spout = ushock;
    }
    const double scr = (spout - spin == 0.0) ? small * cavg : spout - spin;
    const double scrinv = 1.0 / scr;
    double frac = (1.0 + (spout + spin) * scrinv) * HALF;
    frac = fmax(0.0, fmin(1.0, frac));

    double qrho = frac * rstar + (1.0 - frac) * ro;
    double qun = frac * ustar + (1.0 - frac) * uo;
    double qp = frac * pstar + (1.0 - frac) * po;
    double regdnv = frac * estar + (1.0 - frac) * reo;
    if (spout < 0.0) {
        qrho = ro;
        qun = uo;
        qp = po;
        regdnv = reo;
    }
    if (spin >= 0.0) {
        qrho = rstar;
        qun = ustar;
        qp = pstar;
        regdnv = estar;
    }
    const double regdnvinv = 1.0 / regdnv;
    const double qgame = qp * regdnvinv + 1.0;
    qp = fmax(qp, small_pres);

    o.g_rho = qrho;
    o.g_un = qun;
    o.g_p = qp;
    o.g_game = qgame;
    o.g_rhoe = regdnv;

    const double u_adv = qun;
    const double rhoeint = regdnv;
    o.f_rho = qrho * u_adv;
    o.f_mn = o.f_rho * qun;
    o.f_mt = o.f_rho * o.g_ut;
    o.f_mtt = o.f_rho * o.g_utt;
    const double rhoetot = rhoeint + HALF * qrho * (qun * qun + o.g_ut * o.g_ut + o.g_utt * o.g_utt);
    o.f_E = u_adv * (rhoetot + qp);
    o.f_e = u_adv * rhoeint;
    o.fpl = fpl;
    o.fmr = fmr;
}

// Interface state in the order (rho, un, ut, utt, p, rhoe, X[13]) used by
// the Riemann solver: directional velocities already permuted.
struct Face {
    double rho, un, ut, utt, p, rhoe;
    double x[NSPEC];
};

// Riemann solution at one face, Source/riemann.F90:114-325.
struct RiemannOut {
    double f_rho, f_mn, f_mt, f_mtt, f_E, f_e;
    double g_rho, g_un, g_ut, g_utt, g_p, g_game, g_rhoe;
    double gx[NSPEC];  // qint species; species flux = f_rho * gx
};

__device__ __forceinline__ void riemann_cgf(const Face &L, const Face &R, double cL, double cR, double gamcl,
                                            double gamcr, RiemannOut &o)
{
    const FaceH lh = {L.rho, L.un, L.ut, L.utt, L.p, L.rhoe};
    const FaceH rh = {R.rho, R.un, R.ut, R.utt, R.p, R.rhoe};
    RiemannH h;
    riemann_hydro(lh, rh, cL, cR, gamcl, gamcr, h);
    o.f_rho = h.f_rho;
    o.f_mn = h.f_mn;
    o.f_mt = h.f_mt;
    o.f_mtt = h.f_mtt;
    o.f_E = h.f_E;
    o.f_e = h.f_e;
    o.g_rho = h.g_rho;
    o.g_un = h.g_un;
    o.g_ut = h.g_ut;
    o.g_utt = h.g_utt;
    o.g_p = h.g_p;
    o.g_game = h.g_game;
    o.g_rhoe = h.g_rhoe;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) o.gx[n] = h.fpl * L.x[n] + h.fmr * R.x[n];
}


// Interface state as stored in the qm/qp arrays (primitive order).
struct QIn {
    double rho, u, v, w, p, rhoe, game;
    double x[NSPEC];
};
struct QInH {
    double rho, u, v, w, p, rhoe, game;
};

// Interface-state output of a transverse update without species.
struct QStateH {
    double rho, u, v, w, p, rhoe;
    double game;   // only defined when reset (copied through)
    bool reset;
};

// Source/trans.F90:165-229 for one state: lq2 corrected by the transverse
// flux difference fr - fl of URHO..UEDEN and the Godunov p, u at the faces.
__device__ __forceinline__ QStateH trans1_hydro(const QInH &lq2, const double *fr, const double *fl, double pgp,
                                                double pgm, double ugp, double ugm, double gamc, double cdtdx)
{
    QStateH o;
    const double dup = pgp * ugp - pgm * ugm;
    const double du = ugp - ugm;
    const double pav = HALF * (pgp + pgm);

    const double rrl2 = lq2.rho;
    const double rul2 = rrl2 * lq2.u;
    const double rvl2 = rrl2 * lq2.v;
    const double rwl2 = rrl2 * lq2.w;
    const double ekenl2 = HALF * rrl2 * (lq2.u * lq2.u + lq2.v * lq2.v + lq2.w * lq2.w);
    const double rel2 = lq2.rhoe + ekenl2;

    double rrnewl2 = rrl2 - cdtdx * (fr[URHO] - fl[URHO]);
    double runewl2 = rul2 - cdtdx * (fr[UMX] - fl[UMX]);
    double rvnewl2 = rvl2 - cdtdx * (fr[UMY] - fl[UMY]);
    double rwnewl2 = rwl2 - cdtdx * (fr[UMZ] - fl[UMZ]);
    double renewl2 = rel2 - cdtdx * (fr[UEDEN] - fl[UEDEN]);
    o.reset = false;
    if (rrnewl2 < 0.0) {
        rrnewl2 = rrl2;
        runewl2 = rul2;
        rvnewl2 = rvl2;
        rwnewl2 = rwl2;
        renewl2 = rel2;
        o.reset = true;
    }
    o.rho = rrnewl2;
    const double rhoinv = 1.0 / rrnewl2;
    o.u = runewl2 * rhoinv;
    o.v = rvnewl2 * rhoinv;
    o.w = rwnewl2 * rhoinv;
    const double rhoekenl2 = HALF * (runewl2 * runewl2 + rvnewl2 * rvnewl2 + rwnewl2 * rwnewl2) * rhoinv;
    o.rhoe = renewl2 - rhoekenl2;
    if (!o.reset) {
        const double pnewl2 = lq2.p - cdtdx * (dup + pav * du * (gamc - 1.0));
        o.p = fmax(pnewl2, small_pres);
        o.game = 0.0;
    } else {
        o.p = lq2.p;
        o.game = lq2.game;
    }
    return o;
}

// Source/trans.F90:151-158: one species of a trans1-corrected state.
// rrnew = rho - cdtdx*(fr[URHO] - fl[URHO]), without the reset.
__device__ __forceinline__ double trans1_species(double rho, double x, double frs, double fls, double cdtdx,
                                                 double rrnew)
{
    const double compu = rho * x - cdtdx * (frs - fls);
    return compu / rrnew;
}

// Source/trans.F90:407-491 for one state without species: corrected by two
// transverse flux differences (f2r - f2l in direction 2, f3r - f3l in 3).
__device__ __forceinline__ QStateH trans2_hydro(const QInH &lq, const double *f2r, const double *f2l,
                                                const double *f3r, const double *f3l, double pg2p, double pg2m,
                                                double ug2p, double ug2m, double pg3p, double pg3m, double ug3p,
                                                double ug3m, double gamc, double cdtdx2, double cdtdx3)
{
    QStateH o;
    const double du2p = pg2p * ug2p - pg2m * ug2m;
    const double p2av = HALF * (pg2p + pg2m);
    const double du2 = ug2p - ug2m;
    const double p2new = cdtdx2 * (du2p + p2av * du2 * (gamc - 1.0));
    const double du3p = pg3p * ug3p - pg3m * ug3m;
    const double p3av = HALF * (pg3p + pg3m);
    const double du3 = ug3p - ug3m;
    const double p3new = cdtdx3 * (du3p + p3av * du3 * (gamc - 1.0));

    const double rrr = lq.rho;
    const double rur = rrr * lq.u;
    const double rvr = rrr * lq.v;
    const double rwr = rrr * lq.w;
    const double ekenr = HALF * rrr * (lq.u * lq.u + lq.v * lq.v + lq.w * lq.w);
    const double rer = lq.rhoe + ekenr;

    double rrnewr = rrr - cdtdx2 * (f2r[URHO] - f2l[URHO]) - cdtdx3 * (f3r[URHO] - f3l[URHO]);
    double runewr = rur - cdtdx2 * (f2r[UMX] - f2l[UMX]) - cdtdx3 * (f3r[UMX] - f3l[UMX]);
    double rvnewr = rvr - cdtdx2 * (f2r[UMY] - f2l[UMY]) - cdtdx3 * (f3r[UMY] - f3l[UMY]);
    double rwnewr = rwr - cdtdx2 * (f2r[UMZ] - f2l[UMZ]) - cdtdx3 * (f3r[UMZ] - f3l[UMZ]);
    double renewr = rer - cdtdx2 * (f2r[UEDEN] - f2l[UEDEN]) - cdtdx3 * (f3r[UEDEN] - f3l[UEDEN]);
    o.reset = false;
    if (rrnewr < 0.0) {
        rrnewr = rrr;
        runewr = rur;
        rvnewr = rvr;
        rwnewr = rwr;
        renewr = rer;
        o.reset = true;
    }
    o.rho = rrnewr;
    o.u = runewr / rrnewr;
    o.v = rvnewr / rrnewr;
    o.w = rwnewr / rrnewr;
    const double rhoekenr = HALF * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
    o.rhoe = renewr - rhoekenr;
    double pn;
    if (!o.reset) {
        pn = lq.p - p2new - p3new;
        o.game = 0.0;
    } else {
        pn = lq.p;
        o.game = lq.game;
    }
    o.p = fmax(pn, small_pres);
    return o;
}

// Source/trans.F90:393-405: one species of a trans2-corrected state;
// rrnew = rho - cdtdx2*(f2r[URHO]-f2l[URHO]) - cdtdx3*(f3r[URHO]-f3l[URHO]).
__device__ __forceinline__ double trans2_species(double rho, double x, double f2rs, double f2ls, double f3rs,
                                                 double f3ls, double cdtdx2, double cdtdx3, double rrnew)
{
    const double compr = rho * x;
    const double compnr = compr - cdtdx2 * (f2rs - f2ls) - cdtdx3 * (f3rs - f3ls);
    return compnr / rrnew;
}

// Interface-state output of trans1/trans2 in primitive order.
struct QState {
    double rho, u, v, w, p, rhoe;
    double x[NSPEC];
    double game;      // only defined when reset (copied through)
    bool reset;
};

__device__ __forceinline__ QInH qin_hydro(const QIn &q) { return QInH{q.rho, q.u, q.v, q.w, q.p, q.rhoe, q.game}; }

__device__ __forceinline__ void qstate_from(QState &o, const QStateH &h)
{
    o.rho = h.rho;
    o.u = h.u;
    o.v = h.v;
    o.w = h.w;
    o.p = h.p;
    o.rhoe = h.rhoe;
    o.game = h.game;
    o.reset = h.reset;
}

// Source/trans.F90:151-229 for one state (hydro part plus all species).
__device__ __forceinline__ QState trans1_state(const QIn &lq2, const double *fr, const double *fl, double pgp,
                                               double pgm, double ugp, double ugm, double gamc, double cdtdx)
{
    QState o;
    const double rrnew = lq2.rho - cdtdx * (fr[URHO] - fl[URHO]);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) o.x[n] = trans1_species(lq2.rho, lq2.x[n], fr[UFS + n], fl[UFS + n], cdtdx, rrnew);
    qstate_from(o, trans1_hydro(qin_hydro(lq2), fr, fl, pgp, pgm, ugp, ugm, gamc, cdtdx));
    return o;
}

// Source/trans.F90:393-491 for one state (hydro part plus all species).
__device__ __forceinline__ QState trans2_state(const QIn &lq, const double *f2r, const double *f2l, const double *f3r,
                                               const double *f3l, double pg2p, double pg2m, double ug2p, double ug2m,
                                               double pg3p, double pg3m, double ug3p, double ug3m, double gamc,
                                               double cdtdx2, double cdtdx3)
{
    QState o;
    const double rrnew = lq.rho - cdtdx2 * (f2r[URHO] - f2l[URHO]) - cdtdx3 * (f3r[URHO] - f3l[URHO]);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n)
        o.x[n] = trans2_species(lq.rho, lq.x[n], f2r[UFS + n], f2l[UFS + n], f3r[UFS + n], f3l[UFS + n], cdtdx2,
                                cdtdx3, rrnew);
    qstate_from(o, trans2_hydro(qin_hydro(lq), f2r, f2l, f3r, f3l, pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m,
                                gamc, cdtdx2, cdtdx3));
    return o;
}

} // namespace mc

// Device Helmholtz EOS for gfx950.
//
// Same thermodynamics and floating-point evaluation order as the reference
// routine `eos` (Source/eos.F90:141-576), organised for a GPU:
//  * the table is node-interleaved: each (i,j) node holds the 13 columns the
//    hydro path consumes in one 128-byte row (f ft ftt fd fdd fdt fddt fdtt
//    fddtt dpdf dpdft dpdfd dpdfdt + 3 pad), so one evaluation reads four
//    128-byte rows instead of 52 scattered words;
//  * the density-side quantities (iat, xd and all density basis weights)
//    depend only on ye*rho, which is fixed during a Newton solve, so they are
//    computed once per call;
//  * Newton iterations evaluate only e and de/dT (the three h5 sums that feed
//    them); the full set (p, dp/drho, dp/dT, gamma_1, ...) is evaluated once,
//    at the temperature the reference's final pass uses.  Every quantity that
//    is computed is computed with the reference's expression, so results are
//    bit-identical to the reference arithmetic when built without FMA
//    contraction;
//  * the h5 sums of one evaluation run together, two table columns at a
//    time (each sum keeps its own term order), so the four nodes' columns
//    are not all held in registers at once.
// The electron chemical potential and number density tables (ef*, xf*) and
// the entropy terms feed nothing on the hydro path and are not evaluated.
#pragma once

#include "mc_common.h"

namespace mc {

constexpr int EOS_IMAX = 541;
constexpr int EOS_JMAX = 201;
constexpr int EOS_NODE = 16;   // doubles per table node
constexpr int EOS_GRID = 8;    // doubles per grid entry
constexpr int EOS_MAX_NEWTON = 100;  // Source/eos.F90:126

// Node column slots.
enum { NF = 0, NFT, NFTT, NFD, NFDD, NFDT, NFDDT, NFDTT, NFDDTT, NDPDF, NDPDFT, NDPDFD, NDPDFDT };
// Grid entry slots: temperature (t dt dt2 dti dt2i), density (d dd dd2 ddi dd2i).
enum { G_X = 0, G_D, G_D2, G_DI, G_D2I };

struct EosTable {
    const double *node;  // IMAX*JMAX nodes, node (i,j) at (i-1) + IMAX*(j-1)
    const double *tg;    // JMAX entries
    const double *dg;    // IMAX entries
    unsigned long long *newton_fail;  // Newton solves that ended unconverged (may be null)
};

// The reference lets a Newton solve that has not converged after
// EOS_MAX_NEWTON iterations pass silently (Source/eos.F90:193,550); the
// result is the same here, and the solve is also counted.
__device__ __forceinline__ void eos_note_newton_failure(const EosTable &tab)
{
    if (tab.newton_fail) atomicAdd(tab.newton_fail, 1ull);
}

// Source/eos.F90:60-74,129-137 (the entropy constant sioncon is unused)
namespace eosc {
constexpr double tlo = 3.0, thi = 13.0, dlo = -12.0, dhi = 15.0;
constexpr double tstp = (thi - tlo) / 200.0;
constexpr double tstpi = 1.0 / tstp;
constexpr double dstp = (dhi - dlo) / 540.0;
constexpr double dstpi = 1.0 / dstp;
constexpr double mintemp = 1000.0;
constexpr double ttol = 1.0e-8;
constexpr double clight = 2.99792458e10;
constexpr double kerg = 1.380650424e-16;
constexpr double avo_eos = 6.0221417930e23;
constexpr double asol = 4.0 * 5.67051e-5 / clight;
constexpr double kergavo = kerg * avo_eos;
constexpr double asoli3 = asol / 3.0;
} // namespace eosc

// Hermite basis, Source/eos.F90:740-1031
__device__ __forceinline__ double psi0(double z) { return (z * z * z) * (z * (-6.0 * z + 15.0) - 10.0) + 1.0; }
__device__ __forceinline__ double dpsi0(double z) { return (z * z) * (z * (-30.0 * z + 60.0) - 30.0); }
__device__ __forceinline__ double ddpsi0(double z) { return z * (z * (-120.0 * z + 180.0) - 60.0); }
__device__ __forceinline__ double psi1(double z) { return z * ((z * z) * (z * (-3.0 * z + 8.0) - 6.0) + 1.0); }
__device__ __forceinline__ double dpsi1(double z) { return z * z * (z * (-15.0 * z + 32.0) - 18.0) + 1.0; }
__device__ __forceinline__ double ddpsi1(double z) { return z * (z * (-60.0 * z + 96.0) - 36.0); }
__device__ __forceinline__ double psi2(double z) { return 0.5 * z * z * (z * (z * (-z + 3.0) - 3.0) + 1.0); }
__device__ __forceinline__ double dpsi2(double z) { return 0.5 * z * (z * (z * (-5.0 * z + 12.0) - 9.0) + 2.0); }
__device__ __forceinline__ double ddpsi2(double z) { return 0.5 * (z * (z * (-20.0 * z + 36.0) - 18.0) + 2.0); }
__device__ __forceinline__ double xpsi0(double z) { return z * z * (2.0 * z - 3.0) + 1.0; }
__device__ __forceinline__ double xpsi1(double z) { return z * (z * (z - 2.0) + 1.0); }

// Six weights of one variable: value at x and at 1-x for orders 0,1,2.
struct W6 {
    double w0, w1, w2, w0m, w1m, w2m;
};

// The four table nodes of a cell, 13 columns each.
struct Nodes {
    double a[EOS_NODE], b[EOS_NODE], c[EOS_NODE], d[EOS_NODE];  // (i,j) (i+1,j) (i,j+1) (i+1,j+1)
};

__device__ __forceinline__ void load_node(const double *p, double *o, int ncol)
{
    const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int c = 0; c < EOS_NODE / 2; ++c) {
        if (2 * c < ncol) {
            double2 v = q[c];
            o[2 * c] = v.x;
            o[2 * c + 1] = v.y;
        }
    }
}

// One column q of the biquintic sum in the reference's term order:
// fi(4q+1)*wd*wt + fi(4q+2)*wmd*wt + fi(4q+3)*wd*wmt + fi(4q+4)*wmd*wmt
#define MC_H5_TERMS(acc, N, q, wd, wmd, wt, wmt)                                \
    acc = acc + N.a[q] * wd * wt;                                               \
    acc = acc + N.b[q] * wmd * wt;                                              \
    acc = acc + N.c[q] * wd * wmt;                                              \
    acc = acc + N.d[q] * wmd * wmt;

// Source/eos.F90:915-950; t-weights T, d-weights D.
__device__ __forceinline__ double h5(const Nodes &N, const W6 &T, const W6 &D)
{
    double acc = N.a[NF] * D.w0 * T.w0;
    acc = acc + N.b[NF] * D.w0m * T.w0;
    acc = acc + N.c[NF] * D.w0 * T.w0m;
    acc = acc + N.d[NF] * D.w0m * T.w0m;
    MC_H5_TERMS(acc, N, NFT, D.w0, D.w0m, T.w1, T.w1m)
    MC_H5_TERMS(acc, N, NFTT, D.w0, D.w0m, T.w2, T.w2m)
    MC_H5_TERMS(acc, N, NFD, D.w1, D.w1m, T.w0, T.w0m)
    MC_H5_TERMS(acc, N, NFDD, D.w2, D.w2m, T.w0, T.w0m)
    MC_H5_TERMS(acc, N, NFDT, D.w1, D.w1m, T.w1, T.w1m)
    MC_H5_TERMS(acc, N, NFDDT, D.w2, D.w2m, T.w1, T.w1m)
    MC_H5_TERMS(acc, N, NFDTT, D.w1, D.w1m, T.w2, T.w2m)
    MC_H5_TERMS(acc, N, NFDDTT, D.w2, D.w2m, T.w2, T.w2m)
    return acc;
}

// Source/eos.F90:1034-1059 over the dpdf columns; w = (w0, w1, w0m, w1m).
__device__ __forceinline__ double h3_dpdf(const Nodes &N, double t0, double t1, double t0m, double t1m,
                                          double d0, double d1, double d0m, double d1m)
{
    double acc = N.a[NDPDF] * d0 * t0;
    acc = acc + N.b[NDPDF] * d0m * t0;
    acc = acc + N.c[NDPDF] * d0 * t0m;
    acc = acc + N.d[NDPDF] * d0m * t0m;
    MC_H5_TERMS(acc, N, NDPDFT, d0, d0m, t1, t1m)
    MC_H5_TERMS(acc, N, NDPDFD, d1, d1m, t0, t0m)
    MC_H5_TERMS(acc, N, NDPDFDT, d1, d1m, t1, t1m)
    return acc;
}
#undef MC_H5_TERMS

// One column of a biquintic or bicubic sum: the four nodes' values a b c d
// (nodes (i,j) (i+1,j) (i,j+1) (i+1,j+1)) with the density weights (wd, wmd)
// and temperature weights (wt, wmt), added to acc in h5's term order (the
// first column starts the sum).
__device__ __forceinline__ void hcol(double &acc, bool first, double a, double b, double c, double d, double wd,
                                     double wmd, double wt, double wmt)
{
    acc = first ? a * wd * wt : acc + a * wd * wt;
    acc = acc + b * wmd * wt;
    acc = acc + c * wd * wmt;
    acc = acc + d * wmd * wmt;
}

// weight k (0, 1, 2) of a W6, and its m counterpart
__device__ __forceinline__ double wsel(const W6 &w, int k) { return k == 0 ? w.w0 : (k == 1 ? w.w1 : w.w2); }
__device__ __forceinline__ double wselm(const W6 &w, int k) { return k == 0 ? w.w0m : (k == 1 ? w.w1m : w.w2m); }

// Density-side data of one EOS call (constant through the Newton solve).
struct EosDen {
    double den, deni, ytot1, ye, din;
    int iat;
    W6 D;     // quintic weights (value)
    W6 dD;    // quintic weights (d/d din)
    double x0, x1, x0m, x1m;  // cubic weights (value)
};

__device__ __forceinline__ void eos_density(const EosTable &tab, double rho, double abar, double zbar,
                                            EosDen &s)
{
    s.den = rho;
    s.ytot1 = 1.0 / abar;
    s.ye = zbar / abar;
    s.din = s.ye * s.den;
    s.deni = 1.0 / s.den;
    int iat = (int)((log10(s.din) - eosc::dlo) * eosc::dstpi) + 1;
    iat = iat < 1 ? 1 : (iat > EOS_IMAX - 1 ? EOS_IMAX - 1 : iat);
    s.iat = iat;
    const double *g = tab.dg + (long)(iat - 1) * EOS_GRID;
    const double dI = g[G_X], dd = g[G_D], dd2 = g[G_D2], ddi = g[G_DI];
    double xd = fmax((s.din - dI) * ddi, 0.0);
    double mxd = 1.0 - xd;
    s.D.w0 = psi0(xd);
    s.D.w1 = psi1(xd) * dd;
    s.D.w2 = psi2(xd) * dd2;
    s.D.w0m = psi0(mxd);
    s.D.w1m = -psi1(mxd) * dd;
    s.D.w2m = psi2(mxd) * dd2;
    s.dD.w0 = dpsi0(xd) * ddi;
    s.dD.w1 = dpsi1(xd);
    s.dD.w2 = dpsi2(xd) * dd;
    s.dD.w0m = -dpsi0(mxd) * ddi;
    s.dD.w1m = dpsi1(mxd);
    s.dD.w2m = -dpsi2(mxd) * dd;
    s.x0 = xpsi0(xd);
    s.x1 = xpsi1(xd) * dd;
    s.x0m = xpsi0(mxd);
    s.x1m = -xpsi1(mxd) * dd;
}

// Full thermodynamic output of one evaluation.
struct EosFull {
    double p, e, dpdT, dpdr, dedT, dedr, gam1, dpde, dpdr_e;
};

// One evaluation at temperature temp (Source/eos.F90:195-518).  FULL=false
// computes only e and de/dT; FULL=true computes everything the path uses.
template <bool FULL>
__device__ __forceinline__ void eos_eval(const EosTable &tab, const EosDen &s, double temp, int &jat_out,
                                         EosFull &o)
{
    using namespace eosc;
    const double den = s.den, deni = s.deni, ytot1 = s.ytot1, ye = s.ye;
    const double tempi = 1.0 / temp;

    const double prad = asoli3 * temp * temp * temp * temp;
    const double dpraddt = 4.0 * prad * tempi;
    const double erad = 3.0 * prad * deni;
    const double deraddt = 3.0 * dpraddt * deni;

    const double pion = kergavo * ytot1 * den * temp;
    const double dpiondd = kergavo * ytot1 * temp;
    const double dpiondt = kergavo * ytot1 * den;
    const double eion = 1.5 * pion * deni;
    const double deiondt = 1.5 * dpiondt * deni;

    int jat = (int)((log10(temp) - tlo) * tstpi) + 1;
    jat = jat < 1 ? 1 : (jat > EOS_JMAX - 1 ? EOS_JMAX - 1 : jat);
    jat_out = jat;

    const double *g = tab.tg + (long)(jat - 1) * EOS_GRID;
    const double tJ = g[G_X], dt = g[G_D], dt2 = g[G_D2], dti = g[G_DI], dt2i = g[G_D2I];

    const double *n0 = tab.node + ((long)(s.iat - 1) + (long)EOS_IMAX * (jat - 1)) * EOS_NODE;
    const double2 *qa = reinterpret_cast<const double2 *>(n0), *qb = reinterpret_cast<const double2 *>(n0 + EOS_NODE),
                  *qc = reinterpret_cast<const double2 *>(n0 + (long)EOS_IMAX * EOS_NODE),
                  *qd = reinterpret_cast<const double2 *>(n0 + (long)EOS_IMAX * EOS_NODE + EOS_NODE);

    const double xt = fmax((temp - tJ) * dti, 0.0);
    const double mxt = 1.0 - xt;

    W6 T, dT, ddT;
    T.w0 = psi0(xt);
    T.w1 = psi1(xt) * dt;
    T.w2 = psi2(xt) * dt2;
    T.w0m = psi0(mxt);
    T.w1m = -psi1(mxt) * dt;
    T.w2m = psi2(mxt) * dt2;
    dT.w0 = dpsi0(xt) * dti;
    dT.w1 = dpsi1(xt);
    dT.w2 = dpsi2(xt) * dt;
    dT.w0m = -dpsi0(mxt) * dti;
    dT.w1m = dpsi1(mxt);
    dT.w2m = -dpsi2(mxt) * dt;
    ddT.w0 = ddpsi0(xt) * dt2i;
    ddT.w1 = ddpsi1(xt) * dti;
    ddT.w2 = ddpsi2(xt);
    ddT.w0m = ddpsi0(mxt) * dt2i;
    ddT.w1m = -ddpsi1(mxt) * dti;
    ddT.w2m = ddpsi2(mxt);

    // The biquintic sums (Source/eos.F90:915-950) of the free energy and its
    // derivatives, column by column: every sum adds the same terms in the
    // same order as h5, but only two columns of the four nodes are live at a
    // time.  Column q takes density weight dsel and temperature weight tsel.
    constexpr int dsel[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2};
    constexpr int tsel[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
    double free = 0.0, df_t = 0.0, df_tt = 0.0, df_d = 0.0, df_dt = 0.0;
#pragma unroll
    for (int c2 = 0; c2 < 5; ++c2) {
        const double2 va = qa[c2], vb = qb[c2], vc = qc[c2], vd = qd[c2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = 2 * c2 + h;
            if (q >= 9) break;
            const double a = h ? va.y : va.x, b = h ? vb.y : vb.x, c = h ? vc.y : vc.x, d = h ? vd.y : vd.x;
            const double wd = wsel(s.D, dsel[q]), wmd = wselm(s.D, dsel[q]);
            hcol(free, q == 0, a, b, c, d, wd, wmd, wsel(T, tsel[q]), wselm(T, tsel[q]));
            hcol(df_t, q == 0, a, b, c, d, wd, wmd, wsel(dT, tsel[q]), wselm(dT, tsel[q]));
            hcol(df_tt, q == 0, a, b, c, d, wd, wmd, wsel(ddT, tsel[q]), wselm(ddT, tsel[q]));
            if (FULL) {
                const double ed = wsel(s.dD, dsel[q]), emd = wselm(s.dD, dsel[q]);
                hcol(df_d, q == 0, a, b, c, d, ed, emd, wsel(T, tsel[q]), wselm(T, tsel[q]));
                hcol(df_dt, q == 0, a, b, c, d, ed, emd, wsel(dT, tsel[q]), wselm(dT, tsel[q]));
            }
        }
    }

    const double sele = -df_t * ye;
    const double dsepdt = -df_tt * ye;
    const double eele = ye * free + temp * sele;
    const double deepdt = temp * dsepdt;

    o.e = erad + eion + eele;
    o.dedT = deraddt + deiondt + deepdt;

    if (FULL) {
        const double xt0 = xpsi0(xt), xt1 = xpsi1(xt) * dt, xt0m = xpsi0(mxt), xt1m = -xpsi1(mxt) * dt;
        // the bicubic dpdf sum (Source/eos.F90:1034-1059), columns 9..12 in
        // h3_dpdf's order: (d0, t0), (d0, t1), (d1, t0), (d1, t1)
        double dpepdd = 0.0;
        {
            const double2 a0 = qa[4], b0 = qb[4], c0 = qc[4], d0 = qd[4];  // column 9 in .y
            hcol(dpepdd, true, a0.y, b0.y, c0.y, d0.y, s.x0, s.x0m, xt0, xt0m);
            const double2 a1 = qa[5], b1 = qb[5], c1 = qc[5], d1 = qd[5];  // columns 10, 11
            hcol(dpepdd, false, a1.x, b1.x, c1.x, d1.x, s.x0, s.x0m, xt1, xt1m);
            hcol(dpepdd, false, a1.y, b1.y, c1.y, d1.y, s.x1, s.x1m, xt0, xt0m);
            const double2 a2 = qa[6], b2 = qb[6], c2 = qc[6], d2 = qd[6];  // column 12 in .x
            hcol(dpepdd, false, a2.x, b2.x, c2.x, d2.x, s.x1, s.x1m, xt1, xt1m);
        }
        dpepdd = fmax(ye * dpepdd, 0.0);

        const double deraddd = -erad * deni;
        const double deiondd = (1.5 * dpiondd - eion) * deni;
        double x = s.din * s.din;
        const double pele = x * df_d;
        const double dpepdt = x * df_dt;
        x = ye * ye;
        const double dsepdd = -df_dt * x;
        const double deepdd = x * df_d + temp * dsepdd;

        const double pres = prad + pion + pele;
        const double dpresdd = 0.0 + dpiondd + dpepdd;
        const double dpresdt = dpraddt + dpiondt + dpepdt;
        const double denerdd = deraddd + deiondd + deepdd;
        const double denerdt = o.dedT;

        const double zz = pres * deni;
        const double zzi = den / pres;
        const double chit = temp / pres * dpresdt;
        const double chid = dpresdd * zzi;
        o.gam1 = chit * zz * chit / (temp * denerdt) + chid;
        o.p = pres;
        o.dpdT = dpresdt;
        o.dpdr = dpresdd;
        o.dpde = dpresdt / denerdt;
        o.dpdr_e = dpresdd - dpresdt * denerdd / denerdt;
        o.dedr = denerdd;
    }
}

// Result of a (rho, e) solve.
struct EosRE {
    double T, p, gam1, cs, dpde, dpdr_e;
    int iat, jat, niter;
};

// eos(eos_input_re, ...): Newton on T with the reference's step clamps
// (Source/eos.F90:524-550).  FULL=false when the caller only needs T
// (compute_temp restores e and discards everything else).
template <bool FULL>
__device__ __forceinline__ EosRE eos_re(const EosTable &tab, double rho, double T0, double e_want, double abar,
                                        double zbar)
{
    EosDen s;
    eos_density(tab, rho, abar, zbar, s);
    double temp = T0;
    bool converged = false;
    EosFull o;
    int jat = 0, iter;
    for (iter = 1; iter <= EOS_MAX_NEWTON; ++iter) {
        const bool last = converged || iter == EOS_MAX_NEWTON;
        if (FULL && last) {
            eos_eval<true>(tab, s, temp, jat, o);
        } else {
            if (converged) break;
            eos_eval<false>(tab, s, temp, jat, o);
        }
        if (converged) break;
        const double temp_old = temp;
        temp = temp - (o.e - e_want) / o.dedT;
        temp = fmax(0.5 * temp_old, fmin(temp, 2.0 * temp_old));
        temp = fmax(eosc::mintemp, temp);
        const double error = fabs((temp - temp_old) / temp_old);
        if (error < eosc::ttol) converged = true;
    }
    if (!converged) eos_note_newton_failure(tab);
    EosRE r;
    r.T = temp;
    r.iat = s.iat;
    r.jat = jat;
    r.niter = iter > EOS_MAX_NEWTON ? EOS_MAX_NEWTON : iter;
    if (FULL) {
        r.p = o.p;
        r.gam1 = o.gam1;
        r.cs = sqrt(o.gam1 * o.p / rho);
        r.dpde = o.dpde;
        r.dpdr_e = o.dpdr_e;
    }
    return r;
}

// eos(eos_input_rt, ...) for the floors: returns e (Source/eos.F90, converged at entry).
__device__ __forceinline__ double eos_rt_e(const EosTable &tab, double rho, double T, double abar, double zbar)
{
    EosDen s;
    eos_density(tab, rho, abar, zbar, s);
    EosFull o;
    int jat;
    eos_eval<false>(tab, s, T, jat, o);
    return o.e;
}

// eos(eos_input_rp, ...): Newton on T for a target pressure; returns e.
__device__ __forceinline__ double eos_rp_e(const EosTable &tab, double rho, double T0, double p_want, double abar,
                                           double zbar, double *T_out)
{
    EosDen s;
    eos_density(tab, rho, abar, zbar, s);
    double temp = T0;
    bool converged = false;
    EosFull o;
    int jat;
    for (int iter = 1; iter <= EOS_MAX_NEWTON; ++iter) {
        eos_eval<true>(tab, s, temp, jat, o);
        if (converged) break;
        const double temp_old = temp;
        temp = temp - (o.p - p_want) / o.dpdT;
        temp = fmax(0.5 * temp_old, fmin(temp, 2.0 * temp_old));
        temp = fmax(eosc::mintemp, temp);
        const double error = fabs((temp - temp_old) / temp_old);
        if (error < eosc::ttol) converged = true;
    }
    if (!converged) eos_note_newton_failure(tab);
    if (T_out) *T_out = temp;
    return o.e;
}

} // namespace mc

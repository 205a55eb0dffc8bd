/*
 * ORACLE — test infrastructure only.  Helmholtz EOS restated from
 * Source/eos.F90 (module eos_module).  Compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* Source/eos.F90:60-74 */
#define IMAX 541
#define JMAX 201
static const double tlo = 3.0, thi = 13.0;
static const double dlo = -12.0, dhi = 15.0;
static double tstp, tstpi, dstp, dstpi, mintemp;
static const double ttol = 1.0e-8;
static const int max_newton = 100; /* Source/eos.F90:126 */

/* Source/eos.F90:129-137 */
static const double h_ = 6.6260689633e-27;
static const double avo_eos = 6.0221417930e23;
static const double clight = 2.99792458e10;
static const double kerg = 1.380650424e-16;
static const double amu = 1.66053878283e-24;
static double asol, sioncon, kergavo, asoli3;

/* Source/eos.F90:77-94; 1-based Fortran indices mapped with (i-1)+IMAX*(j-1) */
static double d_[IMAX + 1], t_[JMAX + 1];
static double dt_[JMAX + 1], dt2_[JMAX + 1], dti_[JMAX + 1], dt2i_[JMAX + 1];
static double dd_[IMAX + 1], dd2_[IMAX + 1], ddi_[IMAX + 1], dd2i_[IMAX + 1];
static double *tab_ = NULL; /* 21 columns */

enum { F = 0, FD, FT, FDD, FTT, FDT, FDDT, FDTT, FDDTT, DPDF, DPDFD, DPDFT, DPDFDT,
       EF, EFD, EFT, EFDT, XF, XFD, XFT, XFDT, NTAB };

#define TB(c, i, j) tab_[(long)(c) * IMAX * JMAX + (long)((i) - 1) + (long)IMAX * ((j) - 1)]

static void set_constants(void)
{
    const double M_PI_ = 3.141592653589793238462643383279502884197;
    tstp = (thi - tlo) / (double)(float)(JMAX - 1); /* float(jmax-1), eos.F90:65 */
    tstpi = 1.0 / tstp;
    dstp = (dhi - dlo) / (double)(float)(IMAX - 1);
    dstpi = 1.0 / dstp;
    mintemp = pow(10.0, tlo);
    asol = 4.0 * 5.67051e-5 / clight;
    sioncon = (2.0 * M_PI_ * amu * kerg) / (h_ * h_);
    kergavo = kerg * avo_eos;
    asoli3 = asol / 3.0;
}

/* Grids and deltas, Source/eos.F90:626-632 and 699-712 */
static void build_grid(void)
{
    for (int j = 1; j <= JMAX; ++j) t_[j] = pow(10.0, tlo + (j - 1) * tstp);
    for (int i = 1; i <= IMAX; ++i) d_[i] = pow(10.0, dlo + (i - 1) * dstp);
    for (int j = 1; j <= JMAX - 1; ++j) {
        dt_[j] = t_[j + 1] - t_[j];
        dt2_[j] = dt_[j] * dt_[j];
        dti_[j] = 1.0 / dt_[j];
        dt2i_[j] = 1.0 / dt2_[j];
    }
    for (int i = 1; i <= IMAX - 1; ++i) {
        dd_[i] = d_[i + 1] - d_[i];
        dd2_[i] = dd_[i] * dd_[i];
        ddi_[i] = 1.0 / dd_[i];
        dd2i_[i] = 1.0 / dd2_[i];
    }
}

void orc_eos_grid(double *t201, double *d541)
{
    for (int j = 1; j <= JMAX; ++j) t201[j - 1] = t_[j];
    for (int i = 1; i <= IMAX; ++i) d541[i - 1] = d_[i];
}

void orc_eos_init_arrays(const double *tab)
{
    set_constants();
    if (!tab_) tab_ = (double *)malloc(sizeof(double) * NTAB * IMAX * JMAX);
    memcpy(tab_, tab, sizeof(double) * NTAB * IMAX * JMAX);
    build_grid();
}

/* Source/eos.F90:580-734: read the four sections, j outer, i inner */
int orc_eos_init(const char *path)
{
    static const int sec1[9] = {F, FD, FT, FDD, FTT, FDT, FDDT, FDTT, FDDTT};
    static const int sec2[4] = {DPDF, DPDFD, DPDFT, DPDFDT};
    static const int sec3[4] = {EF, EFD, EFT, EFDT};
    static const int sec4[4] = {XF, XFD, XFT, XFDT};
    const int *secs[4] = {sec1, sec2, sec3, sec4};
    const int ncol[4] = {9, 4, 4, 4};
    FILE *fp = fopen(path, "r");
    if (!fp) return -1; /* amrex_error('eos_init: Failed to open helm_table.dat') */
    set_constants();
    if (!tab_) tab_ = (double *)malloc(sizeof(double) * NTAB * IMAX * JMAX);
    for (int s = 0; s < 4; ++s)
        for (int j = 1; j <= JMAX; ++j)
            for (int i = 1; i <= IMAX; ++i)
                for (int c = 0; c < ncol[s]; ++c)
                    if (fscanf(fp, "%lf", &TB(secs[s][c], i, j)) != 1) { fclose(fp); return -2; }
    fclose(fp);
    build_grid();
    return 0;
}

void orc_eos_finalize(void)
{
    free(tab_);
    tab_ = NULL;
}

/* Quintic Hermite basis, Source/eos.F90:740-911 */
static real psi0(real z) { return (z * z * z) * (z * (-6.0 * z + 15.0) - 10.0) + 1.0; }
static real dpsi0(real z) { return (z * z) * (z * (-30.0 * z + 60.0) - 30.0); }
static real ddpsi0(real z) { return z * (z * (-120.0 * z + 180.0) - 60.0); }
static real psi1(real z) { return z * ((z * z) * (z * (-3.0 * z + 8.0) - 6.0) + 1.0); }
static real dpsi1(real z) { return z * z * (z * (-15.0 * z + 32.0) - 18.0) + 1.0; }
static real ddpsi1(real z) { return z * (z * (-60.0 * z + 96.0) - 36.0); }
static real psi2(real z) { return 0.5 * z * z * (z * (z * (-z + 3.0) - 3.0) + 1.0); }
static real dpsi2(real z) { return 0.5 * z * (z * (z * (-5.0 * z + 12.0) - 9.0) + 2.0); }
static real ddpsi2(real z) { return 0.5 * (z * (z * (-20.0 * z + 36.0) - 18.0) + 2.0); }

/* Biquintic Hermite, Source/eos.F90:915-950 */
static real h5(const double *fi, real w0t, real w1t, real w2t, real w0mt, real w1mt,
                 real w2mt, real w0d, real w1d, real w2d, real w0md, real w1md,
                 real w2md)
{
    return fi[0] * w0d * w0t + fi[1] * w0md * w0t + fi[2] * w0d * w0mt + fi[3] * w0md * w0mt +
           fi[4] * w0d * w1t + fi[5] * w0md * w1t + fi[6] * w0d * w1mt + fi[7] * w0md * w1mt +
           fi[8] * w0d * w2t + fi[9] * w0md * w2t + fi[10] * w0d * w2mt + fi[11] * w0md * w2mt +
           fi[12] * w1d * w0t + fi[13] * w1md * w0t + fi[14] * w1d * w0mt + fi[15] * w1md * w0mt +
           fi[16] * w2d * w0t + fi[17] * w2md * w0t + fi[18] * w2d * w0mt + fi[19] * w2md * w0mt +
           fi[20] * w1d * w1t + fi[21] * w1md * w1t + fi[22] * w1d * w1mt + fi[23] * w1md * w1mt +
           fi[24] * w2d * w1t + fi[25] * w2md * w1t + fi[26] * w2d * w1mt + fi[27] * w2md * w1mt +
           fi[28] * w1d * w2t + fi[29] * w1md * w2t + fi[30] * w1d * w2mt + fi[31] * w1md * w2mt +
           fi[32] * w2d * w2t + fi[33] * w2md * w2t + fi[34] * w2d * w2mt + fi[35] * w2md * w2mt;
}

/* Cubic Hermite basis, Source/eos.F90:955-1031 */
static real xpsi0(real z) { return z * z * (2.0 * z - 3.0) + 1.0; }
static real xdpsi0(real z) { return z * (6.0 * z - 6.0); }
static real xpsi1(real z) { return z * (z * (z - 2.0) + 1.0); }
static real xdpsi1(real z) { return z * (3.0 * z - 4.0) + 1.0; }

/* Bicubic Hermite, Source/eos.F90:1034-1059 */
static real h3(const double *fi, real w0t, real w1t, real w0mt, real w1mt, real w0d,
                 real w1d, real w0md, real w1md)
{
    return fi[0] * w0d * w0t + fi[1] * w0md * w0t + fi[2] * w0d * w0mt + fi[3] * w0md * w0mt +
           fi[4] * w0d * w1t + fi[5] * w0md * w1t + fi[6] * w0d * w1mt + fi[7] * w0md * w1mt +
           fi[8] * w1d * w0t + fi[9] * w1md * w0t + fi[10] * w1d * w0mt + fi[11] * w1md * w0mt +
           fi[12] * w1d * w1t + fi[13] * w1md * w1t + fi[14] * w1d * w1mt + fi[15] * w1md * w1mt;
}

static void gather4(double *fi, int c, int iat, int jat)
{
    fi[0] = TB(c, iat, jat);
    fi[1] = TB(c, iat + 1, jat);
    fi[2] = TB(c, iat, jat + 1);
    fi[3] = TB(c, iat + 1, jat + 1);
}

/* Source/eos.F90:141-576 */
void orc_eos(int input, orc_eos_t *state)
{
    int converged, iter, iat = 0, jat = 0, niter = 0;
    real temp_old, v_want = 0.0, v, dvdx, error;
    real temp, den, din, deni, tempi, ytot1, ye;
    real pres, ener, entr, dpresdd, dpresdt, denerdd, denerdt, dentrdd, dentrdt;
    real pele, dpepdt, dpepdd, eele, deepdt, deepdd, sele, dsepdd, dsepdt;
    real prad, dpraddd, dpraddt, erad, deraddd, deraddt, srad, dsraddd, dsraddt;
    real pion, dpiondd, dpiondt, eion, deiondd, deiondt, sion, dsiondd, dsiondt;
    real s, x, y, z, zz, zzi, chit, chid;
    real free, df_d, df_t, df_tt, df_dt;
    real xt, xd, mxt, mxd;
    double fi[36];
    real si0t, si1t, si2t, si0mt, si1mt, si2mt;
    real si0d, si1d, si2d, si0md, si1md, si2md;
    real dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt;
    real dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md;
    real ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt;

    converged = (input == EOS_INPUT_RT); /* :181-185 */
    if (input == EOS_INPUT_RE) v_want = state->e;
    else if (input == EOS_INPUT_RP) v_want = state->p;

    for (iter = 1; iter <= max_newton; ++iter) { /* :193 */
        niter = iter;
        temp = state->T;
        den = state->rho;

        ytot1 = 1.0 / state->abar;
        ye = state->zbar / state->abar;
        din = ye * den;

        deni = 1.0 / den;
        tempi = 1.0 / temp;

        /* radiation, :206-217 */
        prad = asoli3 * temp * temp * temp * temp;
        dpraddd = 0.0;
        dpraddt = 4.0 * prad * tempi;
        erad = 3.0 * prad * deni;
        deraddd = -erad * deni;
        deraddt = 3.0 * dpraddt * deni;
        srad = (prad * deni + erad) * tempi;
        dsraddd = (dpraddd * deni - prad * deni * deni + deraddd) * tempi;
        dsraddt = (dpraddt * deni + deraddt - srad) * tempi;

        /* ions, :219-237 */
        pion = kergavo * ytot1 * den * temp;
        dpiondd = kergavo * ytot1 * temp;
        dpiondt = kergavo * ytot1 * den;
        eion = 1.5 * pion * deni;
        deiondd = (1.5 * dpiondd - eion) * deni;
        deiondt = 1.5 * dpiondt * deni;
        x = state->abar * state->abar * sqrt(state->abar) * deni / avo_eos;
        s = sioncon * temp;
        z = x * s * sqrt(s);
        y = log(z);
        sion = (pion * deni + eion) * tempi + kergavo * ytot1 * y;
        dsiondd = (dpiondd * deni - pion * deni * deni + deiondd) * tempi - kergavo * deni * ytot1;
        dsiondt = (dpiondt * deni + deiondt) * tempi - (pion * deni + eion) * tempi * tempi +
                  1.5 * kergavo * tempi * ytot1;

        /* hash locate, :242-245 (Fortran int() truncates toward zero) */
        jat = (int)((log10(temp) - tlo) * tstpi) + 1;
        jat = jat < 1 ? 1 : (jat > JMAX - 1 ? JMAX - 1 : jat);
        iat = (int)((log10(din) - dlo) * dstpi) + 1;
        iat = iat < 1 ? 1 : (iat > IMAX - 1 ? IMAX - 1 : iat);

        /* :248-283 */
        gather4(fi + 0, F, iat, jat);
        gather4(fi + 4, FT, iat, jat);
        gather4(fi + 8, FTT, iat, jat);
        gather4(fi + 12, FD, iat, jat);
        gather4(fi + 16, FDD, iat, jat);
        gather4(fi + 20, FDT, iat, jat);
        gather4(fi + 24, FDDT, iat, jat);
        gather4(fi + 28, FDTT, iat, jat);
        gather4(fi + 32, FDDTT, iat, jat);

        /* :286-289 */
        xt = fmax((temp - t_[jat]) * dti_[jat], 0.0);
        xd = fmax((din - d_[iat]) * ddi_[iat], 0.0);
        mxt = 1.0 - xt;
        mxd = 1.0 - xd;

        /* :292-332 */
        si0t = psi0(xt);
        si1t = psi1(xt) * dt_[jat];
        si2t = psi2(xt) * dt2_[jat];
        si0mt = psi0(mxt);
        si1mt = -psi1(mxt) * dt_[jat];
        si2mt = psi2(mxt) * dt2_[jat];
        si0d = psi0(xd);
        si1d = psi1(xd) * dd_[iat];
        si2d = psi2(xd) * dd2_[iat];
        si0md = psi0(mxd);
        si1md = -psi1(mxd) * dd_[iat];
        si2md = psi2(mxd) * dd2_[iat];

        dsi0t = dpsi0(xt) * dti_[jat];
        dsi1t = dpsi1(xt);
        dsi2t = dpsi2(xt) * dt_[jat];
        dsi0mt = -dpsi0(mxt) * dti_[jat];
        dsi1mt = dpsi1(mxt);
        dsi2mt = -dpsi2(mxt) * dt_[jat];
        dsi0d = dpsi0(xd) * ddi_[iat];
        dsi1d = dpsi1(xd);
        dsi2d = dpsi2(xd) * dd_[iat];
        dsi0md = -dpsi0(mxd) * ddi_[iat];
        dsi1md = dpsi1(mxd);
        dsi2md = -dpsi2(mxd) * dd_[iat];

        ddsi0t = ddpsi0(xt) * dt2i_[jat];
        ddsi1t = ddpsi1(xt) * dti_[jat];
        ddsi2t = ddpsi2(xt);
        ddsi0mt = ddpsi0(mxt) * dt2i_[jat];
        ddsi1mt = -ddpsi1(mxt) * dti_[jat];
        ddsi2mt = ddpsi2(mxt);

        /* :335-357 */
        free = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, si0d, si1d, si2d, si0md, si1md, si2md);
        df_d = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md,
                  dsi2md);
        df_t = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, si0d, si1d, si2d, si0md, si1md,
                  si2md);
        df_tt = h5(fi, ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt, si0d, si1d, si2d, si0md,
                   si1md, si2md);
        df_dt = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, dsi0d, dsi1d, dsi2d, dsi0md,
                   dsi1md, dsi2md);

        /* cubic weights, :362-385 */
        si0t = xpsi0(xt);
        si1t = xpsi1(xt) * dt_[jat];
        si0mt = xpsi0(mxt);
        si1mt = -xpsi1(mxt) * dt_[jat];
        si0d = xpsi0(xd);
        si1d = xpsi1(xd) * dd_[iat];
        si0md = xpsi0(mxd);
        si1md = -xpsi1(mxd) * dd_[iat];
        dsi0t = xdpsi0(xt) * dti_[jat];
        dsi1t = xdpsi1(xt);
        dsi0mt = -xdpsi0(mxt) * dti_[jat];
        dsi1mt = xdpsi1(mxt);
        dsi0d = xdpsi0(xd) * ddi_[iat];
        dsi1d = xdpsi1(xd);
        dsi0md = -xdpsi0(mxd) * ddi_[iat];
        dsi1md = xdpsi1(mxd);
        (void)dsi0t; (void)dsi1t; (void)dsi0mt; (void)dsi1mt;
        (void)dsi0d; (void)dsi1d; (void)dsi0md; (void)dsi1md;

        /* :388-409 */
        gather4(fi + 0, DPDF, iat, jat);
        gather4(fi + 4, DPDFT, iat, jat);
        gather4(fi + 8, DPDFD, iat, jat);
        gather4(fi + 12, DPDFDT, iat, jat);
        dpepdd = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);
        dpepdd = fmax(ye * dpepdd, 0.0);

        /* :412-432 */
        gather4(fi + 0, EF, iat, jat);
        gather4(fi + 4, EFT, iat, jat);
        gather4(fi + 8, EFD, iat, jat);
        gather4(fi + 12, EFDT, iat, jat);
        state->eta = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);

        /* :435-456 */
        gather4(fi + 0, XF, iat, jat);
        gather4(fi + 4, XFT, iat, jat);
        gather4(fi + 8, XFD, iat, jat);
        gather4(fi + 12, XFDT, iat, jat);
        state->xne = h3(fi, si0t, si1t, si0mt, si1mt, si0d, si1d, si0md, si1md);
        state->xnp = 0.0;

        /* :464-475 */
        x = din * din;
        pele = x * df_d;
        dpepdt = x * df_dt;
        x = ye * ye;
        sele = -df_t * ye;
        dsepdt = -df_tt * ye;
        dsepdd = -df_dt * x;
        eele = ye * free + temp * sele;
        deepdt = temp * dsepdt;
        deepdd = x * df_d + temp * dsepdd;

        /* :478-489 */
        pres = prad + pion + pele;
        ener = erad + eion + eele;
        entr = srad + sion + sele;
        dpresdd = dpraddd + dpiondd + dpepdd;
        dpresdt = dpraddt + dpiondt + dpepdt;
        denerdd = deraddd + deiondd + deepdd;
        denerdt = deraddt + deiondt + deepdt;
        dentrdd = dsraddd + dsiondd + dsepdd;
        dentrdt = dsraddt + dsiondt + dsepdt;

        /* :491-518 */
        zz = pres * deni;
        zzi = den / pres;
        chit = temp / pres * dpresdt;
        chid = dpresdd * zzi;
        state->cv = denerdt;
        state->gam1 = chit * zz * chit / (temp * state->cv) + chid;
        state->cp = state->cv * state->gam1 / chid;
        state->p = pres;
        state->dpdT = dpresdt;
        state->dpdr = dpresdd;
        state->dpde = dpresdt / denerdt;
        state->dpdr_e = dpresdd - dpresdt * denerdd / denerdt;
        state->e = ener;
        state->dedT = denerdt;
        state->dedr = denerdd;
        state->s = entr;
        state->dsdT = dentrdt;
        state->dsdr = dentrdd;
        state->h = ener + pres / den;
        state->dhdr = denerdd + dpresdd / den - pres / (den * den);
        state->dhdT = denerdt + dpresdt / den;
        state->pele = pele;
        state->ppos = 0.0;

        if (converged) break; /* :520-522 */

        /* Newton update, :526-550 */
        temp_old = temp;
        if (input == EOS_INPUT_RE) { v = state->e; dvdx = state->dedT; }
        else { v = state->p; dvdx = state->dpdT; }
        temp = temp - (v - v_want) / dvdx;
        temp = fmax(0.5 * temp_old, fmin(temp, 2.0 * temp_old));
        temp = fmax(mintemp, temp);
        state->T = temp;
        error = fabs((temp - temp_old) / temp_old);
        if (error < ttol) converged = 1;
    }

    state->cs = sqrt(state->gam1 * state->p / state->rho); /* :556 */
    state->dpdA = 0.0;
    state->dpdZ = 0.0;
    state->dedA = 0.0;
    state->dedZ = 0.0;
    if (input == EOS_INPUT_RE) state->e = v_want; /* :566-574 */
    else if (input == EOS_INPUT_RP) state->p = v_want;
    state->iat = iat;
    state->jat = jat;
    state->niter = niter;
}

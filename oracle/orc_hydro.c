/*
 * ORACLE — test infrastructure only.  Per-box kernels restated from
 * Source/Castro_hydro_nd.F90, Source/ppm.F90, Source/riemann.F90,
 * Source/trans.F90, Source/timestep.F90, Source/Castro_nd.F90 and
 * Source/Prob.F90.  Compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>

#include "orc.h"

/* Fortran-ordered array view: a(i,j,k,n), n 0-based. */
typedef struct {
    real *p;
    int lo0, lo1, lo2;
    long s1, s2, s3;
} arr_t;

static arr_t mk(const real *p, const int *lo, const int *hi)
{
    arr_t a;
    a.p = (real *)p;
    a.lo0 = lo[0];
    a.lo1 = lo[1];
    a.lo2 = lo[2];
    a.s1 = hi[0] - lo[0] + 1;
    a.s2 = a.s1 * (hi[1] - lo[1] + 1);
    a.s3 = a.s2 * (hi[2] - lo[2] + 1);
    return a;
}
#define A(a, i, j, k, n) \
    ((a).p[((i) - (a).lo0) + (a).s1 * ((j) - (a).lo1) + (a).s2 * ((k) - (a).lo2) + (a).s3 * (n)])

/* Source/network.F90:7-21 */
static const real aion[ORC_NSPEC] = {4.0, 12.0, 16.0, 20.0, 24.0, 28.0, 32.0,
                                       36.0, 40.0, 44.0, 48.0, 52.0, 56.0};
static const real zion[ORC_NSPEC] = {2.0, 6.0, 8.0, 10.0, 12.0, 14.0, 16.0,
                                       18.0, 20.0, 22.0, 24.0, 26.0, 28.0};
static real aion_inv(int n) { return 1.0 / aion[n]; }

/* Source/Castro_nd.F90:59-65 */
static const real small_dens = 1.0e-12;
static const real small_temp = 1.0e3;
static const real small_pres = 1.e-200;
static const real smallu = 1.e-12;
static const real small = 1.e-8;
static const real dual_energy_eta1 = 1.e0;

/* amrex_constants_module */
static const real ZERO = 0.0, HALF = 0.5, ONE = 1.0, TWO = 2.0, THREE = 3.0, SIX = 6.0;
static const real FOURTH = 0.25;
#define SIXTH (1.0 / 6.0)
#define TWO3RD (2.0 / 3.0)
#define FOUR3RD (4.0 / 3.0)
static const real M_PI_ = 3.141592653589793238462643383279502884197;

/* Fortran SIGN(ONE, x): |1| with the sign bit of x (signed zero honoured) */
static real fsign1(real x) { return copysign(1.0, x); }

#define LOOP3(lo, hi) \
    for (int k = (lo)[2]; k <= (hi)[2]; ++k) \
        for (int j = (lo)[1]; j <= (hi)[1]; ++j) \
            for (int i = (lo)[0]; i <= (hi)[0]; ++i)

/* ------------------------------------------------------------------ */
/* Source/Castro_hydro_nd.F90:11-69 */
void orc_divu(const int *lo, const int *hi, const real *qp_, const int *q_lo, const int *q_hi,
              const real *dx, real *divp, const int *d_lo, const int *d_hi)
{
    arr_t q = mk(qp_, q_lo, q_hi), div = mk(divp, d_lo, d_hi);
    real dxinv = ONE / dx[0], dyinv = ONE / dx[1], dzinv = ONE / dx[2];
    LOOP3(lo, hi)
    {
        real ux = FOURTH *
                    (+A(q, i, j, k, QU) - A(q, i - 1, j, k, QU) + A(q, i, j, k - 1, QU) -
                     A(q, i - 1, j, k - 1, QU) + A(q, i, j - 1, k, QU) - A(q, i - 1, j - 1, k, QU) +
                     A(q, i, j - 1, k - 1, QU) - A(q, i - 1, j - 1, k - 1, QU)) *
                    dxinv;
        real vy = FOURTH *
                    (+A(q, i, j, k, QV) - A(q, i, j - 1, k, QV) + A(q, i, j, k - 1, QV) -
                     A(q, i, j - 1, k - 1, QV) + A(q, i - 1, j, k, QV) - A(q, i - 1, j - 1, k, QV) +
                     A(q, i - 1, j, k - 1, QV) - A(q, i - 1, j - 1, k - 1, QV)) *
                    dyinv;
        real wz = FOURTH *
                    (+A(q, i, j, k, QW) - A(q, i, j, k - 1, QW) + A(q, i, j - 1, k, QW) -
                     A(q, i, j - 1, k - 1, QW) + A(q, i - 1, j, k, QW) - A(q, i - 1, j, k - 1, QW) +
                     A(q, i - 1, j - 1, k, QW) - A(q, i - 1, j - 1, k - 1, QW)) *
                    dzinv;
        A(div, i, j, k, 0) = ux + vy + wz;
    }
}

/* Source/Castro_hydro_nd.F90:73-170 */
void orc_ctoprim(const int *lo, const int *hi, const real *up, const int *u_lo, const int *u_hi,
                 real *qp_, const int *q_lo, const int *q_hi, real *qap, const int *qa_lo,
                 const int *qa_hi)
{
    arr_t u = mk(up, u_lo, u_hi), q = mk(qp_, q_lo, q_hi), qaux = mk(qap, qa_lo, qa_hi);
    LOOP3(lo, hi)
    {
        A(q, i, j, k, QRHO) = A(u, i, j, k, URHO);
        real rhoinv = ONE / A(q, i, j, k, QRHO);
        real vel0 = A(u, i, j, k, UMX) * rhoinv;
        real vel1 = A(u, i, j, k, UMY) * rhoinv;
        real vel2 = A(u, i, j, k, UMZ) * rhoinv;
        A(q, i, j, k, QU) = vel0;
        A(q, i, j, k, QV) = vel1;
        A(q, i, j, k, QW) = vel2;
        real kineng = HALF * A(q, i, j, k, QRHO) *
                        (A(q, i, j, k, QU) * A(q, i, j, k, QU) + A(q, i, j, k, QV) * A(q, i, j, k, QV) +
                         A(q, i, j, k, QW) * A(q, i, j, k, QW));
        if ((A(u, i, j, k, UEDEN) - kineng) / A(u, i, j, k, UEDEN) > dual_energy_eta1)
            A(q, i, j, k, QREINT) = (A(u, i, j, k, UEDEN) - kineng) * rhoinv;
        else
            A(q, i, j, k, QREINT) = A(u, i, j, k, UEINT) * rhoinv;
        A(q, i, j, k, QTEMP) = A(u, i, j, k, UTEMP);
        for (int n = 0; n < ORC_NSPEC; ++n) A(q, i, j, k, QFS + n) = A(u, i, j, k, UFS + n) * rhoinv;

        orc_eos_t st;
        st.T = A(q, i, j, k, QTEMP);
        st.rho = A(q, i, j, k, QRHO);
        st.e = A(q, i, j, k, QREINT);
        real s1 = 0.0, s2 = 0.0;
        for (int n = 0; n < ORC_NSPEC; ++n) s1 = s1 + A(q, i, j, k, QFS + n) * aion_inv(n);
        st.abar = ONE / s1;
        for (int n = 0; n < ORC_NSPEC; ++n) s2 = s2 + A(q, i, j, k, QFS + n) * zion[n] * aion_inv(n);
        st.zbar = st.abar * s2;
        orc_eos(EOS_INPUT_RE, &st);

        A(q, i, j, k, QTEMP) = st.T;
        A(q, i, j, k, QREINT) = st.e * A(q, i, j, k, QRHO);
        A(q, i, j, k, QPRES) = st.p;
        A(q, i, j, k, QGAME) = A(q, i, j, k, QPRES) / A(q, i, j, k, QREINT) + ONE;
        A(qaux, i, j, k, QDPDR) = st.dpdr_e;
        A(qaux, i, j, k, QDPDE) = st.dpde;
        A(qaux, i, j, k, QGAMC) = st.gam1;
        A(qaux, i, j, k, QC) = st.cs;
    }
}

/* Source/Castro_hydro_nd.F90:174-245 (keeps the duplicated div(i,j+1,k) /
   div(i+1,j,k) terms of :216-217 and :223-224) */
void orc_apply_av(const int *lo, const int *hi, int idir, const real *dx, const real *divp,
                  const int *d_lo, const int *d_hi, const real *up, const int *u_lo,
                  const int *u_hi, real *fp, const int *f_lo, const int *f_hi)
{
    arr_t div = mk(divp, d_lo, d_hi), u = mk(up, u_lo, u_hi), flux = mk(fp, f_lo, f_hi);
    const real difmag = 0.1;
    for (int n = 0; n < ORC_NVAR; ++n) {
        if (n == UTEMP) continue;
        LOOP3(lo, hi)
        {
            real div1;
            if (idir == 1) {
                div1 = FOURTH * (A(div, i, j, k, 0) + A(div, i, j + 1, k, 0) + A(div, i, j, k + 1, 0) +
                                 A(div, i, j + 1, k, 0));
                div1 = difmag * fmin(ZERO, div1);
                div1 = div1 * (A(u, i, j, k, n) - A(u, i - 1, j, k, n));
            } else if (idir == 2) {
                div1 = FOURTH * (A(div, i, j, k, 0) + A(div, i + 1, j, k, 0) + A(div, i, j, k + 1, 0) +
                                 A(div, i + 1, j, k, 0));
                div1 = difmag * fmin(ZERO, div1);
                div1 = div1 * (A(u, i, j, k, n) - A(u, i, j - 1, k, n));
            } else {
                div1 = FOURTH * (A(div, i, j, k, 0) + A(div, i + 1, j, k, 0) + A(div, i, j + 1, k, 0) +
                                 A(div, i + 1, j + 1, k, 0));
                div1 = difmag * fmin(ZERO, div1);
                div1 = div1 * (A(u, i, j, k, n) - A(u, i, j, k - 1, n));
            }
            A(flux, i, j, k, n) = A(flux, i, j, k, n) + dx[idir - 1] * div1;
        }
    }
}

/* Source/Castro_hydro_nd.F90:248-297 */
void orc_normalize_species_fluxes(const int *lo, const int *hi, real *fp, const int *f_lo,
                                  const int *f_hi)
{
    arr_t flux = mk(fp, f_lo, f_hi);
    LOOP3(lo, hi)
    {
        real sum = ZERO, fac;
        for (int n = UFS; n < UFS + ORC_NSPEC; ++n) sum = sum + A(flux, i, j, k, n);
        if (sum != ZERO) fac = A(flux, i, j, k, URHO) / sum;
        else fac = ONE;
        for (int n = UFS; n < UFS + ORC_NSPEC; ++n) A(flux, i, j, k, n) = A(flux, i, j, k, n) * fac;
    }
}

/* Source/Castro_hydro_nd.F90:301-340 */
void orc_store_flux(const int *lo, const int *hi, real *fop, const int *fo_lo, const int *fo_hi,
                    const real *fip, const int *fi_lo, const int *fi_hi, const real *ap,
                    const int *a_lo, const int *a_hi, real dt)
{
    arr_t fo = mk(fop, fo_lo, fo_hi), fi = mk(fip, fi_lo, fi_hi), area = mk(ap, a_lo, a_hi);
    for (int n = 0; n < ORC_NVAR; ++n)
        LOOP3(lo, hi) A(fo, i, j, k, n) = dt * A(fi, i, j, k, n) * A(area, i, j, k, 0);
}

/* Source/Castro_hydro_nd.F90:344-441 */
void orc_fill_hydro_source(const int *lo, const int *hi, const real *up, const int *u_lo,
                           const int *u_hi, const real *qp_, const int *q_lo, const int *q_hi,
                           real *sp, const int *s_lo, const int *s_hi, const real *f1p,
                           const int *f1_lo, const int *f1_hi, const real *f2p, const int *f2_lo,
                           const int *f2_hi, const real *f3p, const int *f3_lo, const int *f3_hi,
                           const real *qxp, const int *qx_lo, const int *qx_hi, const real *qyp,
                           const int *qy_lo, const int *qy_hi, const real *qzp, const int *qz_lo,
                           const int *qz_hi, const real *a1p, const int *a1_lo, const int *a1_hi,
                           const real *a2p, const int *a2_lo, const int *a2_hi, const real *a3p,
                           const int *a3_lo, const int *a3_hi, const real *vp, const int *v_lo,
                           const int *v_hi, const real *dx, real dt)
{
    (void)up; (void)u_lo; (void)u_hi; (void)qp_; (void)q_lo; (void)q_hi; (void)dt;
    arr_t src = mk(sp, s_lo, s_hi), f1 = mk(f1p, f1_lo, f1_hi), f2 = mk(f2p, f2_lo, f2_hi),
          f3 = mk(f3p, f3_lo, f3_hi), qx = mk(qxp, qx_lo, qx_hi), qy = mk(qyp, qy_lo, qy_hi),
          qz = mk(qzp, qz_lo, qz_hi), a1 = mk(a1p, a1_lo, a1_hi), a2 = mk(a2p, a2_lo, a2_hi),
          a3 = mk(a3p, a3_lo, a3_hi), vol = mk(vp, v_lo, v_hi);
    for (int n = 0; n < ORC_NVAR; ++n)
        LOOP3(lo, hi)
        {
            real volinv = ONE / A(vol, i, j, k, 0);
            A(src, i, j, k, n) =
                A(src, i, j, k, n) +
                (A(f1, i, j, k, n) * A(a1, i, j, k, 0) - A(f1, i + 1, j, k, n) * A(a1, i + 1, j, k, 0) +
                 A(f2, i, j, k, n) * A(a2, i, j, k, 0) - A(f2, i, j + 1, k, n) * A(a2, i, j + 1, k, 0) +
                 A(f3, i, j, k, n) * A(a3, i, j, k, 0) - A(f3, i, j, k + 1, n) * A(a3, i, j, k + 1, 0)) *
                    volinv;
            if (n == UEINT) {
                real pdivu =
                    HALF * (A(qx, i + 1, j, k, GDPRES) + A(qx, i, j, k, GDPRES)) *
                        (A(qx, i + 1, j, k, GDU) - A(qx, i, j, k, GDU)) / dx[0] +
                    HALF * (A(qy, i, j + 1, k, GDPRES) + A(qy, i, j, k, GDPRES)) *
                        (A(qy, i, j + 1, k, GDV) - A(qy, i, j, k, GDV)) / dx[1] +
                    HALF * (A(qz, i, j, k + 1, GDPRES) + A(qz, i, j, k, GDPRES)) *
                        (A(qz, i, j, k + 1, GDW) - A(qz, i, j, k, GDW)) / dx[2];
                A(src, i, j, k, n) = A(src, i, j, k, n) - pdivu;
            }
        }
}

/* ------------------------------------------------------------------ */
/* Source/ppm.F90:11-128 */
void orc_uflatten(int i, int j, int k, const real *qp_, const int *q_lo, const int *q_hi,
                  real *flatn)
{
    arr_t q = mk(qp_, q_lo, q_hi);
    const real shktst = 0.33, zcut1 = 0.75, zcut2 = 0.85;
    const real dzcut = ONE / (zcut2 - zcut1);
    real pl, pr, dp, dp2, du, z, z2, chi, chi2, denom, denominv, zeta, tst, tmp;
    int ishft;
    *flatn = ONE;
    for (int idir = 1; idir <= 3; ++idir) {
        if (idir == 1) {
            pr = A(q, i + 1, j, k, QPRES);
            pl = A(q, i - 1, j, k, QPRES);
            dp2 = A(q, i + 2, j, k, QPRES) - A(q, i - 2, j, k, QPRES);
            du = A(q, i + 1, j, k, QU) - A(q, i - 1, j, k, QU);
        } else if (idir == 2) {
            pr = A(q, i, j + 1, k, QPRES);
            pl = A(q, i, j - 1, k, QPRES);
            dp2 = A(q, i, j + 2, k, QPRES) - A(q, i, j - 2, k, QPRES);
            du = A(q, i, j + 1, k, QV) - A(q, i, j - 1, k, QV);
        } else {
            pr = A(q, i, j, k + 1, QPRES);
            pl = A(q, i, j, k - 1, QPRES);
            dp2 = A(q, i, j, k + 2, QPRES) - A(q, i, j, k - 2, QPRES);
            du = A(q, i, j, k + 1, QW) - A(q, i, j, k - 1, QW);
        }
        dp = pr - pl;
        ishft = (dp > ZERO) ? 1 : -1;
        denom = fmax(small_pres, fabs(dp2));
        denominv = ONE / denom;
        zeta = fabs(dp) * denominv;
        z = fmin(ONE, fmax(ZERO, dzcut * (zeta - zcut1)));
        tst = (du <= ZERO) ? ONE : ZERO;
        tmp = fmin(pr, pl);
        chi = (fabs(dp) > shktst * tmp) ? tst : ZERO;

        if (idir == 1) {
            pr = A(q, i + 1 - ishft, j, k, QPRES);
            pl = A(q, i - 1 - ishft, j, k, QPRES);
            dp2 = A(q, i + 2 - ishft, j, k, QPRES) - A(q, i - 2 - ishft, j, k, QPRES);
            du = A(q, i + 1 - ishft, j, k, QU) - A(q, i - 1 - ishft, j, k, QU);
        } else if (idir == 2) {
            pr = A(q, i, j + 1 - ishft, k, QPRES);
            pl = A(q, i, j - 1 - ishft, k, QPRES);
            dp2 = A(q, i, j + 2 - ishft, k, QPRES) - A(q, i, j - 2 - ishft, k, QPRES);
            du = A(q, i, j + 1 - ishft, k, QV) - A(q, i, j - 1 - ishft, k, QV);
        } else {
            pr = A(q, i, j, k + 1 - ishft, QPRES);
            pl = A(q, i, j, k - 1 - ishft, QPRES);
            dp2 = A(q, i, j, k + 2 - ishft, QPRES) - A(q, i, j, k - 2 - ishft, QPRES);
            du = A(q, i, j, k + 1 - ishft, QW) - A(q, i, j, k - 1 - ishft, QW);
        }
        dp = pr - pl;
        denom = fmax(small_pres, fabs(dp2));
        denominv = ONE / denom;
        zeta = fabs(dp) * denominv;
        z2 = fmin(ONE, fmax(ZERO, dzcut * (zeta - zcut1)));
        tst = (du <= ZERO) ? ONE : ZERO;
        tmp = fmin(pr, pl);
        chi2 = (fabs(dp) > shktst * tmp) ? tst : ZERO;

        *flatn = fmin(*flatn, ONE - fmax(chi2 * z2, chi * z));
    }
}

/* Source/ppm.F90:132-236; s5[0..4] = s(-2..2) */
void orc_ppm_reconstruct(const real *s5, real flatn, real *smo, real *spo)
{
    const real *s = s5 + 2;
    real dsl, dsr, dsc, dsvl_l, dsvl_r, sm, sp;

    dsl = TWO * (s[-1] - s[-2]);
    dsr = TWO * (s[0] - s[-1]);
    if (dsl * dsr > ZERO) {
        dsc = HALF * (s[0] - s[-2]);
        dsvl_l = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_l = ZERO;
    dsl = TWO * (s[0] - s[-1]);
    dsr = TWO * (s[1] - s[0]);
    if (dsl * dsr > ZERO) {
        dsc = HALF * (s[1] - s[-1]);
        dsvl_r = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_r = ZERO;
    sm = HALF * (s[0] + s[-1]) - SIXTH * (dsvl_r - dsvl_l);
    sm = fmax(sm, fmin(s[0], s[-1]));
    sm = fmin(sm, fmax(s[0], s[-1]));

    dsl = TWO * (s[0] - s[-1]);
    dsr = TWO * (s[1] - s[0]);
    if (dsl * dsr > ZERO) {
        dsc = HALF * (s[1] - s[-1]);
        dsvl_l = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_l = ZERO;
    dsl = TWO * (s[1] - s[0]);
    dsr = TWO * (s[2] - s[1]);
    if (dsl * dsr > ZERO) {
        dsc = HALF * (s[2] - s[0]);
        dsvl_r = fsign1(dsc) * fmin(fabs(dsc), fmin(fabs(dsl), fabs(dsr)));
    } else dsvl_r = ZERO;
    sp = HALF * (s[1] + s[0]) - SIXTH * (dsvl_r - dsvl_l);
    sp = fmax(sp, fmin(s[1], s[0]));
    sp = fmin(sp, fmax(s[1], s[0]));

    sm = flatn * sm + (ONE - flatn) * s[0];
    sp = flatn * sp + (ONE - flatn) * s[0];

    if ((sp - s[0]) * (s[0] - sm) <= ZERO) {
        sp = s[0];
        sm = s[0];
    } else if (fabs(sp - s[0]) >= TWO * fabs(sm - s[0])) {
        sp = THREE * s[0] - TWO * sm;
    } else if (fabs(sm - s[0]) >= TWO * fabs(sp - s[0])) {
        sm = THREE * s[0] - TWO * sp;
    }
    *smo = sm;
    *spo = sp;
}

/* Source/ppm.F90:240-305 */
void orc_ppm_int_profile(real sm, real sp, real sc, real u, real c, real dtdx,
                         real *Ip, real *Im)
{
    real speed, sigma, s6;
    s6 = SIX * sc - THREE * (sm + sp);
    const real sp3[3] = {u - c, u, u + c};
    for (int w = 0; w < 3; ++w) {
        speed = sp3[w];
        sigma = fabs(speed) * dtdx;
        if (speed <= ZERO) {
            Ip[w] = sp;
            Im[w] = sm + HALF * sigma * (sp - sm + (ONE - TWO3RD * sigma) * s6);
        } else {
            Ip[w] = sp - HALF * sigma * (sp - sm - (ONE - TWO3RD * sigma) * s6);
            Im[w] = sm;
        }
    }
}

/* Source/ppm.F90:309-743 */
void orc_trace_ppm(const int *lo, const int *hi, const int *vlo, const int *vhi, int idir,
                   const real *qp_, const int *q_lo, const int *q_hi, const real *qap,
                   const int *qa_lo, const int *qa_hi, real *qmp, const int *qm_lo,
                   const int *qm_hi, real *qpp, const int *qp_lo, const int *qp_hi,
                   const int *domlo, const int *domhi, const real *dx, real dt)
{
    (void)domlo; (void)domhi;
    arr_t q = mk(qp_, q_lo, q_hi), qaux = mk(qap, qa_lo, qa_hi), qm = mk(qmp, qm_lo, qm_hi),
          qp = mk(qpp, qp_lo, qp_hi);
    const real dtdx = dt / dx[idir - 1];
    int QUN, QUT, QUTT;
    if (idir == 1) { QUN = QU; QUT = QV; QUTT = QW; }
    else if (idir == 2) { QUN = QV; QUT = QW; QUTT = QU; }
    else { QUN = QW; QUT = QU; QUTT = QV; }
    const int di = (idir == 1), dj = (idir == 2), dk = (idir == 3);

    LOOP3(lo, hi)
    {
        real Ip[3][ORC_QVAR], Im[3][ORC_QVAR], Ip_gc[3], Im_gc[3], Ip_sp[3], Im_sp[3];
        real s[5], sm, sp, flatn;
        real rho = A(q, i, j, k, QRHO);
        (void)rho;
        real cc = A(qaux, i, j, k, QC);
        real un = A(q, i, j, k, QUN);
        orc_uflatten(i, j, k, qp_, q_lo, q_hi, &flatn);

        for (int n = 0; n <= QTEMP; ++n) {
            if (n == QGAME || n == QTEMP) continue; /* reconstruct_state, :419-421 */
            for (int m = -2; m <= 2; ++m) s[m + 2] = A(q, i + m * di, j + m * dj, k + m * dk, n);
            orc_ppm_reconstruct(s, flatn, &sm, &sp);
            real ip3[3], im3[3];
            orc_ppm_int_profile(sm, sp, s[2], un, cc, dtdx, ip3, im3);
            for (int w = 0; w < 3; ++w) { Ip[w][n] = ip3[w]; Im[w][n] = im3[w]; }
        }
        for (int m = -2; m <= 2; ++m) s[m + 2] = A(qaux, i + m * di, j + m * dj, k + m * dk, QGAMC);
        orc_ppm_reconstruct(s, flatn, &sm, &sp);
        orc_ppm_int_profile(sm, sp, s[2], un, cc, dtdx, Ip_gc, Im_gc);

        const int do_p = (idir == 1 && i >= vlo[0]) || (idir == 2 && j >= vlo[1]) ||
                         (idir == 3 && k >= vlo[2]);
        const int do_m = (idir == 1 && i <= vhi[0]) || (idir == 2 && j <= vhi[1]) ||
                         (idir == 3 && k <= vhi[2]);

        for (int ispec = 0; ispec < ORC_NSPEC; ++ispec) {
            int n = QFS + ispec;
            for (int m = -2; m <= 2; ++m) s[m + 2] = A(q, i + m * di, j + m * dj, k + m * dk, n);
            orc_ppm_reconstruct(s, flatn, &sm, &sp);
            orc_ppm_int_profile(sm, sp, s[2], un, cc, dtdx, Ip_sp, Im_sp);
            if (do_p) A(qp, i, j, k, n) = Im_sp[1];
            if (do_m) A(qm, i + di, j + dj, k + dk, n) = Ip_sp[1];
        }

        if (do_p) { /* :529-625 */
            real rho_ref = Im[0][QRHO], un_ref = Im[0][QUN], p_ref = Im[0][QPRES];
            real rhoe_g_ref = Im[0][QREINT], gam_g_ref = Im_gc[0];
            rho_ref = fmax(rho_ref, small_dens);
            real rho_ref_inv = ONE / rho_ref;
            p_ref = fmax(p_ref, small_pres);
            real csq_ref = gam_g_ref * p_ref * rho_ref_inv;
            real cc_ref = sqrt(csq_ref);
            real cc_ref_inv = ONE / cc_ref;
            real h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv;
            real dum = un_ref - Im[0][QUN];
            real dptotm = p_ref - Im[0][QPRES];
            real drho = rho_ref - Im[1][QRHO];
            real dptot = p_ref - Im[1][QPRES];
            real drhoe_g = rhoe_g_ref - Im[1][QREINT];
            real dup = un_ref - Im[2][QUN];
            real dptotp = p_ref - Im[2][QPRES];
            real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
            real alphap = HALF * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
            real alpha0r = drho - dptot / csq_ref;
            real alpha0e_g = drhoe_g - dptot * h_g_ref / csq_ref;
            alpham = (un - cc > ZERO) ? ZERO : -alpham;
            alphap = (un + cc > ZERO) ? ZERO : -alphap;
            alpha0r = (un > ZERO) ? ZERO : -alpha0r;
            alpha0e_g = (un > ZERO) ? ZERO : -alpha0e_g;
            A(qp, i, j, k, QRHO) = fmax(small_dens, rho_ref + alphap + alpham + alpha0r);
            A(qp, i, j, k, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
            A(qp, i, j, k, QREINT) = rhoe_g_ref + (alphap + alpham) * h_g_ref + alpha0e_g;
            A(qp, i, j, k, QPRES) = fmax(small_pres, p_ref + (alphap + alpham) * csq_ref);
            A(qp, i, j, k, QUT) = Im[1][QUT];
            A(qp, i, j, k, QUTT) = Im[1][QUTT];
        }
        if (do_m) { /* :631-737 */
            real rho_ref = Ip[2][QRHO], un_ref = Ip[2][QUN], p_ref = Ip[2][QPRES];
            real rhoe_g_ref = Ip[2][QREINT], gam_g_ref = Ip_gc[2];
            rho_ref = fmax(rho_ref, small_dens);
            real rho_ref_inv = ONE / rho_ref;
            p_ref = fmax(p_ref, small_pres);
            real csq_ref = gam_g_ref * p_ref * rho_ref_inv;
            real cc_ref = sqrt(csq_ref);
            real cc_ref_inv = ONE / cc_ref;
            real h_g_ref = (p_ref + rhoe_g_ref) * rho_ref_inv;
            real dum = un_ref - Ip[0][QUN];
            real dptotm = p_ref - Ip[0][QPRES];
            real drho = rho_ref - Ip[1][QRHO];
            real dptot = p_ref - Ip[1][QPRES];
            real drhoe_g = rhoe_g_ref - Ip[1][QREINT];
            real dup = un_ref - Ip[2][QUN];
            real dptotp = p_ref - Ip[2][QPRES];
            real alpham = HALF * (dptotm * rho_ref_inv * cc_ref_inv - dum) * rho_ref * cc_ref_inv;
            real alphap = HALF * (dptotp * rho_ref_inv * cc_ref_inv + dup) * rho_ref * cc_ref_inv;
            real alpha0r = drho - dptot / csq_ref;
            real alpha0e_g = drhoe_g - dptot * h_g_ref / csq_ref;
            alpham = (un - cc > ZERO) ? -alpham : ZERO;
            alphap = (un + cc > ZERO) ? -alphap : ZERO;
            alpha0r = (un > ZERO) ? -alpha0r : ZERO;
            alpha0e_g = (un > ZERO) ? -alpha0e_g : ZERO;
            int ii = i + di, jj = j + dj, kk = k + dk;
            A(qm, ii, jj, kk, QRHO) = fmax(small_dens, rho_ref + alphap + alpham + alpha0r);
            A(qm, ii, jj, kk, QUN) = un_ref + (alphap - alpham) * cc_ref * rho_ref_inv;
            A(qm, ii, jj, kk, QREINT) = rhoe_g_ref + (alphap + alpham) * h_g_ref + alpha0e_g;
            A(qm, ii, jj, kk, QPRES) = fmax(small_pres, p_ref + (alphap + alpham) * csq_ref);
            A(qm, ii, jj, kk, QUT) = Ip[1][QUT];
            A(qm, ii, jj, kk, QUTT) = Ip[1][QUTT];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Source/riemann.F90:10-331 */
void orc_compute_flux(const int *lo, const int *hi, const real *qlp, const int *ql_lo,
                      const int *ql_hi, const real *qrp, const int *qr_lo, const int *qr_hi,
                      real *fp, const int *f_lo, const int *f_hi, real *qip, const int *qi_lo,
                      const int *qi_hi, real *qgp, const int *qg_lo, const int *qg_hi,
                      const real *qap, const int *qa_lo, const int *qa_hi, int idir)
{
    arr_t ql = mk(qlp, ql_lo, ql_hi), qr = mk(qrp, qr_lo, qr_hi), flx = mk(fp, f_lo, f_hi),
          qint = mk(qip, qi_lo, qi_hi), qgdnv = mk(qgp, qg_lo, qg_hi), qaux = mk(qap, qa_lo, qa_hi);
    int iu, iv1, iv2, im1, im2, im3;
    if (idir == 1) { iu = QU; iv1 = QV; iv2 = QW; im1 = UMX; im2 = UMY; im3 = UMZ; }
    else if (idir == 2) { iu = QV; iv1 = QU; iv2 = QW; im1 = UMY; im2 = UMX; im3 = UMZ; }
    else { iu = QW; iv1 = QU; iv2 = QV; im1 = UMZ; im2 = UMX; im3 = UMY; }
    const int di = (idir == 1), dj = (idir == 2), dk = (idir == 3);

    LOOP3(lo, hi)
    {
        real rl = fmax(A(ql, i, j, k, QRHO), small_dens);
        real ul = A(ql, i, j, k, iu), v1l = A(ql, i, j, k, iv1), v2l = A(ql, i, j, k, iv2);
        real pl = fmax(A(ql, i, j, k, QPRES), small_pres);
        real rel = A(ql, i, j, k, QREINT);
        real rr = fmax(A(qr, i, j, k, QRHO), small_dens);
        real ur = A(qr, i, j, k, iu), v1r = A(qr, i, j, k, iv1), v2r = A(qr, i, j, k, iv2);
        real pr = fmax(A(qr, i, j, k, QPRES), small_pres);
        real rer = A(qr, i, j, k, QREINT);

        real cR = A(qaux, i, j, k, QC), cL = A(qaux, i - di, j - dj, k - dk, QC);
        real csmall = fmax(small, small * fmax(cR, cL));
        real cavg = HALF * (cR + cL);
        real gamcl = A(qaux, i - di, j - dj, k - dk, QGAMC);
        real gamcr = A(qaux, i, j, k, QGAMC);

        real wsmall = small_dens * csmall;
        real wl = fmax(wsmall, sqrt(fabs(gamcl * pl * rl)));
        real wr = fmax(wsmall, sqrt(fabs(gamcr * pr * rr)));
        real wwinv = ONE / (wl + wr);
        real pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) * wwinv;
        real ustar = ((wl * ul + wr * ur) + (pl - pr)) * wwinv;
        pstar = fmax(pstar, small_pres);
        if (fabs(ustar) < smallu * HALF * (fabs(ul) + fabs(ur))) ustar = ZERO;

        real sgnm = fsign1(ustar);
        if (ustar == ZERO) sgnm = ZERO;
        real fpl = HALF * (ONE + sgnm);
        real fmr = HALF * (ONE - sgnm);
        real ro = fpl * rl + fmr * rr;
        real uo = fpl * ul + fmr * ur;
        real po = fpl * pl + fmr * pr;
        real reo = fpl * rel + fmr * rer;
        real gamco = fpl * gamcl + fmr * gamcr;
        ro = fmax(small_dens, ro);
        real roinv = ONE / ro;
        real co = sqrt(fabs(gamco * po * roinv));
        co = fmax(csmall, co);
        real co2inv = ONE / (co * co);
        A(qint, i, j, k, iv1) = fpl * v1l + fmr * v1r;
        A(qint, i, j, k, iv2) = fpl * v2l + fmr * v2r;

        real drho = (pstar - po) * co2inv;
        real rstar = ro + drho;
        rstar = fmax(small_dens, rstar);
        real rstarinv = ONE / rstar;
        real entho = (reo + po) * roinv * co2inv;
        real estar = reo + (pstar - po) * entho;
        real cstar = sqrt(fabs(gamco * pstar * rstarinv));
        cstar = fmax(cstar, csmall);

        sgnm = fsign1(ustar);
        real spout = co - sgnm * uo;
        real spin = cstar - sgnm * ustar;
        real ushock = HALF * (spin + spout);
        if (pstar - po > ZERO) { spin = ushock; spout = ushock; }
        real scr = (spout - spin == ZERO) ? small * cavg : spout - spin;
        real scrinv = ONE / scr;
        real frac = (ONE + (spout + spin) * scrinv) * HALF;
        frac = fmax(ZERO, fmin(ONE, frac));

        A(qint, i, j, k, QRHO) = frac * rstar + (ONE - frac) * ro;
        A(qint, i, j, k, iu) = frac * ustar + (ONE - frac) * uo;
        A(qint, i, j, k, QPRES) = frac * pstar + (ONE - frac) * po;
        real regdnv = frac * estar + (ONE - frac) * reo;
        if (spout < ZERO) {
            A(qint, i, j, k, QRHO) = ro;
            A(qint, i, j, k, iu) = uo;
            A(qint, i, j, k, QPRES) = po;
            regdnv = reo;
        }
        if (spin >= ZERO) {
            A(qint, i, j, k, QRHO) = rstar;
            A(qint, i, j, k, iu) = ustar;
            A(qint, i, j, k, QPRES) = pstar;
            regdnv = estar;
        }
        real regdnvinv = ONE / regdnv;
        A(qint, i, j, k, QGAME) = A(qint, i, j, k, QPRES) * regdnvinv + ONE;
        A(qint, i, j, k, QPRES) = fmax(A(qint, i, j, k, QPRES), small_pres);
        A(qint, i, j, k, QREINT) = regdnv;
        for (int ispec = 0; ispec < ORC_NSPEC; ++ispec) {
            int nqp = QFS + ispec;
            A(qint, i, j, k, nqp) = fpl * A(ql, i, j, k, nqp) + fmr * A(qr, i, j, k, nqp);
        }

        A(qgdnv, i, j, k, GDRHO) = A(qint, i, j, k, QRHO);
        A(qgdnv, i, j, k, GDU) = A(qint, i, j, k, QU);
        A(qgdnv, i, j, k, GDV) = A(qint, i, j, k, QV);
        A(qgdnv, i, j, k, GDW) = A(qint, i, j, k, QW);
        A(qgdnv, i, j, k, GDPRES) = A(qint, i, j, k, QPRES);
        A(qgdnv, i, j, k, GDGAME) = A(qint, i, j, k, QGAME);

        real u_adv = A(qint, i, j, k, iu);
        real rhoeint = A(qint, i, j, k, QREINT);
        A(flx, i, j, k, URHO) = A(qint, i, j, k, QRHO) * u_adv;
        A(flx, i, j, k, im1) = A(flx, i, j, k, URHO) * A(qint, i, j, k, iu);
        A(flx, i, j, k, im2) = A(flx, i, j, k, URHO) * A(qint, i, j, k, iv1);
        A(flx, i, j, k, im3) = A(flx, i, j, k, URHO) * A(qint, i, j, k, iv2);
        real rhoetot = rhoeint + HALF * A(qint, i, j, k, QRHO) *
                                       (A(qint, i, j, k, iu) * A(qint, i, j, k, iu) +
                                        A(qint, i, j, k, iv1) * A(qint, i, j, k, iv1) +
                                        A(qint, i, j, k, iv2) * A(qint, i, j, k, iv2));
        A(flx, i, j, k, UEDEN) = u_adv * (rhoetot + A(qint, i, j, k, QPRES));
        A(flx, i, j, k, UEINT) = u_adv * rhoeint;
        A(flx, i, j, k, UTEMP) = ZERO;
        for (int ispec = 0; ispec < ORC_NSPEC; ++ispec)
            A(flx, i, j, k, UFS + ispec) = A(flx, i, j, k, URHO) * A(qint, i, j, k, QFS + ispec);
    }
}

/* ------------------------------------------------------------------ */
/* Source/trans.F90:14-243.  QGAME/QTEMP of the outputs are left untouched
   (the reference copies an uninitialised private array there). */
void orc_trans1(const int *lo, const int *hi, int idir1, int idir2, const real *q2mp,
                const int *q2m_lo, const int *q2m_hi, real *q2mop, const int *q2mo_lo,
                const int *q2mo_hi, const real *q2pp, const int *q2p_lo, const int *q2p_hi,
                real *q2pop, const int *q2po_lo, const int *q2po_hi, const real *qap,
                const int *qa_lo, const int *qa_hi, const real *f1p, const int *f1_lo,
                const int *f1_hi, const real *q1p, const int *q1_lo, const int *q1_hi,
                real cdtdx)
{
    arr_t q2m = mk(q2mp, q2m_lo, q2m_hi), q2mo = mk(q2mop, q2mo_lo, q2mo_hi),
          q2p = mk(q2pp, q2p_lo, q2p_hi), q2po = mk(q2pop, q2po_lo, q2po_hi),
          qaux = mk(qap, qa_lo, qa_hi), f1 = mk(f1p, f1_lo, f1_hi), q1 = mk(q1p, q1_lo, q1_hi);
    LOOP3(lo, hi)
    {
        for (int d = -1; d <= 0; ++d) {
            int il = i, jl = j, kl = k, ir, jr, kr;
            if (idir1 == 1) { ir = i + 1; jr = j; kr = k; }
            else if (idir1 == 2) { ir = i; jr = j + 1; kr = k; }
            else { ir = i; jr = j; kr = k + 1; }
            if (idir2 == 1) { il += d; ir += d; }
            else if (idir2 == 2) { jl += d; jr += d; }
            else { kl += d; kr += d; }

            real lq2[ORC_QVAR], lq2o[ORC_QVAR];
            int have_game = 0;
            for (int n = 0; n < ORC_QVAR; ++n)
                lq2[n] = (d == -1) ? A(q2m, i, j, k, n) : A(q2p, i, j, k, n);

            for (int ispec = 0; ispec < ORC_NSPEC; ++ispec) {
                int n = UFS + ispec, nqp = QFS + ispec;
                real rrnew = lq2[QRHO] - cdtdx * (A(f1, ir, jr, kr, URHO) - A(f1, il, jl, kl, URHO));
                real compu = lq2[QRHO] * lq2[nqp] - cdtdx * (A(f1, ir, jr, kr, n) - A(f1, il, jl, kl, n));
                lq2o[nqp] = compu / rrnew;
            }

            real pgp = A(q1, ir, jr, kr, GDPRES), pgm = A(q1, il, jl, kl, GDPRES);
            real ugp = A(q1, ir, jr, kr, GDU + idir1 - 1), ugm = A(q1, il, jl, kl, GDU + idir1 - 1);
            real dup = pgp * ugp - pgm * ugm;
            real du = ugp - ugm;
            real pav = HALF * (pgp + pgm);
            real gamc = A(qaux, il, jl, kl, QGAMC);

            real rrl2 = lq2[QRHO];
            real rul2 = rrl2 * lq2[QU];
            real rvl2 = rrl2 * lq2[QV];
            real rwl2 = rrl2 * lq2[QW];
            real ekenl2 = HALF * rrl2 * (lq2[QU] * lq2[QU] + lq2[QV] * lq2[QV] + lq2[QW] * lq2[QW]);
            real rel2 = lq2[QREINT] + ekenl2;

            real rrnewl2 = rrl2 - cdtdx * (A(f1, ir, jr, kr, URHO) - A(f1, il, jl, kl, URHO));
            real runewl2 = rul2 - cdtdx * (A(f1, ir, jr, kr, UMX) - A(f1, il, jl, kl, UMX));
            real rvnewl2 = rvl2 - cdtdx * (A(f1, ir, jr, kr, UMY) - A(f1, il, jl, kl, UMY));
            real rwnewl2 = rwl2 - cdtdx * (A(f1, ir, jr, kr, UMZ) - A(f1, il, jl, kl, UMZ));
            real renewl2 = rel2 - cdtdx * (A(f1, ir, jr, kr, UEDEN) - A(f1, il, jl, kl, UEDEN));

            int reset_state = 0;
            if (rrnewl2 < ZERO) {
                rrnewl2 = rrl2;
                runewl2 = rul2;
                rvnewl2 = rvl2;
                rwnewl2 = rwl2;
                renewl2 = rel2;
                reset_state = 1;
            }
            lq2o[QRHO] = rrnewl2;
            real rhoinv = ONE / rrnewl2;
            lq2o[QU] = runewl2 * rhoinv;
            lq2o[QV] = rvnewl2 * rhoinv;
            lq2o[QW] = rwnewl2 * rhoinv;
            real rhoekenl2 = HALF * (runewl2 * runewl2 + rvnewl2 * rvnewl2 + rwnewl2 * rwnewl2) * rhoinv;
            lq2o[QREINT] = renewl2 - rhoekenl2;
            if (!reset_state) {
                real pnewl2 = lq2[QPRES] - cdtdx * (dup + pav * du * (gamc - ONE));
                lq2o[QPRES] = fmax(pnewl2, small_pres);
            } else {
                lq2o[QPRES] = lq2[QPRES];
                lq2o[QGAME] = lq2[QGAME];
                have_game = 1;
            }
            arr_t *o = (d == -1) ? &q2mo : &q2po;
            for (int n = 0; n < ORC_QVAR; ++n) {
                if (n == QTEMP || (n == QGAME && !have_game)) continue;
                A(*o, i, j, k, n) = lq2o[n];
            }
        }
    }
}

/* Source/trans.F90:250-505 */
void orc_trans2(const int *lo, const int *hi, int idir1, int idir2, int idir3, const real *qm1p,
                const int *qm1_lo, const int *qm1_hi, real *qm1op, const int *qm1o_lo,
                const int *qm1o_hi, const real *qp1p, const int *qp1_lo, const int *qp1_hi,
                real *qp1op, const int *qp1o_lo, const int *qp1o_hi, const real *qap,
                const int *qa_lo, const int *qa_hi, const real *f2p, const int *f2_lo,
                const int *f2_hi, const real *f3p, const int *f3_lo, const int *f3_hi,
                const real *q2p, const int *q2_lo, const int *q2_hi, const real *q3p,
                const int *q3_lo, const int *q3_hi, real cdtdx1, real cdtdx2, real cdtdx3)
{
    (void)cdtdx1;
    arr_t qm1 = mk(qm1p, qm1_lo, qm1_hi), qm1o = mk(qm1op, qm1o_lo, qm1o_hi),
          qp1 = mk(qp1p, qp1_lo, qp1_hi), qp1o = mk(qp1op, qp1o_lo, qp1o_hi),
          qaux = mk(qap, qa_lo, qa_hi), f2 = mk(f2p, f2_lo, f2_hi), f3 = mk(f3p, f3_lo, f3_hi),
          q2 = mk(q2p, q2_lo, q2_hi), q3 = mk(q3p, q3_lo, q3_hi);
    LOOP3(lo, hi)
    {
        for (int d = -1; d <= 0; ++d) {
            int il1 = i, jl1 = j, kl1 = k, il2 = i, jl2 = j, kl2 = k, il3 = i, jl3 = j, kl3 = k;
            int ir2, jr2, kr2, ir3, jr3, kr3;
            if (idir1 == 1) {
                ir2 = i + d; jr2 = j + 1; kr2 = k;
                ir3 = i + d; jr3 = j; kr3 = k + 1;
                il1 = i + d; il2 = i + d; il3 = i + d;
            } else if (idir1 == 2) {
                ir2 = i + 1; jr2 = j + d; kr2 = k;
                ir3 = i; jr3 = j + d; kr3 = k + 1;
                jl1 = j + d; jl2 = j + d; jl3 = j + d;
            } else {
                ir2 = i + 1; jr2 = j; kr2 = k + d;
                ir3 = i; jr3 = j + 1; kr3 = k + d;
                kl1 = k + d; kl2 = k + d; kl3 = k + d;
            }
            real lq[ORC_QVAR], lqo[ORC_QVAR];
            int have_game = 0;
            for (int n = 0; n < ORC_QVAR; ++n)
                lq[n] = (d == -1) ? A(qm1, i, j, k, n) : A(qp1, i, j, k, n);

            for (int ispec = 0; ispec < ORC_NSPEC; ++ispec) {
                int n = UFS + ispec, nqp = QFS + ispec;
                real rrr = lq[QRHO];
                real compr = rrr * lq[nqp];
                real rrnewr = rrr - cdtdx2 * (A(f2, ir2, jr2, kr2, URHO) - A(f2, il2, jl2, kl2, URHO)) -
                                cdtdx3 * (A(f3, ir3, jr3, kr3, URHO) - A(f3, il3, jl3, kl3, URHO));
                real compnr = compr - cdtdx2 * (A(f2, ir2, jr2, kr2, n) - A(f2, il2, jl2, kl2, n)) -
                                cdtdx3 * (A(f3, ir3, jr3, kr3, n) - A(f3, il3, jl3, kl3, n));
                lqo[nqp] = compnr / rrnewr;
            }

            real pg2p = A(q2, ir2, jr2, kr2, GDPRES), pg2m = A(q2, il2, jl2, kl2, GDPRES);
            real ug2p = A(q2, ir2, jr2, kr2, GDU + idir2 - 1), ug2m = A(q2, il2, jl2, kl2, GDU + idir2 - 1);
            real du2p = pg2p * ug2p - pg2m * ug2m;
            real p2av = HALF * (pg2p + pg2m);
            real du2 = ug2p - ug2m;
            real p2new = cdtdx2 * (du2p + p2av * du2 * (A(qaux, il1, jl1, kl1, QGAMC) - ONE));

            real pg3p = A(q3, ir3, jr3, kr3, GDPRES), pg3m = A(q3, il3, jl3, kl3, GDPRES);
            real ug3p = A(q3, ir3, jr3, kr3, GDU + idir3 - 1), ug3m = A(q3, il3, jl3, kl3, GDU + idir3 - 1);
            real du3p = pg3p * ug3p - pg3m * ug3m;
            real p3av = HALF * (pg3p + pg3m);
            real du3 = ug3p - ug3m;
            real p3new = cdtdx3 * (du3p + p3av * du3 * (A(qaux, il1, jl1, kl1, QGAMC) - ONE));

            real rrr = lq[QRHO];
            real rur = rrr * lq[QU];
            real rvr = rrr * lq[QV];
            real rwr = rrr * lq[QW];
            real ekenr = HALF * rrr * (lq[QU] * lq[QU] + lq[QV] * lq[QV] + lq[QW] * lq[QW]);
            real rer = lq[QREINT] + ekenr;

            real rrnewr = rrr - cdtdx2 * (A(f2, ir2, jr2, kr2, URHO) - A(f2, il2, jl2, kl2, URHO)) -
                            cdtdx3 * (A(f3, ir3, jr3, kr3, URHO) - A(f3, il3, jl3, kl3, URHO));
            real runewr = rur - cdtdx2 * (A(f2, ir2, jr2, kr2, UMX) - A(f2, il2, jl2, kl2, UMX)) -
                            cdtdx3 * (A(f3, ir3, jr3, kr3, UMX) - A(f3, il3, jl3, kl3, UMX));
            real rvnewr = rvr - cdtdx2 * (A(f2, ir2, jr2, kr2, UMY) - A(f2, il2, jl2, kl2, UMY)) -
                            cdtdx3 * (A(f3, ir3, jr3, kr3, UMY) - A(f3, il3, jl3, kl3, UMY));
            real rwnewr = rwr - cdtdx2 * (A(f2, ir2, jr2, kr2, UMZ) - A(f2, il2, jl2, kl2, UMZ)) -
                            cdtdx3 * (A(f3, ir3, jr3, kr3, UMZ) - A(f3, il3, jl3, kl3, UMZ));
            real renewr = rer - cdtdx2 * (A(f2, ir2, jr2, kr2, UEDEN) - A(f2, il2, jl2, kl2, UEDEN)) -
                            cdtdx3 * (A(f3, ir3, jr3, kr3, UEDEN) - A(f3, il3, jl3, kl3, UEDEN));

            int reset_state = 0;
            if (rrnewr < ZERO) {
                rrnewr = rrr;
                runewr = rur;
                rvnewr = rvr;
                rwnewr = rwr;
                renewr = rer;
                reset_state = 1;
            }
            lqo[QRHO] = rrnewr;
            lqo[QU] = runewr / rrnewr;
            lqo[QV] = rvnewr / rrnewr;
            lqo[QW] = rwnewr / rrnewr;
            real rhoekenr = HALF * (runewr * runewr + rvnewr * rvnewr + rwnewr * rwnewr) / rrnewr;
            lqo[QREINT] = renewr - rhoekenr;
            if (!reset_state) {
                real pnewr = lq[QPRES] - p2new - p3new;
                lqo[QPRES] = pnewr;
            } else {
                lqo[QPRES] = lq[QPRES];
                lqo[QGAME] = lq[QGAME];
                have_game = 1;
            }
            lqo[QPRES] = fmax(lqo[QPRES], small_pres);
            arr_t *o = (d == -1) ? &qm1o : &qp1o;
            for (int n = 0; n < ORC_QVAR; ++n) {
                if (n == QTEMP || (n == QGAME && !have_game)) continue;
                A(*o, i, j, k, n) = lqo[n];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* abar, zbar from species partial densities u(UFS..)*rhoInv, as written in
   Source/Castro_nd.F90:285-286 and Source/timestep.F90:49-50 */
static void abar_zbar_rhoinv(const arr_t *u, int i, int j, int k, real rhoInv, real *abar,
                             real *zbar)
{
    real s1 = 0.0, s2 = 0.0;
    for (int n = 0; n < ORC_NSPEC; ++n) s1 = s1 + A(*u, i, j, k, UFS + n) * aion_inv(n);
    *abar = ONE / (s1 * rhoInv);
    for (int n = 0; n < ORC_NSPEC; ++n) s2 = s2 + A(*u, i, j, k, UFS + n) * zion[n] * aion_inv(n);
    *zbar = *abar * (s2 * rhoInv);
}

/* Source/timestep.F90:12-72 (CPU branch of reduce_min, Source/reduction.F90:137-141) */
void orc_estdt(const int *lo, const int *hi, const real *up, const int *u_lo, const int *u_hi,
               const real *dx, real *dt)
{
    arr_t u = mk(up, u_lo, u_hi);
    LOOP3(lo, hi)
    {
        real rhoInv = ONE / A(u, i, j, k, URHO);
        orc_eos_t st;
        st.rho = A(u, i, j, k, URHO);
        st.T = A(u, i, j, k, UTEMP);
        st.e = A(u, i, j, k, UEINT) * rhoInv;
        abar_zbar_rhoinv(&u, i, j, k, rhoInv, &st.abar, &st.zbar);
        orc_eos(EOS_INPUT_RE, &st);
        real ux = A(u, i, j, k, UMX) * rhoInv;
        real uy = A(u, i, j, k, UMY) * rhoInv;
        real uz = A(u, i, j, k, UMZ) * rhoInv;
        real c = st.cs;
        real dt1 = dx[0] / (c + fabs(ux));
        real dt2 = dx[1] / (c + fabs(uy));
        real dt3 = dx[2] / (c + fabs(uz));
        *dt = fmin(*dt, fmin(fmin(dt1, dt2), dt3));
    }
}

/* Source/Castro_nd.F90:69-125 */
void orc_enforce_minimum_density(const int *lo, const int *hi, real *up, const int *u_lo,
                                 const int *u_hi)
{
    arr_t u = mk(up, u_lo, u_hi);
    LOOP3(lo, hi)
    {
        if (A(u, i, j, k, URHO) < small_dens) {
            for (int ispec = 0; ispec < ORC_NSPEC; ++ispec) {
                int n = UFS + ispec;
                A(u, i, j, k, n) = A(u, i, j, k, n) * (small_dens / A(u, i, j, k, URHO));
            }
            orc_eos_t st;
            st.rho = small_dens;
            st.T = small_temp;
            real s1 = 0.0, s2 = 0.0;
            for (int n = 0; n < ORC_NSPEC; ++n) s1 = s1 + A(u, i, j, k, UFS + n) * aion_inv(n);
            st.abar = ONE / (s1 / small_dens);
            for (int n = 0; n < ORC_NSPEC; ++n) s2 = s2 + A(u, i, j, k, UFS + n) * zion[n] * aion_inv(n);
            st.zbar = st.abar * (s2 / small_dens);
            orc_eos(EOS_INPUT_RT, &st);
            A(u, i, j, k, URHO) = st.rho;
            A(u, i, j, k, UTEMP) = st.T;
            A(u, i, j, k, UMX) = ZERO;
            A(u, i, j, k, UMY) = ZERO;
            A(u, i, j, k, UMZ) = ZERO;
            A(u, i, j, k, UEINT) = st.rho * st.e;
            A(u, i, j, k, UEDEN) = A(u, i, j, k, UEINT);
        }
    }
}

/* Source/Castro_nd.F90:129-160: stores X = (rho X)/sum, not rho X */
void orc_normalize_species(const int *lo, const int *hi, real *up, const int *u_lo,
                           const int *u_hi)
{
    arr_t u = mk(up, u_lo, u_hi);
    LOOP3(lo, hi)
    {
        real rho = A(u, i, j, k, URHO);
        real x[ORC_NSPEC], sum = 0.0;
        for (int n = 0; n < ORC_NSPEC; ++n) x[n] = fmax(1.0e-30 * rho, fmin(rho, A(u, i, j, k, UFS + n)));
        for (int n = 0; n < ORC_NSPEC; ++n) sum = sum + x[n];
        for (int n = 0; n < ORC_NSPEC; ++n) A(u, i, j, k, UFS + n) = x[n] / sum;
    }
}

/* Source/Castro_nd.F90:164-249 */
void orc_reset_internal_e(const int *lo, const int *hi, real *up, const int *u_lo, const int *u_hi)
{
    arr_t u = mk(up, u_lo, u_hi);
    const real dual_energy_eta2 = 1.e-4;
    LOOP3(lo, hi)
    {
        real rhoInv = ONE / A(u, i, j, k, URHO);
        real Up = A(u, i, j, k, UMX) * rhoInv;
        real Vp = A(u, i, j, k, UMY) * rhoInv;
        real Wp = A(u, i, j, k, UMZ) * rhoInv;
        real ke = HALF * (Up * Up + Vp * Vp + Wp * Wp);
        if (A(u, i, j, k, UEDEN) < ZERO) {
            if (A(u, i, j, k, UEINT) < ZERO) {
                orc_eos_t st;
                st.rho = A(u, i, j, k, URHO);
                st.T = small_temp;
                abar_zbar_rhoinv(&u, i, j, k, rhoInv, &st.abar, &st.zbar);
                orc_eos(EOS_INPUT_RT, &st);
                A(u, i, j, k, UEINT) = A(u, i, j, k, URHO) * st.e;
            }
            A(u, i, j, k, UEDEN) = A(u, i, j, k, UEINT) + A(u, i, j, k, URHO) * ke;
        } else {
            real rho_eint = A(u, i, j, k, UEDEN) - A(u, i, j, k, URHO) * ke;
            if (rho_eint > ZERO && rho_eint / A(u, i, j, k, UEDEN) > dual_energy_eta2) {
                A(u, i, j, k, UEINT) = rho_eint;
            } else if (A(u, i, j, k, UEINT) <= ZERO) {
                orc_eos_t st;
                st.rho = A(u, i, j, k, URHO);
                st.T = small_temp;
                abar_zbar_rhoinv(&u, i, j, k, rhoInv, &st.abar, &st.zbar);
                orc_eos(EOS_INPUT_RT, &st);
                real eint_new = st.e;
                A(u, i, j, k, UEINT) = A(u, i, j, k, URHO) * eint_new;
            }
        }
    }
}

/* Source/Castro_nd.F90:253-297 */
void orc_compute_temp(const int *lo, const int *hi, real *up, const int *u_lo, const int *u_hi)
{
    arr_t u = mk(up, u_lo, u_hi);
    LOOP3(lo, hi)
    {
        real rhoInv = ONE / A(u, i, j, k, URHO);
        orc_eos_t st;
        st.rho = A(u, i, j, k, URHO);
        st.T = A(u, i, j, k, UTEMP);
        st.e = A(u, i, j, k, UEINT) * rhoInv;
        abar_zbar_rhoinv(&u, i, j, k, rhoInv, &st.abar, &st.zbar);
        orc_eos(EOS_INPUT_RE, &st);
        A(u, i, j, k, UTEMP) = st.T;
        A(u, i, j, k, UEINT) = A(u, i, j, k, URHO) * st.e;
    }
}

/* ------------------------------------------------------------------ */
/* Source/Prob.F90:19-129, with the problem constants of Prob.F90:11-15
   passed in (the reference hard-codes them). */
void orc_initdata(const int *lo, const int *hi, real *up, const int *u_lo, const int *u_hi,
                  const real *dx, const real *problo, const real *probhi,
                  const orc_prob_t *prob)
{
    arr_t u = mk(up, u_lo, u_hi);
    const int nsub = prob->nsub;
    orc_eos_t st;
    st.rho = prob->dens_ambient;
    st.p = prob->p_ambient;
    st.T = 1.e4;
    st.abar = aion[0];
    st.zbar = zion[0];
    orc_eos(EOS_INPUT_RP, &st);
    real e_ambient = st.e;
    real r_init = prob->r_init;
    real vctr = FOUR3RD * M_PI_ * (r_init * r_init * r_init);
    real e_exp = prob->exp_energy / vctr / prob->dens_ambient;
    real center[3];
    for (int d = 0; d < 3; ++d) center[d] = (problo[d] + probhi[d]) / 2.e0;

    LOOP3(lo, hi)
    {
        real xmin = problo[0] + dx[0] * (real)i;
        real ymin = problo[1] + dx[1] * (real)j;
        real zmin = problo[2] + dx[2] * (real)k;
        int npert = 0, nambient = 0;
        for (int kk = 0; kk <= nsub - 1; ++kk) {
            real zz = zmin + (dx[2] / (real)nsub) * (kk + 0.5e0);
            for (int jj = 0; jj <= nsub - 1; ++jj) {
                real yy = ymin + (dx[1] / (real)nsub) * (jj + 0.5e0);
                for (int ii = 0; ii <= nsub - 1; ++ii) {
                    real xx = xmin + (dx[0] / (real)nsub) * (ii + 0.5e0);
                    real dist = (center[0] - xx) * (center[0] - xx) +
                                  (center[1] - yy) * (center[1] - yy) +
                                  (center[2] - zz) * (center[2] - zz);
                    if (dist <= r_init * r_init) npert = npert + 1;
                    else nambient = nambient + 1;
                }
            }
        }
        real eint = (npert * e_exp + nambient * e_ambient) / (real)(nsub * nsub * nsub);
        A(u, i, j, k, URHO) = prob->dens_ambient;
        A(u, i, j, k, UMX) = 0.e0;
        A(u, i, j, k, UMY) = 0.e0;
        A(u, i, j, k, UMZ) = 0.e0;
        A(u, i, j, k, UTEMP) = 1.e9;
        A(u, i, j, k, UEDEN) = A(u, i, j, k, URHO) * eint;
        A(u, i, j, k, UEINT) = A(u, i, j, k, URHO) * eint;
        for (int n = 0; n < ORC_NSPEC; ++n) A(u, i, j, k, UFS + n) = 0.e0;
        A(u, i, j, k, UFS) = A(u, i, j, k, URHO);
    }
}

/* Source/Castro_nd.F90:349-400 (note center = (probhi - problo)/2, :377) */
void orc_calculate_blast_radius(const int *lo, const int *hi, const real *up, const int *u_lo,
                                const int *u_hi, const real *dx, const real *problo,
                                const real *probhi, real *blast_mass, real *blast_radius,
                                real max_density)
{
    arr_t u = mk(up, u_lo, u_hi);
    const real density_tolerance = 0.1;
    real center[3];
    for (int d = 0; d < 3; ++d) center[d] = (probhi[d] - problo[d]) / TWO;
    LOOP3(lo, hi)
    {
        real x = problo[0] + (i + HALF) * dx[0] - center[0];
        real y = problo[1] + (j + HALF) * dx[1] - center[1];
        real z = problo[2] + (k + HALF) * dx[2] - center[2];
        if (fabs(A(u, i, j, k, URHO) - max_density) / max_density <= density_tolerance) {
            *blast_mass = *blast_mass + A(u, i, j, k, URHO);
            *blast_radius = *blast_radius + A(u, i, j, k, URHO) * sqrt(x * x + y * y + z * z);
        }
    }
}

/* Source/Castro_nd.F90:301-345 */
void orc_denerror(const int *lo, const int *hi, signed char *tagp, const int *t_lo, const int *t_hi,
                  const real *denp, const int *d_lo, const int *d_hi, signed char set,
                  signed char clear)
{
    (void)clear;
    arr_t den = mk(denp, d_lo, d_hi);
    long s1 = t_hi[0] - t_lo[0] + 1, s2 = s1 * (t_hi[1] - t_lo[1] + 1);
    const real dengrad_rel = 0.25;
    LOOP3(lo, hi)
    {
        real ax = fabs(A(den, i + 1, j, k, 0) - A(den, i, j, k, 0));
        real ay = fabs(A(den, i, j + 1, k, 0) - A(den, i, j, k, 0));
        real az = fabs(A(den, i, j, k + 1, 0) - A(den, i, j, k, 0));
        ax = fmax(ax, fabs(A(den, i, j, k, 0) - A(den, i - 1, j, k, 0)));
        ay = fmax(ay, fabs(A(den, i, j, k, 0) - A(den, i, j - 1, k, 0)));
        az = fmax(az, fabs(A(den, i, j, k, 0) - A(den, i, j, k - 1, 0)));
        if (fmax(ax, fmax(ay, az)) >= fabs(dengrad_rel * A(den, i, j, k, 0)))
            tagp[(i - t_lo[0]) + s1 * (j - t_lo[1]) + s2 * (k - t_lo[2])] = set;
    }
}

// Reverse-mode (adjoint) device functions of the CTU step, hand-written.
//
// Each `*_b` function re-evaluates its forward counterpart (mc_hydro.h,
// mc_cell.h) keeping the intermediates in registers, then propagates output
// adjoints ("bars") back to its inputs, differentiating the branch the
// forward evaluation took: every PPM limiter case, every flattening clamp,
// every Riemann sampling case and every floor (max/min) contributes the
// derivative of its active argument only.
//
// Scope of the differentiated state: the hydro components rho, rho u,
// rho v, rho w, rho E, rho e.  UTEMP enters the forward map only as the
// initial guess of the EOS Newton solves, whose converged result does not
// depend on it (its derivative is the implicit-function one, below), and the
// species enter the hydro only through abar/zbar after normalize_species,
// whose Jacobian with respect to the species is of the order of the 1e-30
// floor (Source/Castro_nd.F90:148); both are treated as constants.
//
// The EOS is differentiated with a two-component dual number D2 (d/drho,
// d/dT) through the same Helmholtz evaluation as the forward path; the
// Newton solve for T(rho, e) is differentiated by the implicit-function
// theorem: dT/de = 1/(de/dT), dT/drho = -(de/drho)/(de/dT), zero on the
// mintemp floor.
#pragma once

#include "mc_cell.h"
#include "mc_common.h"
#include "mc_eos.h"
#include "mc_hydro.h"

namespace mc {

// ------------------------------------------------------------------ dual numbers
struct D2 {
    double v, a, b;  // value, d/drho, d/dT
};
__host__ __device__ __forceinline__ D2 dc(double v) { return D2{v, 0.0, 0.0}; }
__host__ __device__ __forceinline__ D2 operator+(D2 x, D2 y) { return D2{x.v + y.v, x.a + y.a, x.b + y.b}; }
__host__ __device__ __forceinline__ D2 operator-(D2 x, D2 y) { return D2{x.v - y.v, x.a - y.a, x.b - y.b}; }
__host__ __device__ __forceinline__ D2 operator-(D2 x) { return D2{-x.v, -x.a, -x.b}; }
__host__ __device__ __forceinline__ D2 operator*(D2 x, D2 y)
{
    return D2{x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b};
}
__host__ __device__ __forceinline__ D2 operator/(D2 x, D2 y)
{
    const double q = x.v / y.v, yi = 1.0 / y.v;
    return D2{q, (x.a - q * y.a) * yi, (x.b - q * y.b) * yi};
}
__host__ __device__ __forceinline__ D2 operator+(D2 x, double c) { return D2{x.v + c, x.a, x.b}; }
__host__ __device__ __forceinline__ D2 operator+(double c, D2 x) { return D2{x.v + c, x.a, x.b}; }
__host__ __device__ __forceinline__ D2 operator-(D2 x, double c) { return D2{x.v - c, x.a, x.b}; }
__host__ __device__ __forceinline__ D2 operator-(double c, D2 x) { return D2{c - x.v, -x.a, -x.b}; }
__host__ __device__ __forceinline__ D2 operator*(D2 x, double c) { return D2{x.v * c, x.a * c, x.b * c}; }
__host__ __device__ __forceinline__ D2 operator*(double c, D2 x) { return D2{x.v * c, x.a * c, x.b * c}; }
__host__ __device__ __forceinline__ D2 operator/(D2 x, double c) { return D2{x.v / c, x.a / c, x.b / c}; }
__host__ __device__ __forceinline__ D2 operator/(double c, D2 y)
{
    const double q = c / y.v;
    return D2{q, -q * y.a / y.v, -q * y.b / y.v};
}
__host__ __device__ __forceinline__ D2 dsqrt(D2 x)
{
    const double s = sqrt(x.v), h = 0.5 / s;
    return D2{s, x.a * h, x.b * h};
}
__host__ __device__ __forceinline__ D2 dmax(D2 x, double c) { return x.v > c ? x : dc(c); }
__host__ __device__ __forceinline__ D2 dmax(D2 x, D2 y) { return x.v >= y.v ? x : y; }

// Hermite bases (Source/eos.F90:740-1031) on duals
__host__ __device__ __forceinline__ D2 dpsi0_(D2 z) { return (z * z * z) * (z * (-6.0 * z + 15.0) - 10.0) + 1.0; }
__host__ __device__ __forceinline__ D2 ddpsi0_(D2 z) { return (z * z) * (z * (-30.0 * z + 60.0) - 30.0); }
__host__ __device__ __forceinline__ D2 dddpsi0_(D2 z) { return z * (z * (-120.0 * z + 180.0) - 60.0); }
__host__ __device__ __forceinline__ D2 dpsi1_(D2 z) { return z * ((z * z) * (z * (-3.0 * z + 8.0) - 6.0) + 1.0); }
__host__ __device__ __forceinline__ D2 ddpsi1_(D2 z) { return z * z * (z * (-15.0 * z + 32.0) - 18.0) + 1.0; }
__host__ __device__ __forceinline__ D2 dddpsi1_(D2 z) { return z * (z * (-60.0 * z + 96.0) - 36.0); }
__host__ __device__ __forceinline__ D2 dpsi2_(D2 z) { return 0.5 * z * z * (z * (z * (-1.0 * z + 3.0) - 3.0) + 1.0); }
__host__ __device__ __forceinline__ D2 ddpsi2_(D2 z) { return 0.5 * z * (z * (z * (-5.0 * z + 12.0) - 9.0) + 2.0); }
__host__ __device__ __forceinline__ D2 dddpsi2_(D2 z) { return 0.5 * (z * (z * (-20.0 * z + 36.0) - 18.0) + 2.0); }
__host__ __device__ __forceinline__ D2 dxpsi0_(D2 z) { return z * z * (2.0 * z - 3.0) + 1.0; }
__host__ __device__ __forceinline__ D2 dxpsi1_(D2 z) { return z * (z * (z - 2.0) + 1.0); }

struct DW6 {
    D2 w0, w1, w2, w0m, w1m, w2m;
};

__device__ __forceinline__ D2 dh5(const Nodes &N, const DW6 &T, const DW6 &D)
{
    D2 acc = dc(0.0);
#define MC_DH5(q, wd, wmd, wt, wmt)                                                                      \
    acc = acc + N.a[q] * (wd * wt) + N.b[q] * (wmd * wt) + N.c[q] * (wd * wmt) + N.d[q] * (wmd * wmt);
    MC_DH5(NF, D.w0, D.w0m, T.w0, T.w0m)
    MC_DH5(NFT, D.w0, D.w0m, T.w1, T.w1m)
    MC_DH5(NFTT, D.w0, D.w0m, T.w2, T.w2m)
    MC_DH5(NFD, D.w1, D.w1m, T.w0, T.w0m)
    MC_DH5(NFDD, D.w2, D.w2m, T.w0, T.w0m)
    MC_DH5(NFDT, D.w1, D.w1m, T.w1, T.w1m)
    MC_DH5(NFDDT, D.w2, D.w2m, T.w1, T.w1m)
    MC_DH5(NFDTT, D.w1, D.w1m, T.w2, T.w2m)
    MC_DH5(NFDDTT, D.w2, D.w2m, T.w2, T.w2m)
    return acc;
}

__device__ __forceinline__ D2 dh3(const Nodes &N, D2 t0, D2 t1, D2 t0m, D2 t1m, D2 d0, D2 d1, D2 d0m, D2 d1m)
{
    D2 acc = dc(0.0);
    MC_DH5(NDPDF, d0, d0m, t0, t0m)
    MC_DH5(NDPDFT, d0, d0m, t1, t1m)
    MC_DH5(NDPDFD, d1, d1m, t0, t0m)
    MC_DH5(NDPDFDT, d1, d1m, t1, t1m)
#undef MC_DH5
    return acc;
}

// EOS outputs on the path with their derivatives in (rho, T).
struct EosD {
    D2 p, e, gam1;
};

// Helmholtz evaluation (Source/eos.F90:195-518) on duals; abar and zbar are
// duals too (abar = rho / sum(X/A) carries d/drho).  Table indices come from
// the value parts, as in the forward evaluation.
__device__ __forceinline__ EosD eos_dual(const EosTable &tab, D2 den, D2 temp, D2 abar, D2 zbar)
{
    using namespace eosc;
    const D2 ytot1 = 1.0 / abar;
    const D2 ye = zbar / abar;
    const D2 din = ye * den;
    const D2 deni = 1.0 / den;
    int iat = (int)((log10(din.v) - dlo) * dstpi) + 1;
    iat = iat < 1 ? 1 : (iat > EOS_IMAX - 1 ? EOS_IMAX - 1 : iat);
    int jat = (int)((log10(temp.v) - tlo) * tstpi) + 1;
    jat = jat < 1 ? 1 : (jat > EOS_JMAX - 1 ? EOS_JMAX - 1 : jat);
    const double *gd = tab.dg + (long)(iat - 1) * EOS_GRID;
    const double *gt = tab.tg + (long)(jat - 1) * EOS_GRID;
    const double dI = gd[G_X], dd = gd[G_D], dd2 = gd[G_D2], ddi = gd[G_DI];
    const double tJ = gt[G_X], dt = gt[G_D], dt2 = gt[G_D2], dti = gt[G_DI], dt2i = gt[G_D2I];

    const D2 prad = asoli3 * temp * temp * temp * temp;
    const D2 erad = 3.0 * prad * deni;
    const D2 pion = kergavo * ytot1 * den * temp;
    const D2 eion = 1.5 * pion * deni;

    Nodes N;
    const double *n0 = tab.node + ((long)(iat - 1) + (long)EOS_IMAX * (jat - 1)) * EOS_NODE;
    load_node(n0, N.a, 13);
    load_node(n0 + EOS_NODE, N.b, 13);
    load_node(n0 + (long)EOS_IMAX * EOS_NODE, N.c, 13);
    load_node(n0 + (long)EOS_IMAX * EOS_NODE + EOS_NODE, N.d, 13);

    D2 xd = (din - dI) * ddi;
    if (xd.v < 0.0) xd = dc(0.0);
    const D2 mxd = 1.0 - xd;
    D2 xt = (temp - tJ) * dti;
    if (xt.v < 0.0) xt = dc(0.0);
    const D2 mxt = 1.0 - xt;

    // density weights, then one temperature weight set at a time (keeps the
    // live dual registers to two sets)
    DW6 D, dD, T;
    D.w0 = dpsi0_(xd);
    D.w1 = dpsi1_(xd) * dd;
    D.w2 = dpsi2_(xd) * dd2;
    D.w0m = dpsi0_(mxd);
    D.w1m = -1.0 * dpsi1_(mxd) * dd;
    D.w2m = dpsi2_(mxd) * dd2;
    dD.w0 = ddpsi0_(xd) * ddi;
    dD.w1 = ddpsi1_(xd);
    dD.w2 = ddpsi2_(xd) * dd;
    dD.w0m = -1.0 * ddpsi0_(mxd) * ddi;
    dD.w1m = ddpsi1_(mxd);
    dD.w2m = -1.0 * ddpsi2_(mxd) * dd;
    T.w0 = dpsi0_(xt);
    T.w1 = dpsi1_(xt) * dt;
    T.w2 = dpsi2_(xt) * dt2;
    T.w0m = dpsi0_(mxt);
    T.w1m = -1.0 * dpsi1_(mxt) * dt;
    T.w2m = dpsi2_(mxt) * dt2;
    const D2 free = dh5(N, T, D);
    const D2 df_d = dh5(N, T, dD);
    T.w0 = ddpsi0_(xt) * dti;
    T.w1 = ddpsi1_(xt);
    T.w2 = ddpsi2_(xt) * dt;
    T.w0m = -1.0 * ddpsi0_(mxt) * dti;
    T.w1m = ddpsi1_(mxt);
    T.w2m = -1.0 * ddpsi2_(mxt) * dt;
    const D2 df_t = dh5(N, T, D);
    const D2 df_dt = dh5(N, T, dD);
    T.w0 = dddpsi0_(xt) * dt2i;
    T.w1 = dddpsi1_(xt) * dti;
    T.w2 = dddpsi2_(xt);
    T.w0m = dddpsi0_(mxt) * dt2i;
    T.w1m = -1.0 * dddpsi1_(mxt) * dti;
    T.w2m = dddpsi2_(mxt);
    const D2 df_tt = dh5(N, T, D);
    const D2 xt0 = dxpsi0_(xt), xt1 = dxpsi1_(xt) * dt, xt0m = dxpsi0_(mxt), xt1m = -1.0 * dxpsi1_(mxt) * dt;
    const D2 x0 = dxpsi0_(xd), x1 = dxpsi1_(xd) * dd, x0m = dxpsi0_(mxd), x1m = -1.0 * dxpsi1_(mxd) * dd;
    D2 dpepdd = ye * dh3(N, xt0, xt1, xt0m, xt1m, x0, x1, x0m, x1m);
    if (dpepdd.v < 0.0) dpepdd = dc(0.0);

    // e and its temperature derivative (for Gamma_1)
    const D2 sele = -1.0 * df_t * ye;
    const D2 dsepdt = -1.0 * df_tt * ye;
    const D2 eele = ye * free + temp * sele;
    const D2 deepdt = temp * dsepdt;
    const D2 dpraddt = 4.0 * prad / temp;
    const D2 deraddt = 3.0 * dpraddt * deni;
    const D2 dpiondt = kergavo * ytot1 * den;
    const D2 deiondt = 1.5 * dpiondt * deni;
    const D2 dpiondd = kergavo * ytot1 * temp;
    const D2 x2 = din * din;
    const D2 pele = x2 * df_d;
    const D2 dpepdt = x2 * df_dt;

    EosD o;
    o.e = erad + eion + eele;
    const D2 dedT = deraddt + deiondt + deepdt;
    o.p = prad + pion + pele;
    const D2 dpresdd = dpiondd + dpepdd;
    const D2 dpresdt = dpraddt + dpiondt + dpepdt;
    const D2 chit = temp / o.p * dpresdt;
    const D2 chid = dpresdd * den / o.p;
    o.gam1 = chit * (o.p * deni) * chit / (temp * dedT) + chid;
    return o;
}

// abar, zbar of a cleaned state as duals in rho: abar = rho / sum(u_n / A_n)
// and zbar = sum(u_n Z_n / A_n) / sum(u_n / A_n) with the species slots
// held fixed.
__device__ __forceinline__ void abar_zbar_dual(const double *u, double abar, double zbar, D2 &A, D2 &Z)
{
    A = D2{abar, abar / u[URHO], 0.0};
    Z = dc(zbar);
}

// Derivatives of (p, gamma_1, c) of a (rho, e) state at its converged T,
// with respect to rho at fixed e and to e at fixed rho.
struct EosRED {
    double p_r, p_e, g_r, g_e, c_r, c_e;
};

__device__ __forceinline__ EosRED eos_re_derivs(const EosTable &tab, double rho, double T, double abar, double zbar)
{
    const D2 A = D2{abar, abar / rho, 0.0};
    const D2 Z = dc(zbar);
    const EosD o = eos_dual(tab, D2{rho, 1.0, 0.0}, D2{T, 0.0, 1.0}, A, Z);
    double T_r = 0.0, T_e = 0.0;
    if (T > eosc::mintemp) {
        T_e = 1.0 / o.e.b;
        T_r = -o.e.a / o.e.b;
    }
    EosRED r;
    r.p_r = o.p.a + o.p.b * T_r;
    r.p_e = o.p.b * T_e;
    r.g_r = o.gam1.a + o.gam1.b * T_r;
    r.g_e = o.gam1.b * T_e;
    // c = sqrt(gam1 p / rho)
    const double c = sqrt(o.gam1.v * o.p.v / rho);
    const double h = 0.5 / c;
    r.c_r = h * ((r.g_r * o.p.v + o.gam1.v * r.p_r) / rho - o.gam1.v * o.p.v / (rho * rho));
    r.c_e = h * (r.g_e * o.p.v + o.gam1.v * r.p_e) / rho;
    return r;
}

// d e / d rho at fixed T (the floors of enforce_minimum_density and
// reset_internal_e evaluate e(rho, small_temp)).
__device__ __forceinline__ double eos_de_drho_T(const EosTable &tab, double rho, double T, double abar, double zbar)
{
    const EosD o = eos_dual(tab, D2{rho, 1.0, 0.0}, D2{T, 0.0, 1.0}, D2{abar, abar / rho, 0.0}, dc(zbar));
    return o.e.a;
}

// ------------------------------------------------------------------ state
// Hydro part of a conserved state: rho, mx, my, mz, E, rho e.
enum { HB_RHO = 0, HB_MX, HB_MY, HB_MZ, HB_E, HB_EI, NHB };
__device__ __forceinline__ int hb_comp(int h) { return h < 4 ? h : (h == HB_E ? UEDEN : UEINT); }

// Adjoint of clean_cell (Source/Castro.cpp:561-591) on the hydro part: u is
// the input state of the clean, ub the adjoint of its output on entry and of
// its input on return.
__device__ __forceinline__ void clean_cell_b(const double *uin, const EosTable &tab, double *ub)
{
    double u[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u[n] = uin[n];
    // enforce_minimum_density: a reset state does not depend on the input hydro
    if (u[URHO] < small_dens) {
#pragma unroll
        for (int h = 0; h < NHB; ++h) ub[h] = 0.0;
        return;
    }
    normalize_species_cell(u);
    // reset_internal_e (Source/Castro_nd.F90:164-249)
    const double rho = u[URHO];
    const double rhoInv = 1.0 / rho;
    const double Up = u[UMX] * rhoInv, Vp = u[UMY] * rhoInv, Wp = u[UMZ] * rhoInv;
    const double ke = HALF * (Up * Up + Vp * Vp + Wp * Wp);
    const double E = u[UEDEN];
    // rho * ke = |m|^2 / (2 rho): d/dm = m/rho, d/drho = -ke
    if (E < 0.0) {
        // UEDEN = UEINT' + rho*ke with UEINT' = UEINT or the floor rho e(rho, small_temp)
        const double bE = ub[HB_E];
        ub[HB_EI] += bE;
        ub[HB_E] = 0.0;
        ub[HB_RHO] += bE * (-ke);
        ub[HB_MX] += bE * Up;
        ub[HB_MY] += bE * Vp;
        ub[HB_MZ] += bE * Wp;
        if (u[UEINT] < 0.0) {
            double abar, zbar;
            abar_zbar_u(u, rhoInv, abar, zbar);
            const double e0 = eos_rt_e(tab, rho, small_temp, abar, zbar);
            const double dedr = eos_de_drho_T(tab, rho, small_temp, abar, zbar);
            ub[HB_RHO] += ub[HB_EI] * (e0 + rho * dedr);
            ub[HB_EI] = 0.0;
        }
    } else {
        const double rho_eint = E - rho * ke;
        if (rho_eint > 0.0 && rho_eint / E > 1.e-4) {
            const double b = ub[HB_EI];
            ub[HB_EI] = 0.0;
            ub[HB_E] += b;
            ub[HB_RHO] += b * ke;
            ub[HB_MX] -= b * Up;
            ub[HB_MY] -= b * Vp;
            ub[HB_MZ] -= b * Wp;
        } else if (u[UEINT] <= 0.0) {
            double abar, zbar;
            abar_zbar_u(u, rhoInv, abar, zbar);
            const double e0 = eos_rt_e(tab, rho, small_temp, abar, zbar);
            const double dedr = eos_de_drho_T(tab, rho, small_temp, abar, zbar);
            ub[HB_RHO] += ub[HB_EI] * (e0 + rho * dedr);
            ub[HB_EI] = 0.0;
        }
    }
    // compute_temp: UEINT = rho * (UEINT / rho) is the identity; T is not differentiated
}

// clean_cell without the temperature solve: the hydro components (and the
// species) come out identical, compute_temp's rho*(rho e/rho) rounding
// included; UTEMP keeps its input value.  The adjoint sweeps use it to
// recover the branch of every clean without EOS Newton iterations.
__device__ __forceinline__ void clean_hydro(double *u, const EosTable &tab)
{
    enforce_minimum_density_cell(u, tab);
    normalize_species_cell(u);
    reset_internal_e_cell(u, tab);
    const double rho = u[URHO];
    const double rhoInv = 1.0 / rho;
    const double e = u[UEINT] * rhoInv;
    u[UEINT] = rho * e;
}

// Adjoint of ctoprim_cell (Source/Castro_hydro_nd.F90:115-164): qb holds the
// adjoints of (rho, u, v, w, p, rho e, c, gamma_c); ub (hydro part of the
// cleaned state) is incremented.
enum { QB_RHO = 0, QB_U, QB_V, QB_W, QB_P, QB_RE, QB_C, QB_G, NQB };

__device__ __forceinline__ void ctoprim_b(const double *u, const EosTable &tab, const double *qb, double *ub)
{
    const double rho = u[URHO];
    const double rhoinv = 1.0 / rho;
    const double qu = u[UMX] * rhoinv, qv = u[UMY] * rhoinv, qw = u[UMZ] * rhoinv;
    const double kineng = HALF * rho * (qu * qu + qv * qv + qw * qw);
    const double E = u[UEDEN];
    const bool from_E = (E - kineng) / E > dual_energy_eta1;
    const double e = from_E ? (E - kineng) * rhoinv : u[UEINT] * rhoinv;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s1 = s1 + (u[UFS + n] * rhoinv) * ainv(n);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s2 = s2 + (u[UFS + n] * rhoinv) * zion(n) * ainv(n);
    const double abar = 1.0 / s1;
    const double zbar = abar * s2;
    const EosRE r = eos_re<false>(tab, rho, u[UTEMP], e, abar, zbar);
    const EosRED dr = eos_re_derivs(tab, rho, r.T, abar, zbar);

    // q.rhoe = e * rho; q.p, q.gamc, q.cs from the EOS at (rho, e)
    double e_b = qb[QB_RE] * rho + qb[QB_P] * dr.p_e + qb[QB_G] * dr.g_e + qb[QB_C] * dr.c_e;
    double rho_b = qb[QB_RHO] + qb[QB_RE] * e + qb[QB_P] * dr.p_r + qb[QB_G] * dr.g_r + qb[QB_C] * dr.c_r;
    double qu_b = qb[QB_U], qv_b = qb[QB_V], qw_b = qb[QB_W];
    if (from_E) {
        // e = (E - kineng)/rho, kineng = |m|^2/(2 rho)
        ub[HB_E] += e_b * rhoinv;
        const double kin_b = -e_b * rhoinv;
        rho_b += -e_b * e * rhoinv;
        const double ke_b = kin_b;
        // kineng = 0.5 * rho * (qu^2+qv^2+qw^2)
        rho_b += ke_b * HALF * (qu * qu + qv * qv + qw * qw);
        qu_b += ke_b * rho * qu;
        qv_b += ke_b * rho * qv;
        qw_b += ke_b * rho * qw;
    } else {
        ub[HB_EI] += e_b * rhoinv;
        rho_b += -e_b * e * rhoinv;
    }
    ub[HB_MX] += qu_b * rhoinv;
    ub[HB_MY] += qv_b * rhoinv;
    ub[HB_MZ] += qw_b * rhoinv;
    rho_b += -(qu_b * qu + qv_b * qv + qw_b * qw) * rhoinv;
    ub[HB_RHO] += rho_b;
}

// ------------------------------------------------------------------ PPM
// van Leer slope of Source/ppm.F90:147-168 and its derivative weights with
// respect to (a, b, c) = s(i-1), s(i), s(i+1).
__device__ __forceinline__ double vl_slope(double a, double b, double c, double &wa, double &wb, double &wc)
{
    const double dsl = 2.0 * (b - a), dsr = 2.0 * (c - b);
    wa = wb = wc = 0.0;
    if (!(dsl * dsr > 0.0)) return 0.0;
    const double dsc = HALF * (c - a);
    const double sg = fsign1(dsc);
    const double adsc = fabs(dsc), adsl = fabs(dsl), adsr = fabs(dsr);
    const double m1 = fmin(adsl, adsr);
    const double v = sg * fmin(adsc, m1);
    if (adsc <= m1) {
        // v = sg*|dsc| = dsc
        wa = -HALF;
        wc = HALF;
    } else if (adsl <= adsr) {
        const double s = sg * fsign1(dsl);  // d|dsl| = sign(dsl) d dsl
        wa = -2.0 * s;
        wb = 2.0 * s;
    } else {
        const double s = sg * fsign1(dsr);
        wb = -2.0 * s;
        wc = 2.0 * s;
    }
    return v;
}

// Adjoint of ppm_reconstruct: s5 = s(-2..2), outputs (sm, sp) with bars
// (smb, spb); increments s5b[0..4] and flatb.
__device__ __forceinline__ void ppm_reconstruct_b(const double *s5, double flatn, double smb, double spb, double *s5b,
                                                  double &flatb)
{
    const double *s = s5 + 2;
    double *sb = s5b + 2;
    double wl[3], wr[3], w1[3];
    const double dl = vl_slope(s[-2], s[-1], s[0], wl[0], wl[1], wl[2]);   // slope at -1
    const double dr = vl_slope(s[-1], s[0], s[1], wr[0], wr[1], wr[2]);    // slope at 0
    const double d1 = vl_slope(s[0], s[1], s[2], w1[0], w1[1], w1[2]);     // slope at 1
    // sm0 = 0.5(s0 + s-1) - (dr - dl)/6, clamped to [min, max] of (s0, s-1)
    const double sm0 = HALF * (s[0] + s[-1]) - SIXTH * (dr - dl);
    const double lo_m = fmin(s[0], s[-1]), hi_m = fmax(s[0], s[-1]);
    const double sm1 = fmax(sm0, lo_m);
    const double sm2 = fmin(sm1, hi_m);
    const double sp0 = HALF * (s[1] + s[0]) - SIXTH * (d1 - dr);
    const double lo_p = fmin(s[1], s[0]), hi_p = fmax(s[1], s[0]);
    const double sp1 = fmax(sp0, lo_p);
    const double sp2 = fmin(sp1, hi_p);
    const double smf = flatn * sm2 + (1.0 - flatn) * s[0];
    const double spf = flatn * sp2 + (1.0 - flatn) * s[0];

    // limiter (Source/ppm.F90:219-231)
    double smf_b = 0.0, spf_b = 0.0, s0_b = 0.0;
    if ((spf - s[0]) * (s[0] - smf) <= 0.0) {
        s0_b += smb + spb;
    } else if (fabs(spf - s[0]) >= 2.0 * fabs(smf - s[0])) {
        // sp = 3 s0 - 2 sm
        s0_b += 3.0 * spb;
        smf_b += -2.0 * spb + smb;
    } else if (fabs(smf - s[0]) >= 2.0 * fabs(spf - s[0])) {
        s0_b += 3.0 * smb;
        spf_b += -2.0 * smb + spb;
    } else {
        smf_b += smb;
        spf_b += spb;
    }
    // flattening blend
    flatb += smf_b * (sm2 - s[0]) + spf_b * (sp2 - s[0]);
    s0_b += (1.0 - flatn) * (smf_b + spf_b);
    double sm2_b = flatn * smf_b, sp2_b = flatn * spf_b;
    // clamps of sp
    double sp0_b = 0.0;
    if (sp1 <= hi_p) {
        if (sp0 >= lo_p) sp0_b = sp2_b;
        else {
            if (s[1] <= s[0]) sb[1] += sp2_b;
            else s0_b += sp2_b;
        }
    } else {
        if (s[1] >= s[0]) sb[1] += sp2_b;
        else s0_b += sp2_b;
    }
    double sm0_b = 0.0;
    if (sm1 <= hi_m) {
        if (sm0 >= lo_m) sm0_b = sm2_b;
        else {
            if (s[0] <= s[-1]) s0_b += sm2_b;
            else sb[-1] += sm2_b;
        }
    } else {
        if (s[0] >= s[-1]) s0_b += sm2_b;
        else sb[-1] += sm2_b;
    }
    // sp0 = 0.5(s1 + s0) - (d1 - dr)/6
    sb[1] += HALF * sp0_b;
    s0_b += HALF * sp0_b;
    const double d1_b = -SIXTH * sp0_b;
    double dr_b = SIXTH * sp0_b;
    // sm0 = 0.5(s0 + s-1) - (dr - dl)/6
    s0_b += HALF * sm0_b;
    sb[-1] += HALF * sm0_b;
    dr_b += -SIXTH * sm0_b;
    const double dl_b = SIXTH * sm0_b;
    sb[-2] += dl_b * wl[0];
    sb[-1] += dl_b * wl[1];
    s0_b += dl_b * wl[2];
    sb[-1] += dr_b * wr[0];
    s0_b += dr_b * wr[1];
    sb[1] += dr_b * wr[2];
    s0_b += d1_b * w1[0];
    sb[1] += d1_b * w1[1];
    sb[2] += d1_b * w1[2];
    sb[0] += s0_b;
}

// Adjoint of ppm_int_wave; increments smb, spb, s6b, speedb.
__device__ __forceinline__ void ppm_int_wave_b(double sm, double sp, double s6, double speed, double dtdx, double Ipb,
                                               double Imb, double &smb, double &spb, double &s6b, double &speedb)
{
    const double sigma = fabs(speed) * dtdx;
    double sigb = 0.0;
    if (speed <= 0.0) {
        // Ip = sp; Im = sm + 0.5 sigma (sp - sm + (1 - 2/3 sigma) s6)
        spb += Ipb;
        smb += Imb * (1.0 - HALF * sigma);
        spb += Imb * HALF * sigma;
        s6b += Imb * HALF * sigma * (1.0 - TWO3RD * sigma);
        sigb = Imb * (HALF * (sp - sm + (1.0 - TWO3RD * sigma) * s6) - HALF * sigma * TWO3RD * s6);
    } else {
        // Ip = sp - 0.5 sigma (sp - sm - (1 - 2/3 sigma) s6); Im = sm
        smb += Imb;
        spb += Ipb * (1.0 - HALF * sigma);
        smb += Ipb * HALF * sigma;
        s6b += Ipb * HALF * sigma * (1.0 - TWO3RD * sigma);
        sigb = -Ipb * (HALF * (sp - sm - (1.0 - TWO3RD * sigma) * s6) + HALF * sigma * TWO3RD * s6);
    }
    // sigma = |speed| dtdx
    speedb += sigb * dtdx * (speed < 0.0 ? -1.0 : (speed > 0.0 ? 1.0 : 0.0));
}

// Adjoint of ppm_project<PLUS>: I[w][v] (v = rho, un, p, rhoe), gam_ref;
// output bars (rho, un, rhoe, p); increments Ib and gamb.  un and cc only
// select the active characteristic waves.
template <bool PLUS>
__device__ __forceinline__ void ppm_project_b(const double I[3][4], double gam_ref, double un, double cc, double orb,
                                              double oub, double oreb, double opb, double Ib[3][4], double &gamb)
{
    const int w = PLUS ? 0 : 2;
    const double rho_ref = fmax(I[w][0], small_dens);
    const double un_ref = I[w][1];
    const double p_ref = fmax(I[w][2], small_pres);
    const double rhoe_ref = I[w][3];
    const double rri = 1.0 / rho_ref;
    const double csq = gam_ref * p_ref * rri;
    const double ccr = sqrt(csq);
    const double cci = 1.0 / ccr;
    const double h = (p_ref + rhoe_ref) * rri;
    const double dum = un_ref - I[0][1];
    const double dptotm = p_ref - I[0][2];
    const double drho = rho_ref - I[1][0];
    const double dptot = p_ref - I[1][2];
    const double drhoe = rhoe_ref - I[1][3];
    const double dup = un_ref - I[2][1];
    const double dptotp = p_ref - I[2][2];
    const double Ym = dptotm * rri * cci - dum;
    const double Yp = dptotp * rri * cci + dup;
    const double am0 = HALF * Ym * rho_ref * cci;
    const double ap0 = HALF * Yp * rho_ref * cci;
    const double a0r0 = drho - dptot / csq;
    const double a0e0 = drhoe - dptot * h / csq;
    double fm, fp, f0;
    if (PLUS) {
        fm = (un - cc > 0.0) ? 0.0 : -1.0;
        fp = (un + cc > 0.0) ? 0.0 : -1.0;
        f0 = (un > 0.0) ? 0.0 : -1.0;
    } else {
        fm = (un - cc > 0.0) ? -1.0 : 0.0;
        fp = (un + cc > 0.0) ? -1.0 : 0.0;
        f0 = (un > 0.0) ? -1.0 : 0.0;
    }
    const double am = fm * am0, ap = fp * ap0, a0r = f0 * a0r0, a0e = f0 * a0e0;
    const double A = ap + am;
    double rho_ref_b = 0.0, un_ref_b = 0.0, p_ref_b = 0.0, rhoe_ref_b = 0.0;
    double A_b = 0.0, am_b = 0.0, ap_b = 0.0, a0r_b = 0.0, a0e_b = 0.0, csq_b = 0.0, h_b = 0.0, ccr_b = 0.0,
           rri_b = 0.0;
    // o_p = max(small_pres, p_ref + A csq)
    if (p_ref + A * csq > small_pres) {
        p_ref_b += opb;
        A_b += opb * csq;
        csq_b += opb * A;
    }
    // o_rhoe = rhoe_ref + A h + a0e
    rhoe_ref_b += oreb;
    A_b += oreb * h;
    h_b += oreb * A;
    a0e_b += oreb;
    // o_un = un_ref + (ap - am) ccr rri
    un_ref_b += oub;
    {
        const double D = ap - am;
        const double D_b = oub * ccr * rri;
        ccr_b += oub * D * rri;
        rri_b += oub * D * ccr;
        ap_b += D_b;
        am_b -= D_b;
    }
    // o_rho = max(small_dens, rho_ref + ap + am + a0r)
    if (rho_ref + ap + am + a0r > small_dens) {
        rho_ref_b += orb;
        ap_b += orb;
        am_b += orb;
        a0r_b += orb;
    }
    ap_b += A_b;
    am_b += A_b;
    const double am0_b = fm * am_b, ap0_b = fp * ap_b, a0r0_b = f0 * a0r_b, a0e0_b = f0 * a0e_b;
    double cci_b = 0.0;
    double dptot_b = 0.0, drho_b = 0.0, drhoe_b = 0.0, dum_b = 0.0, dptotm_b = 0.0, dup_b = 0.0, dptotp_b = 0.0;
    // a0e0 = drhoe - dptot h / csq
    drhoe_b += a0e0_b;
    dptot_b -= a0e0_b * h / csq;
    h_b -= a0e0_b * dptot / csq;
    csq_b += a0e0_b * dptot * h / (csq * csq);
    // a0r0 = drho - dptot / csq
    drho_b += a0r0_b;
    dptot_b -= a0r0_b / csq;
    csq_b += a0r0_b * dptot / (csq * csq);
    // ap0 = 0.5 Yp rho_ref cci
    {
        const double Yp_b = ap0_b * HALF * rho_ref * cci;
        rho_ref_b += ap0_b * HALF * Yp * cci;
        cci_b += ap0_b * HALF * Yp * rho_ref;
        dptotp_b += Yp_b * rri * cci;
        rri_b += Yp_b * dptotp * cci;
        cci_b += Yp_b * dptotp * rri;
        dup_b += Yp_b;
    }
    {
        const double Ym_b = am0_b * HALF * rho_ref * cci;
        rho_ref_b += am0_b * HALF * Ym * cci;
        cci_b += am0_b * HALF * Ym * rho_ref;
        dptotm_b += Ym_b * rri * cci;
        rri_b += Ym_b * dptotm * cci;
        cci_b += Ym_b * dptotm * rri;
        dum_b -= Ym_b;
    }
    // differences
    un_ref_b += dup_b;
    Ib[2][1] -= dup_b;
    p_ref_b += dptotp_b;
    Ib[2][2] -= dptotp_b;
    rhoe_ref_b += drhoe_b;
    Ib[1][3] -= drhoe_b;
    p_ref_b += dptot_b;
    Ib[1][2] -= dptot_b;
    rho_ref_b += drho_b;
    Ib[1][0] -= drho_b;
    p_ref_b += dptotm_b;
    Ib[0][2] -= dptotm_b;
    un_ref_b += dum_b;
    Ib[0][1] -= dum_b;
    // h = (p_ref + rhoe_ref) rri
    p_ref_b += h_b * rri;
    rhoe_ref_b += h_b * rri;
    rri_b += h_b * (p_ref + rhoe_ref);
    // cci = 1/ccr, ccr = sqrt(csq)
    ccr_b -= cci_b * cci * cci;
    csq_b += ccr_b * HALF / ccr;
    // csq = gam_ref p_ref rri
    gamb += csq_b * p_ref * rri;
    p_ref_b += csq_b * gam_ref * rri;
    rri_b += csq_b * gam_ref * p_ref;
    rho_ref_b -= rri_b * rri * rri;
    if (I[w][0] > small_dens) Ib[w][0] += rho_ref_b;
    Ib[w][1] += un_ref_b;
    if (I[w][2] > small_pres) Ib[w][2] += p_ref_b;
    Ib[w][3] += rhoe_ref_b;
}

// Adjoint of flatten_dir: increments p7b, u7b (u7 is not differentiable:
// it only selects chi) given the bar of the returned coefficient.
__device__ __forceinline__ void flatten_dir_b(const double *p7, const double *u7, double fb, double *p7b)
{
    const double shktst = 0.33, zcut1 = 0.75, zcut2 = 0.85;
    const double dzcut = 1.0 / (zcut2 - zcut1);
    const double *p = p7 + 3, *u = u7 + 3;
    double *pb = p7b + 3;
    double zz[2], ch[2], zraw[2], zeta[2], dpv[2], dp2v[2], den[2];
    int c0[2];
    const double dp0 = p[1] - p[-1];
    const int ishft = (dp0 > 0.0) ? 1 : -1;
    for (int q = 0; q < 2; ++q) {
        const int o = q == 0 ? 0 : -ishft;
        c0[q] = o;
        const double pr = p[1 + o], pl = p[-1 + o];
        dp2v[q] = p[2 + o] - p[-2 + o];
        const double du = u[1 + o] - u[-1 + o];
        dpv[q] = pr - pl;
        den[q] = fmax(small_pres, fabs(dp2v[q]));
        zeta[q] = fabs(dpv[q]) / den[q];
        zraw[q] = dzcut * (zeta[q] - zcut1);
        zz[q] = fmin(1.0, fmax(0.0, zraw[q]));
        const double tst = (du <= 0.0) ? 1.0 : 0.0;
        ch[q] = (fabs(dpv[q]) > shktst * fmin(pr, pl)) ? tst : 0.0;
    }
    // f = 1 - max(chi2 z2, chi z): q = 1 is the first argument of the max
    const int a = (ch[1] * zz[1] >= ch[0] * zz[0]) ? 1 : 0;
    const double zb = -fb * ch[a];
    if (zb != 0.0 && zraw[a] > 0.0 && zraw[a] < 1.0) {
        const double zetab = dzcut * zb;
        const int o = c0[a];
        // zeta = |dp| / den
        const double dpb = zetab * (dpv[a] > 0.0 ? 1.0 : (dpv[a] < 0.0 ? -1.0 : 0.0)) / den[a];
        pb[1 + o] += dpb;
        pb[-1 + o] -= dpb;
        if (fabs(dp2v[a]) > small_pres) {
            const double denb = -zetab * fabs(dpv[a]) / (den[a] * den[a]);
            const double dp2b = denb * (dp2v[a] > 0.0 ? 1.0 : -1.0);
            pb[2 + o] += dp2b;
            pb[-2 + o] -= dp2b;
        }
    }
}

// ------------------------------------------------------------------ Riemann
struct RiemannHB {   // adjoints of the Riemann outputs used downstream
    double f_rho, f_mn, f_mt, f_mtt, f_E, f_e, g_p, g_un;
};
struct FaceHB {
    double rho, un, ut, utt, p, rhoe;
};

__device__ __forceinline__ double sgn0(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// Adjoint of riemann_hydro (Source/riemann.F90:114-318).  Increments Lb, Rb
// and the bars of cL, cR, gamcl, gamcr.
__device__ __forceinline__ void riemann_hydro_b(const FaceH &L, const FaceH &R, double cL, double cR, double gamcl,
                                                double gamcr, const RiemannHB &ob, FaceHB &Lb, FaceHB &Rb,
                                                double &cLb, double &cRb, double &gamclb, double &gamcrb)
{
    // ---- forward, keeping intermediates
    const double rl = fmax(L.rho, small_dens);
    const double ul = L.un, v1l = L.ut, v2l = L.utt;
    const double pl = fmax(L.p, small_pres);
    const double rel = L.rhoe;
    const double rr = fmax(R.rho, small_dens);
    const double ur = R.un, v1r = R.ut, v2r = R.utt;
    const double pr = fmax(R.p, small_pres);
    const double rer = R.rhoe;
    const double cmax = fmax(cR, cL);
    const double csmall = fmax(small, small * cmax);
    const double cavg = HALF * (cR + cL);
    const double wsmall = small_dens * csmall;
    const double argl = gamcl * pl * rl, argr = gamcr * pr * rr;
    const double sl = sqrt(fabs(argl)), sr = sqrt(fabs(argr));
    const double wl = fmax(wsmall, sl);
    const double wr = fmax(wsmall, sr);
    const double wwinv = 1.0 / (wl + wr);
    const double Bp = (wr * pl + wl * pr) + wl * wr * (ul - ur);
    const double Bu = (wl * ul + wr * ur) + (pl - pr);
    const double pstar0 = Bp * wwinv;
    double ustar = Bu * wwinv;
    const double pstar = fmax(pstar0, small_pres);
    const bool uzero = fabs(ustar) < smallu * HALF * (fabs(ul) + fabs(ur));
    if (uzero) ustar = 0.0;
    double sgnm = fsign1(ustar);
    if (ustar == 0.0) sgnm = 0.0;
    const double fpl = HALF * (1.0 + sgnm);
    const double fmr = HALF * (1.0 - sgnm);
    const double ro0 = fpl * rl + fmr * rr;
    const double uo = fpl * ul + fmr * ur;
    const double po = fpl * pl + fmr * pr;
    const double reo = fpl * rel + fmr * rer;
    const double gamco = fpl * gamcl + fmr * gamcr;
    const double ro = fmax(small_dens, ro0);
    const double roinv = 1.0 / ro;
    const double argo = gamco * po * roinv;
    const double co0 = sqrt(fabs(argo));
    const double co = fmax(csmall, co0);
    const double co2inv = 1.0 / (co * co);
    const double g_ut = fpl * v1l + fmr * v1r;
    const double g_utt = fpl * v2l + fmr * v2r;
    const double drho = (pstar - po) * co2inv;
    const double rstar0 = ro + drho;
    const double rstar = fmax(small_dens, rstar0);
    const double rstarinv = 1.0 / rstar;
    const double entho = (reo + po) * roinv * co2inv;
    const double estar = reo + (pstar - po) * entho;
    const double args = gamco * pstar * rstarinv;
    const double cstar0 = sqrt(fabs(args));
    const double cstar = fmax(cstar0, csmall);
    const double sg2 = fsign1(ustar);
    const double spout0 = co - sg2 * uo;
    const double spin0 = cstar - sg2 * ustar;
    const double ushock = HALF * (spin0 + spout0);
    const bool shock = pstar - po > 0.0;
    const double spout = shock ? ushock : spout0;
    const double spin = shock ? ushock : spin0;
    const bool scr0 = spout - spin == 0.0;
    const double scr = scr0 ? small * cavg : spout - spin;
    const double scrinv = 1.0 / scr;
    const double fraw = (1.0 + (spout + spin) * scrinv) * HALF;
    const double frac = fmax(0.0, fmin(1.0, fraw));
    int mode = 0;  // 0 blend, 1 outer (spout < 0), 2 star (spin >= 0)
    if (spout < 0.0) mode = 1;
    if (spin >= 0.0) mode = 2;
    double qrho, qun, qp0, regdnv;
    if (mode == 0) {
        qrho = frac * rstar + (1.0 - frac) * ro;
        qun = frac * ustar + (1.0 - frac) * uo;
        qp0 = frac * pstar + (1.0 - frac) * po;
        regdnv = frac * estar + (1.0 - frac) * reo;
    } else if (mode == 1) {
        qrho = ro;
        qun = uo;
        qp0 = po;
        regdnv = reo;
    } else {
        qrho = rstar;
        qun = ustar;
        qp0 = pstar;
        regdnv = estar;
    }
    const double qp = fmax(qp0, small_pres);
    const double f_rho = qrho * qun;
    const double ksum = qun * qun + g_ut * g_ut + g_utt * g_utt;
    const double rhoetot = regdnv + HALF * qrho * ksum;

    // ---- reverse
    double qun_b = ob.g_un, qp_b = ob.g_p, qrho_b = 0.0, regdnv_b = 0.0, g_ut_b = 0.0, g_utt_b = 0.0;
    // f_e = qun regdnv
    qun_b += ob.f_e * regdnv;
    regdnv_b += ob.f_e * qun;
    // f_E = qun (rhoetot + qp)
    qun_b += ob.f_E * (rhoetot + qp);
    const double rhoetot_b = ob.f_E * qun;
    qp_b += ob.f_E * qun;
    regdnv_b += rhoetot_b;
    qrho_b += rhoetot_b * HALF * ksum;
    {
        const double ksum_b = rhoetot_b * HALF * qrho;
        qun_b += ksum_b * 2.0 * qun;
        g_ut_b += ksum_b * 2.0 * g_ut;
        g_utt_b += ksum_b * 2.0 * g_utt;
    }
    // momentum fluxes
    const double f_rho_b = ob.f_rho + ob.f_mn * qun + ob.f_mt * g_ut + ob.f_mtt * g_utt;
    qun_b += ob.f_mn * f_rho;
    g_ut_b += ob.f_mt * f_rho;
    g_utt_b += ob.f_mtt * f_rho;
    qrho_b += f_rho_b * qun;
    qun_b += f_rho_b * qrho;
    const double qp0_b = (qp0 > small_pres) ? qp_b : 0.0;

    double rstar_b = 0.0, ro_b = 0.0, ustar_b = 0.0, uo_b = 0.0, pstar_b = 0.0, po_b = 0.0, estar_b = 0.0,
           reo_b = 0.0, frac_b = 0.0;
    if (mode == 0) {
        rstar_b += frac * qrho_b;
        ro_b += (1.0 - frac) * qrho_b;
        ustar_b += frac * qun_b;
        uo_b += (1.0 - frac) * qun_b;
        pstar_b += frac * qp0_b;
        po_b += (1.0 - frac) * qp0_b;
        estar_b += frac * regdnv_b;
        reo_b += (1.0 - frac) * regdnv_b;
        frac_b = qrho_b * (rstar - ro) + qun_b * (ustar - uo) + qp0_b * (pstar - po) + regdnv_b * (estar - reo);
    } else if (mode == 1) {
        ro_b += qrho_b;
        uo_b += qun_b;
        po_b += qp0_b;
        reo_b += regdnv_b;
    } else {
        rstar_b += qrho_b;
        ustar_b += qun_b;
        pstar_b += qp0_b;
        estar_b += regdnv_b;
    }
    double spout_b = 0.0, spin_b = 0.0, cavg_b = 0.0;
    if (frac_b != 0.0 && fraw > 0.0 && fraw < 1.0) {
        const double raw_b = frac_b;
        spout_b += HALF * raw_b * scrinv;
        spin_b += HALF * raw_b * scrinv;
        const double scrinv_b = HALF * raw_b * (spout + spin);
        const double scr_b = -scrinv_b * scrinv * scrinv;
        if (scr0) cavg_b += scr_b * small;
        else {
            spout_b += scr_b;
            spin_b -= scr_b;
        }
    }
    double spout0_b, spin0_b;
    if (shock) {
        const double ushock_b = spout_b + spin_b;
        spout0_b = HALF * ushock_b;
        spin0_b = HALF * ushock_b;
    } else {
        spout0_b = spout_b;
        spin0_b = spin_b;
    }
    double cstar_b = spin0_b, co_b = spout0_b, csmall_b = 0.0;
    ustar_b -= sg2 * spin0_b;
    uo_b -= sg2 * spout0_b;
    // cstar = max(cstar0, csmall)
    double gamco_b = 0.0, rstarinv_b = 0.0;
    if (cstar0 >= csmall) {
        if (cstar0 > 0.0) {
            const double a_b = cstar_b * sgn0(args) * HALF / cstar0;
            gamco_b += a_b * pstar * rstarinv;
            pstar_b += a_b * gamco * rstarinv;
            rstarinv_b += a_b * gamco * pstar;
        }
    } else csmall_b += cstar_b;
    // estar = reo + (pstar - po) entho
    reo_b += estar_b;
    pstar_b += estar_b * entho;
    po_b -= estar_b * entho;
    const double entho_b = estar_b * (pstar - po);
    // entho = (reo + po) roinv co2inv
    double roinv_b = 0.0, co2inv_b = 0.0;
    reo_b += entho_b * roinv * co2inv;
    po_b += entho_b * roinv * co2inv;
    roinv_b += entho_b * (reo + po) * co2inv;
    co2inv_b += entho_b * (reo + po) * roinv;
    // rstarinv = 1/rstar; rstar = max(small_dens, rstar0); rstar0 = ro + drho
    rstar_b -= rstarinv_b * rstarinv * rstarinv;
    const double rstar0_b = (rstar0 > small_dens) ? rstar_b : 0.0;
    ro_b += rstar0_b;
    const double drho_b = rstar0_b;
    // drho = (pstar - po) co2inv
    pstar_b += drho_b * co2inv;
    po_b -= drho_b * co2inv;
    co2inv_b += drho_b * (pstar - po);
    // transverse velocities
    double v1l_b = fpl * g_ut_b, v1r_b = fmr * g_ut_b, v2l_b = fpl * g_utt_b, v2r_b = fmr * g_utt_b;
    // co2inv = 1/co^2
    co_b += co2inv_b * (-2.0) * co2inv / co;
    // co = max(csmall, co0); co0 = sqrt(|gamco po roinv|)
    if (co0 > csmall) {
        if (co0 > 0.0) {
            const double a_b = co_b * sgn0(argo) * HALF / co0;
            gamco_b += a_b * po * roinv;
            po_b += a_b * gamco * roinv;
            roinv_b += a_b * gamco * po;
        }
    } else csmall_b += co_b;
    ro_b -= roinv_b * roinv * roinv;
    const double ro0_b = (ro0 > small_dens) ? ro_b : 0.0;
    double rl_b = fpl * ro0_b, rr_b = fmr * ro0_b;
    double ul_b = fpl * uo_b, ur_b = fmr * uo_b;
    double pl_b = fpl * po_b, pr_b = fmr * po_b;
    double rel_b = fpl * reo_b, rer_b = fmr * reo_b;
    gamclb += fpl * gamco_b;
    gamcrb += fmr * gamco_b;
    // ustar, pstar
    const double ustar0_b = uzero ? 0.0 : ustar_b;
    const double pstar0_b = (pstar0 > small_pres) ? pstar_b : 0.0;
    double wl_b = 0.0, wr_b = 0.0, wwinv_b = 0.0;
    {
        const double Bu_b = ustar0_b * wwinv;
        wwinv_b += ustar0_b * Bu;
        wl_b += Bu_b * ul;
        ul_b += Bu_b * wl;
        wr_b += Bu_b * ur;
        ur_b += Bu_b * wr;
        pl_b += Bu_b;
        pr_b -= Bu_b;
        const double Bp_b = pstar0_b * wwinv;
        wwinv_b += pstar0_b * Bp;
        wr_b += Bp_b * pl;
        pl_b += Bp_b * wr;
        wl_b += Bp_b * pr;
        pr_b += Bp_b * wl;
        wl_b += Bp_b * wr * (ul - ur);
        wr_b += Bp_b * wl * (ul - ur);
        ul_b += Bp_b * wl * wr;
        ur_b -= Bp_b * wl * wr;
    }
    {
        const double s_b = -wwinv_b * wwinv * wwinv;
        wl_b += s_b;
        wr_b += s_b;
    }
    double wsmall_b = 0.0;
    if (sl > wsmall) {
        if (sl > 0.0) {
            const double a_b = wl_b * sgn0(argl) * HALF / sl;
            gamclb += a_b * pl * rl;
            pl_b += a_b * gamcl * rl;
            rl_b += a_b * gamcl * pl;
        }
    } else wsmall_b += wl_b;
    if (sr > wsmall) {
        if (sr > 0.0) {
            const double a_b = wr_b * sgn0(argr) * HALF / sr;
            gamcrb += a_b * pr * rr;
            pr_b += a_b * gamcr * rr;
            rr_b += a_b * gamcr * pr;
        }
    } else wsmall_b += wr_b;
    csmall_b += small_dens * wsmall_b;
    cRb += HALF * cavg_b;
    cLb += HALF * cavg_b;
    if (small * cmax > small) {
        if (cR >= cL) cRb += small * csmall_b;
        else cLb += small * csmall_b;
    }
    if (L.rho > small_dens) Lb.rho += rl_b;
    if (R.rho > small_dens) Rb.rho += rr_b;
    if (L.p > small_pres) Lb.p += pl_b;
    if (R.p > small_pres) Rb.p += pr_b;
    Lb.un += ul_b;
    Rb.un += ur_b;
    Lb.ut += v1l_b;
    Rb.ut += v1r_b;
    Lb.utt += v2l_b;
    Rb.utt += v2r_b;
    Lb.rhoe += rel_b;
    Rb.rhoe += rer_b;
}

// ------------------------------------------------------------------ transverse
struct QStateHB {
    double rho, u, v, w, p, rhoe;
};
struct QInHB {
    double rho, u, v, w, p, rhoe;
};

// Adjoint of trans1_hydro: output bars ob; increments lqb, frb[0..4],
// flb[0..4], the Godunov bars and gamcb.
__device__ __forceinline__ void trans1_hydro_b(const QInH &lq2, const double *fr, const double *fl, double pgp,
                                               double pgm, double ugp, double ugm, double gamc, double cdtdx,
                                               const QStateHB &ob, QInHB &lqb, double *frb, double *flb,
                                               double &pgpb, double &pgmb, double &ugpb, double &ugmb, double &gamcb)
{
    const double dup = pgp * ugp - pgm * ugm;
    const double du = ugp - ugm;
    const double pav = HALF * (pgp + pgm);
    const double rrl2 = lq2.rho;
    const double rul2 = rrl2 * lq2.u, rvl2 = rrl2 * lq2.v, rwl2 = rrl2 * lq2.w;
    double rrn = rrl2 - cdtdx * (fr[URHO] - fl[URHO]);
    double run = rul2 - cdtdx * (fr[UMX] - fl[UMX]);
    double rvn = rvl2 - cdtdx * (fr[UMY] - fl[UMY]);
    double rwn = rwl2 - cdtdx * (fr[UMZ] - fl[UMZ]);
    const bool reset = rrn < 0.0;
    if (reset) {
        rrn = rrl2;
        run = rul2;
        rvn = rvl2;
        rwn = rwl2;
    }
    const double rhoinv = 1.0 / rrn;
    const double ou = run * rhoinv, ov = rvn * rhoinv, ow = rwn * rhoinv;
    // o.rhoe = ren - 0.5 (run^2+rvn^2+rwn^2) rhoinv
    double ren_b = ob.rhoe;
    double run_b = ob.u * rhoinv - ob.rhoe * ou;
    double rvn_b = ob.v * rhoinv - ob.rhoe * ov;
    double rwn_b = ob.w * rhoinv - ob.rhoe * ow;
    double rhoinv_b = ob.u * run + ob.v * rvn + ob.w * rwn - ob.rhoe * HALF * (run * run + rvn * rvn + rwn * rwn);
    double rrn_b = ob.rho - rhoinv_b * rhoinv * rhoinv;
    // pressure
    if (!reset) {
        const double pn = lq2.p - cdtdx * (dup + pav * du * (gamc - 1.0));
        if (pn > small_pres) {
            const double pb = ob.p;
            lqb.p += pb;
            const double X_b = -cdtdx * pb;  // X = dup + pav du (gamc - 1)
            const double dup_b = X_b;
            const double pav_b = X_b * du * (gamc - 1.0);
            const double du_b = X_b * pav * (gamc - 1.0);
            gamcb += X_b * pav * du;
            pgpb += dup_b * ugp + HALF * pav_b;
            ugpb += dup_b * pgp + du_b;
            pgmb += -dup_b * ugm + HALF * pav_b;
            ugmb += -dup_b * pgm - du_b;
        }
    } else {
        lqb.p += ob.p;
    }
    double rrl2_b = 0.0, rul2_b = 0.0, rvl2_b = 0.0, rwl2_b = 0.0, rel2_b = 0.0;
    if (reset) {
        rrl2_b += rrn_b;
        rul2_b += run_b;
        rvl2_b += rvn_b;
        rwl2_b += rwn_b;
        rel2_b += ren_b;
    } else {
        rrl2_b += rrn_b;
        rul2_b += run_b;
        rvl2_b += rvn_b;
        rwl2_b += rwn_b;
        rel2_b += ren_b;
        const double c = cdtdx;
        frb[URHO] -= c * rrn_b;
        flb[URHO] += c * rrn_b;
        frb[UMX] -= c * run_b;
        flb[UMX] += c * run_b;
        frb[UMY] -= c * rvn_b;
        flb[UMY] += c * rvn_b;
        frb[UMZ] -= c * rwn_b;
        flb[UMZ] += c * rwn_b;
        frb[UEDEN] -= c * ren_b;
        flb[UEDEN] += c * ren_b;
    }
    // rel2 = rhoe + 0.5 rho (u^2+v^2+w^2); rul2 = rho u ...
    lqb.rhoe += rel2_b;
    rrl2_b += rel2_b * HALF * (lq2.u * lq2.u + lq2.v * lq2.v + lq2.w * lq2.w);
    lqb.u += rel2_b * rrl2 * lq2.u + rul2_b * rrl2;
    lqb.v += rel2_b * rrl2 * lq2.v + rvl2_b * rrl2;
    lqb.w += rel2_b * rrl2 * lq2.w + rwl2_b * rrl2;
    rrl2_b += rul2_b * lq2.u + rvl2_b * lq2.v + rwl2_b * lq2.w;
    lqb.rho += rrl2_b;
}

// Adjoint of trans2_hydro.
__device__ __forceinline__ void trans2_hydro_b(const QInH &lq, const double *f2r, const double *f2l, const double *f3r,
                                               const double *f3l, double pg2p, double pg2m, double ug2p, double ug2m,
                                               double pg3p, double pg3m, double ug3p, double ug3m, double gamc,
                                               double cdtdx2, double cdtdx3, const QStateHB &ob, QInHB &lqb,
                                               double *f2rb, double *f2lb, double *f3rb, double *f3lb, double *gb,
                                               double &gamcb)
{
    // gb: pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m
    const double du2p = pg2p * ug2p - pg2m * ug2m;
    const double p2av = HALF * (pg2p + pg2m);
    const double du2 = ug2p - ug2m;
    const double p2new = cdtdx2 * (du2p + p2av * du2 * (gamc - 1.0));
    const double du3p = pg3p * ug3p - pg3m * ug3m;
    const double p3av = HALF * (pg3p + pg3m);
    const double du3 = ug3p - ug3m;
    const double p3new = cdtdx3 * (du3p + p3av * du3 * (gamc - 1.0));
    const double rrr = lq.rho;
    const double rur = rrr * lq.u, rvr = rrr * lq.v, rwr = rrr * lq.w;
    double rrn = rrr - cdtdx2 * (f2r[URHO] - f2l[URHO]) - cdtdx3 * (f3r[URHO] - f3l[URHO]);
    double run = rur - cdtdx2 * (f2r[UMX] - f2l[UMX]) - cdtdx3 * (f3r[UMX] - f3l[UMX]);
    double rvn = rvr - cdtdx2 * (f2r[UMY] - f2l[UMY]) - cdtdx3 * (f3r[UMY] - f3l[UMY]);
    double rwn = rwr - cdtdx2 * (f2r[UMZ] - f2l[UMZ]) - cdtdx3 * (f3r[UMZ] - f3l[UMZ]);
    const bool reset = rrn < 0.0;
    if (reset) {
        rrn = rrr;
        run = rur;
        rvn = rvr;
        rwn = rwr;
    }
    const double inv = 1.0 / rrn;
    const double ou = run * inv, ov = rvn * inv, ow = rwn * inv;
    double ren_b = ob.rhoe;
    double run_b = ob.u * inv - ob.rhoe * ou;
    double rvn_b = ob.v * inv - ob.rhoe * ov;
    double rwn_b = ob.w * inv - ob.rhoe * ow;
    // o.u = run / rrn, rhoeken = 0.5 |m|^2 / rrn
    double rrn_b = ob.rho - (ob.u * ou + ob.v * ov + ob.w * ow) * inv +
                   ob.rhoe * HALF * (run * run + rvn * rvn + rwn * rwn) * inv * inv;
    const double pn = reset ? lq.p : lq.p - p2new - p3new;
    if (pn > small_pres) {
        const double pb = ob.p;
        lqb.p += pb;
        if (!reset) {
            const double X2_b = -pb * cdtdx2, X3_b = -pb * cdtdx3;
            const double p2av_b = X2_b * du2 * (gamc - 1.0), du2_b = X2_b * p2av * (gamc - 1.0);
            const double p3av_b = X3_b * du3 * (gamc - 1.0), du3_b = X3_b * p3av * (gamc - 1.0);
            gamcb += X2_b * p2av * du2 + X3_b * p3av * du3;
            gb[0] += X2_b * ug2p + HALF * p2av_b;
            gb[2] += X2_b * pg2p + du2_b;
            gb[1] += -X2_b * ug2m + HALF * p2av_b;
            gb[3] += -X2_b * pg2m - du2_b;
            gb[4] += X3_b * ug3p + HALF * p3av_b;
            gb[6] += X3_b * pg3p + du3_b;
            gb[5] += -X3_b * ug3m + HALF * p3av_b;
            gb[7] += -X3_b * pg3m - du3_b;
        }
    }
    double rrr_b = rrn_b, rur_b = run_b, rvr_b = rvn_b, rwr_b = rwn_b, rer_b = ren_b;
    if (!reset) {
        const double vb[5] = {rrn_b, run_b, rvn_b, rwn_b, ren_b};
#pragma unroll
        for (int n = 0; n <= UEDEN; ++n) {
            f2rb[n] -= cdtdx2 * vb[n];
            f2lb[n] += cdtdx2 * vb[n];
            f3rb[n] -= cdtdx3 * vb[n];
            f3lb[n] += cdtdx3 * vb[n];
        }
    }
    lqb.rhoe += rer_b;
    rrr_b += rer_b * HALF * (lq.u * lq.u + lq.v * lq.v + lq.w * lq.w);
    lqb.u += rer_b * rrr * lq.u + rur_b * rrr;
    lqb.v += rer_b * rrr * lq.v + rvr_b * rrr;
    lqb.w += rer_b * rrr * lq.w + rwr_b * rrr;
    rrr_b += rur_b * lq.u + rvr_b * lq.v + rwr_b * lq.w;
    lqb.rho += rrr_b;
}

} // namespace mc

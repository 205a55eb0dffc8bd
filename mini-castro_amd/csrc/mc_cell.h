// Pointwise per-cell operations on a conserved state held in registers
// (u[NVAR]).  Used by the per-box drop-in kernels and by the fused batched
// kernels, so both paths evaluate identical arithmetic.
#pragma once

#include "mc_common.h"
#include "mc_eos.h"

namespace mc {

__device__ __forceinline__ double ainv(int n) { return 1.0 / aion(n); }

// abar, zbar from species components and 1/rho (Source/Castro_nd.F90:285-286)
__device__ __forceinline__ void abar_zbar_u(const double *u, double rhoInv, double &abar, double &zbar)
{
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s1 = s1 + u[UFS + n] * ainv(n);
    abar = 1.0 / (s1 * rhoInv);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s2 = s2 + u[UFS + n] * zion(n) * ainv(n);
    zbar = abar * (s2 * rhoInv);
}

// Source/Castro_nd.F90:69-125
__device__ __forceinline__ void enforce_minimum_density_cell(double *u, const EosTable &tab)
{
    const double rho0 = u[URHO];
    if (!(rho0 < small_dens)) return;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) u[UFS + n] = u[UFS + n] * (small_dens / rho0);
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s1 = s1 + u[UFS + n] * ainv(n);
    const double abar = 1.0 / (s1 / small_dens);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s2 = s2 + u[UFS + n] * zion(n) * ainv(n);
    const double zbar = abar * (s2 / small_dens);
    const double e = eos_rt_e(tab, small_dens, small_temp, abar, zbar);
    u[URHO] = small_dens;
    u[UTEMP] = small_temp;
    u[UMX] = 0.0;
    u[UMY] = 0.0;
    u[UMZ] = 0.0;
    u[UEINT] = small_dens * e;
    u[UEDEN] = u[UEINT];
}

// Source/Castro_nd.F90:129-160 (stores X = (rho X)/sum, not rho X)
__device__ __forceinline__ void normalize_species_cell(double *u)
{
    const double rho = u[URHO];
    double x[NSPEC], sum = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) x[n] = fmax(1.0e-30 * rho, fmin(rho, u[UFS + n]));
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) sum = sum + x[n];
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) u[UFS + n] = x[n] / sum;
}

// Source/Castro_nd.F90:164-249
__device__ __forceinline__ void reset_internal_e_cell(double *u, const EosTable &tab)
{
    const double rho = u[URHO];
    const double rhoInv = 1.0 / rho;
    const double Up = u[UMX] * rhoInv, Vp = u[UMY] * rhoInv, Wp = u[UMZ] * rhoInv;
    const double ke = HALF * (Up * Up + Vp * Vp + Wp * Wp);
    const double E = u[UEDEN];
    if (E < 0.0) {
        if (u[UEINT] < 0.0) {
            double abar, zbar;
            abar_zbar_u(u, rhoInv, abar, zbar);
            u[UEINT] = rho * eos_rt_e(tab, rho, small_temp, abar, zbar);
        }
        u[UEDEN] = u[UEINT] + rho * ke;
    } else {
        const double rho_eint = E - rho * ke;
        if (rho_eint > 0.0 && rho_eint / E > 1.e-4) {
            u[UEINT] = rho_eint;
        } else if (u[UEINT] <= 0.0) {
            double abar, zbar;
            abar_zbar_u(u, rhoInv, abar, zbar);
            const double eint_new = eos_rt_e(tab, rho, small_temp, abar, zbar);
            u[UEINT] = rho * eint_new;
        }
    }
}

// Source/Castro_nd.F90:253-297
__device__ __forceinline__ void compute_temp_cell(double *u, const EosTable &tab)
{
    const double rho = u[URHO];
    const double rhoInv = 1.0 / rho;
    const double e = u[UEINT] * rhoInv;
    double abar, zbar;
    abar_zbar_u(u, rhoInv, abar, zbar);
    const EosRE r = eos_re<false>(tab, rho, u[UTEMP], e, abar, zbar);
    u[UTEMP] = r.T;
    u[UEINT] = rho * e;
}

// Castro::clean_state's four kernels in order (Source/Castro.cpp:561-591)
__device__ __forceinline__ void clean_cell(double *u, const EosTable &tab)
{
    enforce_minimum_density_cell(u, tab);
    normalize_species_cell(u);
    reset_internal_e_cell(u, tab);
    compute_temp_cell(u, tab);
}

// Primitive state of one cell (Source/Castro_hydro_nd.F90:115-164).
struct Prim {
    double rho, u, v, w, p, rhoe, T, game;
    double x[NSPEC];
    double gamc, cs, dpdr_e, dpde;
};

// ctoprim before the EOS call: density, velocities, internal energy (the
// dual-energy choice), mass fractions and their abar / zbar.
struct PrimPre {
    double rho, u, v, w, e, abar, zbar, T0;
};

__device__ __forceinline__ void ctoprim_pre(const double *u, PrimPre &a, double *xs)
{
    const double rho = u[URHO];
    a.rho = rho;
    const double rhoinv = 1.0 / rho;
    a.u = u[UMX] * rhoinv;
    a.v = u[UMY] * rhoinv;
    a.w = u[UMZ] * rhoinv;
    const double kineng = HALF * rho * (a.u * a.u + a.v * a.v + a.w * a.w);
    const double E = u[UEDEN];
    if ((E - kineng) / E > dual_energy_eta1) a.e = (E - kineng) * rhoinv;
    else a.e = u[UEINT] * rhoinv;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) {
        xs[n] = u[UFS + n] * rhoinv;
        s1 = s1 + xs[n] * ainv(n);
    }
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) s2 = s2 + xs[n] * zion(n) * ainv(n);
    a.abar = 1.0 / s1;
    a.zbar = a.abar * s2;
    a.T0 = u[UTEMP];
}

// ctoprim's EOS call and the fields it gives
__device__ __forceinline__ void ctoprim_eos(const EosTable &tab, const PrimPre &a, Prim &q)
{
    const EosRE r = eos_re<true>(tab, a.rho, a.T0, a.e, a.abar, a.zbar);
    q.T = r.T;
    q.rhoe = a.e * a.rho;
    q.p = r.p;
    q.game = r.p / q.rhoe + 1.0;
    q.gamc = r.gam1;
    q.cs = r.cs;
    q.dpdr_e = r.dpdr_e;
    q.dpde = r.dpde;
}

__device__ __forceinline__ void ctoprim_cell(const double *u, const EosTable &tab, Prim &q)
{
    PrimPre a;
    ctoprim_pre(u, a, q.x);
    q.rho = a.rho;
    q.u = a.u;
    q.v = a.v;
    q.w = a.w;
    ctoprim_eos(tab, a, q);
}

// CFL time step of one cell (Source/timestep.F90:44-64)
__device__ __forceinline__ double estdt_cell(const double *u, double dx0, double dx1, double dx2,
                                             const EosTable &tab)
{
    const double rho = u[URHO];
    const double rhoInv = 1.0 / rho;
    double abar, zbar;
    abar_zbar_u(u, rhoInv, abar, zbar);
    const EosRE r = eos_re<true>(tab, rho, u[UTEMP], u[UEINT] * rhoInv, abar, zbar);
    const double ux = u[UMX] * rhoInv, uy = u[UMY] * rhoInv, uz = u[UMZ] * rhoInv;
    const double c = r.cs;
    const double dt1 = dx0 / (c + fabs(ux));
    const double dt2 = dx1 / (c + fabs(uy));
    const double dt3 = dx2 / (c + fabs(uz));
    return fmin(fmin(dt1, dt2), dt3);
}

// Atomic min of a double through 64-bit compare-and-swap (vector memory).
__device__ __forceinline__ void atomic_min_d(double *addr, double v)
{
    unsigned long long *a = reinterpret_cast<unsigned long long *>(addr);
    unsigned long long old = *a;
    while (v < __longlong_as_double(old)) {
        const unsigned long long assumed = old;
        old = atomicCAS(a, assumed, __double_as_longlong(v));
        if (old == assumed) break;
    }
}

// Minimum over a 1-D block of blockDim.x <= NT threads, blockDim.x a
// multiple of 64 (wavefront shuffles, then one LDS word per wave); the
// result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_min(double v)
{
    __shared__ double red[NT / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int w = 1; w < nw; ++w) v = fmin(v, red[w]);
    }
    return v;
}

} // namespace mc

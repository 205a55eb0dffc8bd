// Discrete adjoint of the fused CTU step (one instance) on gfx950.
//
// Forward pass: the fused engine advances the state with a given dt
// sequence, storing the state at the start of every step (checkpoints).
// Reverse pass, step n = N-1 .. 0: the step is recomputed from its
// checkpoint (every stage's outputs are then resident in the engine's
// arrays), and the stages are reversed in the opposite order:
//
//   A_update  cells   clean_state x2, S_new = S_old + dt src, clean(S_old)
//   A_final   faces   apply_av, final Riemann, trans2
//   A_trans1  faces   Riemann of the trans1 states, trans1
//   A_flux1   faces   Riemann of the PPM states
//   A_trace   cells+1 characteristic projection, PPM integrals, PPM edges
//   A_flat    cells+1 flattening, node-centred div(u)
//   A_prim    cells   ctoprim, clean_state x2
//
// Each thread re-evaluates its forward stencil point and scatters the input
// adjoints with hardware FP64 atomic adds (the forward reads through the
// boundary map, so the adjoint accumulates into the mapped cells).  The
// adjoint carries the hydro components of the state (rho, rho u, rho v,
// rho w, rho E, rho e); see mc_adjoint.h for why T and the species are
// constants of the differentiated map.
#include <float.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/mc_castro.h"
#include "mc_adjoint.h"
#include "mc_cell.h"
#include "mc_common.h"
#include "mc_eos.h"
#include "mc_fused.h"
#include "mc_fused_impl.h"
#include "mc_hydro.h"
#include "mc_internal.h"

namespace mc {

// adjoint component layouts
enum { SB_RHO = 0, SB_U, SB_V, SB_W, SB_P, SB_RE, NSB };            // interface states
enum { FB_RHO = 0, FB_MX, FB_MY, FB_MZ, FB_E, FB_GP, FB_GU, NFB };  // F^d, F^{t|d}
enum { NB_RHO = 0, NB_MX, NB_MY, NB_MZ, NB_E, NB_EI, NB_GP, NB_GU, NNB };  // final fluxes

struct ACtx {
    const double *Uold;  // checkpoint of the step (NVAR, cell layout)
    double *Ub;          // adjoint of the step's output state (NHB, cell layout)
    double *Ubo;         // adjoint of the step's input state (NHB, cell layout)
    double *U1b, *Qb, *FLb, *DVb;
    double *SMb[3], *SPb[3], *FDb[3], *FTb[6], *FNb[3];
};

// device-scope FP64 atomic add (the hardware global_atomic_add_f64)
__device__ __forceinline__ void aadd(gdouble *p, double v)
{
    if (v != 0.0) {
        [[clang::atomic(ignore_denormal_mode)]] {
            (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}
// plain read-add-write, for an entry no other thread of the launch touches
__device__ __forceinline__ void radd(gdouble *p, double v)
{
    if (v != 0.0) *p = *p + v;
}
// add into an adjoint entry: SOLE = this thread is the entry's only writer
// in the launch (no atomic needed)
template <bool SOLE>
__device__ __forceinline__ void badd(gdouble *p, double v)
{
    if (SOLE) radd(p, v);
    else aadd(p, v);
}
__device__ __forceinline__ gdouble *EAp(double *a, const FInst &I, int C, int c, idx_t idx)
{
    return &at(a + ((long)C * I.offE + (long)c * I.nE), idx);
}
__device__ __forceinline__ gdouble *CAp(double *a, const FInst &I, int C, int c, idx_t idx)
{
    return &at(a + ((long)C * I.offD + (long)c * I.nD), idx);
}

// conserved-state adjoint component h -> flux component in the FN array
__device__ __forceinline__ int fn_comp(int h) { return hb_comp(h); }

// ------------------------------------------------------------------ A_update
__global__ __launch_bounds__(256) void A_update(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    if (!plan_cell(p, m, i, j, k)) return;
    const FInst &I = p.inst[m];
    const idx_t c = cidx(I, i, j, k);
    const double dx0 = I.dx[0], dx1 = I.dx[1], dx2 = I.dx[2];
    const double a0 = dx1 * dx2, a1 = dx0 * dx2, a2 = dx0 * dx1;
    const double volinv = 1.0 / (dx0 * dx1 * dx2);
    const idx_t e = eidx(I, i, j, k);
    const idx_t ex = eidx(I, i + 1, j, k), ey = eidx(I, i, j + 1, k), ez = eidx(I, i, j, k + 1);
    double *F0 = x.FN[0], *F1 = x.FN[1], *F2 = x.FN[2];
    const double dt = x.dt_cur[m];
    double u0[NVAR], ua[NVAR], u1[NVAR], u2[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u0[n] = a.Uold[(long)NVAR * I.offD + (long)n * I.nD + c];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) ua[n] = u0[n];
    clean_hydro(ua, x.tab);
    const SpeciesPairs P0 = species_pairs(F0, I, NFN, FN_X), P1 = species_pairs(F1, I, NFN, FN_X),
                       P2 = species_pairs(F2, I, NFN, FN_X);
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const dbl2 f0 = SP2(P0, pr, e), f0x = SP2(P0, pr, ex), f1 = SP2(P1, pr, e), f1y = SP2(P1, pr, ey),
                   f2 = SP2(P2, pr, e), f2z = SP2(P2, pr, ez);
        const int n = UFS + 2 * pr;
        u1[n] = ua[n] + dt * flux_div(f0.x, f0x.x, f1.x, f1y.x, f2.x, f2z.x, a0, a1, a2, volinv);
        if (n + 1 < NVAR) u1[n + 1] = ua[n + 1] + dt * flux_div(f0.y, f0x.y, f1.y, f1y.y, f2.y, f2z.y, a0, a1, a2, volinv);
    }
#pragma unroll
    for (int n = 0; n < UFS; ++n) {
        double s = flux_div(EF(F0, I, NFN, n, e), EF(F0, I, NFN, n, ex), EF(F1, I, NFN, n, e), EF(F1, I, NFN, n, ey),
                            EF(F2, I, NFN, n, e), EF(F2, I, NFN, n, ez), a0, a1, a2, volinv);
        if (n == UEINT) {
            const double pdivu = HALF * (EF(F0, I, NFN, FN_GP, ex) + EF(F0, I, NFN, FN_GP, e)) *
                                     (EF(F0, I, NFN, FN_GU, ex) - EF(F0, I, NFN, FN_GU, e)) / dx0 +
                                 HALF * (EF(F1, I, NFN, FN_GP, ey) + EF(F1, I, NFN, FN_GP, e)) *
                                     (EF(F1, I, NFN, FN_GU, ey) - EF(F1, I, NFN, FN_GU, e)) / dx1 +
                                 HALF * (EF(F2, I, NFN, FN_GP, ez) + EF(F2, I, NFN, FN_GP, e)) *
                                     (EF(F2, I, NFN, FN_GU, ez) - EF(F2, I, NFN, FN_GU, e)) / dx2;
            s = s - pdivu;
        }
        u1[n] = ua[n] + dt * s;
    }
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u2[n] = u1[n];
    clean_hydro(u2, x.tab);

    double ub[NHB];
#pragma unroll
    for (int h = 0; h < NHB; ++h) ub[h] = CA(a.Ub, I, NHB, h, c);
    clean_cell_b(u2, x.tab, ub);  // post_timestep clean
    clean_cell_b(u1, x.tab, ub);  // clean after the update
    // u1 = ua + dt s
    double sb[NHB];
#pragma unroll
    for (int h = 0; h < NHB; ++h) sb[h] = dt * ub[h];
    clean_cell_b(u0, x.tab, ub);  // S_old clean
#pragma unroll
    for (int h = 0; h < NHB; ++h) CA(a.Ubo, I, NHB, h, c) = ub[h];

    const double ad[3] = {a0, a1, a2};
    const idx_t ep[3] = {ex, ey, ez};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double *FB = a.FNb[d];
#pragma unroll
        for (int h = 0; h < NHB; ++h) {
            const double g = sb[h] * ad[d] * volinv;
            aadd(EAp(FB, I, NNB, h, e), g);
            aadd(EAp(FB, I, NNB, h, ep[d]), -g);
        }
        // pdivu term of rho e: s -= 0.5 (GP+ + GP-) (GU+ - GU-) / dx_d
        double *F = x.FN[d];
        const double gpp = EF(F, I, NFN, FN_GP, ep[d]), gpm = EF(F, I, NFN, FN_GP, e);
        const double gup = EF(F, I, NFN, FN_GU, ep[d]), gum = EF(F, I, NFN, FN_GU, e);
        const double b = -sb[HB_EI] / I.dx[d];
        aadd(EAp(FB, I, NNB, NB_GP, ep[d]), b * HALF * (gup - gum));
        aadd(EAp(FB, I, NNB, NB_GP, e), b * HALF * (gup - gum));
        aadd(EAp(FB, I, NNB, NB_GU, ep[d]), b * HALF * (gpp + gpm));
        aadd(EAp(FB, I, NNB, NB_GU, e), -b * HALF * (gpp + gpm));
    }
}

// ------------------------------------------------------------------ helpers
// Riemann adjoint at face (i,j,k) of direction d; the c / gamma_c adjoints go
// to the two cells of the face.
__device__ __forceinline__ void face_riemann_hb(const Ctx &x, const ACtx &a, const FInst &I, int d, int i, int j,
                                                int k, const FaceH &L, const FaceH &R, const RiemannHB &rb,
                                                FaceHB &Lb, FaceHB &Rb)
{
    const idx_t cr = cidx(I, i, j, k);
    const idx_t cl = cidx(I, i - (d == 0), j - (d == 1), k - (d == 2));
    double cLb = 0.0, cRb = 0.0, glb = 0.0, grb = 0.0;
    riemann_hydro_b(L, R, CA(x.Q, I, NFQ, FQ_C, cl), CA(x.Q, I, NFQ, FQ_C, cr), CA(x.Q, I, NFQ, FQ_G, cl),
                    CA(x.Q, I, NFQ, FQ_G, cr), rb, Lb, Rb, cLb, cRb, glb, grb);
    aadd(CAp(a.Qb, I, NQB, QB_C, cl), cLb);
    aadd(CAp(a.Qb, I, NQB, QB_C, cr), cRb);
    aadd(CAp(a.Qb, I, NQB, QB_G, cl), glb);
    aadd(CAp(a.Qb, I, NQB, QB_G, cr), grb);
}

// FaceHB (velocities permuted to direction d) -> absolute-order state adjoint
__device__ __forceinline__ QStateHB from_faceb(const FaceHB &f, int d)
{
    QStateHB o;
    o.rho = f.rho;
    o.p = f.p;
    o.rhoe = f.rhoe;
    // to_face: un = (u,v,w)[d]; ut = d==0 ? v : u; utt = d==2 ? v : w
    double uvw[3] = {0.0, 0.0, 0.0};
    uvw[d] += f.un;
    uvw[d == 0 ? 1 : 0] += f.ut;
    uvw[d == 2 ? 1 : 2] += f.utt;
    o.u = uvw[0];
    o.v = uvw[1];
    o.w = uvw[2];
    return o;
}

template <bool SOLE = false>
__device__ __forceinline__ void add_stateb(double *S, const FInst &I, idx_t e, const QInHB &b)
{
    badd<SOLE>(EAp(S, I, NSB, SB_RHO, e), b.rho);
    badd<SOLE>(EAp(S, I, NSB, SB_U, e), b.u);
    badd<SOLE>(EAp(S, I, NSB, SB_V, e), b.v);
    badd<SOLE>(EAp(S, I, NSB, SB_W, e), b.w);
    badd<SOLE>(EAp(S, I, NSB, SB_P, e), b.p);
    badd<SOLE>(EAp(S, I, NSB, SB_RE, e), b.rhoe);
}

template <bool SOLE = false>
__device__ __forceinline__ void add_faceb(double *S, const FInst &I, idx_t e, int d, const FaceHB &f)
{
    const int un = SB_U + d, ut = d == 0 ? SB_V : SB_U, utt = d == 2 ? SB_V : SB_W;
    badd<SOLE>(EAp(S, I, NSB, SB_RHO, e), f.rho);
    badd<SOLE>(EAp(S, I, NSB, un, e), f.un);
    badd<SOLE>(EAp(S, I, NSB, ut, e), f.ut);
    badd<SOLE>(EAp(S, I, NSB, utt, e), f.utt);
    badd<SOLE>(EAp(S, I, NSB, SB_P, e), f.p);
    badd<SOLE>(EAp(S, I, NSB, SB_RE, e), f.rhoe);
}

// adjoint of a transverse-stage flux entry (URHO..UEDEN + Godunov p, u)
__device__ __forceinline__ void add_tfluxb(double *FB, const FInst &I, idx_t e, const double *vb, double gpb,
                                           double gub)
{
#pragma unroll
    for (int n = 0; n <= UEDEN; ++n) aadd(EAp(FB, I, NFB, n, e), vb[n]);
    aadd(EAp(FB, I, NFB, FB_GP, e), gpb);
    aadd(EAp(FB, I, NFB, FB_GU, e), gub);
}

// Riemann output adjoints stored in a transverse-flux adjoint entry
__device__ __forceinline__ RiemannHB load_tfluxb(double *FB, const FInst &I, idx_t e, int d)
{
    const int mn = FB_MX + d, mt = d == 0 ? FB_MY : FB_MX, mtt = d == 2 ? FB_MY : FB_MZ;
    RiemannHB r;
    r.f_rho = EA(FB, I, NFB, FB_RHO, e);
    r.f_mn = EA(FB, I, NFB, mn, e);
    r.f_mt = EA(FB, I, NFB, mt, e);
    r.f_mtt = EA(FB, I, NFB, mtt, e);
    r.f_E = EA(FB, I, NFB, FB_E, e);
    r.f_e = 0.0;
    r.g_p = EA(FB, I, NFB, FB_GP, e);
    r.g_un = EA(FB, I, NFB, FB_GU, e);
    return r;
}

// ------------------------------------------------------------------ A_final
__global__ __launch_bounds__(256) void A_final(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    int sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    const int d = sub_;
    const FInst &I = p.inst[m];
    const int c[3] = {i, j, k};
    for (int q = 0; q < 3; ++q) {
        if (c[q] < 0) return;
        if (q != d && c[q] > I.n[q] - 1) return;
    }
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int ed[3] = {d == 0, d == 1, d == 2}, e1[3] = {t1 == 0, t1 == 1, t1 == 2}, e2[3] = {t2 == 0, t2 == 1, t2 == 2};
    const double dt = x.dt_cur[m];
    const double h2 = 0.5 * dt / I.dx[t1], h3 = 0.5 * dt / I.dx[t2];
    const idx_t e = eidx(I, i, j, k);
    const int p2 = pair_index(t1, t2), p3 = pair_index(t2, t1);
    double *F2 = x.FT[p2], *F3 = x.FT[p3];
    FaceH side[2];
    idx_t l2[2], r2[2], r3[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int sh = s - 1;
        const int il = i + sh * ed[0], jl = j + sh * ed[1], kl = k + sh * ed[2];
        l2[s] = eidx(I, il, jl, kl);
        r2[s] = eidx(I, il + e1[0], jl + e1[1], kl + e1[2]);
        r3[s] = eidx(I, il + e2[0], jl + e2[1], kl + e2[2]);
        QInH lq;
        load_qin_h(s == 0 ? x.SM[d] : x.SP[d], I, e, lq);
        double f2r[UEDEN + 1], f2l[UEDEN + 1], f3r[UEDEN + 1], f3l[UEDEN + 1], pg2p, ug2p, pg2m, ug2m, pg3p, ug3p,
            pg3m, ug3m;
        load_tflux_h(F2, I, r2[s], f2r, pg2p, ug2p);
        load_tflux_h(F2, I, l2[s], f2l, pg2m, ug2m);
        load_tflux_h(F3, I, r3[s], f3r, pg3p, ug3p);
        load_tflux_h(F3, I, l2[s], f3l, pg3m, ug3m);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cidx(I, il, jl, kl));
        const QStateH o = trans2_hydro(lq, f2r, f2l, f3r, f3l, pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m, gamc,
                                       h2, h3);
        side[s] = to_face(o, d);
    }
    // artificial viscosity (forward: Source/Castro_hydro_nd.F90:174-245)
    idx_t nodes[4];
    if (d == 0) {
        nodes[0] = eidx(I, i, j, k);
        nodes[1] = eidx(I, i, j + 1, k);
        nodes[2] = eidx(I, i, j, k + 1);
        nodes[3] = eidx(I, i, j + 1, k);
    } else if (d == 1) {
        nodes[0] = eidx(I, i, j, k);
        nodes[1] = eidx(I, i + 1, j, k);
        nodes[2] = eidx(I, i, j, k + 1);
        nodes[3] = eidx(I, i + 1, j, k);
    } else {
        nodes[0] = eidx(I, i, j, k);
        nodes[1] = eidx(I, i + 1, j, k);
        nodes[2] = eidx(I, i, j + 1, k);
        nodes[3] = eidx(I, i + 1, j + 1, k);
    }
    const double draw = FOURTH * (EA(x.DV, I, 1, 0, nodes[0]) + EA(x.DV, I, 1, 0, nodes[1]) +
                                  EA(x.DV, I, 1, 0, nodes[2]) + EA(x.DV, I, 1, 0, nodes[3]));
    const double div1 = 0.1 * fmin(0.0, draw);
    const idx_t cr = cidx(I, i, j, k), cl = cidx(I, i - ed[0], j - ed[1], k - ed[2]);
    const double dxd = I.dx[d];
    double fnb[NNB];
#pragma unroll
    for (int q = 0; q < NNB; ++q) fnb[q] = EA(a.FNb[d], I, NNB, q, e);
    double div1_b = 0.0;
#pragma unroll
    for (int h = 0; h < NHB; ++h) {
        const int n = hb_comp(h);
        const double ur = CA(x.U1, I, NU1, n, cr), ul = CA(x.U1, I, NU1, n, cl);
        div1_b += fnb[h] * dxd * (ur - ul);
        const double g = fnb[h] * dxd * div1;
        aadd(CAp(a.U1b, I, NHB, h, cr), g);
        aadd(CAp(a.U1b, I, NHB, h, cl), -g);
    }
    if (draw < 0.0) {
        const double nb = 0.1 * FOURTH * div1_b;
#pragma unroll
        for (int q = 0; q < 4; ++q) aadd(EAp(a.DVb, I, 1, 0, nodes[q]), nb);
    }
    // final Riemann
    const int mn = NB_MX + d, mt = d == 0 ? NB_MY : NB_MX, mtt = d == 2 ? NB_MY : NB_MZ;
    RiemannHB rb;
    rb.f_rho = fnb[NB_RHO];
    rb.f_mn = fnb[mn];
    rb.f_mt = fnb[mt];
    rb.f_mtt = fnb[mtt];
    rb.f_E = fnb[NB_E];
    rb.f_e = fnb[NB_EI];
    rb.g_p = fnb[NB_GP];
    rb.g_un = fnb[NB_GU];
    FaceHB sb[2] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};
    face_riemann_hb(x, a, I, d, i, j, k, side[0], side[1], rb, sb[0], sb[1]);
    // trans2
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int sh = s - 1;
        const int il = i + sh * ed[0], jl = j + sh * ed[1], kl = k + sh * ed[2];
        QInH lq;
        load_qin_h(s == 0 ? x.SM[d] : x.SP[d], I, e, lq);
        double f2r[UEDEN + 1], f2l[UEDEN + 1], f3r[UEDEN + 1], f3l[UEDEN + 1], pg2p, ug2p, pg2m, ug2m, pg3p, ug3p,
            pg3m, ug3m;
        load_tflux_h(F2, I, r2[s], f2r, pg2p, ug2p);
        load_tflux_h(F2, I, l2[s], f2l, pg2m, ug2m);
        load_tflux_h(F3, I, r3[s], f3r, pg3p, ug3p);
        load_tflux_h(F3, I, l2[s], f3l, pg3m, ug3m);
        const idx_t cg = cidx(I, il, jl, kl);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cg);
        const QStateHB ob = from_faceb(sb[s], d);
        QInHB lqb = {0, 0, 0, 0, 0, 0};
        double f2rb[5] = {0, 0, 0, 0, 0}, f2lb[5] = {0, 0, 0, 0, 0}, f3rb[5] = {0, 0, 0, 0, 0},
               f3lb[5] = {0, 0, 0, 0, 0}, gb[8] = {0, 0, 0, 0, 0, 0, 0, 0}, gamcb = 0.0;
        trans2_hydro_b(lq, f2r, f2l, f3r, f3l, pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m, gamc, h2, h3, ob, lqb,
                       f2rb, f2lb, f3rb, f3lb, gb, gamcb);
        // the d-face states at e are this thread's alone in A_final
        add_stateb<true>(s == 0 ? a.SMb[d] : a.SPb[d], I, e, lqb);
        add_tfluxb(a.FTb[p2], I, r2[s], f2rb, gb[0], gb[2]);
        add_tfluxb(a.FTb[p2], I, l2[s], f2lb, gb[1], gb[3]);
        add_tfluxb(a.FTb[p3], I, r3[s], f3rb, gb[4], gb[6]);
        add_tfluxb(a.FTb[p3], I, l2[s], f3lb, gb[5], gb[7]);
        aadd(CAp(a.Qb, I, NQB, QB_G, cg), gamcb);
    }
}

// ------------------------------------------------------------------ A_trans1
__global__ __launch_bounds__(256) void A_trans1(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    int sub_;
    if (!plan_cell_sub(p, 6, sub_, m, i, j, k)) return;
    const int combo = sub_;
    const int t = c_pair_t[combo], d = c_pair_d[combo];
    const FInst &I = p.inst[m];
    const int c[3] = {i, j, k};
    if (c[t] < 0 || c[d] < 0 || c[d] > I.n[d] - 1) return;
    const idx_t e = eidx(I, i, j, k);
    const int et[3] = {t == 0, t == 1, t == 2}, ed[3] = {d == 0, d == 1, d == 2};
    const double cdtdx = x.dt_cur[m] / I.dx[d];
    const double cdt = cdtdx / 3.0;
    FaceH side[2];
    idx_t el[2], er[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int sh = s - 1;
        const int il = i + sh * et[0], jl = j + sh * et[1], kl = k + sh * et[2];
        el[s] = eidx(I, il, jl, kl);
        er[s] = eidx(I, il + ed[0], jl + ed[1], kl + ed[2]);
        QInH lq;
        load_qin_h(s == 0 ? x.SM[t] : x.SP[t], I, e, lq);
        double fr[UEDEN + 1], fl[UEDEN + 1], pgp, ugp, pgm, ugm;
        load_tflux_h(x.FD[d], I, er[s], fr, pgp, ugp);
        load_tflux_h(x.FD[d], I, el[s], fl, pgm, ugm);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cidx(I, il, jl, kl));
        const QStateH o = trans1_hydro(lq, fr, fl, pgp, pgm, ugp, ugm, gamc, cdt);
        side[s] = to_face(o, t);
    }
    const RiemannHB rb = load_tfluxb(a.FTb[combo], I, e, t);
    FaceHB sb[2] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};
    face_riemann_hb(x, a, I, t, i, j, k, side[0], side[1], rb, sb[0], sb[1]);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int sh = s - 1;
        const int il = i + sh * et[0], jl = j + sh * et[1], kl = k + sh * et[2];
        QInH lq;
        load_qin_h(s == 0 ? x.SM[t] : x.SP[t], I, e, lq);
        double fr[UEDEN + 1], fl[UEDEN + 1], pgp, ugp, pgm, ugm;
        load_tflux_h(x.FD[d], I, er[s], fr, pgp, ugp);
        load_tflux_h(x.FD[d], I, el[s], fl, pgm, ugm);
        const idx_t cg = cidx(I, il, jl, kl);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cg);
        const QStateHB ob = from_faceb(sb[s], t);
        QInHB lqb = {0, 0, 0, 0, 0, 0};
        double frb[5] = {0, 0, 0, 0, 0}, flb[5] = {0, 0, 0, 0, 0}, pgpb = 0, pgmb = 0, ugpb = 0, ugmb = 0, gamcb = 0;
        trans1_hydro_b(lq, fr, fl, pgp, pgm, ugp, ugm, gamc, cdt, ob, lqb, frb, flb, pgpb, pgmb, ugpb, ugmb, gamcb);
        add_stateb(s == 0 ? a.SMb[t] : a.SPb[t], I, e, lqb);
        add_tfluxb(a.FDb[d], I, er[s], frb, pgpb, ugpb);
        add_tfluxb(a.FDb[d], I, el[s], flb, pgmb, ugmb);
        aadd(CAp(a.Qb, I, NQB, QB_G, cg), gamcb);
    }
}

// ------------------------------------------------------------------ A_flux1
__global__ __launch_bounds__(256) void A_flux1(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    int sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    const int d = sub_;
    const FInst &I = p.inst[m];
    const int pos = d == 0 ? i : (d == 1 ? j : k);
    if (pos < 0) return;
    const idx_t e = eidx(I, i, j, k);
    FaceH L, R;
    load_fs_h(x.SM[d], I, e, d, L);
    load_fs_h(x.SP[d], I, e, d, R);
    const RiemannHB rb = load_tfluxb(a.FDb[d], I, e, d);
    FaceHB Lb = {0, 0, 0, 0, 0, 0}, Rb = {0, 0, 0, 0, 0, 0};
    face_riemann_hb(x, a, I, d, i, j, k, L, R, rb, Lb, Rb);
    // one thread per d-face: the states at e are this thread's alone here
    add_faceb<true>(a.SMb[d], I, e, d, Lb);
    add_faceb<true>(a.SPb[d], I, e, d, Rb);
}

// ------------------------------------------------------------------ A_trace
__global__ __launch_bounds__(256) void A_trace(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    int sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    const int d = sub_;
    const FInst &I = p.inst[m];
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const int pos = d == 0 ? i : (d == 1 ? j : k);
    const bool do_p = pos >= 0;
    const bool do_m = pos <= I.n[d] - 1;
    const int cun = FQ_U + d, cut = d == 0 ? FQ_V : (d == 1 ? FQ_W : FQ_U), cutt = d == 0 ? FQ_W : (d == 1 ? FQ_U : FQ_V);
    const idx_t c0 = cidx(I, i, j, k);
    const double cc = CA(x.Q, I, NFQ, FQ_C, c0);
    const double un = CA(x.Q, I, NFQ, cun, c0);
    const idx_t ec = eidx(I, i, j, k);
    const double flatn = EA(x.FL, I, 1, 0, ec);
    const double dtdx = x.dt_cur[m] / I.dx[d];
    idx_t cs[5];
#pragma unroll
    for (int s = -2; s <= 2; ++s) cs[s + 2] = cidx(I, i + s * di, j + s * dj, k + s * dk);
    const int vars[7] = {FQ_RHO, cun, FQ_P, FQ_RE, cut, cutt, FQ_G};
    const int qbv[7] = {QB_RHO, QB_U + d, QB_P, QB_RE, QB_U + (cut - FQ_U), QB_U + (cutt - FQ_U), QB_G};
    // Forward: the PPM integrals the projection reads.  Edge values are not
    // kept; the reverse sweep reads each stencil again (L2 hits) and rebuilds
    // them, which keeps the kernel within one register file per wave.
    const double spd[3] = {un - cc, un, un + cc};
    double Ip[3][4], Im[3][4];
    double gam_m, gam_p;
#pragma unroll
    for (int vv = 0; vv < 7; ++vv) {
        if (vv == 4 || vv == 5) continue;
        double s5[5], sm, sp;
#pragma unroll
        for (int q = 0; q < 5; ++q) s5[q] = CA(x.Q, I, NFQ, vars[vv], cs[q]);
        ppm_reconstruct(s5, flatn, sm, sp);
        const double s6 = ppm_s6(sm, sp, s5[2]);
        if (vv < 4) {
#pragma unroll
            for (int w = 0; w < 3; ++w) ppm_int_wave(sm, sp, s6, spd[w], dtdx, Ip[w][vv], Im[w][vv]);
        } else {
            double dummy;
            ppm_int_wave(sm, sp, s6, un - cc, dtdx, dummy, gam_m);
            ppm_int_wave(sm, sp, s6, un + cc, dtdx, gam_p, dummy);
        }
    }

    const idx_t ep = ec, em = eidx(I, i + di, j + dj, k + dk);
    const int sun = SB_U + d, sut = d == 0 ? SB_V : (d == 1 ? SB_W : SB_U), sutt = d == 0 ? SB_W : (d == 1 ? SB_U : SB_V);
    double Ipb[3][4] = {{0}}, Imb[3][4] = {{0}};
    double gam_mb = 0.0, gam_pb = 0.0, Ip_utb = 0.0, Im_utb = 0.0, Ip_uttb = 0.0, Im_uttb = 0.0;
    if (do_p) {
        double *SB = a.SPb[d];
        ppm_project_b<true>(Im, gam_m, un, cc, EA(SB, I, NSB, SB_RHO, ep), EA(SB, I, NSB, sun, ep),
                            EA(SB, I, NSB, SB_RE, ep), EA(SB, I, NSB, SB_P, ep), Imb, gam_mb);
        Im_utb = EA(SB, I, NSB, sut, ep);
        Im_uttb = EA(SB, I, NSB, sutt, ep);
    }
    if (do_m) {
        double *SB = a.SMb[d];
        ppm_project_b<false>(Ip, gam_p, un, cc, EA(SB, I, NSB, SB_RHO, em), EA(SB, I, NSB, sun, em),
                             EA(SB, I, NSB, SB_RE, em), EA(SB, I, NSB, SB_P, em), Ipb, gam_pb);
        Ip_utb = EA(SB, I, NSB, sut, em);
        Ip_uttb = EA(SB, I, NSB, sutt, em);
    }
    double unb = 0.0, ccb = 0.0, flatb = 0.0;
#pragma unroll
    for (int vv = 0; vv < 7; ++vv) {
        double s5[5], SMv, SPv;
#pragma unroll
        for (int q = 0; q < 5; ++q) s5[q] = CA(x.Q, I, NFQ, vars[vv], cs[q]);
        ppm_reconstruct(s5, flatn, SMv, SPv);
        const double S6 = ppm_s6(SMv, SPv, s5[2]);
        double smb = 0.0, spb = 0.0, s6b = 0.0;
        if (vv < 4) {
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                double spdb = 0.0;
                ppm_int_wave_b(SMv, SPv, S6, spd[w], dtdx, Ipb[w][vv], Imb[w][vv], smb, spb, s6b, spdb);
                unb += spdb;
                if (w == 0) ccb -= spdb;
                if (w == 2) ccb += spdb;
            }
        } else if (vv < 6) {
            double spdb = 0.0;
            ppm_int_wave_b(SMv, SPv, S6, un, dtdx, vv == 4 ? Ip_utb : Ip_uttb, vv == 4 ? Im_utb : Im_uttb,
                           smb, spb, s6b, spdb);
            unb += spdb;
        } else {
            double spdb = 0.0;
            ppm_int_wave_b(SMv, SPv, S6, un - cc, dtdx, 0.0, gam_mb, smb, spb, s6b, spdb);
            unb += spdb;
            ccb -= spdb;
            spdb = 0.0;
            ppm_int_wave_b(SMv, SPv, S6, un + cc, dtdx, gam_pb, 0.0, smb, spb, s6b, spdb);
            unb += spdb;
            ccb += spdb;
        }
        // s6 = 6 s0 - 3 (sm + sp)
        double s5b[5] = {0, 0, 0, 0, 0};
        s5b[2] += 6.0 * s6b;
        smb -= 3.0 * s6b;
        spb -= 3.0 * s6b;
        ppm_reconstruct_b(s5, flatn, smb, spb, s5b, flatb);
#pragma unroll
        for (int q = 0; q < 5; ++q) aadd(CAp(a.Qb, I, NQB, qbv[vv], cs[q]), s5b[q]);
    }
    aadd(CAp(a.Qb, I, NQB, QB_U + d, c0), unb);
    aadd(CAp(a.Qb, I, NQB, QB_C, c0), ccb);
    aadd(EAp(a.FLb, I, 1, 0, ec), flatb);
}

// ------------------------------------------------------------------ A_flat
__global__ __launch_bounds__(256) void A_flat(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    if (!plan_cell(p, m, i, j, k)) return;
    const FInst &I = p.inst[m];
    const idx_t e = eidx(I, i, j, k);
    const double fb = EA(a.FLb, I, 1, 0, e);
    if (fb != 0.0) {
        double cur = 1.0;
        int arg = -1;
        for (int d = 0; d < 3; ++d) {
            const int ei = (d == 0), ej = (d == 1), ek = (d == 2);
            double p7[7], u7[7];
#pragma unroll
            for (int s = -3; s <= 3; ++s) {
                const idx_t c = cidx(I, i + s * ei, j + s * ej, k + s * ek);
                p7[s + 3] = CA(x.Q, I, NFQ, FQ_P, c);
                u7[s + 3] = CA(x.Q, I, NFQ, FQ_U + d, c);
            }
            const double f = flatten_dir(p7, u7);
            if (f < cur) {
                cur = f;
                arg = d;
            }
        }
        if (arg >= 0) {
            const int ei = (arg == 0), ej = (arg == 1), ek = (arg == 2);
            double p7[7], u7[7], p7b[7] = {0, 0, 0, 0, 0, 0, 0};
            idx_t cc[7];
#pragma unroll
            for (int s = -3; s <= 3; ++s) {
                cc[s + 3] = cidx(I, i + s * ei, j + s * ej, k + s * ek);
                p7[s + 3] = CA(x.Q, I, NFQ, FQ_P, cc[s + 3]);
                u7[s + 3] = CA_lane(x.Q, I, NFQ, FQ_U + arg, cc[s + 3]);  // arg differs per lane
            }
            flatten_dir_b(p7, u7, fb, p7b);
#pragma unroll
            for (int s = 0; s < 7; ++s) aadd(CAp(a.Qb, I, NQB, QB_P, cc[s]), p7b[s]);
        }
    }
    if (i >= 0 && j >= 0 && k >= 0) {
        const double db = EA(a.DVb, I, 1, 0, e);
        if (db != 0.0) {
            const double gx = FOURTH * db / I.dx[0], gy = FOURTH * db / I.dx[1], gz = FOURTH * db / I.dx[2];
#define QB(comp, a_, b_, c_) CAp(a.Qb, I, NQB, comp, cidx(I, a_, b_, c_))
            aadd(QB(QB_U, i, j, k), gx);
            aadd(QB(QB_U, i - 1, j, k), -gx);
            aadd(QB(QB_U, i, j, k - 1), gx);
            aadd(QB(QB_U, i - 1, j, k - 1), -gx);
            aadd(QB(QB_U, i, j - 1, k), gx);
            aadd(QB(QB_U, i - 1, j - 1, k), -gx);
            aadd(QB(QB_U, i, j - 1, k - 1), gx);
            aadd(QB(QB_U, i - 1, j - 1, k - 1), -gx);
            aadd(QB(QB_V, i, j, k), gy);
            aadd(QB(QB_V, i, j - 1, k), -gy);
            aadd(QB(QB_V, i, j, k - 1), gy);
            aadd(QB(QB_V, i, j - 1, k - 1), -gy);
            aadd(QB(QB_V, i - 1, j, k), gy);
            aadd(QB(QB_V, i - 1, j - 1, k), -gy);
            aadd(QB(QB_V, i - 1, j, k - 1), gy);
            aadd(QB(QB_V, i - 1, j - 1, k - 1), -gy);
            aadd(QB(QB_W, i, j, k), gz);
            aadd(QB(QB_W, i, j, k - 1), -gz);
            aadd(QB(QB_W, i, j - 1, k), gz);
            aadd(QB(QB_W, i, j - 1, k - 1), -gz);
            aadd(QB(QB_W, i - 1, j, k), gz);
            aadd(QB(QB_W, i - 1, j, k - 1), -gz);
            aadd(QB(QB_W, i - 1, j - 1, k), gz);
            aadd(QB(QB_W, i - 1, j - 1, k - 1), -gz);
#undef QB
        }
    }
}

// ------------------------------------------------------------------ A_prim
__global__ __launch_bounds__(256) void A_prim(Plan p, Ctx x, ACtx a)
{
    int m, i, j, k;
    if (!plan_cell(p, m, i, j, k)) return;
    const FInst &I = p.inst[m];
    const idx_t c = cidx(I, i, j, k);
    double u0[NVAR], ua[NVAR], u1[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u0[n] = a.Uold[(long)NVAR * I.offD + (long)n * I.nD + c];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) ua[n] = u0[n];
    clean_hydro(ua, x.tab);
#pragma unroll
    for (int n = 0; n < NU1; ++n) u1[n] = CA(x.U1, I, NU1, n, c);  // clean(clean(U)) of the forward
    const SpeciesPairs XU = cell_species(x.U1X, I);
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const dbl2 v = SP2(XU, pr, c);
        u1[UFS + 2 * pr] = v.x;
        if (2 * pr + 1 < NSPEC) u1[UFS + 2 * pr + 1] = v.y;
    }
    double qb[NQB], ub[NHB];
#pragma unroll
    for (int q = 0; q < NQB; ++q) qb[q] = CA(a.Qb, I, NQB, q, c);
#pragma unroll
    for (int h = 0; h < NHB; ++h) ub[h] = CA(a.U1b, I, NHB, h, c);
    ctoprim_b(u1, x.tab, qb, ub);
    clean_cell_b(ua, x.tab, ub);
    clean_cell_b(u0, x.tab, ub);
#pragma unroll
    for (int h = 0; h < NHB; ++h) CA(a.Ubo, I, NHB, h, c) += ub[h];
}

// seed: hydro components of a full NVAR adjoint (cell layout)
__global__ void A_seed(const double *src, double *Ub, long nD)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nD) return;
#pragma unroll
    for (int h = 0; h < NHB; ++h) Ub[(long)h * nD + t] = src[(long)hb_comp(h) * nD + t];
}

__global__ void A_unseed(const double *Ub, double *dst, long nD)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nD) return;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) dst[(long)n * nD + t] = 0.0;
#pragma unroll
    for (int h = 0; h < NHB; ++h) dst[(long)hb_comp(h) * nD + t] = Ub[(long)h * nD + t];
}

} // namespace mc

using namespace mc;

struct mc_adjoint {
    FusedLevel *F = nullptr;
    int n[3];
    long nD = 0, nE = 0;
    int max_steps = 0;
    int nsteps = 0;
    double *ckpt = nullptr;       // max_steps x NVAR x nD
    std::vector<double> dts;
    double *abuf = nullptr;       // adjoint intermediates (cleared every step)
    long abuf_n = 0;
    double *ub0 = nullptr, *ub1 = nullptr;  // NHB x nD each
    ACtx a;
    // Kept intermediates: step n of the forward pass writes its primitives,
    // traces and fluxes into sets[n], and the reverse pass reads them there
    // instead of recomputing the step.  Empty: recompute mode.
    std::vector<FusedArrays> sets;
    FusedArrays own;  // the engine's own set
};

static void adjoint_drop_sets(mc_adjoint *A)
{
    fused_set_intermediates(A->F, A->own);
    for (auto &S : A->sets) fused_free_intermediates(S);
    A->sets.clear();
    A->nsteps = 0;  // a kept forward pass is gone
}

// One set per step when max_steps sets fit in half of the free device memory.
static int adjoint_keep_sets(mc_adjoint *A, bool report)
{
    if (!A->sets.empty()) return 0;
    size_t free_b = 0, total_b = 0;
    if (check_hip(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) return -1;
    const size_t need = fused_intermediate_bytes(A->F) * (size_t)A->max_steps;
    if (need > free_b / 2) {
        if (report) set_error("mc_adjoint: the intermediates of max_steps steps do not fit in half the free device memory");
        return -1;
    }
    A->sets.resize(A->max_steps);
    for (auto &S : A->sets)
        if (fused_alloc_intermediates(A->F, S)) {
            adjoint_drop_sets(A);
            return -1;
        }
    A->nsteps = 0;
    return 0;
}

static void launch_adj(mc_adjoint *A, int stage, hipStream_t s)
{
    FusedLevel *F = A->F;
    Plan p = get_plan(F, stage);
    const int nb = plan_blocks(F, p);
    const int nt = p.bx * p.by * p.bz;
    if (nb <= 0) return;
    static const int ny[FST_N] = {1, 1, 1, 3, 3, 6, 3, 1};
    // face stages: directions / pairs interleaved brick by brick, as forward
    const bool interleave = ny[stage] > 1;
    dim3 g(interleave ? nb * ny[stage] : nb, interleave ? 1 : ny[stage]), b(nt);
    F->ctx.tab = eos_table();
    switch (stage) {
    case FST_PRIM: hipLaunchKernelGGL(A_prim, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_FLAT: hipLaunchKernelGGL(A_flat, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_TRACE: hipLaunchKernelGGL(A_trace, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_FLUX1: hipLaunchKernelGGL(A_flux1, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_TRANS1: hipLaunchKernelGGL(A_trans1, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_FINAL: hipLaunchKernelGGL(A_final, g, b, 0, s, p, F->ctx, A->a); break;
    case FST_UPDATE: hipLaunchKernelGGL(A_update, g, b, 0, s, p, F->ctx, A->a); break;
    }
}

extern "C" {

void mc_adjoint_destroy(mc_adjoint_t *A);

mc_adjoint_t *mc_adjoint_create(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc, int max_steps)
{
    if (!n_cell || !prob_lo || !prob_hi || max_steps < 1 || (bc != 0 && bc != 1)) {
        set_error("mc_adjoint_create: bad arguments");
        return nullptr;
    }
    for (int d = 0; d < 3; ++d)
        if (n_cell[d] < 4 || !(prob_hi[d] > prob_lo[d])) {
            set_error("mc_adjoint_create: bad domain");
            return nullptr;
        }
    mc_adjoint *A = new mc_adjoint();
    FusedInstanceSpec sp;
    for (int d = 0; d < 3; ++d) {
        sp.n[d] = n_cell[d];
        sp.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        A->n[d] = n_cell[d];
    }
    sp.bc = bc;
    A->F = fused_create_set(std::vector<FusedInstanceSpec>{sp});
    if (!A->F) {
        delete A;
        return nullptr;
    }
    A->nD = (long)n_cell[0] * n_cell[1] * n_cell[2];
    A->nE = fused_next(A->F, 0);  // the reverse kernels index with the engine's stride
    A->max_steps = max_steps;
    // intermediates: U1b, Qb (cells), FLb, DVb, SMb, SPb, FDb, FTb, FNb (extended)
    const long nD = A->nD, nE = A->nE;
    A->abuf_n = (long)(NHB + NQB) * nD + (2 + 6 * NSB + 3 * NFB + 6 * NFB + 3 * NNB) * nE;
    bool ok = !check_hip(hipMalloc(&A->ckpt, sizeof(double) * NVAR * nD * (size_t)max_steps), "hipMalloc(ckpt)") &&
              !check_hip(hipMalloc(&A->abuf, sizeof(double) * A->abuf_n), "hipMalloc(adjoint)") &&
              !check_hip(hipMalloc(&A->ub0, sizeof(double) * NHB * nD), "hipMalloc(adjoint)") &&
              !check_hip(hipMalloc(&A->ub1, sizeof(double) * NHB * nD), "hipMalloc(adjoint)");
    if (!ok) {
        mc_adjoint_destroy(A);
        return nullptr;
    }
    A->own = fused_get_intermediates(A->F);
    (void)adjoint_keep_sets(A, false);  // else recompute mode
    double *q = A->abuf;
    ACtx &a = A->a;
    a.U1b = q;
    q += NHB * nD;
    a.Qb = q;
    q += NQB * nD;
    a.FLb = q;
    q += nE;
    a.DVb = q;
    q += nE;
    for (int d = 0; d < 3; ++d) {
        a.SMb[d] = q;
        q += NSB * nE;
        a.SPb[d] = q;
        q += NSB * nE;
        a.FDb[d] = q;
        q += NFB * nE;
        a.FNb[d] = q;
        q += NNB * nE;
    }
    for (int c = 0; c < 6; ++c) {
        a.FTb[c] = q;
        q += NFB * nE;
    }
    return A;
}

void mc_adjoint_destroy(mc_adjoint_t *A)
{
    if (!A) return;
    if (A->F) {
        (void)hipDeviceSynchronize();
        adjoint_drop_sets(A);
        fused_destroy(A->F);
    }
    (void)hipFree(A->ckpt);
    (void)hipFree(A->abuf);
    (void)hipFree(A->ub0);
    (void)hipFree(A->ub1);
    delete A;
}

// Forward pass: U0 (device, NVAR x cells, Fortran order) advanced nsteps
// with checkpoints; the final state goes to U_out (device).  compute_dt != 0:
// dts[n] receive the CFL-limited sequence dt_0 = cfl * estdt(U0),
// dt_{n+1} = min(cfl * estdt(U_{n+1}), 1.1 dt_n); otherwise dts are used.
int mc_adjoint_forward(mc_adjoint_t *A, const double *U0, double *U_out, int nsteps, double *dts, int compute_dt)
{
    if (!A || !U0 || !U_out || !dts || nsteps < 1 || nsteps > A->max_steps) {
        set_error("mc_adjoint_forward: bad arguments");
        return -1;
    }
    if (!eos_ready()) {
        set_error("mc_adjoint_forward: EOS table not loaded");
        return -1;
    }
    hipStream_t s = cur_stream();
    FusedLevel *F = A->F;
    const size_t bytes = sizeof(double) * NVAR * A->nD;
    double *U = fused_state(F, 0);
    if (check_hip(hipMemcpyAsync(U, U0, bytes, hipMemcpyDeviceToDevice, s), "adjoint forward")) return -1;
    A->dts.assign(nsteps, 0.0);
    double est = 0.0;
    if (compute_dt) {
        if (fused_estimate(F, s) || fused_read_est(F, 0, &est, s)) return -1;
    }
    for (int n = 0; n < nsteps; ++n) {
        double dt;
        if (compute_dt) {
            dt = n == 0 ? cfl * est : fmin(cfl * est, 1.1 * A->dts[n - 1]);
            dts[n] = dt;
        } else dt = dts[n];
        A->dts[n] = dt;
        if (check_hip(hipMemcpyAsync(A->ckpt + (size_t)n * NVAR * A->nD, U, bytes, hipMemcpyDeviceToDevice, s),
                      "checkpoint"))
            return -1;
        if (!A->sets.empty()) fused_set_intermediates(F, A->sets[n]);
        if (fused_advance_dt(F, dt, s)) return -1;
        if (compute_dt && fused_read_est(F, 0, &est, s)) return -1;
    }
    if (!A->sets.empty()) fused_set_intermediates(F, A->own);
    A->nsteps = nsteps;
    if (check_hip(hipMemcpyAsync(U_out, U, bytes, hipMemcpyDeviceToDevice, s), "adjoint forward")) return -1;
    return check_hip(hipStreamSynchronize(s), "adjoint forward");
}

// Reverse pass of the last forward pass: Ub_final (device, NVAR x cells) is
// the adjoint of its final state; Ub0 (device, NVAR x cells) receives the
// adjoint of its initial state (hydro components; UTEMP and species zero).
int mc_adjoint_backward(mc_adjoint_t *A, const double *Ub_final, double *Ub0)
{
    if (!A || !Ub_final || !Ub0 || A->nsteps < 1) {
        set_error("mc_adjoint_backward: bad arguments or no forward pass");
        return -1;
    }
    hipStream_t s = cur_stream();
    FusedLevel *F = A->F;
    const long nD = A->nD;
    const size_t bytes = sizeof(double) * NVAR * nD;
    const unsigned gb = (unsigned)((nD + 255) / 256);
    hipLaunchKernelGGL(A_seed, dim3(gb), dim3(256), 0, s, Ub_final, A->ub0, nD);
    double *cur = A->ub0, *nxt = A->ub1;
    double *U = fused_state(F, 0);
    for (int n = A->nsteps - 1; n >= 0; --n) {
        const double *ck = A->ckpt + (size_t)n * NVAR * nD;
        if (!A->sets.empty()) {
            // the step's intermediates as the forward pass left them
            fused_set_intermediates(F, A->sets[n]);
            fill(F->ctx.dt_cur, 1, A->dts[n], s);
        } else {
            if (check_hip(hipMemcpyAsync(U, ck, bytes, hipMemcpyDeviceToDevice, s), "recompute")) return -1;
            if (fused_recompute_dt(F, A->dts[n], s)) return -1;  // recompute the step's intermediates
        }
        fill(A->abuf, A->abuf_n, 0.0, s);
        A->a.Uold = ck;
        A->a.Ub = cur;
        A->a.Ubo = nxt;
        launch_adj(A, FST_UPDATE, s);
        launch_adj(A, FST_FINAL, s);
        launch_adj(A, FST_TRANS1, s);
        launch_adj(A, FST_FLUX1, s);
        launch_adj(A, FST_TRACE, s);
        launch_adj(A, FST_FLAT, s);
        launch_adj(A, FST_PRIM, s);
        if (check_hip(hipGetLastError(), "adjoint step")) return -1;
        std::swap(cur, nxt);
    }
    if (!A->sets.empty()) fused_set_intermediates(F, A->own);
    hipLaunchKernelGGL(A_unseed, dim3(gb), dim3(256), 0, s, cur, Ub0, nD);
    return check_hip(hipStreamSynchronize(s), "adjoint backward");
}

// keep != 0: keep every step's intermediates (allocated now; fails if
// max_steps sets do not fit in half the free device memory); 0: recompute
// each step from its checkpoint.  Either change drops the last forward pass.
int mc_adjoint_keep_intermediates(mc_adjoint_t *A, int keep)
{
    if (!A) {
        set_error("mc_adjoint_keep_intermediates: null handle");
        return -1;
    }
    (void)hipDeviceSynchronize();
    if (!keep) {
        adjoint_drop_sets(A);
        return 0;
    }
    return adjoint_keep_sets(A, true);
}

int mc_adjoint_keeps_intermediates(const mc_adjoint_t *A) { return A && !A->sets.empty() ? 1 : 0; }

int mc_adjoint_steps(const mc_adjoint_t *A) { return A ? A->nsteps : -1; }

} // extern "C"

// Batched fused CTU step for gfx950.
//
// All boxes of a level -- and all members of an ensemble of independent
// Sedov instances -- advance together, one launch per stage:
//
//   K_prim    cells   clean_state(S_old) (stored in place), clean_state(Sborder),
//                     ctoprim
//   K_flat    cells+1 shock flattening coefficient, node-centred div(u)
//   K_trace   cells+1 PPM edge states in x, y, z
//   K_flux1   faces   CGF fluxes of the PPM states           F^d
//   K_trans1  faces   one-direction transverse update + flux  F^{t|d}
//   K_final   faces   two-direction transverse update + final flux,
//                     artificial viscosity, species flux normalisation
//   K_update  cells   flux divergence and p dV source, S_new = S_old + dt
//                     src, clean_state twice, estdt min-reduction
//   K_dt      1 block the CFL / change_max / init_shrink time-step control
//
// Why the result equals the per-box path bit for bit: every stage applies the
// same device functions (mc_cell.h, mc_hydro.h, mc_eos.h) to the same
// inputs.  The reference's Sborder ghost cells hold clean(S_old) of the
// source cell they are copied from, so every per-cell quantity (cleaned
// state, primitives, sound speed) is a function of the source cell alone: it
// is computed once per domain cell here and read through the boundary map
// (periodic wrap or first-order-extrapolation clamp), where the per-box path
// recomputes it for every ghost cell of every box.  Face quantities are
// computed once per face instead of once per box that touches the face.
// Quantities the reference evaluates on ghost cells of the domain (traces
// and fluxes up to one cell outside) are evaluated on a one-cell ring
// around each instance ("extended" arrays).
//
// The directions (trace, flux1, final) or (t, d) pairs (trans1) of a face
// stage are sub-launches of one launch, interleaved brick by brick
// (plan_cell_sub), so the sub-launches of a brick share its XCD's L2.
//
// Layout: per instance, structure-of-arrays; a cell array with C components
// holds component c of cell (i,j,k) at C*off + c*ncell + i + n0*(j + n1*k);
// extended arrays cover [-1, n] in each direction.  The face arrays hold
// their components in interleaved pairs, and the species of q and of the
// cleaned state sit apart in arrays of pairs (EF, EF2, SP2 in
// mc_fused_impl.h), so one 16-byte access moves two components.  dt lives on
// the device (one slot per instance), so a step can be captured in a hipGraph.
#include <float.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mc_cell.h"
#include "mc_common.h"
#include "mc_eos.h"
#include "mc_fused.h"
#include "mc_fused_impl.h"
#include "mc_hydro.h"
#include "mc_internal.h"

namespace mc {

// ------------------------------------------------------------------ stages
// estdt on the current state (Castro::estTimeStep before the first step)
__global__ __launch_bounds__(256) void K_est(Plan p, Ctx x)
{
    int m, i, j, k;
    double v = DBL_MAX;
    const bool ok = plan_cell(p, m, i, j, k);
    const FInst &I = p.inst[m];
    if (ok) {
        const idx_t c = cidx(I, i, j, k);
        double u[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) u[n] = CA(x.U, I, NVAR, n, c);
        v = estdt_cell(u, I.dx[0], I.dx[1], I.dx[2], x.tab);
    }
    v = block_min<256>(v);
    if (threadIdx.x == 0) atomic_min_d(&x.dt_est[m], v);
}

// S_old clean (Castro_advance.cpp:51), Sborder clean (:66), ctoprim.
// Held to 2 waves per SIMD like K_update (3.85 -> 3.65 ms at 256^3).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void K_prim(Plan p, Ctx x)
{
    int m, i, j, k;
    if (!plan_cell(p, m, i, j, k)) return;
    const FInst &I = p.inst[m];
    const idx_t c = cidx(I, i, j, k);
    double u[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) u[n] = CA(x.U, I, NVAR, n, c);
    clean_cell(u, x.tab); /* S_old clean, Castro_advance.cpp:51 */
    // S_old is replaced by its clean in place: K_update starts from it
    // instead of cleaning S_old a second time (nothing between reads U)
#pragma unroll
    for (int n = 0; n < NVAR; ++n) CA(x.U, I, NVAR, n, c) = u[n];
    clean_cell(u, x.tab); /* Sborder clean, Castro_advance.cpp:66 */
#pragma unroll
    for (int n = 0; n < NU1; ++n) CA(x.U1, I, NU1, n, c) = u[n];
    {
        const SpeciesPairs XU = cell_species(x.U1X, I);
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr)
            SP2(XU, pr, c) = dbl2{u[UFS + 2 * pr], 2 * pr + 1 < NSPEC ? u[UFS + 2 * pr + 1] : 0.0};
    }
    // the EOS-free primitives are stored before the EOS call, so they are
    // not held in registers through it
    PrimPre a;
    {
        double xs[NSPEC];
        ctoprim_pre(u, a, xs);
        const SpeciesPairs XQ = cell_species(x.QX, I);
        CA(x.Q, I, NFQ, FQ_RHO, c) = a.rho;
        CA(x.Q, I, NFQ, FQ_U, c) = a.u;
        CA(x.Q, I, NFQ, FQ_V, c) = a.v;
        CA(x.Q, I, NFQ, FQ_W, c) = a.w;
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr)
            SP2(XQ, pr, c) = dbl2{xs[2 * pr], 2 * pr + 1 < NSPEC ? xs[2 * pr + 1] : 0.0};
    }
    Prim q;
    ctoprim_eos(x.tab, a, q);
    CA(x.Q, I, NFQ, FQ_P, c) = q.p;
    CA(x.Q, I, NFQ, FQ_RE, c) = q.rhoe;
    CA(x.Q, I, NFQ, FQ_C, c) = q.cs;
    CA(x.Q, I, NFQ, FQ_G, c) = q.gamc;
}

__device__ __forceinline__ double qv(const Ctx &x, const FInst &I, int comp, int i, int j, int k)
{
    return CA(x.Q, I, NFQ, comp, cidx(I, i, j, k));
}

// uflatten (Source/ppm.F90:11-128) and divu (Source/Castro_hydro_nd.F90:11-69)
__global__ __launch_bounds__(256) void K_flat(Plan p, Ctx x)
{
    int m, i, j, k;
    if (!plan_cell(p, m, i, j, k)) return;
    const FInst &I = p.inst[m];
    const idx_t e = eidx(I, i, j, k);
    double flatn = 1.0;
    for (int d = 0; d < 3; ++d) {
        const int ei = (d == 0), ej = (d == 1), ek = (d == 2);
        double p7[7], u7[7];
#pragma unroll
        for (int s = -3; s <= 3; ++s) {
            const idx_t c = cidx(I, i + s * ei, j + s * ej, k + s * ek);
            p7[s + 3] = CA(x.Q, I, NFQ, FQ_P, c);
            u7[s + 3] = CA(x.Q, I, NFQ, FQ_U + d, c);
        }
        flatn = fmin(flatn, flatten_dir(p7, u7));
    }
    EA(x.FL, I, 1, 0, e) = flatn;
    if (i >= 0 && j >= 0 && k >= 0) {
        const double dxinv = 1.0 / I.dx[0], dyinv = 1.0 / I.dx[1], dzinv = 1.0 / I.dx[2];
#define QQ(c, a, b, cc) qv(x, I, c, a, b, cc)
        const double ux = FOURTH *
                          (+QQ(FQ_U, i, j, k) - QQ(FQ_U, i - 1, j, k) + QQ(FQ_U, i, j, k - 1) -
                           QQ(FQ_U, i - 1, j, k - 1) + QQ(FQ_U, i, j - 1, k) - QQ(FQ_U, i - 1, j - 1, k) +
                           QQ(FQ_U, i, j - 1, k - 1) - QQ(FQ_U, i - 1, j - 1, k - 1)) *
                          dxinv;
        const double vy = FOURTH *
                          (+QQ(FQ_V, i, j, k) - QQ(FQ_V, i, j - 1, k) + QQ(FQ_V, i, j, k - 1) -
                           QQ(FQ_V, i, j - 1, k - 1) + QQ(FQ_V, i - 1, j, k) - QQ(FQ_V, i - 1, j - 1, k) +
                           QQ(FQ_V, i - 1, j, k - 1) - QQ(FQ_V, i - 1, j - 1, k - 1)) *
                          dyinv;
        const double wz = FOURTH *
                          (+QQ(FQ_W, i, j, k) - QQ(FQ_W, i, j, k - 1) + QQ(FQ_W, i, j - 1, k) -
                           QQ(FQ_W, i, j - 1, k - 1) + QQ(FQ_W, i - 1, j, k) - QQ(FQ_W, i - 1, j, k - 1) +
                           QQ(FQ_W, i - 1, j - 1, k) - QQ(FQ_W, i - 1, j - 1, k - 1)) *
                          dzinv;
#undef QQ
        EA(x.DV, I, 1, 0, e) = ux + vy + wz;
    }
}

// What the species part of a cell's trace takes from its hydro part: the
// stencil, the wave speed and dt/dx, which states the cell stores, and the
// species pair bases (taken before any store, see SpeciesPairs).
struct TraceCtx {
    idx_t cs[5], ep, em;
    double flatn, un, dtdx;
    bool do_p, do_m;
    SpeciesPairs XP, XM, XQ;
};

// hydro part of a traced state, velocities in absolute order
struct StateH {
    double rho, u, v, w, p, rhoe;
};

// Doubles per component of K_tflux's LDS tile of the primitives: a brick
// of n threads with b cells along its direction stages (b + 4) n / b cells
// (the stencil's two cells either side), at most QTILE.
constexpr int QTILE = 448;

// PPM reconstruction + characteristic tracing (Source/ppm.F90:309-743) of
// cell (i,j,k) of instance m in direction d, hydro part: the plus state of
// face pos (stored when own_p and pos >= 0) and the minus state of face
// pos + 1, both also returned as the Riemann solver of direction d takes
// them (fp, fm: the values load_fs_h reads back from the stored states).
// STAGED: the primitives of cells pos-2 .. pos+2 come from the block's LDS
// tile (component c of cell pos-2+q at tile[c * QTILE + self + q * sr])
// instead of the cell arrays.
template <bool STAGED>
__device__ __forceinline__ void trace_hydro(const Ctx &x, int m, const FInst &I, int d, int i, int j, int k,
                                            bool own_p, TraceCtx &t, FaceH &fp, FaceH &fm,
                                            const double *tile = nullptr, int self = 0, int sr = 0)
{
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const int pos = d == 0 ? i : (d == 1 ? j : k);
    t.do_p = own_p && pos >= 0;
    // the minus state of face pos + 1; in z it must lie in the slab's faces
    t.do_m = pos <= I.n[d] - 1 && (d != 2 || k + 1 <= I.zlo + I.zn);
    const int cun = FQ_U + d, cut = d == 0 ? FQ_V : (d == 1 ? FQ_W : FQ_U), cutt = d == 0 ? FQ_W : (d == 1 ? FQ_U : FQ_V);
#pragma unroll
    for (int s = -2; s <= 2; ++s) t.cs[s + 2] = cidx(I, i + s * di, j + s * dj, k + s * dk);
    // component c of the primitives of cell pos - 2 + q
    auto qs = [&](int c, int q) -> double {
        if (STAGED) return tile[c * QTILE + self + q * sr];
        return CA(x.Q, I, NFQ, c, t.cs[q]);
    };
    const double cc = qs(FQ_C, 2);
    const double un = qs(cun, 2);
    const double flatn = EA(x.FL, I, 1, 0, eidx(I, i, j, k));
    const double dtdx = x.dt_cur[m] / I.dx[d];
    t.un = un;
    t.flatn = flatn;
    t.dtdx = dtdx;

    double Ip[3][4], Im[3][4], Ip_ut = 0, Im_ut = 0, Ip_utt = 0, Im_utt = 0;
    const int vars[6] = {FQ_RHO, cun, FQ_P, FQ_RE, cut, cutt};
    for (int vv = 0; vv < 6; ++vv) {
        double s[5], sm_, sp_;
#pragma unroll
        for (int q = 0; q < 5; ++q) s[q] = qs(vars[vv], q);
        ppm_reconstruct(s, flatn, sm_, sp_);
        const double s6 = ppm_s6(sm_, sp_, s[2]);
        if (vv < 4) {
            ppm_int_wave(sm_, sp_, s6, un - cc, dtdx, Ip[0][vv], Im[0][vv]);
            ppm_int_wave(sm_, sp_, s6, un, dtdx, Ip[1][vv], Im[1][vv]);
            ppm_int_wave(sm_, sp_, s6, un + cc, dtdx, Ip[2][vv], Im[2][vv]);
        } else if (vv == 4) {
            ppm_int_wave(sm_, sp_, s6, un, dtdx, Ip_ut, Im_ut);
        } else {
            ppm_int_wave(sm_, sp_, s6, un, dtdx, Ip_utt, Im_utt);
        }
    }
    double gam_m, gam_p;
    {
        double s[5], sm_, sp_, a, b;
#pragma unroll
        for (int q = 0; q < 5; ++q) s[q] = qs(FQ_G, q);
        ppm_reconstruct(s, flatn, sm_, sp_);
        const double s6 = ppm_s6(sm_, sp_, s[2]);
        ppm_int_wave(sm_, sp_, s6, un - cc, dtdx, a, gam_m);
        ppm_int_wave(sm_, sp_, s6, un + cc, dtdx, gam_p, b);
        (void)a;
        (void)b;
    }
    t.ep = eidx(I, i, j, k);
    t.em = eidx(I, i + di, j + dj, k + dk);
    double *SP = x.SP[d], *SM = x.SM[d];
    t.XP = species_pairs(SP, I, NFS, FS_X);
    t.XM = species_pairs(SM, I, NFS, FS_X);
    t.XQ = cell_species(x.QX, I);
    // velocities back in absolute order: the normal one is FS_U + d, the
    // first transverse FS_V, FS_W, FS_U and the second FS_W, FS_U, FS_V for
    // d = 0, 1, 2; stored as the pairs (rho, u), (v, w), (p, rhoe)
    // the solver's transverse velocities are (v, w), (u, w), (u, v) for
    // d = 0, 1, 2: the traced (ut, utt) except at d = 1, where they swap
    auto to_abs = [&](StateH &o, FaceH &f, double r, double un_, double ut_, double utt_, double e_, double p_) {
        o.rho = r;
        o.u = d == 0 ? un_ : (d == 1 ? utt_ : ut_);
        o.v = d == 0 ? ut_ : (d == 1 ? un_ : utt_);
        o.w = d == 0 ? utt_ : (d == 1 ? ut_ : un_);
        o.p = p_;
        o.rhoe = e_;
        f.rho = r;
        f.un = un_;
        f.ut = d == 1 ? utt_ : ut_;
        f.utt = d == 1 ? ut_ : utt_;
        f.p = p_;
        f.rhoe = e_;
    };
    auto store_state = [&](double *S, idx_t ee, const StateH &o) {
        EF2(S, I, NFS, FS_RHO, ee) = dbl2{o.rho, o.u};
        EF2(S, I, NFS, FS_V, ee) = dbl2{o.v, o.w};
        EF2(S, I, NFS, FS_P, ee) = dbl2{o.p, o.rhoe};
    };
    {
        double r, u_, e, pp;
        ppm_project<true>(Im, gam_m, un, cc, r, u_, e, pp);
        StateH sp;
        to_abs(sp, fp, r, u_, Im_ut, Im_utt, e, pp);
        if (t.do_p) store_state(SP, t.ep, sp);
    }
    {
        double r, u_, e, pp;
        ppm_project<false>(Ip, gam_p, un, cc, r, u_, e, pp);
        StateH sm;
        to_abs(sm, fm, r, u_, Ip_ut, Ip_utt, e, pp);
        if (t.do_m) store_state(SM, t.em, sm);
    }
}

// Species pair pr of the trace (after trace_hydro): the plus-state pair of
// face pos in xp, the minus-state pair of face pos + 1 in xm, both stored
// where the cell stores its states.  The last pair of an odd species count
// carries 0 in its second slot.
__device__ __forceinline__ void trace_species(const TraceCtx &t, int pr, dbl2 &xp, dbl2 &xm)
{
    double ipv[2] = {0.0, 0.0}, imv[2] = {0.0, 0.0};
    dbl2 s2[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) s2[q] = SP2(t.XQ, pr, t.cs[q]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ns = 2 * pr + h;
        if (ns >= NSPEC) break;
        double s[5], sm, sp;
#pragma unroll
        for (int q = 0; q < 5; ++q) s[q] = h == 0 ? s2[q].x : s2[q].y;
        ppm_reconstruct(s, t.flatn, sm, sp);
        const double s6 = ppm_s6(sm, sp, s[2]);
        ppm_int_wave(sm, sp, s6, t.un, t.dtdx, ipv[h], imv[h]);
    }
    xp = dbl2{imv[0], imv[1]};
    xm = dbl2{ipv[0], ipv[1]};
    if (t.do_p) SP2(t.XP, pr, t.ep) = xp;
    if (t.do_m) SP2(t.XM, pr, t.em) = xm;
}

__global__ __launch_bounds__(256) void K_trace(Plan p, Ctx x)
{
    int m, i, j, k, sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    TraceCtx t;
    FaceH fp, fm;
    trace_hydro<false>(x, m, p.inst[m], sub_, i, j, k, true, t, fp, fm);
    // species last: their stores come after every read of the instance table
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        dbl2 xp, xm;
        trace_species(t, pr, xp, xm);
    }
}

// Source/riemann.F90:10-331 on the PPM states: F^d on d-face (i,j,k)
// (d-face index 0..n_d, transverse cells -1..n).  Species are streamed in
// pairs.
__device__ __forceinline__ void flux1_face(const Ctx &x, const FInst &I, int d, int i, int j, int k)
{
    const idx_t e = eidx(I, i, j, k);
    FaceH L, R;
    load_fs_h(x.SM[d], I, e, d, L);
    load_fs_h(x.SP[d], I, e, d, R);
    RiemannH o;
    face_riemann_h(x, I, d, i, j, k, L, R, o);
    store_tflux_h(x.FD[d], I, e, d, o);
    double *F = x.FD[d];
    // a side whose weight is zero contributes +-0 to fpl*xl + fmr*xr: its
    // species are not read
    const bool useL = o.fpl != 0.0, useR = o.fmr != 0.0;
    const SpeciesPairs XL = species_pairs(x.SM[d], I, NFS, FS_X), XR = species_pairs(x.SP[d], I, NFS, FS_X),
                       XF = species_pairs(F, I, NFF, FF_X);
#pragma unroll 1
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const dbl2 xl = useL ? (dbl2)SP2(XL, pr, e) : dbl2{0.0, 0.0};
        const dbl2 xr = useR ? (dbl2)SP2(XR, pr, e) : dbl2{0.0, 0.0};
        const double gx0 = o.fpl * xl.x + o.fmr * xr.x;
        const double gx1 = o.fpl * xl.y + o.fmr * xr.y;
        SP2(XF, pr, e) = dbl2{o.f_rho * gx0, o.f_rho * gx1};
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void K_flux1(Plan p, Ctx x)
{
    int m, i, j, k, sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    const int d = sub_;
    const FInst &I = p.inst[m];
    const int pos = d == 0 ? i : (d == 1 ? j : k);
    if (pos < lo_index(I, d)) return;
    flux1_face(x, I, d, i, j, k);
}

// the thread of the layer below along a brick's overlap direction (stride st
// in threads); only read for threads of layer l >= 1
__device__ __forceinline__ int below_of(int tid, int st) { return tid - st; }

// K_trace and K_flux1 in one launch.  The bricks of direction d overlap
// along d by one cell layer (plan_cell_ovl).  The block first copies the
// primitives its stencils reach (its brick and two cells either side along
// d, every component) into an LDS tile, one global load per cell and
// component; every thread then traces its cell from the tile and stores
// both states as K_trace does.  The minus state (for the cell's high d-face)
// also goes to LDS, and after the block's barrier the thread of each layer
// l >= 1 solves the Riemann problem on its cell's low d-face with the minus
// state the layer below left in LDS and its own plus state, kept in
// registers, and the sound speeds and Gamma_1 of the tile: the first-stage
// flux reads no traced state from memory.  The species follow the same way,
// all pairs at once.
// A layer-0 cell belongs to the brick below as its top layer: it stores its
// minus state (the same values that brick stores) but not its plus state,
// and solves no Riemann problem.  The sub-launches are the directions,
// interleaved block by block like the other face stages.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void K_tflux(Plan3 p3, Ctx x)
{
    // the minus states' hydro part; the primitives' tile, then (after the
    // barrier that ends its reads) the minus states' species in its memory
    __shared__ FaceH lds_h[256];
    __shared__ union {
        double q[NFQ * QTILE];
        dbl2 s[NSPEC_PAIRS][256];
    } lds;
    double *tile = lds.q;
    dbl2(&lds_s)[NSPEC_PAIRS][256] = lds.s;
    int lb = logical_block(), d;
    if (gridDim.y > 1) {
        d = blockIdx.y;
    } else {
        d = lb % 3;
        lb /= 3;
    }
    const Plan p = d == 0 ? p3.p[0] : (d == 1 ? p3.p[1] : p3.p[2]);
    if (lb >= (d == 0 ? p3.nb[0] : (d == 1 ? p3.nb[1] : p3.nb[2]))) return;  // the whole block
    int m, i, j, k, l;
    const bool ok = plan_cell_ovl(p, d, lb, m, i, j, k, l);
    const FInst &I = p.inst[m];
    const int tid = threadIdx.x;
    const int below = tid - (d == 0 ? 1 : (d == 1 ? p.bx : p.bx * p.by));  // the thread of layer l - 1
    const int pos = d == 0 ? i : (d == 1 ? j : k);
    const bool fx = ok && l > 0 && pos >= lo_index(I, d);
    // the tile: bd + 4 rows along d (row r: cell pos0 - 2 + r of the brick's
    // layer-0 position pos0) of cw cells across; tile element e is row
    // e % (bd + 4), column e / (bd + 4) in x (rows contiguous, as the x
    // bricks' lanes run along x) and row e / cw, column e % cw in y and z
    // (columns contiguous, x fastest), so that both the global loads and the
    // threads' LDS reads are consecutive across lanes
    const int nthr = p.bx * p.by * p.bz;
    const int bd = d == 0 ? p.bx : (d == 1 ? p.by : p.bz);
    const int cw = nthr / bd, rows = bd + 4;
    const int ti = tid % p.bx, tj = (tid / p.bx) % p.by, tk = tid / (p.bx * p.by);
    const int i0 = i - ti, j0 = j - tj, k0 = k - tk;  // the brick's first cell
    for (int e = tid; e < rows * cw; e += nthr) {
        const int r = d == 0 ? e % rows : e / cw, c = d == 0 ? e / rows : e % cw;
        int a, b, cz;
        if (d == 0) {
            a = i0 - 2 + r;
            b = j0 + c % p.by;
            cz = k0 + c / p.by;
        } else if (d == 1) {
            a = i0 + c % p.bx;
            b = j0 - 2 + r;
            cz = k0 + c / p.bx;
        } else {
            a = i0 + c % p.bx;
            b = j0 + c / p.bx;
            cz = k0 - 2 + r;
        }
        const idx_t ce = cidx(I, a, b, cz);
#pragma unroll
        for (int n = 0; n < NFQ; ++n) tile[n * QTILE + e] = CA(x.Q, I, NFQ, n, ce);
    }
    const int sr = d == 0 ? 1 : cw;  // tile stride of one row
    const int self = d == 0 ? l + (tj + p.by * tk) * rows : l * cw + (d == 1 ? ti + p.bx * tk : ti + p.bx * tj);
    __syncthreads();
    TraceCtx t;
    FaceH fp, fm;
    if (ok) {
        trace_hydro<true>(x, m, I, d, i, j, k, l > 0, t, fp, fm, tile, self, sr);
        lds_h[tid] = fm;
    }
    const SpeciesPairs XF = species_pairs(x.FD[d], I, NFF, FF_X);
    __syncthreads();
    RiemannH o;
    bool useL = false, useR = false;
    if (fx) {
        const FaceH L = lds_h[below];
        // cells pos - 1 and pos: tile offsets 1 and 2 from the thread's row
        riemann_hydro(L, fp, tile[FQ_C * QTILE + self + sr], tile[FQ_C * QTILE + self + 2 * sr],
                      tile[FQ_G * QTILE + self + sr], tile[FQ_G * QTILE + self + 2 * sr], o);
        store_tflux_h(x.FD[d], I, t.ep, d, o);
        // a side whose weight is zero contributes +-0 to fpl*xl + fmr*xr
        useL = o.fpl != 0.0;
        useR = o.fmr != 0.0;
    }
    __syncthreads();  // the tile is read: its memory takes the species
    // every pair's species trace first (the minus states to LDS, the plus
    // states kept), so that their loads are in flight together; one barrier
    dbl2 xpv[NSPEC_PAIRS];
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        dbl2 xp = dbl2{0.0, 0.0}, xm = dbl2{0.0, 0.0};
        if (ok) trace_species(t, pr, xp, xm);
        lds_s[pr][tid] = xm;
        xpv[pr] = xp;
    }
    __syncthreads();
    if (fx) {
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
            const dbl2 xl = useL ? lds_s[pr][below] : dbl2{0.0, 0.0};
            const dbl2 xr = useR ? xpv[pr] : dbl2{0.0, 0.0};
            const double gx0 = o.fpl * xl.x + o.fmr * xr.x;
            const double gx1 = o.fpl * xl.y + o.fmr * xr.y;
            SP2(XF, pr, t.ep) = dbl2{o.f_rho * gx0, o.f_rho * gx1};
        }
    }
}

// trans1 (Source/trans.F90:14-243) of the t-face states by the d-flux, then
// the flux F^{t|d} of the corrected states on the t-faces of tbx[t][d].
// The hydro part of the d-fluxes the brick's faces read (on the d-faces of
// the cells either side of each t-face: the brick and one cell more below
// along t and above along d) is first copied into LDS, one global load per
// face and pair of components; the block's LDS is sized by the launch
// (trans1_tile_cells).
// Held to 5 waves per SIMD (96 VGPRs, no spills): 12.50 against 13.02 ms
// at the 4 waves its 99 VGPRs would give.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void K_trans1(Plan p, Ctx x)
{
    extern __shared__ double ftile[];  // 7 components x nT cells
    int m, i, j, k, sub_;
    const bool inr = plan_cell_sub(p, 6, sub_, m, i, j, k);
    const int combo = sub_;
    const int t = c_pair_t[combo], d = c_pair_d[combo];
    const FInst &I = p.inst[m];
    const int c[3] = {i, j, k};
    // range: t-face 0..n_t, d-cells 0..n_d-1, the third direction -1..n
    const bool act = inr && !(c[t] < lo_index(I, t) || c[d] < lo_index(I, d) || c[d] > hi_cell(I, d));
    // the d-flux tile: the brick's faces widened by one below along t and one
    // above along d, x fastest; cell (ax, ay, az) is face (i0, j0, k0) + (ax,
    // ay, az) - e_t of the extended d-face array
    const int tid = threadIdx.x, nthr = p.bx * p.by * p.bz;
    const int ti = tid % p.bx, tj = (tid / p.bx) % p.by, tk = tid / (p.bx * p.by);
    const int et0 = t == 0, et1 = t == 1, et2 = t == 2, ed0 = d == 0, ed1 = d == 1, ed2 = d == 2;
    const int Tx = p.bx + et0 + ed0, Ty = p.by + et1 + ed1, Tz = p.bz + et2 + ed2, nT = Tx * Ty * Tz;
    {
        double *FDd = x.FD[d];
        const int gx0 = i - ti - et0, gy0 = j - tj - et1, gz0 = k - tk - et2;
        for (int e = tid; e < nT; e += nthr) {
            const int gx = gx0 + e % Tx, gy = gy0 + (e / Tx) % Ty, gz = gz0 + e / (Tx * Ty);
            // only faces of the extended array (the brick may reach past it)
            if (gx < -1 || gx > I.n[0] || gy < -1 || gy > I.n[1] || gz < I.zlo - 1 || gz > I.zlo + I.zn) continue;
            const idx_t ge = eidx(I, gx, gy, gz);
            const dbl2 a = EF2(FDd, I, NFF, FF_RHO, ge), b = EF2(FDd, I, NFF, FF_MY, ge), cc = EF2(FDd, I, NFF, FF_E, ge);
            ftile[e] = a.x;
            ftile[nT + e] = a.y;
            ftile[2 * nT + e] = b.x;
            ftile[3 * nT + e] = b.y;
            ftile[4 * nT + e] = cc.x;
            ftile[5 * nT + e] = cc.y;
            ftile[6 * nT + e] = EF(FDd, I, NFF, FF_GU, ge);
        }
    }
    __syncthreads();
    if (!act) return;
    // the tile's copy of load_tflux_h at tile cell q
    auto tile_tflux = [&](int q, double *v, double &gp, double &gu) {
        v[URHO] = ftile[q];
        v[UMX] = ftile[nT + q];
        v[UMY] = ftile[2 * nT + q];
        v[UMZ] = ftile[3 * nT + q];
        v[UEDEN] = ftile[4 * nT + q];
        gp = ftile[5 * nT + q];
        gu = ftile[6 * nT + q];
    };
    const idx_t e = eidx(I, i, j, k);
    const int et[3] = {t == 0, t == 1, t == 2}, ed[3] = {d == 0, d == 1, d == 2};
    const double cdtdx = x.dt_cur[m] / I.dx[d];
    const double cdt = cdtdx / 3.0;
    FaceH side[2];
    idx_t el[2], er[2];
    double rho[2], rrnew[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int sh = s - 1;  // -1: minus state (zone t-1), 0: plus state (zone t)
        const int il = i + sh * et[0], jl = j + sh * et[1], kl = k + sh * et[2];
        el[s] = eidx(I, il, jl, kl);
        er[s] = eidx(I, il + ed[0], jl + ed[1], kl + ed[2]);
        QInH lq;
        load_qin_h(s == 0 ? x.SM[t] : x.SP[t], I, e, lq);
        double fr[UEDEN + 1], fl[UEDEN + 1], pgp, ugp, pgm, ugm;
        // the side's cell in the tile: the thread's face + s e_t (the tile
        // starts one cell below along t), and its high d-face one above
        const int ql = (ti + s * et0) + Tx * ((tj + s * et1) + Ty * (tk + s * et2));
        tile_tflux(ql + (ed0 + Tx * (ed1 + Ty * ed2)), fr, pgp, ugp);
        tile_tflux(ql, fl, pgm, ugm);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cidx(I, il, jl, kl));
        const QStateH o = trans1_hydro(lq, fr, fl, pgp, pgm, ugp, ugm, gamc, cdt);
        rho[s] = lq.rho;
        rrnew[s] = lq.rho - cdt * (fr[URHO] - fl[URHO]);
        side[s] = to_face(o, t);
    }
    RiemannH o;
    face_riemann_h(x, I, t, i, j, k, side[0], side[1], o);
    double *F = x.FT[combo], *FDd = x.FD[d];
    store_tflux_h(F, I, e, t, o);
    // only the side(s) with a nonzero Riemann weight are corrected and read;
    // species go in pairs (SP2)
    const bool useL = o.fpl != 0.0, useR = o.fmr != 0.0;
    const SpeciesPairs XL = species_pairs(x.SM[t], I, NFS, FS_X), XR = species_pairs(x.SP[t], I, NFS, FS_X),
                       XD = species_pairs(FDd, I, NFF, FF_X), XF = species_pairs(F, I, NFF, FF_X);
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const bool two = 2 * pr + 1 < NSPEC;  // the last pair holds one species
        double x0[2] = {0.0, 0.0}, x1[2] = {0.0, 0.0};
        if (useL) {
            const dbl2 s = SP2(XL, pr, e), fr = SP2(XD, pr, er[0]), fl = SP2(XD, pr, el[0]);
            x0[0] = trans1_species(rho[0], s.x, fr.x, fl.x, cdt, rrnew[0]);
            if (two) x0[1] = trans1_species(rho[0], s.y, fr.y, fl.y, cdt, rrnew[0]);
        }
        if (useR) {
            const dbl2 s = SP2(XR, pr, e), fr = SP2(XD, pr, er[1]), fl = SP2(XD, pr, el[1]);
            x1[0] = trans1_species(rho[1], s.x, fr.x, fl.x, cdt, rrnew[1]);
            if (two) x1[1] = trans1_species(rho[1], s.y, fr.y, fl.y, cdt, rrnew[1]);
        }
        SP2(XF, pr, e) = dbl2{o.f_rho * (o.fpl * x0[0] + o.fmr * x1[0]),
                                            two ? o.f_rho * (o.fpl * x0[1] + o.fmr * x1[1]) : 0.0};
    }
}

// K_trans1 by cells: the thread of cell c (bricks elongated along the
// state direction t, overlapping by one cell, as K_tflux) applies trans1 to
// both t-states of its cell, the plus state of t-face c and the minus state
// of t-face c + 1, with the d-fluxes on its cell's two d-faces, which it
// reads once for both (K_trans1 reads them once per face side).  The minus
// side goes through LDS to the thread above, which solves the Riemann
// problem on its t-face as K_trans1 does.  Species: a face's weights decide
// which sides it reads (the minus side's thread learns it through LDS), and
// the corrected minus-side species follow pair by pair through LDS.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void K_trans1c(Plan3 p3, Ctx x)
{
    // the minus sides' hydro part, then (after the barrier that ends its
    // reads) their species in the same memory
    __shared__ union {
        FaceH h[256];
        dbl2 s[NSPEC_PAIRS][256];
    } lds;
    __shared__ unsigned char lds_f[256];
    FaceH(&lds_h)[256] = lds.h;
    dbl2(&lds_s)[NSPEC_PAIRS][256] = lds.s;
    int lb = logical_block(), combo;
    if (gridDim.y > 1) {
        combo = blockIdx.y;
    } else {
        combo = lb % 6;
        lb /= 6;
    }
    const int t = c_pair_t[combo], d = c_pair_d[combo];
    const Plan p = t == 0 ? p3.p[0] : (t == 1 ? p3.p[1] : p3.p[2]);
    if (lb >= (t == 0 ? p3.nb[0] : (t == 1 ? p3.nb[1] : p3.nb[2]))) return;  // the whole block
    int m, i, j, k, l;
    const bool ok = plan_cell_ovl(p, t, lb, m, i, j, k, l);
    const FInst &I = p.inst[m];
    const int tid = threadIdx.x;
    const int bt = t == 0 ? p.bx : (t == 1 ? p.by : p.bz);
    const int st = t == 0 ? 1 : (t == 1 ? p.bx : p.bx * p.by);
    const int ct = t == 0 ? i : (t == 1 ? j : k), cd = d == 0 ? i : (d == 1 ? j : k);
    // t-faces lo_index..n_t (the slab's top in z), d-cells lo_index..hi_cell
    const int tmax = t == 2 ? I.zlo + I.zn : I.n[t];
    const bool act = ok && cd >= lo_index(I, d) && cd <= hi_cell(I, d);
    const bool has_p = act && l > 0 && ct >= lo_index(I, t);  // t-face c (the plus side)
    const bool has_m = act && l + 1 < bt && ct + 1 <= tmax;    // t-face c + 1 (the minus side)
    const int et0 = t == 0, et1 = t == 1, et2 = t == 2, ed0 = d == 0, ed1 = d == 1, ed2 = d == 2;
    const idx_t ep = eidx(I, i, j, k), em = eidx(I, i + et0, j + et1, k + et2);
    const idx_t fl_i = ep, fr_i = eidx(I, i + ed0, j + ed1, k + ed2);  // the cell's d-faces
    const double cdtdx = x.dt_cur[m] / I.dx[d];
    const double cdt = cdtdx / 3.0;
    double *F = x.FT[combo], *FDd = x.FD[d];
    const SpeciesPairs XL = species_pairs(x.SM[t], I, NFS, FS_X), XR = species_pairs(x.SP[t], I, NFS, FS_X),
                       XD = species_pairs(FDd, I, NFF, FF_X), XF = species_pairs(F, I, NFF, FF_X);
    FaceH side_p;
    double rho_p = 0.0, rrnew_p = 0.0, rho_m = 0.0, rrnew_m = 0.0;
    if (has_p || has_m) {
        double fr[UEDEN + 1], fl[UEDEN + 1], pgp, ugp, pgm, ugm;
        load_tflux_h(FDd, I, fr_i, fr, pgp, ugp);
        load_tflux_h(FDd, I, fl_i, fl, pgm, ugm);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cidx(I, i, j, k));
        if (has_m) {
            QInH lq;
            load_qin_h(x.SM[t], I, em, lq);
            const QStateH o = trans1_hydro(lq, fr, fl, pgp, pgm, ugp, ugm, gamc, cdt);
            rho_m = lq.rho;
            rrnew_m = lq.rho - cdt * (fr[URHO] - fl[URHO]);
            lds_h[tid] = to_face(o, t);
        }
        if (has_p) {
            QInH lq;
            load_qin_h(x.SP[t], I, ep, lq);
            const QStateH o = trans1_hydro(lq, fr, fl, pgp, pgm, ugp, ugm, gamc, cdt);
            rho_p = lq.rho;
            rrnew_p = lq.rho - cdt * (fr[URHO] - fl[URHO]);
            side_p = to_face(o, t);
        }
    }
    __syncthreads();
    RiemannH o;
    bool useL = false, useR = false;
    if (has_p) {
        face_riemann_h(x, I, t, i, j, k, lds_h[below_of(tid, st)], side_p, o);
        store_tflux_h(F, I, ep, t, o);
        useL = o.fpl != 0.0;
        useR = o.fmr != 0.0;
    }
    lds_f[tid] = useL;
    __syncthreads();
    const bool need_m = has_m && lds_f[tid + st];  // the face above reads this minus side
    // every pair's corrected species first (the minus side to LDS, the plus
    // side kept), so that their loads are in flight together; one barrier
    dbl2 xpv[NSPEC_PAIRS];
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const bool two = 2 * pr + 1 < NSPEC;  // the last pair holds one species
        double xm[2] = {0.0, 0.0}, xp[2] = {0.0, 0.0};
        if (need_m || useR) {
            const dbl2 fr = SP2(XD, pr, fr_i), fl = SP2(XD, pr, fl_i);
            if (need_m) {
                const dbl2 s = SP2(XL, pr, em);
                xm[0] = trans1_species(rho_m, s.x, fr.x, fl.x, cdt, rrnew_m);
                if (two) xm[1] = trans1_species(rho_m, s.y, fr.y, fl.y, cdt, rrnew_m);
            }
            if (useR) {
                const dbl2 s = SP2(XR, pr, ep);
                xp[0] = trans1_species(rho_p, s.x, fr.x, fl.x, cdt, rrnew_p);
                if (two) xp[1] = trans1_species(rho_p, s.y, fr.y, fl.y, cdt, rrnew_p);
            }
        }
        lds_s[pr][tid] = dbl2{xm[0], xm[1]};
        xpv[pr] = dbl2{xp[0], xp[1]};
    }
    __syncthreads();
    if (has_p) {
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
            const bool two = 2 * pr + 1 < NSPEC;
            const dbl2 x0 = useL ? lds_s[pr][below_of(tid, st)] : dbl2{0.0, 0.0};
            SP2(XF, pr, ep) = dbl2{o.f_rho * (o.fpl * x0.x + o.fmr * xpv[pr].x),
                                   two ? o.f_rho * (o.fpl * x0.y + o.fmr * xpv[pr].y) : 0.0};
        }
    }
}

// The artificial viscosity of d-face (i,j,k) and the hydro part of its final
// flux (apply_av, Source/Castro_hydro_nd.F90:174-245, with the reference's
// corner-node pattern): stores the conserved fluxes and the Godunov p of
// face e and returns what the species fluxes need.
struct FinalAV {
    double div1, dxd, frho;
    idx_t cr, cl;
};
__device__ __forceinline__ FinalAV final_hydro(const Ctx &x, const FInst &I, int d, int i, int j, int k, idx_t e,
                                               const RiemannH &o)
{
    FinalAV v;
    double div1;
#define DVN(a, b, cc) EA(x.DV, I, 1, 0, eidx(I, a, b, cc))
    if (d == 0) div1 = FOURTH * (DVN(i, j, k) + DVN(i, j + 1, k) + DVN(i, j, k + 1) + DVN(i, j + 1, k));
    else if (d == 1) div1 = FOURTH * (DVN(i, j, k) + DVN(i + 1, j, k) + DVN(i, j, k + 1) + DVN(i + 1, j, k));
    else div1 = FOURTH * (DVN(i, j, k) + DVN(i + 1, j, k) + DVN(i, j + 1, k) + DVN(i + 1, j + 1, k));
#undef DVN
    v.div1 = 0.1 * fmin(0.0, div1);
    v.cr = cidx(I, i, j, k);
    v.cl = cidx(I, i - (d == 0), j - (d == 1), k - (d == 2));
    v.dxd = I.dx[d];
#define AVN(n, fval) ((fval) + v.dxd * (v.div1 * (CA(x.U1, I, NU1, n, v.cr) - CA(x.U1, I, NU1, n, v.cl))))
    const int mn = UMX + d, mt = d == 0 ? UMY : UMX, mtt = d == 2 ? UMY : UMZ;
    double *FN = x.FN[d];
    v.frho = AVN(URHO, o.f_rho);
    // momenta back in absolute order (mn = UMX + d, mt, mtt as above)
    const double vmn = AVN(mn, o.f_mn), vmt = AVN(mt, o.f_mt), vmtt = AVN(mtt, o.f_mtt);
    const double mx = d == 0 ? vmn : vmt;
    const double my = d == 0 ? vmt : (d == 1 ? vmn : vmtt);
    const double mz = d == 2 ? vmn : vmtt;
    EF2(FN, I, NFN, URHO, e) = dbl2{v.frho, mx};
    EF2(FN, I, NFN, UMY, e) = dbl2{my, mz};
    EF2(FN, I, NFN, UEDEN, e) = dbl2{AVN(UEDEN, o.f_E), AVN(UEINT, o.f_e)};
    EF2(FN, I, NFN, UTEMP, e) = dbl2{0.0, o.g_p};  // UTEMP, FN_GP
#undef AVN
    return v;
}

// species flux of species pair pr with its artificial viscosity (the
// corrected species x0 of the minus side, x1 of the plus side), added to
// the running sum in species order
__device__ __forceinline__ void final_species_pair(const FinalAV &v, const RiemannH &o, const SpeciesPairs &XU,
                                                   int pr, const double *x0, const double *x1, double *fs,
                                                   double &sum)
{
    const dbl2 ur = SP2(XU, pr, v.cr), ul = SP2(XU, pr, v.cl);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int n = 2 * pr + h;
        if (n >= NSPEC) break;
        const double f = o.f_rho * (o.fpl * x0[h] + o.fmr * x1[h]);
        fs[n] = f + v.dxd * (v.div1 * ((h == 0 ? ur.x : ur.y) - (h == 0 ? ul.x : ul.y)));
        sum = sum + fs[n];
    }
}

// normalize_species_fluxes (Source/Castro_hydro_nd.F90:248-297) and the
// species and Godunov-u stores of face e
__device__ __forceinline__ void final_species_store(const SpeciesPairs &XN, double *FN, const FInst &I, idx_t e,
                                                    const double *fs, double sum, double frho, double g_un)
{
    const double fac = (sum != 0.0) ? frho / sum : 1.0;
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr)
        SP2(XN, pr, e) = dbl2{fs[2 * pr] * fac, 2 * pr + 1 < NSPEC ? fs[2 * pr + 1] * fac : 0.0};
    EF(FN, I, NFN, FN_GU, e) = g_un;
}

// trans2 (Source/trans.F90:250-505) + final flux + apply_av
// (Source/Castro_hydro_nd.F90:174-245) + normalize_species_fluxes (:248-297)
// on the d-faces of the valid region.
__global__ __launch_bounds__(256) void K_final(Plan p, Ctx x)
{
    int m, i, j, k, sub_;
    if (!plan_cell_sub(p, 3, sub_, m, i, j, k)) return;
    const int d = sub_;
    const FInst &I = p.inst[m];
    const int c[3] = {i, j, k};
    for (int a = 0; a < 3; ++a) {
        if (c[a] < lo_index(I, a)) return;
        if (a != d && c[a] > hi_cell(I, a)) return;
    }
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int ed[3] = {d == 0, d == 1, d == 2}, e1[3] = {t1 == 0, t1 == 1, t1 == 2}, e2[3] = {t2 == 0, t2 == 1, t2 == 2};
    const double dt = x.dt_cur[m];
    const double h2 = 0.5 * dt / I.dx[t1], h3 = 0.5 * dt / I.dx[t2];
    const idx_t e = eidx(I, i, j, k);
    double *F2 = x.FT[pair_index(t1, t2)], *F3 = x.FT[pair_index(t2, t1)];
    FaceH side[2];
    idx_t l2[2], r2[2], r3[2];
    double rho[2], rrnew[2];
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
        rho[s] = lq.rho;
        rrnew[s] = lq.rho - h2 * (f2r[URHO] - f2l[URHO]) - h3 * (f3r[URHO] - f3l[URHO]);
        side[s] = to_face(o, d);
    }
    RiemannH o;
    face_riemann_h(x, I, d, i, j, k, side[0], side[1], o);
    double *FN = x.FN[d];
    const SpeciesPairs XL = species_pairs(x.SM[d], I, NFS, FS_X), XR = species_pairs(x.SP[d], I, NFS, FS_X),
                       X2 = species_pairs(F2, I, NFF, FF_X), X3 = species_pairs(F3, I, NFF, FF_X),
                       XN = species_pairs(FN, I, NFN, FN_X);
    const SpeciesPairs XU = cell_species(x.U1X, I);
    const FinalAV v = final_hydro(x, I, d, i, j, k, e, o);
    double fs[NSPEC], sum = 0.0;
    const bool useL = o.fpl != 0.0, useR = o.fmr != 0.0;
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const bool two = 2 * pr + 1 < NSPEC;  // the last pair holds one species
        double x0[2] = {0.0, 0.0}, x1[2] = {0.0, 0.0};
        if (useL) {
            const dbl2 s = SP2(XL, pr, e);
            const dbl2 a = SP2(X2, pr, r2[0]), b = SP2(X2, pr, l2[0]);
            const dbl2 c3 = SP2(X3, pr, r3[0]), d3 = SP2(X3, pr, l2[0]);
            x0[0] = trans2_species(rho[0], s.x, a.x, b.x, c3.x, d3.x, h2, h3, rrnew[0]);
            if (two) x0[1] = trans2_species(rho[0], s.y, a.y, b.y, c3.y, d3.y, h2, h3, rrnew[0]);
        }
        if (useR) {
            const dbl2 s = SP2(XR, pr, e);
            const dbl2 a = SP2(X2, pr, r2[1]), b = SP2(X2, pr, l2[1]);
            const dbl2 c3 = SP2(X3, pr, r3[1]), d3 = SP2(X3, pr, l2[1]);
            x1[0] = trans2_species(rho[1], s.x, a.x, b.x, c3.x, d3.x, h2, h3, rrnew[1]);
            if (two) x1[1] = trans2_species(rho[1], s.y, a.y, b.y, c3.y, d3.y, h2, h3, rrnew[1]);
        }
        final_species_pair(v, o, XU, pr, x0, x1, fs, sum);
    }
    final_species_store(XN, FN, I, e, fs, sum, v.frho, o.g_un);
}

// K_final by cells: the thread of cell c (bricks elongated along d,
// overlapping by one cell, as K_tflux) applies trans2 to both d-states of
// its cell, the plus state of d-face c and the minus state of d-face c + 1,
// with the four transverse fluxes on its cell's t1- and t2-faces, which it
// reads once for both (K_final reads them once per face side).  The minus
// side goes through LDS to the thread above, which solves the Riemann
// problem on its d-face and applies the artificial viscosity and the
// species normalisation as K_final does; species sides as in K_trans1c.
__global__ __launch_bounds__(256) void K_finalc(Plan3 p3, Ctx x)
{
    // the minus sides' hydro part, then (after the barrier that ends its
    // reads) their species in the same memory
    __shared__ union {
        FaceH h[256];
        dbl2 s[NSPEC_PAIRS][256];
    } lds;
    __shared__ unsigned char lds_f[256];
    FaceH(&lds_h)[256] = lds.h;
    dbl2(&lds_s)[NSPEC_PAIRS][256] = lds.s;
    int lb = logical_block(), d;
    if (gridDim.y > 1) {
        d = blockIdx.y;
    } else {
        d = lb % 3;
        lb /= 3;
    }
    const Plan p = d == 0 ? p3.p[0] : (d == 1 ? p3.p[1] : p3.p[2]);
    if (lb >= (d == 0 ? p3.nb[0] : (d == 1 ? p3.nb[1] : p3.nb[2]))) return;  // the whole block
    int m, i, j, k, l;
    const bool ok = plan_cell_ovl(p, d, lb, m, i, j, k, l);
    const FInst &I = p.inst[m];
    const int tid = threadIdx.x;
    const int bd = d == 0 ? p.bx : (d == 1 ? p.by : p.bz);
    const int st = d == 0 ? 1 : (d == 1 ? p.bx : p.bx * p.by);
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int cd = d == 0 ? i : (d == 1 ? j : k), c1 = t1 == 0 ? i : j, c2 = t2 == 1 ? j : k;
    // d-faces lo_index..n_d (the slab's top in z), transverse cells lo_index..hi_cell
    const int dmax = d == 2 ? I.zlo + I.zn : I.n[d];
    const bool act = ok && c1 >= lo_index(I, t1) && c1 <= hi_cell(I, t1) && c2 >= lo_index(I, t2) &&
                     c2 <= hi_cell(I, t2);
    const bool has_p = act && l > 0 && cd >= lo_index(I, d);  // d-face c (the plus side)
    const bool has_m = act && l + 1 < bd && cd + 1 <= dmax;   // d-face c + 1 (the minus side)
    const idx_t ep = eidx(I, i, j, k), em = eidx(I, i + (d == 0), j + (d == 1), k + (d == 2));
    const idx_t l2 = ep, r2 = eidx(I, i + (t1 == 0), j + (t1 == 1), k), r3 = eidx(I, i, j + (t2 == 1), k + (t2 == 2));
    const double dt = x.dt_cur[m];
    const double h2 = 0.5 * dt / I.dx[t1], h3 = 0.5 * dt / I.dx[t2];
    double *F2 = x.FT[pair_index(t1, t2)], *F3 = x.FT[pair_index(t2, t1)], *FN = x.FN[d];
    const SpeciesPairs XL = species_pairs(x.SM[d], I, NFS, FS_X), XR = species_pairs(x.SP[d], I, NFS, FS_X),
                       X2 = species_pairs(F2, I, NFF, FF_X), X3 = species_pairs(F3, I, NFF, FF_X),
                       XN = species_pairs(FN, I, NFN, FN_X);
    const SpeciesPairs XU = cell_species(x.U1X, I);
    FaceH side_p;
    double rho_p = 0.0, rrnew_p = 0.0, rho_m = 0.0, rrnew_m = 0.0;
    if (has_p || has_m) {
        double f2r[UEDEN + 1], f2l[UEDEN + 1], f3r[UEDEN + 1], f3l[UEDEN + 1], pg2p, ug2p, pg2m, ug2m, pg3p, ug3p,
            pg3m, ug3m;
        load_tflux_h(F2, I, r2, f2r, pg2p, ug2p);
        load_tflux_h(F2, I, l2, f2l, pg2m, ug2m);
        load_tflux_h(F3, I, r3, f3r, pg3p, ug3p);
        load_tflux_h(F3, I, l2, f3l, pg3m, ug3m);
        const double gamc = CA(x.Q, I, NFQ, FQ_G, cidx(I, i, j, k));
        if (has_m) {
            QInH lq;
            load_qin_h(x.SM[d], I, em, lq);
            const QStateH o = trans2_hydro(lq, f2r, f2l, f3r, f3l, pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m,
                                           gamc, h2, h3);
            rho_m = lq.rho;
            rrnew_m = lq.rho - h2 * (f2r[URHO] - f2l[URHO]) - h3 * (f3r[URHO] - f3l[URHO]);
            lds_h[tid] = to_face(o, d);
        }
        if (has_p) {
            QInH lq;
            load_qin_h(x.SP[d], I, ep, lq);
            const QStateH o = trans2_hydro(lq, f2r, f2l, f3r, f3l, pg2p, pg2m, ug2p, ug2m, pg3p, pg3m, ug3p, ug3m,
                                           gamc, h2, h3);
            rho_p = lq.rho;
            rrnew_p = lq.rho - h2 * (f2r[URHO] - f2l[URHO]) - h3 * (f3r[URHO] - f3l[URHO]);
            side_p = to_face(o, d);
        }
    }
    __syncthreads();
    RiemannH o;
    FinalAV v;
    bool useL = false, useR = false;
    if (has_p) {
        face_riemann_h(x, I, d, i, j, k, lds_h[below_of(tid, st)], side_p, o);
        v = final_hydro(x, I, d, i, j, k, ep, o);
        useL = o.fpl != 0.0;
        useR = o.fmr != 0.0;
    }
    lds_f[tid] = useL;
    __syncthreads();
    const bool need_m = has_m && lds_f[tid + st];  // the face above reads this minus side
    // every pair's corrected species first (the minus side to LDS, the plus
    // side kept), so that their loads are in flight together; one barrier
    dbl2 xpv[NSPEC_PAIRS];
#pragma unroll
    for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
        const bool two = 2 * pr + 1 < NSPEC;  // the last pair holds one species
        double xm[2] = {0.0, 0.0}, xp[2] = {0.0, 0.0};
        if (need_m || useR) {
            const dbl2 a = SP2(X2, pr, r2), b = SP2(X2, pr, l2);
            const dbl2 c3 = SP2(X3, pr, r3), d3 = SP2(X3, pr, l2);
            if (need_m) {
                const dbl2 s = SP2(XL, pr, em);
                xm[0] = trans2_species(rho_m, s.x, a.x, b.x, c3.x, d3.x, h2, h3, rrnew_m);
                if (two) xm[1] = trans2_species(rho_m, s.y, a.y, b.y, c3.y, d3.y, h2, h3, rrnew_m);
            }
            if (useR) {
                const dbl2 s = SP2(XR, pr, ep);
                xp[0] = trans2_species(rho_p, s.x, a.x, b.x, c3.x, d3.x, h2, h3, rrnew_p);
                if (two) xp[1] = trans2_species(rho_p, s.y, a.y, b.y, c3.y, d3.y, h2, h3, rrnew_p);
            }
        }
        lds_s[pr][tid] = dbl2{xm[0], xm[1]};
        xpv[pr] = dbl2{xp[0], xp[1]};
    }
    __syncthreads();
    if (has_p) {
        double fs[NSPEC], sum = 0.0;
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
            const dbl2 x0v = useL ? lds_s[pr][below_of(tid, st)] : dbl2{0.0, 0.0};
            const double x0[2] = {x0v.x, x0v.y}, x1[2] = {xpv[pr].x, xpv[pr].y};
            final_species_pair(v, o, XU, pr, x0, x1, fs, sum);
        }
        final_species_store(XN, FN, I, ep, fs, sum, v.frho, o.g_un);
    }
}

// fill_hydro_source (Source/Castro_hydro_nd.F90:344-441), Saxpy and the two
// clean_state calls of Castro_advance.cpp:74-78 / Castro.cpp:369, then the
// next step's estdt (Source/timestep.F90:12-72) min-reduced per instance.
// Held to 256 registers (2 waves per SIMD): the few spilled values cost less
// than the second wave gains on the EOS table gathers (6.1 -> 5.5 ms at 256^3;
// the same limit slows K_prim, which spills more).
// POST: post_timestep's clean_state and the estdt of the new state follow
// the advance's clean (the step of a level that is not the coarse side of a
// flux register); without it the state is left as Castro::advance leaves
// it, for the reflux and average-down that come before that clean.
template <bool POST>
__device__ __forceinline__ void update_cells(Plan p, Ctx x)
{
    int m, i, j, k;
    double v = DBL_MAX;
    const bool ok = plan_cell(p, m, i, j, k);
    const FInst &I = p.inst[m];
    if (ok) {
        const idx_t c = cidx(I, i, j, k);
        const double dx0 = I.dx[0], dx1 = I.dx[1], dx2 = I.dx[2];
        const double a0 = dx1 * dx2, a1 = dx0 * dx2, a2 = dx0 * dx1;
        const double volinv = 1.0 / (dx0 * dx1 * dx2);
        const idx_t e = eidx(I, i, j, k);
        const idx_t ex = eidx(I, i + 1, j, k), ey = eidx(I, i, j + 1, k), ez = eidx(I, i, j, k + 1);
        double *F0 = x.FN[0], *F1 = x.FN[1], *F2 = x.FN[2];
        const SpeciesPairs P0 = species_pairs(F0, I, NFN, FN_X), P1 = species_pairs(F1, I, NFN, FN_X),
                           P2 = species_pairs(F2, I, NFN, FN_X);
        const double dt = x.dt_cur[m];
        double u[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) u[n] = CA(x.U, I, NVAR, n, c);  // clean(S_old), stored by K_prim
#pragma unroll
        for (int pr = 0; pr < NSPEC_PAIRS; ++pr) {
            const dbl2 f0 = SP2(P0, pr, e), f0x = SP2(P0, pr, ex), f1 = SP2(P1, pr, e), f1y = SP2(P1, pr, ey),
                       f2 = SP2(P2, pr, e), f2z = SP2(P2, pr, ez);
            const int n = UFS + 2 * pr;
            u[n] = u[n] + dt * flux_div(f0.x, f0x.x, f1.x, f1y.x, f2.x, f2z.x, a0, a1, a2, volinv);
            if (n + 1 < NVAR) u[n + 1] = u[n + 1] + dt * flux_div(f0.y, f0x.y, f1.y, f1y.y, f2.y, f2z.y, a0, a1, a2, volinv);
        }
        // URHO..UTEMP and FN_GP as the four pairs (0,1) (2,3) (4,5) (6,7) of
        // the six faces; the last pair (UTEMP, Godunov p) first, for p div u
        double pdivu;
        {
            const dbl2 g0 = EF2(F0, I, NFN, UTEMP, e), g0x = EF2(F0, I, NFN, UTEMP, ex), g1 = EF2(F1, I, NFN, UTEMP, e),
                       g1y = EF2(F1, I, NFN, UTEMP, ey), g2 = EF2(F2, I, NFN, UTEMP, e),
                       g2z = EF2(F2, I, NFN, UTEMP, ez);
            u[UTEMP] = u[UTEMP] + dt * flux_div(g0.x, g0x.x, g1.x, g1y.x, g2.x, g2z.x, a0, a1, a2, volinv);
            pdivu = HALF * (g0x.y + g0.y) * (EF(F0, I, NFN, FN_GU, ex) - EF(F0, I, NFN, FN_GU, e)) / dx0 +
                    HALF * (g1y.y + g1.y) * (EF(F1, I, NFN, FN_GU, ey) - EF(F1, I, NFN, FN_GU, e)) / dx1 +
                    HALF * (g2z.y + g2.y) * (EF(F2, I, NFN, FN_GU, ez) - EF(F2, I, NFN, FN_GU, e)) / dx2;
        }
#pragma unroll
        for (int n = 0; n < UTEMP; n += 2) {
            const dbl2 f0 = EF2(F0, I, NFN, n, e), f0x = EF2(F0, I, NFN, n, ex), f1 = EF2(F1, I, NFN, n, e),
                       f1y = EF2(F1, I, NFN, n, ey), f2 = EF2(F2, I, NFN, n, e), f2z = EF2(F2, I, NFN, n, ez);
            const double s0 = flux_div(f0.x, f0x.x, f1.x, f1y.x, f2.x, f2z.x, a0, a1, a2, volinv);
            double s1 = flux_div(f0.y, f0x.y, f1.y, f1y.y, f2.y, f2z.y, a0, a1, a2, volinv);
            if (n + 1 == UEINT) s1 = s1 - pdivu;
            u[n] = u[n] + dt * s0;
            u[n + 1] = u[n + 1] + dt * s1;
        }
        clean_cell(u, x.tab); /* Castro_advance.cpp:78 */
        if (POST) clean_cell(u, x.tab); /* post_timestep, Castro.cpp:369 */
#pragma unroll
        for (int n = 0; n < NVAR; ++n) CA(x.U, I, NVAR, n, c) = u[n];
        if (POST) v = estdt_cell(u, dx0, dx1, dx2, x.tab);
    }
    if (POST) {
        v = block_min<256>(v);
        if (threadIdx.x == 0) atomic_min_d(&x.dt_est[m], v);
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void K_update(Plan p, Ctx x)
{
    update_cells<true>(p, x);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void K_update_adv(Plan p, Ctx x)
{
    update_cells<false>(p, x);
}

// Time-step control of Castro::computeInitialDt / computeNewDt
// (Source/Castro.cpp:162-173, 175-216, 218-341) for every instance.
__global__ void K_dt(Ctx x, int ninst, double stop_time)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ninst) return;
    const double est = x.dt_est[m] * cfl;
    double dt_0;
    if (x.steps[m] > 0) {
        double dt_min = fmin(est, 1.1e0 * x.dt_cur[m]);
        dt_0 = fmin(1.0e+100, dt_min);
    } else {
        dt_0 = fmin(1.0e+100, 0.01 * est);
    }
    const double eps = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((x.time[m] + dt_0) > (stop_time - eps)) dt_0 = stop_time - x.time[m];
    x.dt_cur[m] = dt_0;
    x.dt_est[m] = DBL_MAX;
}

__global__ void K_tick(Ctx x, int ninst)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ninst) return;
    x.time[m] += x.dt_cur[m];
    x.steps[m] += 1;
}

__global__ void K_fill(double *p, long n, double v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ __launch_bounds__(256) void K_arena(const ArenaBox *boxes, int nbox, const long *bstart, double *arena,
                                               double *U, long nD, int n0, int n1, int to_arena)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bstart[nbox]) return;
    int b = 0;
    while (b + 1 < nbox && bstart[b + 1] <= t) ++b;
    const ArenaBox B = boxes[b];
    const long lt = t - bstart[b];
    const int bx = B.hi[0] - B.lo[0] + 1, by = B.hi[1] - B.lo[1] + 1;
    const long nb = lt < 0 ? 0 : (long)bx * by * (B.hi[2] - B.lo[2] + 1);
    const int i = B.lo[0] + (int)(lt % bx), j = B.lo[1] + (int)((lt / bx) % by), k = B.lo[2] + (int)(lt / ((long)bx * by));
    const long c = (long)i + (long)n0 * (j + (long)n1 * k);
    for (int n = 0; n < NVAR; ++n) {
        double *a = arena + B.off + nb * n + lt;
        double *u = U + nD * n + c;
        if (to_arena) *a = *u;
        else *u = *a;
    }
}

// ------------------------------------------------------------------ host side
// index range of each stage: cells [0,n), faces [0,n], or the ring [-1,n]
static const int stage_shift[NSTAGE] = {0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 1};
static const int stage_add[NSTAGE] = {0, 0, 2, 2, 2, 2, 1, 0, 2, 2, 2};
// sub-launches (directions or pairs) of each stage
static const int stage_ny[NSTAGE] = {1, 1, 1, 3, 3, 6, 3, 1, 3, 6, 3};
// stages whose bricks overlap along their direction (plan_cell_ovl)
static bool overlap_stage(int stage) { return stage == ST_TFLUX || stage == ST_TRANS1C || stage == ST_FINALC; }

// the stages one slab of a step runs, in order (after K_prim)
static std::vector<int> step_stages(const FusedLevel *F)
{
    std::vector<int> v = {ST_FLAT};
    if (F->variant & FV_SPLIT_TRACE) {
        v.push_back(ST_TRACE);
        v.push_back(ST_FLUX1);
    } else {
        v.push_back(ST_TFLUX);
    }
    v.push_back(F->variant & FV_FACE_TRANS1 ? ST_TRANS1 : ST_TRANS1C);
    v.push_back(F->variant & FV_CELL_FINAL ? ST_FINALC : ST_FINAL);
    v.push_back(ST_UPDATE);
    return v;
}

// the cell stages before and after the face stages of a step (estdt,
// K_prim) run over whole instances; the face stages and K_update run slab
// by slab
static bool whole_instance_stage(int stage) { return stage == ST_EST || stage == ST_PRIM; }

// brick of an overlapping stage's direction d for its tiling (W, L, X)
static KernelTiling ovl_brick(const KernelTiling &t, int d)
{
    if (d == 0) return KernelTiling{t.bz, t.bx * t.by / t.bz, 1, t.zg};
    if (d == 1) return KernelTiling{t.bx, t.by, 1, t.zg};
    return KernelTiling{t.bx, 1, t.by, t.zg};
}

// an overlapping stage's tiling is valid when every direction's brick spans
// at least two cells along its direction and X divides the thread count
static bool ovl_tiling_ok(const KernelTiling &t)
{
    const int nt = t.bx * t.by;
    return t.bx > 0 && t.by >= 2 && t.bz >= 2 && nt <= 256 && nt % 64 == 0 && nt % t.bz == 0;
}

// block table of a brick tiling t over the range (shift, add) of a slab;
// ovl >= 0: the bricks overlap along that direction (plan_cell_ovl)
static Plan plan_for(FusedLevel *F, const KernelTiling &t, int sh, int add, int slab, int ovl)
{
    const FInst *tab = slab < 0 ? F->inst : F->sinst[slab];
    for (auto &pb : F->plans)
        if (pb.t.bx == t.bx && pb.t.by == t.by && pb.t.bz == t.bz && pb.t.zg == t.zg && pb.shift == sh &&
            pb.add == add && pb.slab == slab && pb.ovl == ovl)
            return Plan{tab, pb.blk0, F->ninst, t.bx, t.by, t.bz, sh, add, t.zg};
    FusedLevel::PlanBuf pb;
    pb.t = t;
    pb.shift = sh;
    pb.add = add;
    pb.slab = slab;
    pb.ovl = ovl;
    const std::vector<FInst> &hs = slab < 0 ? F->hinst : F->hsinst[slab];
    std::vector<int> b0(F->ninst + 1);
    const int bb[3] = {t.bx, t.by, t.bz};
    int acc = 0;
    for (int m = 0; m < F->ninst; ++m) {
        b0[m] = acc;
        const FInst &I = hs[m];
        const int e[3] = {I.n[0] + add, I.n[1] + add, I.zn + add};
        // an instance with no cells in this slab launches no bricks
        if (I.zn <= 0) continue;
        int nb = 1;
        for (int a = 0; a < 3; ++a) {
            const int st = bb[a] - (a == ovl);  // brick stride (plan_cell_ovl)
            nb *= (e[a] - (a == ovl) + st - 1) / st;
        }
        acc += nb;
    }
    b0[F->ninst] = acc;
    pb.nblocks = acc;
    if (check_hip(hipMalloc(&pb.blk0, sizeof(int) * (F->ninst + 1)), "hipMalloc(plan)") ||
        check_hip(hipMemcpy(pb.blk0, b0.data(), sizeof(int) * (F->ninst + 1), hipMemcpyHostToDevice), "plan h2d"))
        return Plan{nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0};
    F->plans.push_back(pb);
    return Plan{tab, pb.blk0, F->ninst, t.bx, t.by, t.bz, sh, add, t.zg};
}

Plan get_plan(FusedLevel *F, int stage, int slab)
{
    if (whole_instance_stage(stage)) slab = -1;
    return plan_for(F, F->tiling[stage], stage_shift[stage], stage_add[stage], slab, -1);
}

// K_trans1's LDS for a brick: its d-flux tile (the brick widened by one cell
// along t and one along d) of 7 components, for the largest of the six pairs
static size_t trans1_tile_bytes(int bx, int by, int bz)
{
    static const int pt[6] = {1, 2, 0, 2, 0, 1}, pd[6] = {0, 0, 1, 1, 2, 2};
    long cells = 0;
    for (int c = 0; c < 6; ++c) {
        long n = 1;
        const int b[3] = {bx, by, bz};
        for (int a = 0; a < 3; ++a) n *= b[a] + (a == pt[c]) + (a == pd[c]);
        cells = std::max(cells, n);
    }
    return (size_t)cells * 7 * sizeof(double);
}

// the three direction plans of an overlapping stage
static Plan3 get_plan3(FusedLevel *F, int stage, int slab)
{
    Plan3 q;
    for (int d = 0; d < 3; ++d) {
        q.p[d] = plan_for(F, ovl_brick(F->tiling[stage], d), stage_shift[stage], stage_add[stage], slab, d);
        q.nb[d] = plan_blocks(F, q.p[d]);
    }
    return q;
}

int plan_blocks(FusedLevel *F, const Plan &p)
{
    for (auto &pb : F->plans)
        if (pb.blk0 == p.blk0) return pb.nblocks;
    return 0;
}

} // namespace mc

using namespace mc;

namespace mc {

static void launch_stage(FusedLevel *F, int stage, hipStream_t s, int ny = 1, int slab = 0)
{
    F->ctx.tab = eos_table();
    if (overlap_stage(stage)) {
        // the directions / pairs interleaved block by block; one with fewer
        // bricks than the largest leaves its last blocks idle
        const Plan3 q = get_plan3(F, stage, slab);
        const int nb = std::max(q.nb[0], std::max(q.nb[1], q.nb[2]));
        if (nb <= 0) return;
        const KernelTiling &t = F->tiling[stage];
        if (stage == ST_TFLUX) hipLaunchKernelGGL(K_tflux, dim3(3 * nb), dim3(t.bx * t.by), 0, s, q, F->ctx);
        else if (stage == ST_TRANS1C) hipLaunchKernelGGL(K_trans1c, dim3(6 * nb), dim3(t.bx * t.by), 0, s, q, F->ctx);
        else hipLaunchKernelGGL(K_finalc, dim3(3 * nb), dim3(t.bx * t.by), 0, s, q, F->ctx);
        return;
    }
    Plan p = get_plan(F, stage, slab);
    const int nb = plan_blocks(F, p);
    const int nt = p.bx * p.by * p.bz;
    if (nb <= 0) return;
    // the directions / pairs of a face stage are interleaved brick by brick
    // (plan_cell_sub): the sub-launches of a brick share inputs and run
    // together on one XCD, so the shared reads hit its L2
    const bool interleave = stage == ST_TRACE || stage == ST_FLUX1 || stage == ST_TRANS1 || stage == ST_FINAL;
    if (interleave) ny = stage == ST_TRANS1 ? 6 : 3;  // the count the kernels assume
    dim3 g(interleave ? nb * ny : nb, interleave ? 1 : ny), b(nt);
    switch (stage) {
    case ST_EST: hipLaunchKernelGGL(K_est, g, b, 0, s, p, F->ctx); break;
    case ST_PRIM: hipLaunchKernelGGL(K_prim, g, b, 0, s, p, F->ctx); break;
    case ST_FLAT: hipLaunchKernelGGL(K_flat, g, b, 0, s, p, F->ctx); break;
    case ST_TRACE: hipLaunchKernelGGL(K_trace, g, b, 0, s, p, F->ctx); break;
    case ST_FLUX1: hipLaunchKernelGGL(K_flux1, g, b, 0, s, p, F->ctx); break;
    case ST_TRANS1: hipLaunchKernelGGL(K_trans1, g, b, trans1_tile_bytes(p.bx, p.by, p.bz), s, p, F->ctx); break;
    case ST_FINAL: hipLaunchKernelGGL(K_final, g, b, 0, s, p, F->ctx); break;
    case ST_UPDATE:
        if (F->defer_post) hipLaunchKernelGGL(K_update_adv, g, b, 0, s, p, F->ctx);
        else hipLaunchKernelGGL(K_update, g, b, 0, s, p, F->ctx);
        break;
    }
}

void fill(double *p, long n, double v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(K_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

FusedLevel *fused_create_set(const std::vector<FusedInstanceSpec> &spec, int slab)
{
    FusedLevel *F = new FusedLevel();
    F->ninst = (int)spec.size();
    int nzmax = 0;
    for (const auto &sp : spec) nzmax = std::max(nzmax, sp.n[2]);
    // slab planes: 0 (or a slab at least as deep as every instance) runs the
    // face stages over whole instances
    F->slab = (slab > 0 && slab < nzmax) ? slab : 0;
    F->nslab = F->slab ? (nzmax + F->slab - 1) / F->slab : 1;
    long offD = 0, offE = 0, offQ = 0;
    for (const auto &sp : spec) {
        FInst I;
        for (int d = 0; d < 3; ++d) {
            I.n[d] = sp.n[d];
            I.dx[d] = sp.dx[d];
            I.ext[d] = sp.n[d] + 2;
        }
        if (F->slab) I.ext[2] = std::min(F->slab, sp.n[2]) + 2;  // one slab and its ring
        I.zlo = 0;
        I.zn = sp.n[2];
        I.periodic = sp.bc == 0;
        I.nD = (long)I.n[0] * I.n[1] * I.n[2];
        I.nE = (long)I.ext[0] * I.ext[1] * I.ext[2];
        I.nE += I.nE & 1;  // even: the species pairs (SP2) stay 16-byte aligned
        if (I.nE >= (1L << 28)) {  // 16-byte lane offsets of the face pairs are 32-bit
            set_error("fused engine: an instance may hold at most 2^28 extended cells (run it in z-slabs)");
            delete F;
            return nullptr;
        }
        if (I.nD >= (1L << 29)) {  // 8-byte lane offsets of the cell arrays are 32-bit
            set_error("fused engine: an instance may hold at most 2^29 cells");
            delete F;
            return nullptr;
        }
        I.nQ = I.nD + (I.nD & 1);  // even, like nE
        I.offD = offD;
        I.offE = offE;
        I.offQ = offQ;
        offD += I.nD;
        offE += I.nE;
        offQ += I.nQ;
        F->hinst.push_back(I);
    }
    // one instance table per slab: cells [s * slab, (s + 1) * slab) in z,
    // clipped to each instance (an instance may have no cells in a slab)
    F->hsinst.assign(F->nslab, F->hinst);
    for (int sl = 0; sl < F->nslab && F->slab; ++sl)
        for (FInst &I : F->hsinst[sl]) {
            I.zlo = sl * F->slab;
            I.zn = std::max(0, std::min(F->slab, I.n[2] - I.zlo));
        }
    F->totD = offD;
    F->totE = offE;
    F->totQ = offQ;
    for (int s = 0; s < NSTAGE; ++s) F->tiling[s] = KernelTiling{64, 2, 2};
    F->tiling[ST_EST] = F->tiling[ST_PRIM] = F->tiling[ST_UPDATE] = KernelTiling{64, 4, 1};
    F->tiling[ST_TFLUX] = F->tiling[ST_TRANS1C] = F->tiling[ST_FINALC] = KernelTiling{16, 16, 64};
    bool ok = !check_hip(hipMalloc(&F->inst, sizeof(FInst) * F->ninst), "hipMalloc(inst)") &&
              !check_hip(hipMemcpy(F->inst, F->hinst.data(), sizeof(FInst) * F->ninst, hipMemcpyHostToDevice), "h2d");
    F->sinst.assign(F->nslab, nullptr);
    if (!F->slab) {
        F->sinst[0] = F->inst;
    } else {
        for (int sl = 0; sl < F->nslab && ok; ++sl)
            ok = !check_hip(hipMalloc(&F->sinst[sl], sizeof(FInst) * F->ninst), "hipMalloc(slab inst)") &&
                 !check_hip(hipMemcpy(F->sinst[sl], F->hsinst[sl].data(), sizeof(FInst) * F->ninst,
                                      hipMemcpyHostToDevice), "h2d");
    }
    auto A = [&](double *&p, long n) {
        if (ok && check_hip(hipMalloc(&p, sizeof(double) * std::max(n, 1L)), "hipMalloc(fused)")) ok = false;
    };
    Ctx &x = F->ctx;
    memset(&x, 0, sizeof(x));
    x.inst = F->inst;
    A(x.U, NVAR * offD);
    A(x.U1, NU1 * offD);
    A(x.U1X, 2L * NSPEC_PAIRS * offQ);
    A(x.Q, NFQ * offD);
    A(x.QX, 2L * NSPEC_PAIRS * offQ);
    A(x.FL, offE);
    A(x.DV, offE);
    for (int d = 0; d < 3; ++d) {
        A(x.SM[d], NFS * offE);
        A(x.SP[d], NFS * offE);
        A(x.FD[d], NFF * offE);
        A(x.FN[d], NFN * offE);
    }
    for (int c = 0; c < 6; ++c) A(x.FT[c], NFF * offE);
    A(x.dt_cur, F->ninst);
    A(x.dt_est, F->ninst);
    A(x.time, F->ninst);
    if (ok && check_hip(hipMalloc(&x.steps, sizeof(int) * F->ninst), "hipMalloc(steps)")) ok = false;
    if (!ok) {
        fused_destroy(F);
        return nullptr;
    }
    x.tab = eos_table();
    std::vector<int> z(F->ninst, 0);
    (void)hipMemcpy(x.steps, z.data(), sizeof(int) * F->ninst, hipMemcpyHostToDevice);
    fill(x.time, F->ninst, 0.0, nullptr);
    fill(x.dt_cur, F->ninst, 0.0, nullptr);
    fill(x.dt_est, F->ninst, DBL_MAX, nullptr);
    (void)hipDeviceSynchronize();
    return F;
}

FusedLevel *fused_create(const std::vector<Box> &boxes, const int *n_cell, int bc, const double *dx, int slab)
{
    FusedInstanceSpec sp;
    for (int d = 0; d < 3; ++d) {
        sp.n[d] = n_cell[d];
        sp.dx[d] = dx[d];
    }
    sp.bc = bc;
    FusedLevel *F = fused_create_set({sp}, slab);
    if (!F) return nullptr;
    // box arena description for import / export
    std::vector<ArenaBox> ab;
    std::vector<long> bs;
    long off = 0, acc = 0;
    for (const Box &b : boxes) {
        ArenaBox a;
        for (int d = 0; d < 3; ++d) {
            a.lo[d] = b.lo[d];
            a.hi[d] = b.hi[d];
        }
        a.off = off;
        off += box_npts(b) * NVAR;
        ab.push_back(a);
        bs.push_back(acc);
        acc += box_npts(b);
    }
    bs.push_back(acc);
    F->nabox = (int)boxes.size();
    F->arena_n = off;
    if (check_hip(hipMalloc(&F->abox, sizeof(ArenaBox) * ab.size()), "hipMalloc(abox)") ||
        check_hip(hipMalloc(&F->abstart, sizeof(long) * bs.size()), "hipMalloc(abstart)") ||
        check_hip(hipMemcpy(F->abox, ab.data(), sizeof(ArenaBox) * ab.size(), hipMemcpyHostToDevice), "h2d") ||
        check_hip(hipMemcpy(F->abstart, bs.data(), sizeof(long) * bs.size(), hipMemcpyHostToDevice), "h2d")) {
        fused_destroy(F);
        return nullptr;
    }
    return F;
}

void fused_destroy(FusedLevel *F)
{
    if (!F) return;
    (void)hipDeviceSynchronize();
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    Ctx &x = F->ctx;
    double *ps[] = {x.U, x.U1, x.U1X, x.Q, x.QX, x.FL, x.DV, x.dt_cur, x.dt_est, x.time};
    for (double *p : ps)
        if (p) (void)hipFree(p);
    for (int d = 0; d < 3; ++d) {
        double *pd[] = {x.SM[d], x.SP[d], x.FD[d], x.FN[d]};
        for (double *p : pd)
            if (p) (void)hipFree(p);
    }
    for (int c = 0; c < 6; ++c)
        if (x.FT[c]) (void)hipFree(x.FT[c]);
    if (x.steps) (void)hipFree(x.steps);
    for (auto &pb : F->plans)
        if (pb.blk0) (void)hipFree(pb.blk0);
    if (F->slab)
        for (FInst *p : F->sinst)
            if (p) (void)hipFree(p);
    if (F->inst) (void)hipFree(F->inst);
    if (F->abox) (void)hipFree(F->abox);
    if (F->abstart) (void)hipFree(F->abstart);
    delete F;
}

int fused_import(FusedLevel *F, const double *arena, hipStream_t s)
{
    const long n = F->hinst[0].nD;
    hipLaunchKernelGGL(K_arena, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F->abox, F->nabox, F->abstart,
                       const_cast<double *>(arena), F->ctx.U, n, F->hinst[0].n[0], F->hinst[0].n[1], 0);
    return check_hip(hipGetLastError(), "fused_import");
}

int fused_export(FusedLevel *F, double *arena, hipStream_t s)
{
    const long n = F->hinst[0].nD;
    hipLaunchKernelGGL(K_arena, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F->abox, F->nabox, F->abstart,
                       arena, F->ctx.U, n, F->hinst[0].n[0], F->hinst[0].n[1], 1);
    return check_hip(hipGetLastError(), "fused_export");
}

// ---- intermediate sets
template <class Fn>
static void for_each_intermediate(FusedArrays &S, long totD, long totE, long totQ, Fn f)
{
    f(S.U1, NU1 * totD);
    f(S.U1X, 2L * NSPEC_PAIRS * totQ);
    f(S.Q, NFQ * totD);
    f(S.QX, 2L * NSPEC_PAIRS * totQ);
    f(S.FL, totE);
    f(S.DV, totE);
    for (int d = 0; d < 3; ++d) {
        f(S.SM[d], NFS * totE);
        f(S.SP[d], NFS * totE);
        f(S.FD[d], NFF * totE);
        f(S.FN[d], NFN * totE);
    }
    for (int c = 0; c < 6; ++c) f(S.FT[c], NFF * totE);
}

// device bytes of an engine for these instances: the state, the cell arrays
// and the face arrays (one slab deep in z when slab > 0)
size_t fused_bytes_for(const std::vector<FusedInstanceSpec> &spec, int slab)
{
    size_t cells = 0, ext = 0, q = 0;
    for (const auto &sp : spec) {
        const long nD = (long)sp.n[0] * sp.n[1] * sp.n[2];
        const int ez = (slab > 0 ? std::min(slab, sp.n[2]) : sp.n[2]) + 2;
        long nE = (long)(sp.n[0] + 2) * (sp.n[1] + 2) * ez;
        nE += nE & 1;
        cells += nD;
        ext += nE;
        q += nD + (nD & 1);
    }
    const size_t per_cell = NVAR + NU1 + NFQ, per_q = 2 * 2 * NSPEC_PAIRS;
    const size_t per_ext = 2 + 3 * (2 * NFS + NFF + NFN) + 6 * NFF;
    return sizeof(double) * (per_cell * cells + per_q * q + per_ext * ext);
}

// Whole instances when their arrays take at most half the free device
// memory; else the deepest slab (halving from the deepest instance) whose
// arrays take at most half of it, but no shallower than 8 planes.
int fused_auto_slab(const std::vector<FusedInstanceSpec> &spec)
{
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const size_t budget = freeb / 2;
    if (fused_bytes_for(spec, 0) <= budget) return 0;
    int nz = 0;
    for (const auto &sp : spec) nz = std::max(nz, sp.n[2]);
    for (int slab = nz / 2; slab >= 8; slab /= 2)
        if (fused_bytes_for(spec, slab) <= budget) return slab;
    return 8;
}

int fused_slab(const FusedLevel *F) { return F->slab; }

size_t fused_intermediate_bytes(const FusedLevel *F)
{
    FusedArrays S{};
    size_t n = 0;
    for_each_intermediate(S, F->totD, F->totE, F->totQ, [&](double *&, long c) { n += sizeof(double) * std::max(c, 1L); });
    return n;
}

int fused_alloc_intermediates(const FusedLevel *F, FusedArrays &S)
{
    S = FusedArrays{};
    bool ok = true;
    for_each_intermediate(S, F->totD, F->totE, F->totQ, [&](double *&p, long c) {
        if (ok && check_hip(hipMalloc(&p, sizeof(double) * std::max(c, 1L)), "hipMalloc(intermediates)")) ok = false;
    });
    if (!ok) fused_free_intermediates(S);
    return ok ? 0 : -1;
}

void fused_free_intermediates(FusedArrays &S)
{
    FusedArrays T = S;
    for_each_intermediate(T, 0, 0, 0, [&](double *&p, long) {
        if (p) (void)hipFree(p);
        p = nullptr;
    });
    S = FusedArrays{};
}

FusedArrays fused_get_intermediates(const FusedLevel *F)
{
    const Ctx &x = F->ctx;
    FusedArrays S;
    S.U1 = x.U1;
    S.U1X = x.U1X;
    S.Q = x.Q;
    S.QX = x.QX;
    S.FL = x.FL;
    S.DV = x.DV;
    for (int d = 0; d < 3; ++d) {
        S.SM[d] = x.SM[d];
        S.SP[d] = x.SP[d];
        S.FD[d] = x.FD[d];
        S.FN[d] = x.FN[d];
    }
    for (int c = 0; c < 6; ++c) S.FT[c] = x.FT[c];
    return S;
}

void fused_set_intermediates(FusedLevel *F, const FusedArrays &S)
{
    Ctx &x = F->ctx;
    x.U1 = S.U1;
    x.U1X = S.U1X;
    x.Q = S.Q;
    x.QX = S.QX;
    x.FL = S.FL;
    x.DV = S.DV;
    for (int d = 0; d < 3; ++d) {
        x.SM[d] = S.SM[d];
        x.SP[d] = S.SP[d];
        x.FD[d] = S.FD[d];
        x.FN[d] = S.FN[d];
    }
    for (int c = 0; c < 6; ++c) x.FT[c] = S.FT[c];
    // a captured graph holds the old pointers
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
}

double *fused_state(FusedLevel *F, int m) { return F->ctx.U + (long)NVAR * F->hinst[m].offD; }
long fused_ncell(FusedLevel *F, int m) { return F->hinst[m].nD; }
long fused_next(const FusedLevel *F, int m) { return F->hinst[m].nE; }
int fused_ninst(FusedLevel *F) { return F->ninst; }

// estdt of the current state of every instance into dt_est (no cleaning)
int fused_estimate(FusedLevel *F, hipStream_t s)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    fill(F->ctx.dt_est, F->ninst, DBL_MAX, s);
    launch_stage(F, ST_EST, s);
    F->have_est = true;
    return check_hip(hipGetLastError(), "fused_estimate");
}

// the face stages and the update of one z-slab
static void enqueue_slab(FusedLevel *F, int sl, hipStream_t s)
{
    for (int st : step_stages(F)) launch_stage(F, st, s, stage_ny[st], sl);
}

// enqueue one coarse step of every instance.  K_prim runs over whole
// instances; then, slab after slab, the face stages and K_update: a slab's
// update rewrites only its own cells of the state, which no later stage of
// the step reads (the face stages read the primitives and the cleaned
// state K_prim stored).
static void enqueue_step(FusedLevel *F, double stop_time, hipStream_t s)
{
    F->ctx.tab = eos_table();
    hipLaunchKernelGGL(K_dt, dim3(1), dim3(256), 0, s, F->ctx, F->ninst, stop_time);
    launch_stage(F, ST_PRIM, s);
    for (int sl = 0; sl < F->nslab; ++sl) enqueue_slab(F, sl, s);
    hipLaunchKernelGGL(K_tick, dim3(1), dim3(256), 0, s, F->ctx, F->ninst);
}

// Eager steps with a HIP event pair around every stage launch on stream s;
// stage_ms[st] receives the summed duration of stage st (over its slabs),
// launches[st] the number of launches (the dt kernels are counted under
// FST_N).
int fused_steps_timed(FusedLevel *F, int nsteps, double stop_time, hipStream_t s, double *stage_ms, int *launches)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    if (!F->have_est && fused_estimate(F, s)) return -1;
    const std::vector<int> slab_order = step_stages(F);
    const int nso = (int)slab_order.size();
    const int *ny = stage_ny;
    std::vector<int> stage_of;  // stage of each timed launch of one step (NSTAGE: dt control)
    stage_of.push_back(NSTAGE);
    stage_of.push_back(ST_PRIM);
    for (int sl = 0; sl < F->nslab; ++sl)
        for (int q = 0; q < nso; ++q) stage_of.push_back(slab_order[q]);
    const size_t per = stage_of.size();
    std::vector<hipEvent_t> ev(2 * per * (size_t)nsteps, nullptr);
    int rc = 0;
    for (auto &e : ev)
        if ((rc = check_hip(hipEventCreate(&e), "hipEventCreate"))) break;
    F->ctx.tab = eos_table();
    for (int n = 0; n < nsteps && !rc; ++n) {
        hipEvent_t *E = ev.data() + 2 * per * (size_t)n;
        size_t q = 0;
        (void)hipEventRecord(E[2 * q], s);
        hipLaunchKernelGGL(K_dt, dim3(1), dim3(256), 0, s, F->ctx, F->ninst, stop_time);
        (void)hipEventRecord(E[2 * q + 1], s);
        ++q;
        (void)hipEventRecord(E[2 * q], s);
        launch_stage(F, ST_PRIM, s);
        (void)hipEventRecord(E[2 * q + 1], s);
        ++q;
        for (int sl = 0; sl < F->nslab; ++sl)
            for (int k = 0; k < nso; ++k, ++q) {
                (void)hipEventRecord(E[2 * q], s);
                launch_stage(F, slab_order[k], s, ny[slab_order[k]], sl);
                (void)hipEventRecord(E[2 * q + 1], s);
            }
        hipLaunchKernelGGL(K_tick, dim3(1), dim3(256), 0, s, F->ctx, F->ninst);
    }
    if (!rc) rc = check_hip(hipStreamSynchronize(s), "timed steps");
    for (int st = 0; st <= NSTAGE; ++st) {
        stage_ms[st] = 0.0;
        launches[st] = 0;
    }
    for (int n = 0; n < nsteps && rc == 0; ++n) {
        hipEvent_t *E = ev.data() + 2 * per * (size_t)n;
        for (size_t q = 0; q < per; ++q) {
            float ms = 0.f;
            if (check_hip(hipEventElapsedTime(&ms, E[2 * q], E[2 * q + 1]), "hipEventElapsedTime")) rc = -1;
            stage_ms[stage_of[q]] += ms;
            launches[stage_of[q]] += 1;
        }
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int fused_steps(FusedLevel *F, int nsteps, double stop_time, hipStream_t s, bool use_graph)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    if (!F->have_est && fused_estimate(F, s)) return -1;
    if (!use_graph) {
        for (int n = 0; n < nsteps; ++n) enqueue_step(F, stop_time, s);
        return check_hip(hipGetLastError(), "fused_steps");
    }
    // a captured graph holds the EOS table pointer: capture again when the
    // table was loaded anew since (eos_finalize + a new load may move it)
    if (!F->gexec || F->graph_stop != stop_time || F->graph_tab != eos_table().node ||
        F->graph_fail != eos_table().newton_fail) {
        if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
        if (F->graph) (void)hipGraphDestroy(F->graph);
        F->gexec = nullptr;
        F->graph = nullptr;
        // make sure every plan table exists before capture
        for (int sl = 0; sl < F->nslab; ++sl)
            for (int st = 0; st < NSTAGE; ++st) {
                if (overlap_stage(st)) (void)get_plan3(F, st, sl);
                else (void)get_plan(F, st, sl);
            }
        if (check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture")) return -1;
        enqueue_step(F, stop_time, s);
        if (check_hip(hipStreamEndCapture(s, &F->graph), "end capture")) return -1;
        if (check_hip(hipGraphInstantiate(&F->gexec, F->graph, nullptr, nullptr, 0), "graph instantiate")) return -1;
        F->graph_stop = stop_time;
        F->graph_tab = eos_table().node;
        F->graph_fail = eos_table().newton_fail;
    }
    for (int n = 0; n < nsteps; ++n)
        if (check_hip(hipGraphLaunch(F->gexec, s), "graph launch")) return -1;
    return 0;
}

int fused_advance_dt(FusedLevel *F, double dt, hipStream_t s)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    fill(F->ctx.dt_cur, F->ninst, dt, s);
    fill(F->ctx.dt_est, F->ninst, DBL_MAX, s);
    launch_stage(F, ST_PRIM, s);
    for (int sl = 0; sl < F->nslab; ++sl) enqueue_slab(F, sl, s);
    F->have_est = true;
    return check_hip(hipGetLastError(), "fused_advance_dt");
}

// The stages of a step up to the final fluxes with a given dt: recomputes
// every intermediate array of the step.  K_prim replaces the state by its
// clean in place (the clean(S_old) a step starts from); no later stage of a
// step is run, so the state is not advanced.  The intermediates must cover
// whole instances (no z-slabs).
int fused_recompute_dt(FusedLevel *F, double dt, hipStream_t s)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    if (F->slab) {
        set_error("fused_recompute_dt: the engine runs in z-slabs; a step's intermediates are not kept whole");
        return -1;
    }
    fill(F->ctx.dt_cur, F->ninst, dt, s);
    launch_stage(F, ST_PRIM, s);
    launch_stage(F, ST_FLAT, s);
    for (int st : step_stages(F))
        if (st != ST_UPDATE) launch_stage(F, st, s, stage_ny[st]);
    return check_hip(hipGetLastError(), "fused_recompute_dt");
}

__global__ void K_setclock(Ctx x, int m, double dt, double time, int steps)
{
    x.dt_cur[m] = dt;
    x.time[m] = time;
    x.steps[m] = steps;
}

int fused_set_clock(FusedLevel *F, int m, double dt, double time, int steps, hipStream_t s)
{
    hipLaunchKernelGGL(K_setclock, dim3(1), dim3(1), 0, s, F->ctx, m, dt, time, steps);
    return check_hip(hipGetLastError(), "fused_set_clock");
}

int fused_read_clock(FusedLevel *F, int m, double *dt, double *time, int *steps, hipStream_t s)
{
    if (check_hip(hipMemcpyAsync(dt, F->ctx.dt_cur + m, sizeof(double), hipMemcpyDeviceToHost, s), "d2h") ||
        check_hip(hipMemcpyAsync(time, F->ctx.time + m, sizeof(double), hipMemcpyDeviceToHost, s), "d2h") ||
        check_hip(hipMemcpyAsync(steps, F->ctx.steps + m, sizeof(int), hipMemcpyDeviceToHost, s), "d2h"))
        return -1;
    return check_hip(hipStreamSynchronize(s), "clock sync");
}

int fused_read_est(FusedLevel *F, int m, double *est, hipStream_t s)
{
    if (check_hip(hipMemcpyAsync(est, F->ctx.dt_est + m, sizeof(double), hipMemcpyDeviceToHost, s), "d2h"))
        return -1;
    return check_hip(hipStreamSynchronize(s), "est sync");
}

void fused_set_variant(FusedLevel *F, int flags)
{
    F->variant = flags & FV_ALL;
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
}

int fused_variant(const FusedLevel *F) { return F->variant; }

void fused_set_defer_post(FusedLevel *F, bool defer)
{
    F->defer_post = defer;
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
}

// K_tflux's bricks must fit their primitives' tile (bd + 4 rows of the
// brick's cells across, at most QTILE)
static bool tflux_tile_fits(const KernelTiling &t)
{
    for (int d = 0; d < 3; ++d) {
        const KernelTiling b = ovl_brick(t, d);
        const int bd = d == 0 ? b.bx : (d == 1 ? b.by : b.bz);
        if ((bd + 4) * (b.bx * b.by * b.bz / bd) > QTILE) return false;
    }
    return true;
}

bool fused_tiling_ok(int stage, int bx, int by, int bz)
{
    if (stage < 0 || stage >= NSTAGE) return false;
    if (overlap_stage(stage))
        return ovl_tiling_ok(KernelTiling{bx, by, bz}) &&
               (stage != ST_TFLUX || tflux_tile_fits(KernelTiling{bx, by, bz}));
    if (!(bx > 0 && by > 0 && bz > 0 && (bx * by * bz) % 64 == 0 && bx * by * bz <= 256)) return false;
    // K_trans1's LDS tile stays within a workgroup's 64 KB
    if (stage == ST_TRANS1) return trans1_tile_bytes(bx, by, bz) <= 65536;
    return true;
}

void fused_set_tiling(FusedLevel *F, int stage, int bx, int by, int bz)
{
    if (!fused_tiling_ok(stage, bx, by, bz)) return;
    F->tiling[stage] = KernelTiling{bx, by, bz, F->tiling[stage].zg};
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
}

void fused_get_tiling(FusedLevel *F, int stage, int *bxyz)
{
    bxyz[0] = F->tiling[stage].bx;
    bxyz[1] = F->tiling[stage].by;
    bxyz[2] = F->tiling[stage].bz;
}

bool fused_brick_order_ok(int zg) { return zg == 0 || (zg >= 2 && zg <= 64); }

void fused_set_brick_order(FusedLevel *F, int stage, int zg)
{
    if (stage < 0 || stage >= NSTAGE || !fused_brick_order_ok(zg)) return;
    F->tiling[stage].zg = zg;
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
}

int fused_brick_order(const FusedLevel *F, int stage) { return F->tiling[stage].zg; }

int fused_run_stage(FusedLevel *F, int stage, hipStream_t s)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    const int *ny = stage_ny;
    if (stage < 0 || stage >= NSTAGE) {
        set_error("fused_run_stage: no such stage");
        return -1;
    }
    if (whole_instance_stage(stage)) launch_stage(F, stage, s, ny[stage]);
    else
        for (int sl = 0; sl < F->nslab; ++sl) launch_stage(F, stage, s, ny[stage], sl);
    return check_hip(hipGetLastError(), "fused_run_stage");
}

// Brick candidates of the tiling autotuner: 64, 128 and 256 threads, from
// i-pencils (64,1,1) to cubes (8,8,4).
static const KernelTiling tune_cand[] = {
    {64, 1, 1}, {32, 2, 1}, {16, 4, 1}, {8, 8, 1},  {16, 2, 2}, {128, 1, 1}, {64, 2, 1}, {32, 4, 1},
    {16, 8, 1}, {32, 2, 2}, {16, 4, 2}, {8, 8, 2},  {64, 4, 1}, {64, 2, 2},  {32, 8, 1}, {32, 4, 2},
    {16, 16, 1}, {16, 8, 2}, {16, 4, 4}, {8, 8, 4}, {128, 2, 1}};
// candidates (W, L, X) of the overlapping stages: y bricks (W, L, 1), z
// bricks (W, 1, L), x bricks (X, W*L/X, 1); longer bricks along the
// direction recompute fewer overlap layers, wider ones read longer rows
static const KernelTiling tflux_cand[] = {
    {16, 16, 64}, {32, 8, 64}, {8, 32, 64}, {16, 16, 128}, {32, 8, 128}, {64, 4, 64}, {16, 8, 64},
    {32, 4, 64}, {8, 16, 64}, {16, 16, 256}, {8, 8, 64}};

// Tiling autotuner: times every candidate brick of every stage of every
// kernel variant on the data one coarse step leaves behind and keeps the
// fastest, then picks the variant: K_tflux or K_trace + K_flux1, K_trans1c
// or K_trans1, K_final or K_finalc, whichever is faster.  The stages between
// K_prim and K_update read no array they write, so repeating one of their
// launches recomputes the same outputs (and either kernel of an alternative
// writes the same outputs).  K_prim (its clean) and K_update rewrite the
// state in place, so each of their timed launches starts from the state the
// step started from (K_update's from K_prim's clean of it, as in a step),
// restored by untimed copies.  The state, the clocks and the EOS
// Newton-failure counter are restored afterwards, so results do not change
// (tilings and variants never change results).  bricks (NSTAGE x 3) and
// best_ms (NSTAGE) receive every stage's brick and the tuned stages' time
// per launch; either may be null.
int fused_autotune(FusedLevel *F, int reps, hipStream_t s, int *bricks, double *best_ms)
{
    if (!eos_ready()) {
        set_error("fused engine: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return -1;
    }
    if (reps < 1) reps = 1;
    if (!F->have_est && fused_estimate(F, s)) return -1;
    const int ni = F->ninst;
    double *saveU = nullptr, *saveC = nullptr;
    int *saveS = nullptr;
    unsigned long long *saveN = nullptr;
    unsigned long long *cnt = eos_table().newton_fail;
    const size_t nU = sizeof(double) * NVAR * (size_t)F->totD;
    if (check_hip(hipMalloc(&saveU, nU), "hipMalloc(autotune)") ||
        check_hip(hipMalloc(&saveC, sizeof(double) * 3 * ni), "hipMalloc(autotune)") ||
        check_hip(hipMalloc(&saveS, sizeof(int) * ni), "hipMalloc(autotune)") ||
        check_hip(hipMalloc(&saveN, sizeof(unsigned long long)), "hipMalloc(autotune)")) {
        (void)hipFree(saveU);
        (void)hipFree(saveC);
        (void)hipFree(saveS);
        return -1;
    }
    Ctx &x = F->ctx;
    int rc = check_hip(hipMemcpyAsync(saveU, x.U, nU, hipMemcpyDeviceToDevice, s), "save") ||
             check_hip(hipMemcpyAsync(saveC, x.dt_cur, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "save") ||
             check_hip(hipMemcpyAsync(saveC + ni, x.dt_est, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "save") ||
             check_hip(hipMemcpyAsync(saveC + 2 * ni, x.time, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "save") ||
             check_hip(hipMemcpyAsync(saveS, x.steps, sizeof(int) * ni, hipMemcpyDeviceToDevice, s), "save");
    if (!rc && cnt)
        rc = check_hip(hipMemcpyAsync(saveN, cnt, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s), "save");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = check_hip(hipEventCreate(&e0), "hipEventCreate") || check_hip(hipEventCreate(&e1), "hipEventCreate");
    if (!rc) enqueue_step(F, -1.0, s);  // every intermediate array now holds data of a real step
    const int *ny = stage_ny;
    auto restore_state = [&]() {
        return check_hip(hipMemcpyAsync(x.U, saveU, nU, hipMemcpyDeviceToDevice, s), "autotune restore");
    };
    // every stage of every kernel variant; the faster kernel of each
    // alternative is chosen below
    const std::vector<int> tuned = {ST_PRIM,   ST_FLAT,  ST_TRACE,  ST_FLUX1,  ST_TFLUX,
                                    ST_TRANS1, ST_TRANS1C, ST_FINAL, ST_FINALC, ST_UPDATE};
    double tb[NSTAGE] = {0.0};
    if (bricks)
        for (int st = 0; st < NSTAGE; ++st) fused_get_tiling(F, st, bricks + 3 * st);
    if (best_ms)
        for (int st = 0; st < NSTAGE; ++st) best_ms[st] = 0.0;
    for (int st : tuned) {
        if (rc) break;
        const bool writes_state = st == ST_PRIM || st == ST_UPDATE;
        KernelTiling best = F->tiling[st];
        double tbest = 1e300;
        std::vector<KernelTiling> cands(std::begin(tune_cand), std::end(tune_cand));
        if (overlap_stage(st)) cands.assign(std::begin(tflux_cand), std::end(tflux_cand));
        // reps launches of stage st with tiling c; the time per launch, or a
        // negative value on an error (rc set)
        // the stage a step runs before st (already tuned), launched untimed
        // before each timed launch so that st finds its inputs in the caches
        // as a step leaves them rather than as its own previous launch did
        const int pre = st == ST_TRACE || st == ST_TFLUX ? ST_FLAT
                        : st == ST_FLUX1                ? ST_TRACE
                        : st == ST_TRANS1 || st == ST_TRANS1C ? ST_TFLUX
                        : st == ST_FINAL || st == ST_FINALC   ? ST_TRANS1
                                                              : -1;
        auto time_tiling = [&](const KernelTiling &c) -> double {
            F->tiling[st] = c;
            float ms = 0.f;
            if (!writes_state) {
                launch_stage(F, st, s, ny[st]);  // creates the block table outside the timed region
                for (int r = 0; r < reps && !rc; ++r) {
                    if (pre >= 0) launch_stage(F, pre, s, ny[pre]);
                    (void)hipEventRecord(e0, s);
                    launch_stage(F, st, s, ny[st]);
                    (void)hipEventRecord(e1, s);
                    float one = 0.f;
                    rc = check_hip(hipEventSynchronize(e1), "autotune sync") ||
                         check_hip(hipEventElapsedTime(&one, e0, e1), "hipEventElapsedTime");
                    ms += one;
                }
                if (rc) return -1.0;
            } else {
                (void)get_plan(F, st);  // the block table, outside the timed region
                for (int r = 0; r < reps && !rc; ++r) {
                    if ((rc = restore_state())) break;
                    if (st == ST_UPDATE) {
                        launch_stage(F, ST_PRIM, s);  // clean(S_old), as in a step
                        launch_stage(F, ST_FINAL, s, ny[ST_FINAL]);  // the final fluxes it reads, as in a step
                    }
                    (void)hipEventRecord(e0, s);
                    launch_stage(F, st, s, ny[st]);
                    (void)hipEventRecord(e1, s);
                    float one = 0.f;
                    rc = check_hip(hipEventSynchronize(e1), "autotune sync") ||
                         check_hip(hipEventElapsedTime(&one, e0, e1), "hipEventElapsedTime");
                    ms += one;
                }
                if (rc) return -1.0;
            }
            return ms / reps;
        };
        for (const KernelTiling &c : cands) {
            if (!fused_tiling_ok(st, c.bx, c.by, c.bz)) continue;
            const double t = time_tiling(c);
            if (t < 0) break;
            if (t < tbest) {
                tbest = t;
                best = c;
            }
        }
        // then the brick order of the best brick: groups of brick planes in z
        for (int zg : {2, 4, 8}) {
            if (rc) break;
            KernelTiling c = best;
            c.zg = zg;
            const double t = time_tiling(c);
            if (t >= 0 && t < tbest) {
                tbest = t;
                best = c;
            }
        }
        F->tiling[st] = best;
        tb[st] = tbest;
        if (bricks) {
            bricks[3 * st] = best.bx;
            bricks[3 * st + 1] = best.by;
            bricks[3 * st + 2] = best.bz;
        }
        if (best_ms) best_ms[st] = tbest;
    }
    if (!rc) {
        int v = 0;
        if (tb[ST_TRACE] + tb[ST_FLUX1] < tb[ST_TFLUX]) v |= FV_SPLIT_TRACE;
        if (tb[ST_TRANS1] < tb[ST_TRANS1C]) v |= FV_FACE_TRANS1;
        if (tb[ST_FINALC] < tb[ST_FINAL]) v |= FV_CELL_FINAL;
        F->variant = v;
    }
    // restore the state, the clocks and the Newton-failure counter
    rc = restore_state() ||
         check_hip(hipMemcpyAsync(x.dt_cur, saveC, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "restore") ||
         check_hip(hipMemcpyAsync(x.dt_est, saveC + ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "restore") ||
         check_hip(hipMemcpyAsync(x.time, saveC + 2 * ni, sizeof(double) * ni, hipMemcpyDeviceToDevice, s), "restore") ||
         check_hip(hipMemcpyAsync(x.steps, saveS, sizeof(int) * ni, hipMemcpyDeviceToDevice, s), "restore") ||
         (cnt && check_hip(hipMemcpyAsync(cnt, saveN, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s),
                           "restore")) ||
         check_hip(hipStreamSynchronize(s), "autotune restore") || rc;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(saveU);
    (void)hipFree(saveC);
    (void)hipFree(saveS);
    (void)hipFree(saveN);
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    F->gexec = nullptr;
    F->graph = nullptr;
    return rc ? -1 : 0;
}

} // namespace mc

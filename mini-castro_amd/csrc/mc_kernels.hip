// Per-box CTU hydro kernels for gfx950 behind the reference's C ABI.
//
// Every extern "C" entry point below has the name and argument list of the
// Fortran kernel it replaces (Source/Castro_F.H:11-161) and the
// OpenACC-build semantics (Source/Castro.H:23-24): host-callable, one call
// per (box, lo, hi), arrays are device pointers with Fortran layout, scalars
// by value.  Launches are asynchronous on the stream set by mc_set_stream;
// only estdt synchronises (the reference host reads the reduced dt right
// after the launch, Source/Castro.cpp:205).
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "helmgen.h"
#include "mc_common.h"
#include "mc_eos.h"
#include "mc_hydro.h"
#include "mc_cell.h"
#include "mc_internal.h"
#include "../../include/mc_castro.h"

namespace mc {

// ------------------------------------------------------------------ state
static thread_local hipStream_t t_stream = nullptr;
static EosTable g_eos = {nullptr, nullptr, nullptr, nullptr};
static unsigned long long *g_eos_fail = nullptr;
static double *g_eos_dev = nullptr;
static std::mutex g_mu;
static thread_local char t_err[512];

hipStream_t cur_stream() { return t_stream; }
const EosTable &eos_table() { return g_eos; }
bool eos_ready() { return g_eos.node != nullptr; }

void set_error(const char *msg)
{
    snprintf(t_err, sizeof(t_err), "%s", msg);
    fprintf(stderr, "mini-castro-amd: %s\n", msg);
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    char buf[400];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return -1;
}

// Build the device table from 21 reference columns (Source/eos.F90:82-94)
// and the grids of Source/eos.F90:626-632,699-712.
static int eos_upload(const double *tab)
{
    const long nn = (long)EOS_IMAX * EOS_JMAX;
    const long n_node = nn * EOS_NODE, n_tg = (long)EOS_JMAX * EOS_GRID, n_dg = (long)EOS_IMAX * EOS_GRID;
    std::vector<double> h(n_node + n_tg + n_dg, 0.0);
    static const int col[13] = {MC_HT_F,    MC_HT_FT,   MC_HT_FTT,   MC_HT_FD,    MC_HT_FDD,
                                MC_HT_FDT,  MC_HT_FDDT, MC_HT_FDTT,  MC_HT_FDDTT, MC_HT_DPDF,
                                MC_HT_DPDFT, MC_HT_DPDFD, MC_HT_DPDFDT};
    for (long o = 0; o < nn; ++o)
        for (int c = 0; c < 13; ++c) h[o * EOS_NODE + c] = tab[(long)col[c] * nn + o];
    double *tg = h.data() + n_node, *dg = tg + n_tg;
    const double tstp = (13.0 - 3.0) / 200.0, dstp = (15.0 - (-12.0)) / 540.0;
    std::vector<double> t(EOS_JMAX + 1), d(EOS_IMAX + 1);
    for (int j = 1; j <= EOS_JMAX; ++j) t[j] = pow(10.0, 3.0 + (j - 1) * tstp);
    for (int i = 1; i <= EOS_IMAX; ++i) d[i] = pow(10.0, -12.0 + (i - 1) * dstp);
    for (int j = 1; j <= EOS_JMAX; ++j) {
        double *g = tg + (long)(j - 1) * EOS_GRID;
        g[G_X] = t[j];
        if (j < EOS_JMAX) {
            g[G_D] = t[j + 1] - t[j];
            g[G_D2] = g[G_D] * g[G_D];
            g[G_DI] = 1.0 / g[G_D];
            g[G_D2I] = 1.0 / g[G_D2];
        }
    }
    for (int i = 1; i <= EOS_IMAX; ++i) {
        double *g = dg + (long)(i - 1) * EOS_GRID;
        g[G_X] = d[i];
        if (i < EOS_IMAX) {
            g[G_D] = d[i + 1] - d[i];
            g[G_D2] = g[G_D] * g[G_D];
            g[G_DI] = 1.0 / g[G_D];
            g[G_D2I] = 1.0 / g[G_D2];
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_eos_dev) {
        if (check_hip(hipMalloc(&g_eos_dev, sizeof(double) * h.size()), "hipMalloc(eos table)")) return -1;
    }
    if (check_hip(hipMemcpy(g_eos_dev, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice),
                  "hipMemcpy(eos table)"))
        return -1;
    if (!g_eos_fail) {
        if (check_hip(hipMalloc(&g_eos_fail, sizeof(unsigned long long)), "hipMalloc(newton counter)") ||
            check_hip(hipMemset(g_eos_fail, 0, sizeof(unsigned long long)), "hipMemset(newton counter)"))
            return -1;
    }
    g_eos.node = g_eos_dev;
    g_eos.tg = g_eos_dev + n_node;
    g_eos.dg = g_eos_dev + n_node + n_tg;
    g_eos.newton_fail = g_eos_fail;
    return 0;
}

// ------------------------------------------------------------------ kernels
#define MC_CELL(it)                                                      \
    int i, j, k;                                                         \
    {                                                                    \
        long t_ = (long)logical_block() * blockDim.x + threadIdx.x;      \
        if (!iter_cell(it, t_, i, j, k)) return;                         \
    }

// Source/Castro_hydro_nd.F90:73-170
__global__ __launch_bounds__(kBlock) void k_ctoprim(BoxIter it, Arr u, Arr q, Arr qaux, EosTable tab)
{
    MC_CELL(it)
    double uc[NVAR];
#pragma unroll
    for (int n = 0; n < NVAR; ++n) uc[n] = u(i, j, k, n);
    Prim p;
    ctoprim_cell(uc, tab, p);
    q(i, j, k, QRHO) = p.rho;
    q(i, j, k, QU) = p.u;
    q(i, j, k, QV) = p.v;
    q(i, j, k, QW) = p.w;
    q(i, j, k, QGAME) = p.game;
    q(i, j, k, QPRES) = p.p;
    q(i, j, k, QREINT) = p.rhoe;
    q(i, j, k, QTEMP) = p.T;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) q(i, j, k, QFS + n) = p.x[n];
    qaux(i, j, k, QGAMC) = p.gamc;
    qaux(i, j, k, QC) = p.cs;
    qaux(i, j, k, QDPDR) = p.dpdr_e;
    qaux(i, j, k, QDPDE) = p.dpde;
}

// Source/Castro_hydro_nd.F90:11-69
__global__ __launch_bounds__(kBlock) void k_divu(BoxIter it, Arr q, double dxinv, double dyinv, double dzinv, Arr div)
{
    MC_CELL(it)
    const double ux = FOURTH *
                      (+q(i, j, k, QU) - q(i - 1, j, k, QU) + q(i, j, k - 1, QU) - q(i - 1, j, k - 1, QU) +
                       q(i, j - 1, k, QU) - q(i - 1, j - 1, k, QU) + q(i, j - 1, k - 1, QU) -
                       q(i - 1, j - 1, k - 1, QU)) *
                      dxinv;
    const double vy = FOURTH *
                      (+q(i, j, k, QV) - q(i, j - 1, k, QV) + q(i, j, k - 1, QV) - q(i, j - 1, k - 1, QV) +
                       q(i - 1, j, k, QV) - q(i - 1, j - 1, k, QV) + q(i - 1, j, k - 1, QV) -
                       q(i - 1, j - 1, k - 1, QV)) *
                      dyinv;
    const double wz = FOURTH *
                      (+q(i, j, k, QW) - q(i, j, k - 1, QW) + q(i, j - 1, k, QW) - q(i, j - 1, k - 1, QW) +
                       q(i - 1, j, k, QW) - q(i - 1, j, k - 1, QW) + q(i - 1, j - 1, k, QW) -
                       q(i - 1, j - 1, k - 1, QW)) *
                      dzinv;
    div(i, j, k, 0) = ux + vy + wz;
}

// Source/ppm.F90:309-743 (with uflatten, :11-128)
__global__ __launch_bounds__(kBlock) void k_trace_ppm(BoxIter it, int idir, int vlo_d, int vhi_d, Arr q, Arr qaux,
                                                      Arr qm, Arr qp, double dtdx)
{
    MC_CELL(it)
    const int di = (idir == 1), dj = (idir == 2), dk = (idir == 3);
    const int QUN = idir == 1 ? QU : (idir == 2 ? QV : QW);
    const int QUT = idir == 1 ? QV : (idir == 2 ? QW : QU);
    const int QUTT = idir == 1 ? QW : (idir == 2 ? QU : QV);
    const double cc = qaux(i, j, k, QC);
    const double un = q(i, j, k, QUN);

    double flatn = 1.0;
    for (int d = 0; d < 3; ++d) {
        const int ei = (d == 0), ej = (d == 1), ek = (d == 2);
        const int qv = QU + d;
        double p7[7], u7[7];
#pragma unroll
        for (int m = -3; m <= 3; ++m) {
            p7[m + 3] = q(i + m * ei, j + m * ej, k + m * ek, QPRES);
            u7[m + 3] = q(i + m * ei, j + m * ej, k + m * ek, qv);
        }
        flatn = fmin(flatn, flatten_dir(p7, u7));
    }

    const int pos = idir == 1 ? i : (idir == 2 ? j : k);
    const bool do_p = pos >= vlo_d;
    const bool do_m = pos <= vhi_d;

    // I[w][v]: v = rho, un, p, rhoe; plus transverse velocities under the u wave
    double Ip[3][4], Im[3][4], Ip_ut = 0, Im_ut = 0, Ip_utt = 0, Im_utt = 0;
    const int vars[6] = {QRHO, QUN, QPRES, QREINT, QUT, QUTT};
    for (int vv = 0; vv < 6; ++vv) {
        const int n = vars[vv];
        double s[5], sm, sp;
#pragma unroll
        for (int m = -2; m <= 2; ++m) s[m + 2] = q(i + m * di, j + m * dj, k + m * dk, n);
        ppm_reconstruct(s, flatn, sm, sp);
        const double s6 = ppm_s6(sm, sp, s[2]);
        if (vv < 4) {
            ppm_int_wave(sm, sp, s6, un - cc, dtdx, Ip[0][vv], Im[0][vv]);
            ppm_int_wave(sm, sp, s6, un, dtdx, Ip[1][vv], Im[1][vv]);
            ppm_int_wave(sm, sp, s6, un + cc, dtdx, Ip[2][vv], Im[2][vv]);
        } else if (vv == 4) {
            ppm_int_wave(sm, sp, s6, un, dtdx, Ip_ut, Im_ut);
        } else {
            ppm_int_wave(sm, sp, s6, un, dtdx, Ip_utt, Im_utt);
        }
    }
    double gam_m = 0, gam_p = 0;
    {
        double s[5], sm, sp, a, b;
#pragma unroll
        for (int m = -2; m <= 2; ++m) s[m + 2] = qaux(i + m * di, j + m * dj, k + m * dk, QGAMC);
        ppm_reconstruct(s, flatn, sm, sp);
        const double s6 = ppm_s6(sm, sp, s[2]);
        ppm_int_wave(sm, sp, s6, un - cc, dtdx, a, gam_m);   // Im_gc(1)
        ppm_int_wave(sm, sp, s6, un + cc, dtdx, gam_p, b);   // Ip_gc(3)
        (void)a;
        (void)b;
    }
    const int im = i + di, jm = j + dj, km = k + dk;
    for (int ns = 0; ns < NSPEC; ++ns) {
        const int n = QFS + ns;
        double s[5], sm, sp, ipv, imv;
#pragma unroll
        for (int m = -2; m <= 2; ++m) s[m + 2] = q(i + m * di, j + m * dj, k + m * dk, n);
        ppm_reconstruct(s, flatn, sm, sp);
        const double s6 = ppm_s6(sm, sp, s[2]);
        ppm_int_wave(sm, sp, s6, un, dtdx, ipv, imv);
        if (do_p) qp(i, j, k, n) = imv;
        if (do_m) qm(im, jm, km, n) = ipv;
    }
    if (do_p) {
        double r, u_, e, p;
        ppm_project<true>(Im, gam_m, un, cc, r, u_, e, p);
        qp(i, j, k, QRHO) = r;
        qp(i, j, k, QUN) = u_;
        qp(i, j, k, QREINT) = e;
        qp(i, j, k, QPRES) = p;
        qp(i, j, k, QUT) = Im_ut;
        qp(i, j, k, QUTT) = Im_utt;
    }
    if (do_m) {
        double r, u_, e, p;
        ppm_project<false>(Ip, gam_p, un, cc, r, u_, e, p);
        qm(im, jm, km, QRHO) = r;
        qm(im, jm, km, QUN) = u_;
        qm(im, jm, km, QREINT) = e;
        qm(im, jm, km, QPRES) = p;
        qm(im, jm, km, QUT) = Ip_ut;
        qm(im, jm, km, QUTT) = Ip_utt;
    }
}

__device__ __forceinline__ void load_face(const Arr &a, int i, int j, int k, int iu, int iv1, int iv2, Face &f)
{
    f.rho = a(i, j, k, QRHO);
    f.un = a(i, j, k, iu);
    f.ut = a(i, j, k, iv1);
    f.utt = a(i, j, k, iv2);
    f.p = a(i, j, k, QPRES);
    f.rhoe = a(i, j, k, QREINT);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) f.x[n] = a(i, j, k, QFS + n);
}

// Source/riemann.F90:10-331
__global__ __launch_bounds__(kBlock) void k_compute_flux(BoxIter it, int idir, Arr ql, Arr qr, Arr flx, Arr qint,
                                                         Arr qg, Arr qaux)
{
    MC_CELL(it)
    const int di = (idir == 1), dj = (idir == 2), dk = (idir == 3);
    const int iu = idir == 1 ? QU : (idir == 2 ? QV : QW);
    const int iv1 = idir == 1 ? QV : QU;
    const int iv2 = idir == 3 ? QV : QW;
    const int im1 = idir == 1 ? UMX : (idir == 2 ? UMY : UMZ);
    const int im2 = idir == 1 ? UMY : UMX;
    const int im3 = idir == 3 ? UMY : UMZ;
    Face L, R;
    load_face(ql, i, j, k, iu, iv1, iv2, L);
    load_face(qr, i, j, k, iu, iv1, iv2, R);
    const double cR = qaux(i, j, k, QC), cL = qaux(i - di, j - dj, k - dk, QC);
    const double gR = qaux(i, j, k, QGAMC), gL = qaux(i - di, j - dj, k - dk, QGAMC);
    RiemannOut o;
    riemann_cgf(L, R, cL, cR, gL, gR, o);

    qint(i, j, k, QRHO) = o.g_rho;
    qint(i, j, k, iu) = o.g_un;
    qint(i, j, k, iv1) = o.g_ut;
    qint(i, j, k, iv2) = o.g_utt;
    qint(i, j, k, QPRES) = o.g_p;
    qint(i, j, k, QGAME) = o.g_game;
    qint(i, j, k, QREINT) = o.g_rhoe;
    flx(i, j, k, URHO) = o.f_rho;
    flx(i, j, k, im1) = o.f_mn;
    flx(i, j, k, im2) = o.f_mt;
    flx(i, j, k, im3) = o.f_mtt;
    flx(i, j, k, UEDEN) = o.f_E;
    flx(i, j, k, UEINT) = o.f_e;
    flx(i, j, k, UTEMP) = 0.0;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) {
        qint(i, j, k, QFS + n) = o.gx[n];
        flx(i, j, k, UFS + n) = o.f_rho * o.gx[n];
    }
    // Godunov state in absolute velocity order (GDU = qint(QU), ...)
    double gu[3];
    gu[iu - QU] = o.g_un;
    gu[iv1 - QU] = o.g_ut;
    gu[iv2 - QU] = o.g_utt;
    qg(i, j, k, GDRHO) = o.g_rho;
    qg(i, j, k, GDU) = gu[0];
    qg(i, j, k, GDV) = gu[1];
    qg(i, j, k, GDW) = gu[2];
    qg(i, j, k, GDPRES) = o.g_p;
    qg(i, j, k, GDGAME) = o.g_game;
}


__device__ __forceinline__ void load_qin(const Arr &a, int i, int j, int k, QIn &q)
{
    q.rho = a(i, j, k, QRHO);
    q.u = a(i, j, k, QU);
    q.v = a(i, j, k, QV);
    q.w = a(i, j, k, QW);
    q.p = a(i, j, k, QPRES);
    q.rhoe = a(i, j, k, QREINT);
    q.game = a(i, j, k, QGAME);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) q.x[n] = a(i, j, k, QFS + n);
}

__device__ __forceinline__ void store_qstate(const Arr &a, int i, int j, int k, const QState &o)
{
    a(i, j, k, QRHO) = o.rho;
    a(i, j, k, QU) = o.u;
    a(i, j, k, QV) = o.v;
    a(i, j, k, QW) = o.w;
    a(i, j, k, QPRES) = o.p;
    a(i, j, k, QREINT) = o.rhoe;
    if (o.reset) a(i, j, k, QGAME) = o.game;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) a(i, j, k, QFS + n) = o.x[n];
}

// conserved components a transverse update reads: URHO..UEDEN and species
__device__ __forceinline__ void load_flux(const Arr &f, int i, int j, int k, double *v)
{
#pragma unroll
    for (int n = URHO; n <= UEDEN; ++n) v[n] = f(i, j, k, n);
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) v[UFS + n] = f(i, j, k, UFS + n);
}

// Source/trans.F90:14-243
__global__ __launch_bounds__(kBlock) void k_trans1(BoxIter it, int idir1, int idir2, Arr q2m, Arr q2mo, Arr q2p,
                                                   Arr q2po, Arr qaux, Arr f1, Arr q1, double cdtdx)
{
    MC_CELL(it)
    const int e1i = (idir1 == 1), e1j = (idir1 == 2), e1k = (idir1 == 3);
    const int e2i = (idir2 == 1), e2j = (idir2 == 2), e2k = (idir2 == 3);
    for (int d = -1; d <= 0; ++d) {
        const int il = i + d * e2i, jl = j + d * e2j, kl = k + d * e2k;
        const int ir = il + e1i, jr = jl + e1j, kr = kl + e1k;
        QIn lq;
        load_qin(d == -1 ? q2m : q2p, i, j, k, lq);
        double fr[NVAR], fl[NVAR];
        load_flux(f1, ir, jr, kr, fr);
        load_flux(f1, il, jl, kl, fl);
        const int gdu = GDU + idir1 - 1;
        const QState o = trans1_state(lq, fr, fl, q1(ir, jr, kr, GDPRES), q1(il, jl, kl, GDPRES), q1(ir, jr, kr, gdu),
                                      q1(il, jl, kl, gdu), qaux(il, jl, kl, QGAMC), cdtdx);
        store_qstate(d == -1 ? q2mo : q2po, i, j, k, o);
    }
}

// Source/trans.F90:250-505
__global__ __launch_bounds__(kBlock) void k_trans2(BoxIter it, int idir1, int idir2, int idir3, Arr qm, Arr qmo,
                                                   Arr qp, Arr qpo, Arr qaux, Arr f2, Arr f3, Arr q2, Arr q3,
                                                   double cdtdx2, double cdtdx3)
{
    MC_CELL(it)
    const int e1i = (idir1 == 1), e1j = (idir1 == 2), e1k = (idir1 == 3);
    for (int d = -1; d <= 0; ++d) {
        // the zone on the d side of face (i,j,k) in direction idir1
        const int il = i + d * e1i, jl = j + d * e1j, kl = k + d * e1k;
        int ir2 = il, jr2 = jl, kr2 = kl, ir3 = il, jr3 = jl, kr3 = kl;
        if (idir1 == 1) { jr2 += 1; kr3 += 1; }
        else if (idir1 == 2) { ir2 += 1; kr3 += 1; }
        else { ir2 += 1; jr3 += 1; }
        QIn lq;
        load_qin(d == -1 ? qm : qp, i, j, k, lq);
        double f2r[NVAR], f2l[NVAR], f3r[NVAR], f3l[NVAR];
        load_flux(f2, ir2, jr2, kr2, f2r);
        load_flux(f2, il, jl, kl, f2l);
        load_flux(f3, ir3, jr3, kr3, f3r);
        load_flux(f3, il, jl, kl, f3l);
        const int g2 = GDU + idir2 - 1, g3 = GDU + idir3 - 1;
        const QState o = trans2_state(lq, f2r, f2l, f3r, f3l, q2(ir2, jr2, kr2, GDPRES), q2(il, jl, kl, GDPRES),
                                      q2(ir2, jr2, kr2, g2), q2(il, jl, kl, g2), q3(ir3, jr3, kr3, GDPRES),
                                      q3(il, jl, kl, GDPRES), q3(ir3, jr3, kr3, g3), q3(il, jl, kl, g3),
                                      qaux(il, jl, kl, QGAMC), cdtdx2, cdtdx3);
        store_qstate(d == -1 ? qmo : qpo, i, j, k, o);
    }
}

// Source/Castro_hydro_nd.F90:174-245, keeping the reference's duplicated
// corner terms for idir = 1, 2 (:216-217, :223-224).
__global__ __launch_bounds__(kBlock) void k_apply_av(BoxIter it, int idir, double dxd, Arr div, Arr u, Arr flux)
{
    MC_CELL(it)
    const double difmag = 0.1;
    double div1;
    int im = i, jm = j, km = k;
    if (idir == 1) {
        div1 = FOURTH * (div(i, j, k, 0) + div(i, j + 1, k, 0) + div(i, j, k + 1, 0) + div(i, j + 1, k, 0));
        im = i - 1;
    } else if (idir == 2) {
        div1 = FOURTH * (div(i, j, k, 0) + div(i + 1, j, k, 0) + div(i, j, k + 1, 0) + div(i + 1, j, k, 0));
        jm = j - 1;
    } else {
        div1 = FOURTH * (div(i, j, k, 0) + div(i + 1, j, k, 0) + div(i, j + 1, k, 0) + div(i + 1, j + 1, k, 0));
        km = k - 1;
    }
    div1 = difmag * fmin(0.0, div1);
    for (int n = 0; n < NVAR; ++n) {
        if (n == UTEMP) continue;
        const double d1 = div1 * (u(i, j, k, n) - u(im, jm, km, n));
        flux(i, j, k, n) = flux(i, j, k, n) + dxd * d1;
    }
}

// Source/Castro_hydro_nd.F90:248-297
__global__ __launch_bounds__(kBlock) void k_normalize_species_fluxes(BoxIter it, Arr flux)
{
    MC_CELL(it)
    double sum = 0.0;
#pragma unroll
    for (int n = UFS; n < UFS + NSPEC; ++n) sum = sum + flux(i, j, k, n);
    const double fac = (sum != 0.0) ? flux(i, j, k, URHO) / sum : 1.0;
#pragma unroll
    for (int n = UFS; n < UFS + NSPEC; ++n) flux(i, j, k, n) = flux(i, j, k, n) * fac;
}

// Source/Castro_hydro_nd.F90:301-340
__global__ __launch_bounds__(kBlock) void k_store_flux(BoxIter it, Arr fo, Arr fi, Arr area, double dt)
{
    MC_CELL(it)
    const double a = area(i, j, k, 0);
    for (int n = 0; n < NVAR; ++n) fo(i, j, k, n) = dt * fi(i, j, k, n) * a;
}

// Source/Castro_hydro_nd.F90:344-441
__global__ __launch_bounds__(kBlock) void k_fill_hydro_source(BoxIter it, Arr src, Arr f1, Arr f2, Arr f3, Arr qx,
                                                              Arr qy, Arr qz, Arr a1, Arr a2, Arr a3, Arr vol,
                                                              double dx0, double dx1, double dx2)
{
    MC_CELL(it)
    const double volinv = 1.0 / vol(i, j, k, 0);
    const double a1l = a1(i, j, k, 0), a1r = a1(i + 1, j, k, 0);
    const double a2l = a2(i, j, k, 0), a2r = a2(i, j + 1, k, 0);
    const double a3l = a3(i, j, k, 0), a3r = a3(i, j, k + 1, 0);
    for (int n = 0; n < NVAR; ++n) {
        double s = src(i, j, k, n) + (f1(i, j, k, n) * a1l - f1(i + 1, j, k, n) * a1r + f2(i, j, k, n) * a2l -
                                      f2(i, j + 1, k, n) * a2r + f3(i, j, k, n) * a3l - f3(i, j, k + 1, n) * a3r) *
                                         volinv;
        if (n == UEINT) {
            const double pdivu = HALF * (qx(i + 1, j, k, GDPRES) + qx(i, j, k, GDPRES)) *
                                     (qx(i + 1, j, k, GDU) - qx(i, j, k, GDU)) / dx0 +
                                 HALF * (qy(i, j + 1, k, GDPRES) + qy(i, j, k, GDPRES)) *
                                     (qy(i, j + 1, k, GDV) - qy(i, j, k, GDV)) / dx1 +
                                 HALF * (qz(i, j, k + 1, GDPRES) + qz(i, j, k, GDPRES)) *
                                     (qz(i, j, k + 1, GDW) - qz(i, j, k, GDW)) / dx2;
            s = s - pdivu;
        }
        src(i, j, k, n) = s;
    }
}

// Source/timestep.F90:12-72 with the reduce_min of Source/reduction.F90:115-143
__global__ __launch_bounds__(kBlock) void k_estdt(BoxIter it, Arr u, double dx0, double dx1, double dx2, double *dt,
                                                  EosTable tab)
{
    double v = DBL_MAX;
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (iter_cell(it, t, i, j, k)) {
        double uc[NVAR];
#pragma unroll
        for (int n = 0; n < NVAR; ++n) uc[n] = u(i, j, k, n);
        v = estdt_cell(uc, dx0, dx1, dx2, tab);
    }
    v = block_min<kBlock>(v);
    if (threadIdx.x == 0) atomic_min_d(dt, v);
}

#define MC_POINTWISE_KERNEL(kname, body)                                                  \
    __global__ __launch_bounds__(kBlock) void kname(BoxIter it, Arr u, EosTable tab)      \
    {                                                                                     \
        MC_CELL(it)                                                                       \
        double uc[NVAR];                                                                  \
        _Pragma("unroll") for (int n = 0; n < NVAR; ++n) uc[n] = u(i, j, k, n);           \
        body;                                                                             \
        _Pragma("unroll") for (int n = 0; n < NVAR; ++n) u(i, j, k, n) = uc[n];           \
    }

// Source/Castro_nd.F90:69-125, 129-160, 164-249, 253-297
MC_POINTWISE_KERNEL(k_enforce_minimum_density, enforce_minimum_density_cell(uc, tab))
MC_POINTWISE_KERNEL(k_normalize_species, (void)tab; normalize_species_cell(uc))
MC_POINTWISE_KERNEL(k_reset_internal_e, reset_internal_e_cell(uc, tab))
MC_POINTWISE_KERNEL(k_compute_temp, compute_temp_cell(uc, tab))

// Source/Prob.F90:19-129 with the problem constants as arguments
struct ProbDev {
    double dens_ambient, e_ambient, e_exp, r_init;
    double problo[3], dx[3], center[3];
    int nsub;
};

__global__ __launch_bounds__(kBlock) void k_initdata(BoxIter it, Arr u, ProbDev P)
{
    MC_CELL(it)
    const int nsub = P.nsub;
    const double xmin = P.problo[0] + P.dx[0] * (double)i;
    const double ymin = P.problo[1] + P.dx[1] * (double)j;
    const double zmin = P.problo[2] + P.dx[2] * (double)k;
    int npert = 0, nambient = 0;
    const double r2 = P.r_init * P.r_init;
    for (int kk = 0; kk < nsub; ++kk) {
        const double zz = zmin + (P.dx[2] / (double)nsub) * (kk + 0.5);
        for (int jj = 0; jj < nsub; ++jj) {
            const double yy = ymin + (P.dx[1] / (double)nsub) * (jj + 0.5);
            for (int ii = 0; ii < nsub; ++ii) {
                const double xx = xmin + (P.dx[0] / (double)nsub) * (ii + 0.5);
                const double dist = (P.center[0] - xx) * (P.center[0] - xx) + (P.center[1] - yy) * (P.center[1] - yy) +
                                    (P.center[2] - zz) * (P.center[2] - zz);
                if (dist <= r2) ++npert;
                else ++nambient;
            }
        }
    }
    const double eint = (npert * P.e_exp + nambient * P.e_ambient) / (double)(nsub * nsub * nsub);
    const double rho = P.dens_ambient;
    u(i, j, k, URHO) = rho;
    u(i, j, k, UMX) = 0.0;
    u(i, j, k, UMY) = 0.0;
    u(i, j, k, UMZ) = 0.0;
    u(i, j, k, UTEMP) = 1.e9;
    u(i, j, k, UEDEN) = rho * eint;
    u(i, j, k, UEINT) = rho * eint;
#pragma unroll
    for (int n = 0; n < NSPEC; ++n) u(i, j, k, UFS + n) = 0.0;
    u(i, j, k, UFS) = rho;
}

// e_ambient of Source/Prob.F90:196-204 (one EOS rp solve)
__global__ void k_prob_eamb(double rho, double p, double *e_out, EosTable tab)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *e_out = eos_rp_e(tab, rho, 1.e4, p, 4.0, 2.0, nullptr);
}

// Source/Castro_nd.F90:349-400: per-block partial sums, reduced on the host
__global__ __launch_bounds__(kBlock) void k_blast_radius(BoxIter it, Arr u, double dx0, double dx1, double dx2,
                                                         double lo0, double lo1, double lo2, double c0, double c1,
                                                         double c2, double max_density, double *part)
{
    double m = 0.0, r = 0.0;
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (iter_cell(it, t, i, j, k)) {
        const double x = lo0 + (i + HALF) * dx0 - c0;
        const double y = lo1 + (j + HALF) * dx1 - c1;
        const double z = lo2 + (k + HALF) * dx2 - c2;
        const double rho = u(i, j, k, URHO);
        if (fabs(rho - max_density) / max_density <= 0.1) {
            m = rho;
            r = rho * sqrt(x * x + y * y + z * z);
        }
    }
    __shared__ double sm[kBlock], sr[kBlock];
    sm[threadIdx.x] = m;
    sr[threadIdx.x] = r;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            sr[threadIdx.x] += sr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0];
        part[2 * blockIdx.x + 1] = sr[0];
    }
}

// Source/Castro_nd.F90:301-345
__global__ __launch_bounds__(kBlock) void k_denerror(BoxIter it, int8_t *tag, int tlo0, int tlo1, int tlo2, int tnx,
                                                     int tny, Arr den, int8_t set)
{
    MC_CELL(it)
    const double d0 = den(i, j, k, 0);
    double ax = fabs(den(i + 1, j, k, 0) - d0);
    double ay = fabs(den(i, j + 1, k, 0) - d0);
    double az = fabs(den(i, j, k + 1, 0) - d0);
    ax = fmax(ax, fabs(d0 - den(i - 1, j, k, 0)));
    ay = fmax(ay, fabs(d0 - den(i, j - 1, k, 0)));
    az = fmax(az, fabs(d0 - den(i, j, k - 1, 0)));
    if (fmax(ax, fmax(ay, az)) >= fabs(0.25 * d0))
        tag[(long)(i - tlo0) + (long)tnx * ((j - tlo1) + (long)tny * (k - tlo2))] = set;
}


// Batch EOS evaluation for diagnostics and parity tests: mode 1 (rho,T),
// 2 (rho,e), 3 (rho,p).  out[10*c + ...] = T p e cs gam1 dpde dpdr_e dedT iat jat.
__global__ __launch_bounds__(kBlock) void k_eos_batch(int n, int mode, const double *rho, const double *T,
                                                      const double *ev, const double *abar, const double *zbar,
                                                      double *out, EosTable tab)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double *o = out + 10 * (long)c;
    if (mode == 2) {
        const EosRE r = eos_re<true>(tab, rho[c], T[c], ev[c], abar[c], zbar[c]);
        o[0] = r.T; o[1] = r.p; o[2] = ev[c]; o[3] = r.cs; o[4] = r.gam1; o[5] = r.dpde; o[6] = r.dpdr_e;
        o[7] = 0.0; o[8] = r.iat; o[9] = r.jat;
    } else if (mode == 1) {
        EosDen s;
        eos_density(tab, rho[c], abar[c], zbar[c], s);
        EosFull f;
        int jat;
        eos_eval<true>(tab, s, T[c], jat, f);
        o[0] = T[c]; o[1] = f.p; o[2] = f.e; o[3] = sqrt(f.gam1 * f.p / rho[c]); o[4] = f.gam1; o[5] = f.dpde;
        o[6] = f.dpdr_e; o[7] = f.dedT; o[8] = s.iat; o[9] = jat;
    } else {
        double Tf;
        const double e = eos_rp_e(tab, rho[c], T[c], ev[c], abar[c], zbar[c], &Tf);
        o[0] = Tf; o[1] = ev[c]; o[2] = e; o[3] = 0; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = 0; o[8] = 0; o[9] = 0;
    }
}

// Problem constants for initdata (Source/Prob.F90:11-15); the reference
// values are the defaults.
static double g_p_ambient = 1.e21, g_dens_ambient = 1.e4, g_exp_energy = 1.e52, g_r_init = 1.25e8;
static int g_nsub = 10;

// true if p can be written by device code (device or managed memory)
static bool device_accessible(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged || a.isManaged;
}

static thread_local double *t_scratch = nullptr;
static double *scratch_dev(size_t n)
{
    static thread_local size_t cap = 0;
    if (cap < n) {
        if (t_scratch) (void)hipFree(t_scratch);
        t_scratch = nullptr;
        cap = 0;
        if (check_hip(hipMalloc(&t_scratch, n * sizeof(double)), "hipMalloc(scratch)")) return nullptr;
        cap = n;
    }
    return t_scratch;
}

void launch_estdt(const Box &bx, const double *state, const Box &sbox, const double *dx, double *dt_dev,
                  hipStream_t s)
{
    BoxIter it = make_iter(bx);
    if (it.n > 0)
        hipLaunchKernelGGL(k_estdt, dim3(grid_for(it.n)), dim3(kBlock), 0, s, it, make_arr(state, sbox), dx[0],
                           dx[1], dx[2], dt_dev, g_eos);
}

} // namespace mc

using namespace mc;

#define ARR(p) make_arr(p, p##_lo, p##_hi)
#define REQUIRE_EOS()                                                          \
    do {                                                                       \
        if (!eos_ready()) {                                                    \
            set_error("EOS table not initialised (call eos_init first)");     \
            return;                                                            \
        }                                                                      \
    } while (0)

extern "C" {

// ---- extensions (no counterpart in the reference) ----
void mc_set_stream(void *stream) { t_stream = (hipStream_t)stream; }
void *mc_get_stream(void) { return (void *)t_stream; }
const char *mc_last_error(void) { return t_err; }

int mc_eos_init_arrays(const double *tab21) { return eos_upload(tab21); }

int mc_eos_init_from(const char *path)
{
    std::vector<double> tab((size_t)MC_HELM_NTAB * MC_HELM_IMAX * MC_HELM_JMAX);
    int rc = mc_helm_read(tab.data(), path);
    if (rc != 0) {
        char buf[300];
        snprintf(buf, sizeof(buf), "eos_init: failed to read %s (%d)", path, rc);
        set_error(buf);
        return -1;
    }
    return eos_upload(tab.data());
}

int mc_eos_init_generated(void)
{
    std::vector<double> tab((size_t)MC_HELM_NTAB * MC_HELM_IMAX * MC_HELM_JMAX);
    mc_helm_fill(tab.data());
    return eos_upload(tab.data());
}

int mc_eos_eval_batch(int n, int mode, const double *rho, const double *T, const double *ev, const double *abar,
                      const double *zbar, double *out)
{
    if (!eos_ready()) {
        set_error("EOS table not initialised");
        return -1;
    }
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_eos_batch, dim3(grid_for(n)), dim3(kBlock), 0, t_stream, n, mode, rho, T, ev, abar, zbar,
                       out, g_eos);
    return check_hip(hipGetLastError(), "k_eos_batch");
}

void mc_set_problem(double p_ambient, double dens_ambient, double exp_energy, double r_init, int nsub)
{
    g_p_ambient = p_ambient;
    g_dens_ambient = dens_ambient;
    g_exp_energy = exp_energy;
    g_r_init = r_init;
    g_nsub = nsub;
}

// ---- Source/Castro_F.H:11-17 ----
void network_init(void) {}
void network_finalize(void) {}

// eos_init reads helm_table.dat from the working directory, as the
// reference does (Source/eos.F90:637-641); failure is reported loudly.
void eos_init(void)
{
    if (mc_eos_init_from("helm_table.dat") != 0) {
        fprintf(stderr, "eos_init: Failed to open helm_table.dat\n");
        abort();
    }
}

void eos_finalize(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_eos_dev) (void)hipFree(g_eos_dev);
    if (g_eos_fail) (void)hipFree(g_eos_fail);
    g_eos_dev = nullptr;
    g_eos_fail = nullptr;
    g_eos = EosTable{nullptr, nullptr, nullptr, nullptr};
}

int mc_eos_newton_failures(unsigned long long *count, int reset)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_eos_fail) {
        set_error("mc_eos_newton_failures: EOS table not initialised");
        return -1;
    }
    if (check_hip(hipDeviceSynchronize(), "newton counter sync")) return -1;
    unsigned long long v = 0;
    if (check_hip(hipMemcpy(&v, g_eos_fail, sizeof(v), hipMemcpyDeviceToHost), "newton counter d2h")) return -1;
    if (count) *count = v;
    if (reset && check_hip(hipMemset(g_eos_fail, 0, sizeof(v)), "newton counter reset")) return -1;
    return 0;
}

// ---- Source/Castro_F.H:19-161 ----
void ctoprim(const int *lo, const int *hi, const double *u, const int *u_lo, const int *u_hi, double *q,
             const int *q_lo, const int *q_hi, double *qaux, const int *qa_lo, const int *qa_hi)
{
    REQUIRE_EOS();
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_ctoprim, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(u), ARR(q), make_arr(qaux, qa_lo, qa_hi),
                       g_eos);
}

void enforce_minimum_density(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi)
{
    REQUIRE_EOS();
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_enforce_minimum_density, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state),
                       g_eos);
}

void normalize_species(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_normalize_species, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state), g_eos);
}

void reset_internal_e(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi)
{
    REQUIRE_EOS();
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_reset_internal_e, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state), g_eos);
}

void compute_temp(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi)
{
    REQUIRE_EOS();
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_compute_temp, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state), g_eos);
}

// dt: min-reduced in place.  Device or managed memory is updated by the
// kernel; host memory goes through a device scratch word.  Either way the
// call returns after the stream has drained, like the reference's host
// read of the managed dt_loc (Source/Castro.cpp:184-205).
void estdt(const int *lo, const int *hi, const double *state, const int *state_lo, const int *state_hi,
           const double *dx, double *dt)
{
    REQUIRE_EOS();
    BoxIter it = make_iter(make_box(lo, hi));
    const bool dev = device_accessible(dt);
    double *target = dt;
    if (!dev) {
        target = scratch_dev(1);
        if (!target) return;
        if (check_hip(hipMemcpyAsync(target, dt, sizeof(double), hipMemcpyHostToDevice, t_stream), "estdt h2d"))
            return;
    }
    if (it.n > 0)
        hipLaunchKernelGGL(k_estdt, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state), dx[0], dx[1],
                           dx[2], target, g_eos);
    if (!dev) (void)check_hip(hipMemcpyAsync(dt, target, sizeof(double), hipMemcpyDeviceToHost, t_stream), "estdt d2h");
    (void)check_hip(hipStreamSynchronize(t_stream), "estdt sync");
}

void store_flux(const int *lo, const int *hi, double *flux_out, const int *flux_out_lo, const int *flux_out_hi,
                double *flux_in, const int *flux_in_lo, const int *flux_in_hi, double *area, const int *area_lo,
                const int *area_hi, const double dt)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_store_flux, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(flux_out),
                       ARR(flux_in), ARR(area), dt);
}

void fill_hydro_source(const int *lo, const int *hi, const double *state, const int *state_lo, const int *state_hi,
                       const double *q, const int *q_lo, const int *q_hi, double *source, const int *source_lo,
                       const int *source_hi, const double *flux0, const int *flux0_lo, const int *flux0_hi,
                       const double *flux1, const int *flux1_lo, const int *flux1_hi, const double *flux2,
                       const int *flux2_lo, const int *flux2_hi, const double *qe0, const int *qe0_lo,
                       const int *qe0_hi, const double *qe1, const int *qe1_lo, const int *qe1_hi, const double *qe2,
                       const int *qe2_lo, const int *qe2_hi, const double *area0, const int *area0_lo,
                       const int *area0_hi, const double *area1, const int *area1_lo, const int *area1_hi,
                       const double *area2, const int *area2_lo, const int *area2_hi, const double *volume,
                       const int *volume_lo, const int *volume_hi, const double *dx, const double dt)
{
    (void)state; (void)state_lo; (void)state_hi; (void)q; (void)q_lo; (void)q_hi; (void)dt;
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_fill_hydro_source, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(source),
                       ARR(flux0), ARR(flux1), ARR(flux2), ARR(qe0), ARR(qe1), ARR(qe2), ARR(area0), ARR(area1),
                       ARR(area2), ARR(volume), dx[0], dx[1], dx[2]);
}

void normalize_species_fluxes(const int *lo, const int *hi, double *flux, const int *flux_lo, const int *flux_hi)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_normalize_species_fluxes, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(flux));
}

void apply_av(const int *lo, const int *hi, const int idir, const double *dx, double *div, const int *div_lo,
              const int *div_hi, double *Sborder, const int *Sborder_lo, const int *Sborder_hi, double *flux,
              const int *flux_lo, const int *flux_hi)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_apply_av, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, idir, dx[idir - 1], ARR(div),
                       ARR(Sborder), ARR(flux));
}

void trans1(const int *lo, const int *hi, const int idir1, const int idir2, const double *q2m, const int *q2m_lo,
            const int *q2m_hi, double *q2mo, const int *q2mo_lo, const int *q2mo_hi, const double *q2p,
            const int *q2p_lo, const int *q2p_hi, double *q2po, const int *q2po_lo, const int *q2po_hi,
            const double *qaux, const int *qaux_lo, const int *qaux_hi, const double *f1, const int *f1_lo,
            const int *f1_hi, const double *q1, const int *q1_lo, const int *q1_hi, const double cdtdx)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_trans1, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, idir1, idir2, ARR(q2m),
                       ARR(q2mo), ARR(q2p), ARR(q2po), ARR(qaux), ARR(f1), ARR(q1), cdtdx);
}

void trans2(const int *lo, const int *hi, const int idir1, const int idir2, const int idir3, const double *qm,
            const int *qm_lo, const int *qm_hi, double *qmo, const int *qmo_lo, const int *qmo_hi, const double *qp,
            const int *qp_lo, const int *qp_hi, double *qpo, const int *qpo_lo, const int *qpo_hi,
            const double *qaux, const int *qaux_lo, const int *qaux_hi, const double *fyz, const int *fyz_lo,
            const int *fyz_hi, const double *fzy, const int *fzy_lo, const int *fzy_hi, const double *qy,
            const int *qy_lo, const int *qy_hi, const double *qz, const int *qz_lo, const int *qz_hi,
            const double hdtdx, const double cdtdy, const double cdtdz)
{
    (void)hdtdx;
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_trans2, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, idir1, idir2, idir3, ARR(qm),
                       ARR(qmo), ARR(qp), ARR(qpo), ARR(qaux), ARR(fyz), ARR(fzy), ARR(qy), ARR(qz), cdtdy, cdtdz);
}

void compute_flux(const int *lo, const int *hi, const double *ql, const int *ql_lo, const int *ql_hi,
                  const double *qr, const int *qr_lo, const int *qr_hi, double *flux, const int *flux_lo,
                  const int *flux_hi, double *q_int, const int *q_int_lo, const int *q_int_hi, double *qe,
                  const int *qe_lo, const int *qe_hi, const double *qaux, const int *qaux_lo, const int *qaux_hi,
                  const int idir)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_compute_flux, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, idir, ARR(ql), ARR(qr),
                       ARR(flux), ARR(q_int), ARR(qe), ARR(qaux));
}

void divu(const int *lo, const int *hi, double *q, const int *q_lo, const int *q_hi, const double *dx, double *div,
          const int *div_lo, const int *div_hi)
{
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_divu, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(q), 1.0 / dx[0], 1.0 / dx[1],
                       1.0 / dx[2], ARR(div));
}

void trace_ppm(const int *lo, const int *hi, const int *vlo, const int *vhi, const int idir, const double *q,
               const int *q_lo, const int *q_hi, const double *qaux, const int *qaux_lo, const int *qaux_hi,
               double *qm, const int *qm_lo, const int *qm_hi, double *qp, const int *qp_lo, const int *qp_hi,
               const int *domlo, const int *domhi, const double *dx, const double dt)
{
    (void)domlo;
    (void)domhi;
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_trace_ppm, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, idir, vlo[idir - 1],
                       vhi[idir - 1], ARR(q), ARR(qaux), ARR(qm), ARR(qp), dt / dx[idir - 1]);
}

// ---- next rows: problem set-up and diagnostics ----
void initdata(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi,
              const double *dx, const double *problo, const double *probhi)
{
    const ProblemConsts pc = {g_p_ambient, g_dens_ambient, g_exp_energy, g_r_init, g_nsub};
    initdata_with(pc, lo, hi, state, state_lo, state_hi, dx, problo, probhi);
}
} // extern "C"

namespace mc {
void initdata_with(const ProblemConsts &pc, const int *lo, const int *hi, double *state, const int *state_lo,
                   const int *state_hi, const double *dx, const double *problo, const double *probhi)
{
    if (!eos_ready()) {
        set_error("initdata: EOS table not initialised (call eos_init or mc_eos_init_*)");
        return;
    }
    double *e_dev = scratch_dev(1);
    if (!e_dev) return;
    hipLaunchKernelGGL(k_prob_eamb, dim3(1), dim3(64), 0, t_stream, pc.dens_ambient, pc.p_ambient, e_dev, g_eos);
    double e_amb = 0;
    if (check_hip(hipMemcpyAsync(&e_amb, e_dev, sizeof(double), hipMemcpyDeviceToHost, t_stream), "initdata d2h"))
        return;
    if (check_hip(hipStreamSynchronize(t_stream), "initdata sync")) return;
    ProbDev P;
    P.dens_ambient = pc.dens_ambient;
    P.e_ambient = e_amb;
    const double vctr = FOUR3RD * M_PI_D * (pc.r_init * pc.r_init * pc.r_init);
    P.e_exp = pc.exp_energy / vctr / pc.dens_ambient;
    P.r_init = pc.r_init;
    P.nsub = pc.nsub;
    for (int d = 0; d < 3; ++d) {
        P.problo[d] = problo[d];
        P.dx[d] = dx[d];
        P.center[d] = (problo[d] + probhi[d]) / 2.e0;
    }
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_initdata, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, ARR(state), P);
}
} // namespace mc

extern "C" {

void denerror(const int *lo, const int *hi, int8_t *tag, const int *tag_lo, const int *tag_hi, const double *den,
              const int *den_lo, const int *den_hi, const int8_t tagval, const int8_t clearval)
{
    (void)clearval;
    BoxIter it = make_iter(make_box(lo, hi));
    hipLaunchKernelGGL(k_denerror, dim3(grid_for(it.n)), dim3(kBlock), 0, t_stream, it, tag, tag_lo[0], tag_lo[1],
                       tag_lo[2], tag_hi[0] - tag_lo[0] + 1, tag_hi[1] - tag_lo[1] + 1, ARR(den), tagval);
}

// blast_mass / blast_radius are accumulated (reduce_add semantics) and must
// be host or managed memory; the call synchronises.
void calculate_blast_radius(const int *lo, const int *hi, double *state, const int *state_lo, const int *state_hi,
                            const double *dx, const double *problo, const double *probhi, double *blast_mass,
                            double *blast_radius, const double max_density)
{
    BoxIter it = make_iter(make_box(lo, hi));
    const unsigned nb = grid_for(it.n);
    if (nb == 0) return;
    double *part = scratch_dev(2 * (size_t)nb);
    if (!part) return;
    double c[3];
    for (int d = 0; d < 3; ++d) c[d] = (probhi[d] - problo[d]) / 2.0;
    hipLaunchKernelGGL(k_blast_radius, dim3(nb), dim3(kBlock), 0, t_stream, it, ARR(state), dx[0], dx[1], dx[2],
                       problo[0], problo[1], problo[2], c[0], c[1], c[2], max_density, part);
    std::vector<double> h(2 * (size_t)nb);
    if (check_hip(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, t_stream),
                  "blast d2h"))
        return;
    if (check_hip(hipStreamSynchronize(t_stream), "blast sync")) return;
    double m = 0, r = 0;
    for (unsigned b = 0; b < nb; ++b) {
        m += h[2 * b];
        r += h[2 * b + 1];
    }
    *blast_mass += m;
    *blast_radius += r;
}

} // extern "C"

// ---- allocation-free blast radius ----
// The block partials of calculate_blast_radius and their two sums live in
// `scratch`, a device array the library owns for the life of the process
// (one copy per GPU), instead of memory allocated on the first call.
namespace mc {

constexpr long MC_NOALLOC_MAX_BLOCKS = MC_NOALLOC_MAX_CELLS / kBlock;
static_assert(MC_NOALLOC_MAX_CELLS % kBlock == 0, "whole blocks");
// 256 KiB + 16 B per GPU
static __device__ __align__(256) double scratch[2 * MC_NOALLOC_MAX_BLOCKS + 2];

// Sums the nb block partials in block order, as the host loop of
// calculate_blast_radius does, so both entry points return the same bits.
__global__ void k_blast_sum(const double *part, unsigned nb, double *out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double m = 0, r = 0;
    for (unsigned b = 0; b < nb; ++b) {
        m += part[2 * b];
        r += part[2 * b + 1];
    }
    out[0] = m;
    out[1] = r;
}

// Device address of `scratch` per device, taken once per device with that
// device current (hipGetSymbolAddress); host code uses only these.
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[64];

static double *scratch_ptr()
{
    int dev = 0;
    if (check_hip(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= 64) {
        set_error("mc_calculate_blast_radius_noalloc: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        if (check_hip(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace mc

extern "C" int mc_calculate_blast_radius_noalloc(const int *lo, const int *hi, double *state, const int *state_lo,
                                                 const int *state_hi, const double *dx, const double *problo,
                                                 const double *probhi, double *blast_mass, double *blast_radius,
                                                 const double max_density)
{
    BoxIter it = make_iter(make_box(lo, hi));
    const unsigned nb = grid_for(it.n);
    if (nb == 0) return 0;
    if (nb > MC_NOALLOC_MAX_BLOCKS) {
        char buf[200];
        snprintf(buf, sizeof(buf),
                 "mc_calculate_blast_radius_noalloc: box of %ld cells exceeds the limit of %ld cells", (long)it.n,
                 (long)MC_NOALLOC_MAX_BLOCKS * kBlock);
        set_error(buf);
        return -1;
    }
    double *part = scratch_ptr();
    if (!part) return -1;
    double *sums = part + 2 * (size_t)MC_NOALLOC_MAX_BLOCKS;
    double c[3];
    for (int d = 0; d < 3; ++d) c[d] = (probhi[d] - problo[d]) / 2.0;
    // every partial the sum reads is written by this launch first: nothing to clear
    hipLaunchKernelGGL(k_blast_radius, dim3(nb), dim3(kBlock), 0, t_stream, it, ARR(state), dx[0], dx[1], dx[2],
                       problo[0], problo[1], problo[2], c[0], c[1], c[2], max_density, part);
    hipLaunchKernelGGL(k_blast_sum, dim3(1), dim3(64), 0, t_stream, part, nb, sums);
    double h[2] = {0.0, 0.0};
    if (check_hip(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, t_stream), "blast d2h")) return -1;
    if (check_hip(hipStreamSynchronize(t_stream), "blast sync")) return -1;
    *blast_mass += h[0];
    *blast_radius += h[1];
    return 0;
}

// Level driver: the role of Castro::advance / construct_hydro_source /
// clean_state / estTimeStep on one AMR level of boxes (single GPU).
//
// Plain mode replays the reference's per-box kernel sequence through the
// drop-in C ABI (Source/Castro_hydro.cpp:31-397: 36 launches per box, plus
// the four clean_state kernels at each of its four call sites and estdt).
// Fused mode hands the whole box list to the batched kernels (mc_fused.hip).
//
// Device memory layout: every MultiFab (S_new, S_old, Sborder, hydro
// source) is one arena holding the boxes' FABs back to back, each in
// Fortran order with 20 components; the per-box temporaries of the plain
// path are allocated once for the largest box and reused box after box.
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/mc_castro.h"
#include "mc_common.h"
#include "mc_fused.h"
#include "mc_internal.h"
#include "mc_amr.h"

namespace mc {

static Box grow(Box b, int g0, int g1, int g2)
{
    b.lo[0] -= g0; b.lo[1] -= g1; b.lo[2] -= g2;
    b.hi[0] += g0; b.hi[1] += g1; b.hi[2] += g2;
    return b;
}
static Box grow1(Box b, int g) { return grow(b, g, g, g); }
static Box nodes(Box b, int d) { b.hi[d] += 1; return b; }

struct FabDesc {
    int lo[3], hi[3];
    long off;  // offset in doubles inside an arena
};

// FillPatch of Sborder (grown by 4) from S_old: exact copies of the owning
// box's value; periodic images, or the nearest domain cell (first-order
// extrapolation) outside the domain.
__global__ __launch_bounds__(kBlock) void k_fill_patch(BoxIter it, double *dst, long dst_sn, const double *src_arena,
                                                       const FabDesc *fabs, const int *owner, int gran, int mx,
                                                       int my, int n0, int n1, int n2, int periodic)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(it, t, i, j, k)) return;
    int g0 = i, g1 = j, g2 = k;
    if (periodic) {
        g0 = ((g0 % n0) + n0) % n0;
        g1 = ((g1 % n1) + n1) % n1;
        g2 = ((g2 % n2) + n2) % n2;
    } else {
        g0 = min(max(g0, 0), n0 - 1);
        g1 = min(max(g1, 0), n1 - 1);
        g2 = min(max(g2, 0), n2 - 1);
    }
    const int ob = owner[(g0 / gran) + mx * ((g1 / gran) + my * (g2 / gran))];
    const FabDesc f = fabs[ob];
    const long fx = f.hi[0] - f.lo[0] + 1, fxy = fx * (f.hi[1] - f.lo[1] + 1);
    const long fn = fxy * (f.hi[2] - f.lo[2] + 1);
    const double *s = src_arena + f.off + (g0 - f.lo[0]) + fx * (g1 - f.lo[1]) + fxy * (g2 - f.lo[2]);
#pragma unroll 4
    for (int c = 0; c < NVAR; ++c) dst[t + dst_sn * c] = s[fn * c];
}

// MultiFab::Saxpy(S_new, dt, hydro_source): s = s + dt*h
__global__ __launch_bounds__(kBlock) void k_saxpy(long n, double *s, const double *h, double dt)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) s[t] = s[t] + dt * h[t];
}

// MultiFab::max of one component over a box arena: block max, then a
// compare-and-swap max on *out (values are densities: finite, positive).
__global__ __launch_bounds__(kBlock) void k_max_comp(long nvalid_cells_total, const double *arena, const long *cell0,
                                                     const long *fab_off, const long *fab_n, int nbox, int comp,
                                                     double *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = -DBL_MAX;
    if (t < nvalid_cells_total) {
        int b = 0;
        while (b + 1 < nbox && cell0[b + 1] <= t) ++b;
        v = arena[fab_off[b] + fab_n[b] * comp + (t - cell0[b])];
    }
    __shared__ double red[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) v = fmax(v, red[w]);
        unsigned long long *a = reinterpret_cast<unsigned long long *>(out);
        unsigned long long old = *a;
        while (v > __longlong_as_double(old)) {
            const unsigned long long assumed = old;
            old = atomicCAS(a, assumed, __double_as_longlong(v));
            if (old == assumed) break;
        }
    }
}

__global__ void k_fill(long n, double *p, double v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// the boxes lie in [0, n) and cover it exactly once
static bool tiles_domain(const std::vector<Box> &boxes, const int *n)
{
    long v = 0;
    for (const Box &b : boxes) {
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < 0 || b.hi[d] >= n[d] || b.hi[d] < b.lo[d]) return false;
        v += box_npts(b);
    }
    if (v != (long)n[0] * n[1] * n[2]) return false;
    for (size_t a = 0; a < boxes.size(); ++a)
        for (size_t b = a + 1; b < boxes.size(); ++b) {
            bool apart = false;
            for (int d = 0; d < 3; ++d)
                apart = apart || boxes[a].hi[d] < boxes[b].lo[d] || boxes[b].hi[d] < boxes[a].lo[d];
            if (!apart) return false;
        }
    return true;
}

} // namespace mc

using namespace mc;

struct mc_level {
    int n_cell[3];
    double prob_lo[3], prob_hi[3], dx[3];
    int bc;
    int mode;
    std::vector<Box> boxes;
    std::vector<long> off_valid, off_grown;  // arena offsets (doubles)
    long n_valid = 0, n_grown = 0;
    double *S_new = nullptr, *S_old = nullptr, *Sborder = nullptr, *src = nullptr;
    FabDesc *fabs_dev = nullptr;
    int *owner_dev = nullptr;
    int gran = 1, mx = 0, my = 0;
    // plain-path temporaries, sized for the largest box
    double *q = nullptr, *qaux = nullptr, *divb = nullptr, *q_int = nullptr, *ftmp1 = nullptr, *ftmp2 = nullptr;
    double *qg1 = nullptr, *qg2 = nullptr, *ql = nullptr, *qr = nullptr, *vol = nullptr;
    double *flux[3] = {}, *qe[3] = {}, *area[3] = {}, *fluxes_out[3] = {};
    double *qm[3][3] = {}, *qp[3][3] = {};
    double *dt_dev = nullptr;
    hipStream_t stream = nullptr;
    int steps = 0;
    double time = 0.0, dt = 0.0;
    FusedLevel *fused = nullptr;
    bool fused_ahead = false;  // the fused engine's state is newer than S_new
    // AMR coupling (per-box path): the finer level's register this level
    // initialises with its coarse fluxes, this level's own register it adds
    // its fine fluxes to, and the register's fine box of each level box (-1)
    mc_flux_register *reg_crse = nullptr, *reg_fine = nullptr;
    std::vector<int> fine_index;
    int fused_slab_req = -1;  // z-slab planes of the fused engine: -1 chosen by memory, 0 whole level
    int fused_variant = 0;  // kernel variant flags of the fused engine (FV_*)
    // the fused step stops after the advance's clean_state; post_timestep
    // runs separately (a coarse level of a flux register)
    bool defer_post = false;
};

static double *dalloc(size_t n, bool &ok)
{
    double *p = nullptr;
    if (check_hip(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)), "hipMalloc(level)")) ok = false;
    return p;
}

static void fill_const(hipStream_t s, double *p, long n, double v)
{
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, n, p, v);
}

extern "C" {

mc_level_t *mc_level_create(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc, int nbox,
                            const int *boxes)
{
    if (nbox <= 0) {
        set_error("mc_level_create: no boxes");
        return nullptr;
    }
    mc_level_t *L = new mc_level_t();
    long ncell = 1;
    for (int d = 0; d < 3; ++d) {
        L->n_cell[d] = n_cell[d];
        L->prob_lo[d] = prob_lo[d];
        L->prob_hi[d] = prob_hi[d];
        L->dx[d] = (prob_hi[d] - prob_lo[d]) / (double)n_cell[d];
        ncell *= n_cell[d];
    }
    L->bc = bc;
    L->mode = MC_MODE_PLAIN;
    long vsum = 0;
    int g = 0;
    for (int b = 0; b < nbox; ++b) {
        Box bx;
        for (int d = 0; d < 3; ++d) {
            bx.lo[d] = boxes[6 * b + d];
            bx.hi[d] = boxes[6 * b + 3 + d];
            if (bx.lo[d] < 0 || bx.hi[d] >= n_cell[d] || bx.hi[d] < bx.lo[d]) {
                set_error("mc_level_create: box outside the domain");
                delete L;
                return nullptr;
            }
            g = std::gcd(g, bx.lo[d]);
            g = std::gcd(g, bx.hi[d] + 1);
        }
        L->boxes.push_back(bx);
        L->off_valid.push_back(L->n_valid);
        L->off_grown.push_back(L->n_grown);
        L->n_valid += box_npts(bx) * NVAR;
        L->n_grown += box_npts(grow1(bx, 4)) * NVAR;
        vsum += box_npts(bx);
    }
    if (vsum != ncell || !tiles_domain(L->boxes, n_cell)) {
        set_error("mc_level_create: boxes do not tile the domain exactly once");
        delete L;
        return nullptr;
    }
    L->gran = g > 0 ? g : 1;
    for (int d = 0; d < 3; ++d)
        if (n_cell[d] % L->gran) L->gran = 1;
    L->mx = n_cell[0] / L->gran;
    L->my = n_cell[1] / L->gran;
    const int mz = n_cell[2] / L->gran;
    std::vector<int> owner((size_t)L->mx * L->my * mz, -1);
    std::vector<FabDesc> fd(nbox);
    for (int b = 0; b < nbox; ++b) {
        const Box &bx = L->boxes[b];
        for (int d = 0; d < 3; ++d) {
            fd[b].lo[d] = bx.lo[d];
            fd[b].hi[d] = bx.hi[d];
        }
        fd[b].off = L->off_valid[b];
        for (int k = bx.lo[2] / L->gran; k <= bx.hi[2] / L->gran; ++k)
            for (int j = bx.lo[1] / L->gran; j <= bx.hi[1] / L->gran; ++j)
                for (int i = bx.lo[0] / L->gran; i <= bx.hi[0] / L->gran; ++i) {
                    int &o = owner[i + (size_t)L->mx * (j + (size_t)L->my * k)];
                    if (o != -1) {
                        set_error("mc_level_create: overlapping boxes");
                        delete L;
                        return nullptr;
                    }
                    o = b;
                }
    }
    bool ok = true;
    ok = ok && !check_hip(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking), "hipStreamCreate");
    L->S_new = dalloc(L->n_valid, ok);
    L->S_old = dalloc(L->n_valid, ok);
    L->src = dalloc(L->n_valid, ok);
    L->Sborder = dalloc(L->n_grown, ok);
    L->dt_dev = dalloc(1, ok);
    if (ok) ok = !check_hip(hipMalloc(&L->fabs_dev, sizeof(FabDesc) * nbox), "hipMalloc(fabs)");
    if (ok) ok = !check_hip(hipMalloc(&L->owner_dev, sizeof(int) * owner.size()), "hipMalloc(owner)");
    if (ok) ok = !check_hip(hipMemcpy(L->fabs_dev, fd.data(), sizeof(FabDesc) * nbox, hipMemcpyHostToDevice), "h2d");
    if (ok) ok = !check_hip(hipMemcpy(L->owner_dev, owner.data(), sizeof(int) * owner.size(), hipMemcpyHostToDevice), "h2d");
    if (!ok) {
        mc_level_destroy(L);
        return nullptr;
    }
    return L;
}

static void free_plain(mc_level_t *L)
{
    double **ps[] = {&L->q, &L->qaux, &L->divb, &L->q_int, &L->ftmp1, &L->ftmp2, &L->qg1, &L->qg2, &L->ql, &L->qr,
                     &L->vol};
    for (double **p : ps) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    for (int d = 0; d < 3; ++d) {
        double **pd[] = {&L->flux[d], &L->qe[d], &L->area[d], &L->fluxes_out[d]};
        for (double **p : pd) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        for (int e = 0; e < 3; ++e) {
            if (L->qm[d][e]) (void)hipFree(L->qm[d][e]);
            if (L->qp[d][e]) (void)hipFree(L->qp[d][e]);
            L->qm[d][e] = L->qp[d][e] = nullptr;
        }
    }
}

// Per-box temporaries of Source/Castro_hydro.cpp:54-170 for the largest box.
static int alloc_plain(mc_level_t *L)
{
    if (L->q) return 0;
    long mv = 0;
    int mlen[3] = {0, 0, 0};
    for (const Box &b : L->boxes)
        for (int d = 0; d < 3; ++d) mlen[d] = std::max(mlen[d], b.hi[d] - b.lo[d] + 1);
    Box m;
    for (int d = 0; d < 3; ++d) {
        m.lo[d] = 0;
        m.hi[d] = mlen[d] - 1;
    }
    (void)mv;
    const long nq = box_npts(grow1(m, 4)), no = box_npts(grow1(m, 1));
    bool ok = true;
    L->q = dalloc(nq * QVAR, ok);
    L->qaux = dalloc(nq * NQAUX, ok);
    L->divb = dalloc(no, ok);
    L->q_int = dalloc(no * QVAR, ok);
    L->ftmp1 = dalloc(no * NVAR, ok);
    L->ftmp2 = dalloc(no * NVAR, ok);
    L->qg1 = dalloc(no * NGDNV, ok);
    L->qg2 = dalloc(no * NGDNV, ok);
    L->ql = dalloc(no * QVAR, ok);
    L->qr = dalloc(no * QVAR, ok);
    L->vol = dalloc(nq, ok);
    for (int d = 0; d < 3; ++d) {
        const long ne = box_npts(grow1(nodes(m, d), 1));
        L->flux[d] = dalloc(ne * NVAR, ok);
        L->qe[d] = dalloc(ne * NGDNV, ok);
        L->area[d] = dalloc(box_npts(grow1(nodes(m, d), 4)), ok);
        L->fluxes_out[d] = dalloc(box_npts(nodes(m, d)) * NVAR, ok);
        for (int e = 0; e < 3; ++e) {
            // tbx[d][e] is contained in grow(ebx[d], 1)
            L->qm[d][e] = dalloc(ne * QVAR, ok);
            L->qp[d][e] = dalloc(ne * QVAR, ok);
        }
    }
    if (!ok) {
        free_plain(L);
        return -1;
    }
    // Geometry::GetVolume / GetFaceArea, Cartesian
    const double *dx = L->dx;
    fill_const(L->stream, L->vol, nq, dx[0] * dx[1] * dx[2]);
    const double ar[3] = {dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]};
    for (int d = 0; d < 3; ++d) fill_const(L->stream, L->area[d], box_npts(grow1(nodes(m, d), 4)), ar[d]);
    return check_hip(hipGetLastError(), "alloc_plain fill");
}

void mc_level_destroy(mc_level_t *L)
{
    if (!L) return;
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    if (L->fused) fused_destroy(L->fused);
    free_plain(L);
    double *ps[] = {L->S_new, L->S_old, L->src, L->Sborder, L->dt_dev};
    for (double *p : ps)
        if (p) (void)hipFree(p);
    if (L->fabs_dev) (void)hipFree(L->fabs_dev);
    if (L->owner_dev) (void)hipFree(L->owner_dev);
    if (L->stream) (void)hipStreamDestroy(L->stream);
    delete L;
}

// the level's fused engine, created on first use with the requested z-slabs
static FusedLevel *level_engine(mc_level_t *L)
{
    if (L->fused) return L->fused;
    int slab = L->fused_slab_req;
    if (slab < 0) {
        FusedInstanceSpec sp;
        for (int d = 0; d < 3; ++d) {
            sp.n[d] = L->n_cell[d];
            sp.dx[d] = L->dx[d];
        }
        sp.bc = L->bc;
        slab = fused_auto_slab({sp});
    }
    L->fused = fused_create(L->boxes, L->n_cell, L->bc, L->dx, slab);
    if (L->fused) {
        fused_set_variant(L->fused, L->fused_variant);
        fused_set_defer_post(L->fused, L->defer_post);
    }
    return L->fused;
}

static int sync_from_fused(mc_level_t *L)
{
    if (!L->fused_ahead) return 0;
    if (fused_export(L->fused, L->S_new, L->stream)) return -1;
    L->fused_ahead = false;
    return check_hip(hipStreamSynchronize(L->stream), "fused export sync");
}

// Hand S_new and the step clock to the fused engine.
static int sync_to_fused(mc_level_t *L)
{
    if (!level_engine(L)) return -1;
    if (fused_import(L->fused, L->S_new, L->stream)) return -1;
    if (fused_set_clock(L->fused, 0, L->dt, L->time, L->steps, L->stream)) return -1;
    if (fused_estimate(L->fused, L->stream)) return -1;
    L->fused_ahead = true;
    return 0;
}

int mc_level_set_mode(mc_level_t *L, int mode)
{
    if (mode != MC_MODE_PLAIN && mode != MC_MODE_FUSED) {
        set_error("mc_level_set_mode: unknown mode");
        return -1;
    }
    if (mode == L->mode) return 0;
    if (mode == MC_MODE_PLAIN && sync_from_fused(L)) return -1;
    L->mode = mode;
    if (mode == MC_MODE_FUSED && eos_ready()) return sync_to_fused(L);
    return 0;
}

int mc_level_set_fused_slab(mc_level_t *L, int planes)
{
    if (planes < -1) {
        set_error("mc_level_set_fused_slab: planes must be -1 (by memory), 0 (whole level) or positive");
        return -1;
    }
    if (L->fused) {  // rebuild the engine, keeping the state and the clock
        if (sync_from_fused(L)) return -1;
        fused_destroy(L->fused);
        L->fused = nullptr;
    }
    L->fused_slab_req = planes;
    if (L->mode == MC_MODE_FUSED && eos_ready()) return sync_to_fused(L);
    return 0;
}

int mc_level_fused_slab(mc_level_t *L)
{
    if (!level_engine(L)) return -1;
    return fused_slab(L->fused);
}

int mc_level_set_tiling(mc_level_t *L, int stage, int bx, int by, int bz)
{
    if (!level_engine(L)) return -1;
    if (!fused_tiling_ok(stage, bx, by, bz)) {
        set_error(stage == FST_TFLUX || stage == FST_TRANS1C || stage == FST_FINALC
                      ? "mc_level_set_tiling: the tflux / trans1c / finalc tiling (W, L, X) needs L, X >= 2, W*L a "
                        "multiple of 64 threads, at most 256, and X dividing W*L; tflux also (L + 4) W <= 448 and "
                        "(X + 4) W L / X <= 448 (its LDS tile)"
                      : "mc_level_set_tiling: tile must be a multiple of 64 threads, at most 256; trans1 bricks "
                        "also keep their LDS tile (the brick widened by one cell in two directions, 7 doubles a "
                        "cell) within 64 KB");
        return -1;
    }
    fused_set_tiling(L->fused, stage, bx, by, bz);
    return 0;
}

int mc_level_set_brick_order(mc_level_t *L, int stage, int zgroup)
{
    if (!level_engine(L)) return -1;
    if (stage < 0 || stage >= FST_N || !fused_brick_order_ok(zgroup)) {
        set_error("mc_level_set_brick_order: zgroup must be 0 (raster) or 2..64 brick planes");
        return -1;
    }
    fused_set_brick_order(L->fused, stage, zgroup);
    return 0;
}

int mc_level_brick_order(mc_level_t *L, int stage)
{
    if (!level_engine(L)) return -1;
    if (stage < 0 || stage >= FST_N) return -1;
    return fused_brick_order(L->fused, stage);
}

int mc_level_get_tiling(mc_level_t *L, int stage, int *bxyz)
{
    if (!level_engine(L)) return -1;
    if (stage < 0 || stage >= FST_N) return -1;
    fused_get_tiling(L->fused, stage, bxyz);
    return 0;
}

int mc_level_run_stage(mc_level_t *L, int stage)
{
    if (L->mode != MC_MODE_FUSED || !L->fused) {
        set_error("mc_level_run_stage: fused mode only");
        return -1;
    }
    return fused_run_stage(L->fused, stage, L->stream);
}

int mc_level_set_stage_variant(mc_level_t *L, int flags)
{
    if (flags & ~FV_ALL) {
        set_error("mc_level_set_stage_variant: unknown flags");
        return -1;
    }
    L->fused_variant = flags;
    if (L->fused) fused_set_variant(L->fused, flags);
    return 0;
}

int mc_level_autotune(mc_level_t *L, int reps, int *bricks, double *best_ms)
{
    if (L->mode != MC_MODE_FUSED) {
        set_error("mc_level_autotune: fused mode only");
        return -1;
    }
    if (!eos_ready()) {
        set_error("mc_level_autotune: EOS table not loaded");
        return -1;
    }
    if (!L->fused_ahead && sync_to_fused(L)) return -1;
    if (fused_autotune(L->fused, reps, L->stream, bricks, best_ms)) return -1;
    L->fused_variant = fused_variant(L->fused);  // the variant the autotuner chose
    return 0;
}

int mc_level_stage_variant(mc_level_t *L)
{
    if (!level_engine(L)) return -1;
    return fused_variant(L->fused);
}

int mc_level_nbox(const mc_level_t *L) { return (int)L->boxes.size(); }

double *mc_level_state_ptr(mc_level_t *L, int b)
{
    if (sync_from_fused(L)) return nullptr;
    return L->S_new + L->off_valid[b];
}

int mc_level_set_state(mc_level_t *L, int b, const double *host)
{
    if (b < 0 || b >= (int)L->boxes.size()) return -1;
    const long n = box_npts(L->boxes[b]) * NVAR;
    if (sync_from_fused(L)) return -1;
    if (check_hip(hipMemcpyAsync(L->S_new + L->off_valid[b], host, n * sizeof(double), hipMemcpyHostToDevice,
                                 L->stream), "set_state"))
        return -1;
    if (L->mode == MC_MODE_FUSED && sync_to_fused(L)) return -1;
    return check_hip(hipStreamSynchronize(L->stream), "set_state sync");
}

int mc_level_get_state(mc_level_t *L, int b, double *host)
{
    if (b < 0 || b >= (int)L->boxes.size()) return -1;
    const long n = box_npts(L->boxes[b]) * NVAR;
    if (sync_from_fused(L)) return -1;
    if (check_hip(hipMemcpyAsync(host, L->S_new + L->off_valid[b], n * sizeof(double), hipMemcpyDeviceToHost,
                                 L->stream), "get_state"))
        return -1;
    return check_hip(hipStreamSynchronize(L->stream), "get_state sync");
}

int mc_level_sync(mc_level_t *L) { return check_hip(hipStreamSynchronize(L->stream), "level sync"); }

int mc_level_init_data(mc_level_t *L, double p_ambient, double dens_ambient, double exp_energy, double r_init,
                       int nsub)
{
    if (!cubic_cells(L->dx)) {
        set_error("mc_level_init_data: We don't support dx != dy != dz");
        return -1;
    }
    if (!eos_ready()) {
        set_error("mc_level_init_data: EOS table not initialised");
        return -1;
    }
    const ProblemConsts pc = {p_ambient, dens_ambient, exp_energy, r_init, nsub};
    void *prev = mc_get_stream();
    mc_set_stream(L->stream);
    fill_const(L->stream, L->S_new, L->n_valid, 0.0); /* S_new.setVal(0.), Castro.cpp:93 */
    for (size_t b = 0; b < L->boxes.size(); ++b) {
        const Box &bx = L->boxes[b];
        initdata_with(pc, bx.lo, bx.hi, L->S_new + L->off_valid[b], bx.lo, bx.hi, L->dx, L->prob_lo, L->prob_hi);
    }
    mc_set_stream(prev);
    L->steps = 0;
    L->time = 0.0;
    L->dt = 0.0;
    L->fused_ahead = false;
    if (L->mode == MC_MODE_FUSED && sync_to_fused(L)) return -1;
    return check_hip(hipStreamSynchronize(L->stream), "init_data");
}

// Castro::clean_state over every box of an arena (ng ghost cells).
static void clean_state(mc_level_t *L, double *arena, const std::vector<long> &off, int ng)
{
    for (size_t b = 0; b < L->boxes.size(); ++b) {
        const Box fb = grow1(L->boxes[b], ng);
        double *s = arena + off[b];
        enforce_minimum_density(fb.lo, fb.hi, s, fb.lo, fb.hi);
        normalize_species(fb.lo, fb.hi, s, fb.lo, fb.hi);
        reset_internal_e(fb.lo, fb.hi, s, fb.lo, fb.hi);
        compute_temp(fb.lo, fb.hi, s, fb.lo, fb.hi);
    }
}

static void fill_patch(mc_level_t *L)
{
    for (size_t b = 0; b < L->boxes.size(); ++b) {
        const Box gb = grow1(L->boxes[b], 4);
        BoxIter it = make_iter(gb);
        hipLaunchKernelGGL(k_fill_patch, dim3(grid_for(it.n)), dim3(kBlock), 0, L->stream, it,
                           L->Sborder + L->off_grown[b], (long)it.n, L->S_old, L->fabs_dev, L->owner_dev, L->gran,
                           L->mx, L->my, L->n_cell[0], L->n_cell[1], L->n_cell[2], L->bc == MC_BC_PERIODIC);
    }
}

// Source/Castro_hydro.cpp:31-397 for one box, through the C ABI.
static void hydro_box(mc_level_t *L, int ib, double dt)
{
    const double *dx = L->dx;
    const int domlo[3] = {0, 0, 0};
    const int domhi[3] = {L->n_cell[0] - 1, L->n_cell[1] - 1, L->n_cell[2] - 1};
    const Box bx = L->boxes[ib];
    const Box fb = grow1(bx, 4);
    double *state = L->Sborder + L->off_grown[ib];
    double *source = L->src + L->off_valid[ib];
    const Box obx = grow1(bx, 1), qbx = grow1(bx, 4), volb = grow1(bx, 4);
    Box ebx[3], gebx[3], tbx[3][3], areab[3];
    for (int d = 0; d < 3; ++d) {
        ebx[d] = nodes(bx, d);
        gebx[d] = grow1(ebx[d], 1);
        areab[d] = grow1(nodes(bx, d), 4);
    }
    tbx[0][0] = grow(ebx[0], 0, 1, 1);
    tbx[0][1] = grow(ebx[0], 0, 0, 1);
    tbx[0][2] = grow(ebx[0], 0, 1, 0);
    tbx[1][0] = grow(ebx[1], 0, 0, 1);
    tbx[1][1] = grow(ebx[1], 1, 0, 1);
    tbx[1][2] = grow(ebx[1], 1, 0, 0);
    tbx[2][0] = grow(ebx[2], 0, 1, 0);
    tbx[2][1] = grow(ebx[2], 1, 0, 0);
    tbx[2][2] = grow(ebx[2], 1, 1, 0);
    double *q = L->q, *qaux = L->qaux;

    ctoprim(qbx.lo, qbx.hi, state, fb.lo, fb.hi, q, qbx.lo, qbx.hi, qaux, qbx.lo, qbx.hi);
    divu(obx.lo, obx.hi, q, qbx.lo, qbx.hi, dx, L->divb, obx.lo, obx.hi);

    const double hdtdx[3] = {0.5 * dt / dx[0], 0.5 * dt / dx[1], 0.5 * dt / dx[2]};
    const double cdtdx[3] = {dt / dx[0] / 3.0, dt / dx[1] / 3.0, dt / dx[2] / 3.0};

    for (int d = 0; d < 3; ++d)
        trace_ppm(obx.lo, obx.hi, bx.lo, bx.hi, d + 1, q, qbx.lo, qbx.hi, qaux, qbx.lo, qbx.hi, L->qm[d][d],
                  tbx[d][d].lo, tbx[d][d].hi, L->qp[d][d], tbx[d][d].lo, tbx[d][d].hi, domlo, domhi, dx, dt);

    for (int d = 0; d < 3; ++d) {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        compute_flux(tbx[d][d].lo, tbx[d][d].hi, L->qm[d][d], tbx[d][d].lo, tbx[d][d].hi, L->qp[d][d], tbx[d][d].lo,
                     tbx[d][d].hi, L->ftmp1, obx.lo, obx.hi, L->q_int, obx.lo, obx.hi, L->qg1, obx.lo, obx.hi, qaux,
                     qbx.lo, qbx.hi, d + 1);
        const int ts[2] = {t1, t2};
        for (int s = 0; s < 2; ++s) {
            const int t = ts[s];
            trans1(tbx[t][d].lo, tbx[t][d].hi, d + 1, t + 1, L->qm[t][t], tbx[t][t].lo, tbx[t][t].hi, L->qm[t][d],
                   tbx[t][d].lo, tbx[t][d].hi, L->qp[t][t], tbx[t][t].lo, tbx[t][t].hi, L->qp[t][d], tbx[t][d].lo,
                   tbx[t][d].hi, qaux, qbx.lo, qbx.hi, L->ftmp1, obx.lo, obx.hi, L->qg1, obx.lo, obx.hi, cdtdx[d]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        compute_flux(tbx[t1][t2].lo, tbx[t1][t2].hi, L->qm[t1][t2], tbx[t1][t2].lo, tbx[t1][t2].hi, L->qp[t1][t2],
                     tbx[t1][t2].lo, tbx[t1][t2].hi, L->ftmp1, obx.lo, obx.hi, L->q_int, obx.lo, obx.hi, L->qg1,
                     obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, t1 + 1);
        compute_flux(tbx[t2][t1].lo, tbx[t2][t1].hi, L->qm[t2][t1], tbx[t2][t1].lo, tbx[t2][t1].hi, L->qp[t2][t1],
                     tbx[t2][t1].lo, tbx[t2][t1].hi, L->ftmp2, obx.lo, obx.hi, L->q_int, obx.lo, obx.hi, L->qg2,
                     obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, t2 + 1);
        trans2(ebx[d].lo, ebx[d].hi, d + 1, t1 + 1, t2 + 1, L->qm[d][d], tbx[d][d].lo, tbx[d][d].hi, L->ql, obx.lo,
               obx.hi, L->qp[d][d], tbx[d][d].lo, tbx[d][d].hi, L->qr, obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, L->ftmp1,
               obx.lo, obx.hi, L->ftmp2, obx.lo, obx.hi, L->qg1, obx.lo, obx.hi, L->qg2, obx.lo, obx.hi, hdtdx[d],
               hdtdx[t1], hdtdx[t2]);
        compute_flux(ebx[d].lo, ebx[d].hi, L->ql, obx.lo, obx.hi, L->qr, obx.lo, obx.hi, L->flux[d], gebx[d].lo,
                     gebx[d].hi, L->q_int, obx.lo, obx.hi, L->qe[d], gebx[d].lo, gebx[d].hi, qaux, qbx.lo, qbx.hi,
                     d + 1);
    }
    for (int d = 0; d < 3; ++d) {
        apply_av(ebx[d].lo, ebx[d].hi, d + 1, dx, L->divb, obx.lo, obx.hi, state, fb.lo, fb.hi, L->flux[d],
                 gebx[d].lo, gebx[d].hi);
        normalize_species_fluxes(ebx[d].lo, ebx[d].hi, L->flux[d], gebx[d].lo, gebx[d].hi);
        store_flux(ebx[d].lo, ebx[d].hi, L->fluxes_out[d], ebx[d].lo, ebx[d].hi, L->flux[d], gebx[d].lo, gebx[d].hi,
                   L->area[d], areab[d].lo, areab[d].hi, dt);
    }
    // the flux registers take this box's store_flux output before the next
    // box reuses the buffer (Castro_advance.cpp:82-88 reads the level's
    // fluxes after the advance: the same values)
    for (int d = 0; d < 3; ++d) {
        if (L->reg_crse) (void)flux_register_crse_init(L->reg_crse, L->fluxes_out[d], ebx[d], d, -1.0, L->stream);
        if (L->reg_fine && L->fine_index[ib] >= 0)
            (void)flux_register_fine_add(L->reg_fine, L->fine_index[ib], L->fluxes_out[d], ebx[d], d, 1.0, L->stream);
    }
    fill_hydro_source(bx.lo, bx.hi, state, fb.lo, fb.hi, q, qbx.lo, qbx.hi, source, bx.lo, bx.hi, L->flux[0],
                      gebx[0].lo, gebx[0].hi, L->flux[1], gebx[1].lo, gebx[1].hi, L->flux[2], gebx[2].lo, gebx[2].hi,
                      L->qe[0], gebx[0].lo, gebx[0].hi, L->qe[1], gebx[1].lo, gebx[1].hi, L->qe[2], gebx[2].lo,
                      gebx[2].hi, L->area[0], areab[0].lo, areab[0].hi, L->area[1], areab[1].lo, areab[1].hi,
                      L->area[2], areab[2].lo, areab[2].hi, L->vol, volb.lo, volb.hi, dx, dt);
}

double mc_level_est_time_step(mc_level_t *L)
{
    if (!eos_ready()) {
        set_error("mc_level_est_time_step: EOS table not initialised");
        return -1.0;
    }
    double dt = DBL_MAX;
    if (L->mode == MC_MODE_FUSED) {
        if (!L->fused_ahead && sync_to_fused(L)) return -1.0;
        if (fused_estimate(L->fused, L->stream) || fused_read_est(L->fused, 0, &dt, L->stream)) return -1.0;
        return dt * cfl;
    } else {
        fill_const(L->stream, L->dt_dev, 1, DBL_MAX);
        for (size_t b = 0; b < L->boxes.size(); ++b)
            launch_estdt(L->boxes[b], L->S_new + L->off_valid[b], L->boxes[b], L->dx, L->dt_dev, L->stream);
    }
    if (check_hip(hipMemcpyAsync(&dt, L->dt_dev, sizeof(double), hipMemcpyDeviceToHost, L->stream), "dt d2h") ||
        check_hip(hipStreamSynchronize(L->stream), "dt sync"))
        return -1.0;
    dt *= cfl; /* Castro.cpp:212-213 */
    return dt;
}

double mc_level_initial_dt(mc_level_t *L)
{
    const double init_shrink = 0.01;
    const double e = mc_level_est_time_step(L);
    return e < 0 ? e : init_shrink * e;
}

int mc_level_advance(mc_level_t *L, double dt)
{
    if (!eos_ready()) {
        set_error("mc_level_advance: EOS table not initialised");
        return -1;
    }
    if (L->mode == MC_MODE_FUSED) {
        // the fused update includes post_timestep's clean_state unless the
        // level is the coarse side of a flux register
        if (!L->fused_ahead && sync_to_fused(L)) return -1;
        if (fused_advance_dt(L->fused, dt, L->stream)) return -1;
        // CrseInit / FineAdd (Source/Castro_advance.cpp:82-88) from the
        // step's final fluxes, with store_flux's face areas
        // (Geometry::GetFaceArea)
        const double *dx = L->dx;
        const double ar[3] = {dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]};
        if (L->reg_crse && flux_register_crse_init_fused(L->reg_crse, L->fused, dt, ar, -1.0, L->stream)) return -1;
        if (L->reg_fine && flux_register_fine_add_fused(L->reg_fine, L->fused, dt, ar, 1.0, L->stream)) return -1;
        return 0;
    }
    std::swap(L->S_old, L->S_new); /* swapTimeLevels, Castro_advance.cpp:32-33 */
    if (alloc_plain(L)) return -1;
    void *prev = mc_get_stream();
    mc_set_stream(L->stream);
    clean_state(L, L->S_old, L->off_valid, 0); /* :51 */
    (void)check_hip(hipMemcpyAsync(L->S_new, L->S_old, L->n_valid * sizeof(double), hipMemcpyDeviceToDevice,
                                   L->stream), "Copy"); /* :55 */
    fill_patch(L);                                 /* :62 */
    clean_state(L, L->Sborder, L->off_grown, 4);   /* :66 */
    fill_const(L->stream, L->src, L->n_valid, 0.0); /* hydro_source.setVal(0), Castro_hydro.cpp:17 */
    for (size_t b = 0; b < L->boxes.size(); ++b) hydro_box(L, (int)b, dt); /* :70 */
    hipLaunchKernelGGL(k_saxpy, dim3(grid_for(L->n_valid)), dim3(kBlock), 0, L->stream, L->n_valid, L->S_new, L->src,
                       dt); /* :74 */
    clean_state(L, L->S_new, L->off_valid, 0); /* :78 */
    mc_set_stream(prev);
    return check_hip(hipGetLastError(), "advance");
}

int mc_level_post_timestep(mc_level_t *L)
{
    if (L->mode == MC_MODE_FUSED && !L->defer_post) return 0; /* folded into the fused update */
    // a fused coarse level of a register: the clean runs on S_new, which the
    // engine takes back at the next step
    if (L->mode == MC_MODE_FUSED && sync_from_fused(L)) return -1;
    void *prev = mc_get_stream();
    mc_set_stream(L->stream);
    clean_state(L, L->S_new, L->off_valid, 0); /* Castro.cpp:369 */
    mc_set_stream(prev);
    return check_hip(hipGetLastError(), "post_timestep");
}

int mc_level_run(mc_level_t *L, int nsteps, double stop_time, int use_graph)
{
    if (L->reg_crse || L->reg_fine) {
        set_error("mc_level_run: a level with flux registers steps through mc_level_advance and "
                  "mc_level_post_timestep (the AMR driver's order)");
        return -1;
    }
    if (L->mode == MC_MODE_FUSED) {
        if (!L->fused_ahead && sync_to_fused(L)) return -1;
        if (fused_steps(L->fused, nsteps, stop_time, L->stream, use_graph != 0)) return -1;
        return fused_read_clock(L->fused, 0, &L->dt, &L->time, &L->steps, L->stream);
    }
    for (int n = 0; n < nsteps; ++n)
        if (mc_level_coarse_step(L, stop_time) < 0) return -1;
    return 0;
}

int mc_level_run_timed(mc_level_t *L, int nsteps, double stop_time, double *stage_ms, int *launches)
{
    if (L->mode != MC_MODE_FUSED) {
        set_error("mc_level_run_timed: fused mode only");
        return -1;
    }
    if (L->reg_crse || L->reg_fine) {
        set_error("mc_level_run_timed: not with flux registers attached");
        return -1;
    }
    if (!L->fused_ahead && sync_to_fused(L)) return -1;
    if (fused_steps_timed(L->fused, nsteps, stop_time, L->stream, stage_ms, launches)) return -1;
    return fused_read_clock(L->fused, 0, &L->dt, &L->time, &L->steps, L->stream);
}

double mc_level_coarse_step(mc_level_t *L, double stop_time)
{
    if (L->mode == MC_MODE_FUSED) return mc_level_run(L, 1, stop_time, 0) ? -1.0 : L->dt;
    double dt_0;
    if (L->steps > 0) { /* Castro.cpp:218-295 */
        const double change_max = 1.1e0;
        double dt_min = mc_level_est_time_step(L);
        if (dt_min < 0) return -1.0;
        dt_min = fmin(dt_min, change_max * L->dt);
        dt_0 = fmin(1.0e+100, dt_min);
    } else { /* Castro.cpp:297-341 */
        const double d0 = mc_level_initial_dt(L);
        if (d0 < 0) return -1.0;
        dt_0 = fmin(1.0e+100, d0);
    }
    const double eps = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((L->time + dt_0) > (stop_time - eps)) dt_0 = stop_time - L->time;
    L->dt = dt_0;
    if (mc_level_advance(L, L->dt)) return -1.0;
    if (mc_level_post_timestep(L)) return -1.0;
    L->time += L->dt;
    L->steps += 1;
    return L->dt;
}

// Castro::post_timestep's diagnostic (Source/Castro.cpp:371-424): the
// density maximum of S_new, then calculate_blast_radius over every box.
int mc_level_blast_radius(mc_level_t *L, double *blast_mass, double *blast_radius, double *max_density)
{
    if (sync_from_fused(L)) return -1;
    const int nbox = (int)L->boxes.size();
    std::vector<long> cell0(nbox + 1), off(nbox), nn(nbox);
    long acc = 0;
    for (int b = 0; b < nbox; ++b) {
        cell0[b] = acc;
        nn[b] = box_npts(L->boxes[b]);
        off[b] = L->off_valid[b];
        acc += nn[b];
    }
    cell0[nbox] = acc;
    long *dv = nullptr;
    if (check_hip(hipMalloc(&dv, sizeof(long) * (3 * nbox + 1)), "hipMalloc(max)")) return -1;
    std::vector<long> h(3 * nbox + 1);
    for (int b = 0; b <= nbox; ++b) h[b] = cell0[b];
    for (int b = 0; b < nbox; ++b) {
        h[nbox + 1 + b] = off[b];
        h[2 * nbox + 1 + b] = nn[b];
    }
    double mx = -DBL_MAX;
    int rc = check_hip(hipMemcpyAsync(dv, h.data(), sizeof(long) * h.size(), hipMemcpyHostToDevice, L->stream), "h2d");
    fill_const(L->stream, L->dt_dev, 1, -DBL_MAX);
    if (!rc)
        hipLaunchKernelGGL(k_max_comp, dim3(grid_for(acc)), dim3(kBlock), 0, L->stream, acc, L->S_new, dv,
                           dv + nbox + 1, dv + 2 * nbox + 1, nbox, (int)URHO, L->dt_dev);
    if (!rc) rc = check_hip(hipMemcpyAsync(&mx, L->dt_dev, sizeof(double), hipMemcpyDeviceToHost, L->stream), "d2h");
    if (!rc) rc = check_hip(hipStreamSynchronize(L->stream), "max sync");
    (void)hipFree(dv);
    if (rc) return -1;
    double m = 0.0, r = 0.0;
    void *prev = mc_get_stream();
    mc_set_stream(L->stream);
    for (int b = 0; b < nbox; ++b) {
        const Box &bx = L->boxes[b];
        calculate_blast_radius(bx.lo, bx.hi, L->S_new + L->off_valid[b], bx.lo, bx.hi, L->dx, L->prob_lo,
                               L->prob_hi, &m, &r, mx);
    }
    mc_set_stream(prev);
    *blast_mass = m;
    *blast_radius = r;
    *max_density = mx;
    return check_hip(hipGetLastError(), "blast radius");
}

int mc_level_steps(const mc_level_t *L) { return L->steps; }
double mc_level_time(const mc_level_t *L) { return L->time; }
double mc_level_dt(const mc_level_t *L) { return L->dt; }

// ---- AMR coupling between two levels (per-box path) ----
int mc_level_set_flux_registers(mc_level_t *L, mc_flux_register_t *crse_side, mc_flux_register_t *fine_side)
{

    for (int d = 0; d < 3; ++d) {
        if (crse_side && crse_side->crse_n[d] != L->n_cell[d]) {
            set_error("mc_level_set_flux_registers: the register's coarse domain is not this level's");
            return -1;
        }
        if (fine_side && fine_side->crse_n[d] * fine_side->ratio[d] != L->n_cell[d]) {
            set_error("mc_level_set_flux_registers: the register's refined domain is not this level's");
            return -1;
        }
    }
    L->reg_crse = crse_side;
    L->reg_fine = fine_side;
    // a coarse level's post_timestep clean follows its reflux and
    // average-down (Castro::post_timestep, Source/Castro.cpp:343-369): the
    // fused step of such a level stops after the advance's clean
    L->defer_post = crse_side != nullptr;
    if (L->fused) fused_set_defer_post(L->fused, L->defer_post);
    L->fine_index.assign(L->boxes.size(), -1);
    if (fine_side) {
        for (size_t b = 0; b < L->boxes.size(); ++b)
            for (size_t j = 0; j < fine_side->fine.size(); ++j) {
                bool same = true;
                for (int d = 0; d < 3; ++d)
                    same = same && L->boxes[b].lo[d] == fine_side->fine[j].lo[d] &&
                           L->boxes[b].hi[d] == fine_side->fine[j].hi[d];
                if (same) L->fine_index[b] = (int)j;
            }
        for (size_t j = 0; j < fine_side->fine.size(); ++j)
            if (std::find(L->fine_index.begin(), L->fine_index.end(), (int)j) == L->fine_index.end()) {
                set_error("mc_level_set_flux_registers: a fine box of the register is not a box of this level");
                L->reg_crse = L->reg_fine = nullptr;
                return -1;
            }
    }
    return 0;
}

// Castro::reflux (Source/Castro.cpp:451-486) into this (coarse) level's S_new
int mc_level_reflux(mc_level_t *L, mc_flux_register_t *R)
{
    for (int d = 0; d < 3; ++d)
        if (R->crse_n[d] != L->n_cell[d]) {
            set_error("mc_level_reflux: the register's coarse domain is not this level's");
            return -1;
        }
    // a fused level hands its state to S_new (the engine takes it back at
    // its next step); the cell volume is Geometry::GetVolume's dx dy dz
    if (sync_from_fused(L)) return -1;
    const double vol = L->dx[0] * L->dx[1] * L->dx[2];
    if (flux_register_clear_internal(R, L->stream)) return -1;
    for (size_t b = 0; b < L->boxes.size(); ++b)
        if (flux_register_reflux(R, L->S_new + L->off_valid[b], L->boxes[b], L->boxes[b], nullptr, L->boxes[b],
                                 1.0, L->stream, vol))
            return -1;
    if (flux_register_fill(R, 0.0, L->stream)) return -1;
    return check_hip(hipStreamSynchronize(L->stream), "reflux");
}

// Castro::avgDown (Source/Castro.cpp:488-505): the coarse cells under the
// given fine boxes (all of them when nfine < 0) become the mean of their
// fine cells
int mc_level_average_down(mc_level_t *fine, mc_level_t *crse, const int *ratio, int nfine, const int *fine_box_ids)
{
    // a fused level hands its state to S_new first; a fused coarse level
    // takes the averaged S_new back at its next step
    if (sync_from_fused(fine) || sync_from_fused(crse)) return -1;
    for (int d = 0; d < 3; ++d)
        if (ratio[d] < 1 || crse->n_cell[d] * ratio[d] != fine->n_cell[d]) {
            set_error("mc_level_average_down: the levels' domains do not differ by the ratio");
            return -1;
        }
    const int n = nfine < 0 ? (int)fine->boxes.size() : nfine;
    if (check_hip(hipStreamSynchronize(fine->stream), "average_down: fine level sync")) return -1;
    for (int q = 0; q < n; ++q) {
        const int fb = nfine < 0 ? q : fine_box_ids[q];
        if (fb < 0 || fb >= (int)fine->boxes.size()) {
            set_error("mc_level_average_down: bad fine box index");
            return -1;
        }
        const Box F = fine->boxes[fb];
        Box C;
        for (int d = 0; d < 3; ++d) {
            if (F.lo[d] % ratio[d] || (F.hi[d] + 1) % ratio[d]) {
                set_error("mc_level_average_down: a fine box is not coarsenable by the ratio");
                return -1;
            }
            C.lo[d] = F.lo[d] / ratio[d];
            C.hi[d] = (F.hi[d] + 1) / ratio[d] - 1;
        }
        for (size_t cb = 0; cb < crse->boxes.size(); ++cb) {
            Box is;
            bool any = true;
            for (int d = 0; d < 3; ++d) {
                is.lo[d] = std::max(C.lo[d], crse->boxes[cb].lo[d]);
                is.hi[d] = std::min(C.hi[d], crse->boxes[cb].hi[d]);
                any = any && is.lo[d] <= is.hi[d];
            }
            if (!any) continue;
            if (average_down(fine->S_new + fine->off_valid[fb], F, crse->S_new + crse->off_valid[cb],
                             crse->boxes[cb], is, ratio, NVAR, crse->stream))
                return -1;
        }
    }
    return check_hip(hipStreamSynchronize(crse->stream), "average_down");
}

} // extern "C"

// ------------------------------------------------------------------ batched advance
// mc_advance_batch: Castro::advance + post_timestep over caller-owned FABs,
// every box of the level in one launch per stage (the fused engine).  The
// engine for the last geometry and box list is kept between calls.
namespace mc {

struct BatchFab {
    int lo[3], hi[3];   // valid box
    int flo[3], fhi[3]; // FAB extent
    double *p;
};

// caller FABs <-> the engine's domain layout (component c of cell (i,j,k) at
// c*nD + i + n0*(j + n1*k)); one thread per valid cell of the box list
__global__ __launch_bounds__(kBlock) void k_batch_copy(const BatchFab *fabs, const long *bstart, int nbox, double *U,
                                                     long nD, int n0, int n1, int to_domain)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bstart[nbox]) return;
    int lo = 0, hi = nbox - 1;  // box b with bstart[b] <= t < bstart[b+1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const BatchFab F = fabs[lo];
    const long lt = t - bstart[lo];
    const int bx = F.hi[0] - F.lo[0] + 1, by = F.hi[1] - F.lo[1] + 1;
    const int i = F.lo[0] + (int)(lt % bx), j = F.lo[1] + (int)((lt / bx) % by), k = F.lo[2] + (int)(lt / ((long)bx * by));
    const long fx = F.fhi[0] - F.flo[0] + 1, fy = F.fhi[1] - F.flo[1] + 1, fz = F.fhi[2] - F.flo[2] + 1;
    const long fn = fx * fy * fz;
    const long f0 = (i - F.flo[0]) + fx * ((j - F.flo[1]) + fy * (long)(k - F.flo[2]));
    const long c = (long)i + (long)n0 * (j + (long)n1 * k);
    for (int n = 0; n < NVAR; ++n) {
        if (to_domain) U[nD * n + c] = F.p[fn * n + f0];
        else F.p[fn * n + f0] = U[nD * n + c];
    }
}

struct BatchEngine {
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {0, 0, 0};
    int bc = -1;
    std::vector<Box> boxes;
    FusedLevel *F = nullptr;
    BatchFab *fabs = nullptr;  // 2 * nbox entries: S_old FABs, then S_new FABs
    long *bstart = nullptr;
    std::vector<BatchFab> hfabs;
};

static BatchEngine g_batch;
static std::mutex g_batch_mutex;

static void batch_release()
{
    if (g_batch.F) fused_destroy(g_batch.F);
    if (g_batch.fabs) (void)hipFree(g_batch.fabs);
    if (g_batch.bstart) (void)hipFree(g_batch.bstart);
    g_batch = BatchEngine();
}

static bool same_geometry(const BatchEngine &E, const int *n_cell, const double *prob_lo, const double *prob_hi,
                          int bc, const std::vector<Box> &boxes)
{
    if (!E.F || E.bc != bc || E.boxes.size() != boxes.size()) return false;
    for (int d = 0; d < 3; ++d)
        if (E.n_cell[d] != n_cell[d] || E.prob_lo[d] != prob_lo[d] || E.prob_hi[d] != prob_hi[d]) return false;
    for (size_t b = 0; b < boxes.size(); ++b)
        for (int d = 0; d < 3; ++d)
            if (E.boxes[b].lo[d] != boxes[b].lo[d] || E.boxes[b].hi[d] != boxes[b].hi[d]) return false;
    return true;
}

} // namespace mc

extern "C" int mc_advance_batch(const int *n_cell, const double *prob_lo, const double *prob_hi, int bc, int nbox,
                                const mc_batch_box_t *work, double dt, double *dt_est)
{
    if (!eos_ready()) {
        set_error("mc_advance_batch: EOS table not initialised");
        return -1;
    }
    if (nbox <= 0 || !work || !n_cell || !prob_lo || !prob_hi || (bc != MC_BC_PERIODIC && bc != MC_BC_OUTFLOW) ||
        !(dt > 0.0)) {
        set_error("mc_advance_batch: bad arguments");
        return -1;
    }
    std::vector<Box> boxes(nbox);
    for (int b = 0; b < nbox; ++b) {
        const mc_batch_box_t &w = work[b];
        if (!w.s_old || !w.s_new) {
            set_error("mc_advance_batch: null FAB pointer");
            return -1;
        }
        for (int d = 0; d < 3; ++d) {
            boxes[b].lo[d] = w.lo[d];
            boxes[b].hi[d] = w.hi[d];
            if (w.old_lo[d] > w.lo[d] || w.old_hi[d] < w.hi[d] || w.new_lo[d] > w.lo[d] || w.new_hi[d] < w.hi[d]) {
                set_error("mc_advance_batch: a FAB does not contain its box");
                return -1;
            }
        }
    }
    if (!tiles_domain(boxes, n_cell)) {
        set_error("mc_advance_batch: the boxes do not tile the domain exactly once");
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    BatchEngine &E = g_batch;
    hipStream_t s = cur_stream();
    if (!same_geometry(E, n_cell, prob_lo, prob_hi, bc, boxes)) {
        batch_release();
        double dx[3];
        for (int d = 0; d < 3; ++d) {
            E.n_cell[d] = n_cell[d];
            E.prob_lo[d] = prob_lo[d];
            E.prob_hi[d] = prob_hi[d];
            dx[d] = (prob_hi[d] - prob_lo[d]) / (double)n_cell[d];
        }
        E.bc = bc;
        E.boxes = boxes;
        std::vector<long> bs(nbox + 1, 0);
        for (int b = 0; b < nbox; ++b) bs[b + 1] = bs[b] + box_npts(boxes[b]);
        FusedInstanceSpec sp;
        for (int d = 0; d < 3; ++d) {
            sp.n[d] = n_cell[d];
            sp.dx[d] = dx[d];
        }
        sp.bc = bc;
        if (!(E.F = fused_create(boxes, n_cell, bc, dx, fused_auto_slab({sp}))) ||
            check_hip(hipMalloc(&E.fabs, sizeof(BatchFab) * 2 * nbox), "hipMalloc(batch fabs)") ||
            check_hip(hipMalloc(&E.bstart, sizeof(long) * (nbox + 1)), "hipMalloc(batch starts)") ||
            check_hip(hipMemcpy(E.bstart, bs.data(), sizeof(long) * (nbox + 1), hipMemcpyHostToDevice), "h2d")) {
            batch_release();
            return -1;
        }
        E.hfabs.resize(2 * nbox);
    }
    for (int b = 0; b < nbox; ++b) {
        const mc_batch_box_t &w = work[b];
        BatchFab &o = E.hfabs[b], &nw = E.hfabs[nbox + b];
        for (int d = 0; d < 3; ++d) {
            o.lo[d] = nw.lo[d] = w.lo[d];
            o.hi[d] = nw.hi[d] = w.hi[d];
            o.flo[d] = w.old_lo[d];
            o.fhi[d] = w.old_hi[d];
            nw.flo[d] = w.new_lo[d];
            nw.fhi[d] = w.new_hi[d];
        }
        o.p = const_cast<double *>(w.s_old);
        nw.p = w.s_new;
    }
    const long ncell = (long)n_cell[0] * n_cell[1] * n_cell[2];
    const unsigned grid = (unsigned)((ncell + kBlock - 1) / kBlock);
    double *U = fused_state(E.F, 0);
    if (check_hip(hipMemcpyAsync(E.fabs, E.hfabs.data(), sizeof(BatchFab) * 2 * nbox, hipMemcpyHostToDevice, s),
                  "h2d batch fabs"))
        return -1;
    hipLaunchKernelGGL(k_batch_copy, dim3(grid), dim3(kBlock), 0, s, E.fabs, E.bstart, nbox, U, ncell, n_cell[0],
                       n_cell[1], 1);
    if (fused_advance_dt(E.F, dt, s)) return -1;
    hipLaunchKernelGGL(k_batch_copy, dim3(grid), dim3(kBlock), 0, s, E.fabs + nbox, E.bstart, nbox, U, ncell,
                       n_cell[0], n_cell[1], 0);
    double est = DBL_MAX;
    if (fused_read_est(E.F, 0, &est, s)) return -1;  // synchronises the stream (the FAB table is reused)
    if (dt_est) *dt_est = est * cfl;                  /* Castro.cpp:212-213 */
    return check_hip(hipGetLastError(), "mc_advance_batch");
}

extern "C" void mc_advance_batch_release(void)
{
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    batch_release();
}

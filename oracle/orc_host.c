/*
 * ORACLE — test infrastructure only.  Single-level driver restating
 * Source/Castro_advance.cpp:11-100 (advance), Source/Castro_hydro.cpp:6-399
 * (construct_hydro_source, reference kernel order and box algebra),
 * Source/Castro.cpp:162-341 (time step control), :343-369 (post_timestep)
 * and :546-593 (clean_state), plus the AMReX arithmetic those call
 * (FillPatch as exact copies with periodic or first-order-extrapolation
 * boundaries, MultiFab::Copy, MultiFab::Saxpy, Geometry volume / face area).
 *
 * CPU tiling follows the reference CPU build: MFIter tiles of (1024,16,16)
 * (Source/Castro.cpp:26) distributed over OpenMP threads.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc.h"

typedef struct { int lo[3], hi[3]; } box_t;

static box_t grow(box_t b, int g0, int g1, int g2)
{
    b.lo[0] -= g0; b.lo[1] -= g1; b.lo[2] -= g2;
    b.hi[0] += g0; b.hi[1] += g1; b.hi[2] += g2;
    return b;
}
static box_t grow1(box_t b, int g) { return grow(b, g, g, g); }
/* amrex::surroundingNodes(b, d) */
static box_t nodes(box_t b, int d) { b.hi[d] += 1; return b; }
static long npts(box_t b)
{
    return (long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}
static real *alloc_fab(box_t b, int nc) { return (real *)calloc((size_t)npts(b) * nc, sizeof(real)); }

struct orc_level {
    int n_cell[3], bc, nbox, nthreads;
    real prob_lo[3], prob_hi[3], dx[3];
    box_t *box;
    real **S_new, **S_old, **Sborder, **src;
    int *owner;          /* owner box of each domain cell */
    int steps;
    real time, dt;
    int tile[3];
    /* AMR coupling: the level's fluxes MultiFab (store_flux output, one FAB
       per box and direction, allocated when a register is attached), the
       finer level's register this level initialises and its own register */
    real **fluxes[3];
    orc_flux_register_t *reg_crse, *reg_fine;
};

static box_t vbox(const orc_level_t *L, int b) { return L->box[b]; }

orc_level_t *orc_level_create(const int *n_cell, const real *prob_lo, const real *prob_hi,
                              int bc, int nbox, const int *boxes, int nthreads)
{
    orc_level_t *L = (orc_level_t *)calloc(1, sizeof(orc_level_t));
    for (int d = 0; d < 3; ++d) {
        L->n_cell[d] = n_cell[d];
        L->prob_lo[d] = prob_lo[d];
        L->prob_hi[d] = prob_hi[d];
        L->dx[d] = (prob_hi[d] - prob_lo[d]) / (real)n_cell[d]; /* Geometry::CellSize */
    }
    L->bc = bc;
    L->nbox = nbox;
    L->nthreads = nthreads > 0 ? nthreads : 1;
    L->tile[0] = 1024; L->tile[1] = 16; L->tile[2] = 16;
    L->box = (box_t *)malloc(sizeof(box_t) * nbox);
    L->S_new = (real **)calloc(nbox, sizeof(real *));
    L->S_old = (real **)calloc(nbox, sizeof(real *));
    L->Sborder = (real **)calloc(nbox, sizeof(real *));
    L->src = (real **)calloc(nbox, sizeof(real *));
    long ncell = (long)n_cell[0] * n_cell[1] * n_cell[2];
    L->owner = (int *)malloc(sizeof(int) * ncell);
    for (long c = 0; c < ncell; ++c) L->owner[c] = -1;
    for (int b = 0; b < nbox; ++b) {
        for (int d = 0; d < 3; ++d) { L->box[b].lo[d] = boxes[6 * b + d]; L->box[b].hi[d] = boxes[6 * b + 3 + d]; }
        L->S_new[b] = alloc_fab(L->box[b], ORC_NVAR);
        L->S_old[b] = alloc_fab(L->box[b], ORC_NVAR);
        L->Sborder[b] = alloc_fab(grow1(L->box[b], 4), ORC_NVAR);
        L->src[b] = alloc_fab(L->box[b], ORC_NVAR);
        box_t bb = L->box[b];
        for (int k = bb.lo[2]; k <= bb.hi[2]; ++k)
            for (int j = bb.lo[1]; j <= bb.hi[1]; ++j)
                for (int i = bb.lo[0]; i <= bb.hi[0]; ++i)
                    L->owner[i + (long)n_cell[0] * (j + (long)n_cell[1] * k)] = b;
    }
    return L;
}

void orc_level_destroy(orc_level_t *L)
{
    for (int b = 0; b < L->nbox; ++b) {
        free(L->S_new[b]); free(L->S_old[b]); free(L->Sborder[b]); free(L->src[b]);
    }
    free(L->S_new); free(L->S_old); free(L->Sborder); free(L->src);
    for (int d = 0; d < 3; ++d)
        if (L->fluxes[d]) {
            for (int b = 0; b < L->nbox; ++b) free(L->fluxes[d][b]);
            free(L->fluxes[d]);
        }
    free(L->box); free(L->owner); free(L);
}

int orc_level_nbox(const orc_level_t *L) { return L->nbox; }
real *orc_level_state(orc_level_t *L, int b) { return L->S_new[b]; }
real *orc_level_source(orc_level_t *L, int b) { return L->src[b]; }
int orc_level_steps(const orc_level_t *L) { return L->steps; }
real orc_level_time(const orc_level_t *L) { return L->time; }
real orc_level_dt(const orc_level_t *L) { return L->dt; }

/* MFIter tiles of a box: tile t of nt */
static int ntiles(const orc_level_t *L, box_t b, int *nt3)
{
    int n = 1;
    for (int d = 0; d < 3; ++d) {
        int len = b.hi[d] - b.lo[d] + 1;
        nt3[d] = (len + L->tile[d] - 1) / L->tile[d];
        n *= nt3[d];
    }
    return n;
}
static box_t tilebox(const orc_level_t *L, box_t b, const int *nt3, int t)
{
    int tc[3] = {t % nt3[0], (t / nt3[0]) % nt3[1], t / (nt3[0] * nt3[1])};
    box_t r;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = b.lo[d] + tc[d] * L->tile[d];
        r.hi[d] = r.lo[d] + L->tile[d] - 1;
        if (r.hi[d] > b.hi[d]) r.hi[d] = b.hi[d];
    }
    return r;
}

typedef struct { int b, t; box_t tb; } work_t;

static int make_work(const orc_level_t *L, work_t **out)
{
    int total = 0, nt3[3];
    for (int b = 0; b < L->nbox; ++b) total += ntiles(L, L->box[b], nt3);
    work_t *w = (work_t *)malloc(sizeof(work_t) * total);
    int c = 0;
    for (int b = 0; b < L->nbox; ++b) {
        int n = ntiles(L, L->box[b], nt3);
        for (int t = 0; t < n; ++t) { w[c].b = b; w[c].t = t; w[c].tb = tilebox(L, L->box[b], nt3, t); ++c; }
    }
    *out = w;
    return total;
}

/* Castro::clean_state, Source/Castro.cpp:546-593: the four kernels over
   growntilebox(ng).  All four are pointwise, so the growth of the edge tiles
   is restated by cleaning the whole grown box in slabs. */
static void clean_state(orc_level_t *L, real **S, int ng)
{
#pragma omp parallel for schedule(dynamic) num_threads(L->nthreads)
    for (int b = 0; b < L->nbox; ++b) {
        box_t fb = grow1(L->box[b], ng);
        orc_enforce_minimum_density(fb.lo, fb.hi, S[b], fb.lo, fb.hi);
        orc_normalize_species(fb.lo, fb.hi, S[b], fb.lo, fb.hi);
        orc_reset_internal_e(fb.lo, fb.hi, S[b], fb.lo, fb.hi);
        orc_compute_temp(fb.lo, fb.hi, S[b], fb.lo, fb.hi);
    }
}

void orc_level_clean_state_new(orc_level_t *L) { clean_state(L, L->S_new, 0); }

/* AmrLevel::FillPatch of Sborder from S_old: exact copies of the owning
   box's data, periodic images or first-order extrapolation (clamp to the
   domain) outside it. */
static void fill_patch(orc_level_t *L)
{
    const int *n = L->n_cell;
#pragma omp parallel for schedule(dynamic) num_threads(L->nthreads)
    for (int b = 0; b < L->nbox; ++b) {
        box_t gb = grow1(L->box[b], 4);
        long sx = gb.hi[0] - gb.lo[0] + 1, sxy = sx * (gb.hi[1] - gb.lo[1] + 1), s3 = npts(gb);
        for (int k = gb.lo[2]; k <= gb.hi[2]; ++k)
            for (int j = gb.lo[1]; j <= gb.hi[1]; ++j)
                for (int i = gb.lo[0]; i <= gb.hi[0]; ++i) {
                    int g[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d) {
                        if (L->bc == ORC_BC_PERIODIC) g[d] = ((g[d] % n[d]) + n[d]) % n[d];
                        else g[d] = g[d] < 0 ? 0 : (g[d] >= n[d] ? n[d] - 1 : g[d]);
                    }
                    int ob = L->owner[g[0] + (long)n[0] * (g[1] + (long)n[1] * g[2])];
                    box_t bb = L->box[ob];
                    long bx = bb.hi[0] - bb.lo[0] + 1, bxy = bx * (bb.hi[1] - bb.lo[1] + 1), b3 = npts(bb);
                    long so = (g[0] - bb.lo[0]) + bx * (g[1] - bb.lo[1]) + bxy * (g[2] - bb.lo[2]);
                    long d0 = (i - gb.lo[0]) + sx * (j - gb.lo[1]) + sxy * (k - gb.lo[2]);
                    for (int c = 0; c < ORC_NVAR; ++c) L->Sborder[b][d0 + s3 * c] = L->S_old[ob][so + b3 * c];
                }
    }
}

static real *const_fab(box_t b, real v)
{
    long n = npts(b);
    real *p = (real *)malloc(sizeof(real) * n);
    for (long c = 0; c < n; ++c) p[c] = v;
    return p;
}

/* Source/Castro_hydro.cpp:31-397 for one tile (bx) of box ib */
static void hydro_tile(orc_level_t *L, int ib, box_t bx, real dt)
{
    const real *dx = L->dx;
    const int *domlo = (const int[3]){0, 0, 0};
    const int domhi[3] = {L->n_cell[0] - 1, L->n_cell[1] - 1, L->n_cell[2] - 1};
    box_t fb = grow1(L->box[ib], 4); /* Sborder FAB */
    real *state = L->Sborder[ib];
    real *source = L->src[ib];
    box_t sb = L->box[ib];
    box_t obx = grow1(bx, 1), qbx = grow1(bx, 4);
    box_t ebx[3], gebx[3], tbx[3][3];

    real *q = alloc_fab(qbx, ORC_QVAR), *qaux = alloc_fab(qbx, ORC_NQAUX);
    orc_ctoprim(qbx.lo, qbx.hi, state, fb.lo, fb.hi, q, qbx.lo, qbx.hi, qaux, qbx.lo, qbx.hi);

    real *flux[3], *qe[3], *area[3], *fluxes_out[3];
    box_t areab[3], volb = grow1(L->box[ib], 4);
    /* Geometry::GetFaceArea / GetVolume for Cartesian coordinates */
    const real ar[3] = {dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]};
    real *vol = const_fab(volb, dx[0] * dx[1] * dx[2]);
    for (int d = 0; d < 3; ++d) {
        ebx[d] = nodes(bx, d);
        gebx[d] = grow1(ebx[d], 1);
        flux[d] = alloc_fab(gebx[d], ORC_NVAR);
        qe[d] = alloc_fab(gebx[d], ORC_NGDNV);
        areab[d] = grow1(nodes(L->box[ib], d), 4);
        area[d] = const_fab(areab[d], ar[d]);
        fluxes_out[d] = alloc_fab(ebx[d], ORC_NVAR);
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

    real *qm[3][3], *qp[3][3];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
            qm[a][c] = alloc_fab(tbx[a][c], ORC_QVAR);
            qp[a][c] = alloc_fab(tbx[a][c], ORC_QVAR);
        }
    real *div = alloc_fab(obx, 1);
    orc_divu(obx.lo, obx.hi, q, qbx.lo, qbx.hi, dx, div, obx.lo, obx.hi);

    real *q_int = alloc_fab(obx, ORC_QVAR);
    real *ftmp1 = alloc_fab(obx, ORC_NVAR), *ftmp2 = alloc_fab(obx, ORC_NVAR);
    real *qg1 = alloc_fab(obx, ORC_NGDNV), *qg2 = alloc_fab(obx, ORC_NGDNV);
    real *ql = alloc_fab(obx, ORC_QVAR), *qr = alloc_fab(obx, ORC_QVAR);

    const real hdtdx[3] = {0.5 * dt / dx[0], 0.5 * dt / dx[1], 0.5 * dt / dx[2]};
    const real cdtdx[3] = {dt / dx[0] / 3.0, dt / dx[1] / 3.0, dt / dx[2] / 3.0};

    for (int d = 0; d < 3; ++d)
        orc_trace_ppm(obx.lo, obx.hi, bx.lo, bx.hi, d + 1, q, qbx.lo, qbx.hi, qaux, qbx.lo, qbx.hi,
                      qm[d][d], tbx[d][d].lo, tbx[d][d].hi, qp[d][d], tbx[d][d].lo, tbx[d][d].hi,
                      domlo, domhi, dx, dt);

    for (int d = 0; d < 3; ++d) {
        int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        orc_compute_flux(tbx[d][d].lo, tbx[d][d].hi, qm[d][d], tbx[d][d].lo, tbx[d][d].hi, qp[d][d],
                         tbx[d][d].lo, tbx[d][d].hi, ftmp1, obx.lo, obx.hi, q_int, obx.lo, obx.hi,
                         qg1, obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, d + 1);
        int ts[2] = {t1, t2};
        for (int s = 0; s < 2; ++s) {
            int t = ts[s];
            orc_trans1(tbx[t][d].lo, tbx[t][d].hi, d + 1, t + 1, qm[t][t], tbx[t][t].lo, tbx[t][t].hi,
                       qm[t][d], tbx[t][d].lo, tbx[t][d].hi, qp[t][t], tbx[t][t].lo, tbx[t][t].hi,
                       qp[t][d], tbx[t][d].lo, tbx[t][d].hi, qaux, qbx.lo, qbx.hi, ftmp1, obx.lo,
                       obx.hi, qg1, obx.lo, obx.hi, cdtdx[d]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        orc_compute_flux(tbx[t1][t2].lo, tbx[t1][t2].hi, qm[t1][t2], tbx[t1][t2].lo, tbx[t1][t2].hi,
                         qp[t1][t2], tbx[t1][t2].lo, tbx[t1][t2].hi, ftmp1, obx.lo, obx.hi, q_int,
                         obx.lo, obx.hi, qg1, obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, t1 + 1);
        orc_compute_flux(tbx[t2][t1].lo, tbx[t2][t1].hi, qm[t2][t1], tbx[t2][t1].lo, tbx[t2][t1].hi,
                         qp[t2][t1], tbx[t2][t1].lo, tbx[t2][t1].hi, ftmp2, obx.lo, obx.hi, q_int,
                         obx.lo, obx.hi, qg2, obx.lo, obx.hi, qaux, qbx.lo, qbx.hi, t2 + 1);
        orc_trans2(ebx[d].lo, ebx[d].hi, d + 1, t1 + 1, t2 + 1, qm[d][d], tbx[d][d].lo, tbx[d][d].hi,
                   ql, obx.lo, obx.hi, qp[d][d], tbx[d][d].lo, tbx[d][d].hi, qr, obx.lo, obx.hi, qaux,
                   qbx.lo, qbx.hi, ftmp1, obx.lo, obx.hi, ftmp2, obx.lo, obx.hi, qg1, obx.lo, obx.hi,
                   qg2, obx.lo, obx.hi, hdtdx[d], hdtdx[t1], hdtdx[t2]);
        orc_compute_flux(ebx[d].lo, ebx[d].hi, ql, obx.lo, obx.hi, qr, obx.lo, obx.hi, flux[d],
                         gebx[d].lo, gebx[d].hi, q_int, obx.lo, obx.hi, qe[d], gebx[d].lo,
                         gebx[d].hi, qaux, qbx.lo, qbx.hi, d + 1);
    }
    for (int d = 0; d < 3; ++d) {
        orc_apply_av(ebx[d].lo, ebx[d].hi, d + 1, dx, div, obx.lo, obx.hi, state, fb.lo, fb.hi,
                     flux[d], gebx[d].lo, gebx[d].hi);
        orc_normalize_species_fluxes(ebx[d].lo, ebx[d].hi, flux[d], gebx[d].lo, gebx[d].hi);
        if (L->fluxes[d]) { /* (*fluxes[d])[mfi], Source/Castro_hydro.cpp:380-386 */
            box_t nb = nodes(L->box[ib], d);
            orc_store_flux(ebx[d].lo, ebx[d].hi, L->fluxes[d][ib], nb.lo, nb.hi, flux[d],
                           gebx[d].lo, gebx[d].hi, area[d], areab[d].lo, areab[d].hi, dt);
        } else {
            orc_store_flux(ebx[d].lo, ebx[d].hi, fluxes_out[d], ebx[d].lo, ebx[d].hi, flux[d],
                           gebx[d].lo, gebx[d].hi, area[d], areab[d].lo, areab[d].hi, dt);
        }
    }
    orc_fill_hydro_source(bx.lo, bx.hi, state, fb.lo, fb.hi, q, qbx.lo, qbx.hi, source, sb.lo, sb.hi,
                          flux[0], gebx[0].lo, gebx[0].hi, flux[1], gebx[1].lo, gebx[1].hi, flux[2],
                          gebx[2].lo, gebx[2].hi, qe[0], gebx[0].lo, gebx[0].hi, qe[1], gebx[1].lo,
                          gebx[1].hi, qe[2], gebx[2].lo, gebx[2].hi, area[0], areab[0].lo,
                          areab[0].hi, area[1], areab[1].lo, areab[1].hi, area[2], areab[2].lo,
                          areab[2].hi, vol, volb.lo, volb.hi, dx, dt);

    free(q); free(qaux); free(div); free(q_int); free(ftmp1); free(ftmp2); free(qg1); free(qg2);
    free(ql); free(qr); free(vol);
    for (int d = 0; d < 3; ++d) { free(flux[d]); free(qe[d]); free(area[d]); free(fluxes_out[d]); }
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) { free(qm[a][c]); free(qp[a][c]); }
}

/* Castro::construct_hydro_source, Source/Castro_hydro.cpp:6-399 */
static void construct_hydro_source(orc_level_t *L, real dt)
{
    for (int b = 0; b < L->nbox; ++b) memset(L->src[b], 0, sizeof(real) * npts(L->box[b]) * ORC_NVAR);
    work_t *w;
    int nw = make_work(L, &w);
#pragma omp parallel for schedule(dynamic) num_threads(L->nthreads)
    for (int t = 0; t < nw; ++t) hydro_tile(L, w[t].b, w[t].tb, dt);
    free(w);
}

/* Castro::advance, Source/Castro_advance.cpp:11-100 (single level) */
void orc_level_advance(orc_level_t *L, real dt)
{
    real **tmp = L->S_old; /* swapTimeLevels, :32-33 */
    L->S_old = L->S_new;
    L->S_new = tmp;
    clean_state(L, L->S_old, 0); /* :51 */
    for (int b = 0; b < L->nbox; ++b) /* MultiFab::Copy, :55 */
        memcpy(L->S_new[b], L->S_old[b], sizeof(real) * npts(L->box[b]) * ORC_NVAR);
    fill_patch(L);                /* :62 */
    clean_state(L, L->Sborder, 4); /* :66 */
    construct_hydro_source(L, dt); /* :70 */
    for (int b = 0; b < L->nbox; ++b) { /* MultiFab::Saxpy, :74 */
        long n = npts(L->box[b]) * ORC_NVAR;
        real *s = L->S_new[b], *h = L->src[b];
        for (long c = 0; c < n; ++c) s[c] = s[c] + dt * h[c];
    }
    clean_state(L, L->S_new, 0); /* :78 */
    /* flux registers, :82-88 */
    if (L->reg_crse)
        for (int b = 0; b < L->nbox; ++b)
            for (int d = 0; d < 3; ++d) {
                box_t nb = nodes(L->box[b], d);
                orc_flux_register_crse_init(L->reg_crse, L->fluxes[d][b], nb.lo, nb.hi, d, -1.0);
            }
    if (L->reg_fine)
        for (int b = 0; b < L->nbox; ++b) {
            int j = orc_flux_register_find(L->reg_fine, L->box[b].lo, L->box[b].hi);
            if (j < 0) continue;
            for (int d = 0; d < 3; ++d) {
                box_t nb = nodes(L->box[b], d);
                orc_flux_register_fine_add(L->reg_fine, j, L->fluxes[d][b], nb.lo, nb.hi, d, 1.0);
            }
        }
}

void orc_level_set_flux_registers(orc_level_t *L, orc_flux_register_t *crse_side, orc_flux_register_t *fine_side)
{
    L->reg_crse = crse_side;
    L->reg_fine = fine_side;
    if ((crse_side || fine_side) && !L->fluxes[0])
        for (int d = 0; d < 3; ++d) {
            L->fluxes[d] = (real **)calloc(L->nbox, sizeof(real *));
            for (int b = 0; b < L->nbox; ++b) L->fluxes[d][b] = alloc_fab(nodes(L->box[b], d), ORC_NVAR);
        }
}

/* Castro::reflux, Source/Castro.cpp:451-486 */
void orc_level_reflux(orc_level_t *L, orc_flux_register_t *R)
{
    orc_flux_register_clear_internal_borders(R);
    const real v = L->dx[0] * L->dx[1] * L->dx[2]; /* Geometry::GetVolume */
    for (int b = 0; b < L->nbox; ++b) {
        real *vol = const_fab(L->box[b], v);
        orc_flux_register_reflux(R, L->S_new[b], L->box[b].lo, L->box[b].hi, L->box[b].lo, L->box[b].hi, vol,
                                 L->box[b].lo, L->box[b].hi, 1.0);
        free(vol);
    }
    orc_flux_register_set_val(R, 0.0);
}

/* Castro::avgDown, Source/Castro.cpp:488-505, from the given fine boxes
   (all when nfine < 0) */
void orc_level_average_down(orc_level_t *F, orc_level_t *Cl, const int *ratio, int nfine, const int *ids)
{
    int n = nfine < 0 ? F->nbox : nfine;
    for (int q = 0; q < n; ++q) {
        int fb = nfine < 0 ? q : ids[q];
        box_t fbox = F->box[fb], cc;
        for (int d = 0; d < 3; ++d) {
            cc.lo[d] = fbox.lo[d] / ratio[d];
            cc.hi[d] = (fbox.hi[d] + 1) / ratio[d] - 1;
        }
        for (int cb = 0; cb < Cl->nbox; ++cb) {
            box_t is;
            int any = 1;
            for (int d = 0; d < 3; ++d) {
                is.lo[d] = cc.lo[d] > Cl->box[cb].lo[d] ? cc.lo[d] : Cl->box[cb].lo[d];
                is.hi[d] = cc.hi[d] < Cl->box[cb].hi[d] ? cc.hi[d] : Cl->box[cb].hi[d];
                any = any && is.lo[d] <= is.hi[d];
            }
            if (!any) continue;
            orc_average_down(F->S_new[fb], fbox.lo, fbox.hi, Cl->S_new[cb], Cl->box[cb].lo, Cl->box[cb].hi, is.lo,
                             is.hi, ratio, ORC_NVAR);
        }
    }
}

void orc_level_post_timestep(orc_level_t *L) { clean_state(L, L->S_new, 0); } /* Castro.cpp:369 */

/* Castro::estTimeStep, Source/Castro.cpp:175-216 */
real orc_level_est_time_step(orc_level_t *L)
{
    real dt = DBL_MAX;
    real *dtb = (real *)malloc(sizeof(real) * L->nbox);
#pragma omp parallel for schedule(dynamic) num_threads(L->nthreads)
    for (int b = 0; b < L->nbox; ++b) {
        box_t bb = vbox(L, b);
        dtb[b] = DBL_MAX;
        orc_estdt(bb.lo, bb.hi, L->S_new[b], bb.lo, bb.hi, L->dx, &dtb[b]);
    }
    for (int b = 0; b < L->nbox; ++b) dt = fmin(dt, dtb[b]);
    free(dtb);
    const real cfl = 0.5;
    dt *= cfl;
    return dt;
}

/* Castro::initialTimeStep, Source/Castro.cpp:162-173 */
real orc_level_initial_dt(orc_level_t *L)
{
    const real init_shrink = 0.01;
    return init_shrink * orc_level_est_time_step(L);
}

/* Castro::initData, Source/Castro.cpp:80-117 */
void orc_level_init_data(orc_level_t *L, const orc_prob_t *prob)
{
#pragma omp parallel for schedule(dynamic) num_threads(L->nthreads)
    for (int b = 0; b < L->nbox; ++b) {
        box_t bb = vbox(L, b);
        memset(L->S_new[b], 0, sizeof(real) * npts(bb) * ORC_NVAR);
        orc_initdata(bb.lo, bb.hi, L->S_new[b], bb.lo, bb.hi, L->dx, L->prob_lo, L->prob_hi, prob);
    }
    L->steps = 0;
    L->time = 0.0;
    L->dt = 0.0;
}

/* Amr::coarseTimeStep with Castro::computeInitialDt / computeNewDt
   (Source/Castro.cpp:218-341) on a single level, then Amr::timeStep:
   advance + post_timestep. */
real orc_level_coarse_step(orc_level_t *L, real stop_time)
{
    real dt_0;
    if (L->steps > 0) {
        const real change_max = 1.1e0;
        real dt_min = orc_level_est_time_step(L);
        dt_min = fmin(dt_min, change_max * L->dt);
        dt_0 = fmin(1.0e+100, dt_min);
    } else {
        dt_0 = fmin(1.0e+100, orc_level_initial_dt(L));
    }
    const real eps = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((L->time + dt_0) > (stop_time - eps)) dt_0 = stop_time - L->time;
    L->dt = dt_0;
    orc_level_advance(L, L->dt);
    orc_level_post_timestep(L);
    L->time += L->dt;
    L->steps += 1;
    return L->dt;
}

#ifdef ORC_CNT
/* op-count build: the operation counters of orc_count.h */
void orc_opcount_read(unsigned long long *out5)
{
    out5[0] = orc_opc.add;
    out5[1] = orc_opc.mul;
    out5[2] = orc_opc.div;
    out5[3] = orc_opc.sqrt;
    out5[4] = orc_opc.log;
}

void orc_opcount_reset(void) { orc_opc = orc_opcount_t{0, 0, 0, 0, 0}; }
#endif

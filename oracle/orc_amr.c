/*
 * ORACLE — test infrastructure only.  The AMR coupling of a coarse and a
 * fine level that mini-Castro takes from AMReX, restated on host arrays:
 *
 *   FluxRegister of a fine level (Source/Castro.H:143, defined at
 *   Source/Castro.cpp:69): CrseInit with mult -1 after each coarse step and
 *   FineAdd with mult +1 after each fine step (Source/Castro_advance.cpp:
 *   82-88); Castro::reflux (Source/Castro.cpp:451-486) = ClearInternalBorders,
 *   Reflux with scale 1, setVal(0); Castro::avgDown (Source/Castro.cpp:
 *   488-505) = amrex::average_down.
 *
 * AMReX is an empty submodule of the reference (.gitmodules:1-3), so these
 * follow its published algorithm and the loop order of its 3-D kernels
 * (fluxreg_fineadd, fluxreg_reflux, amrex_avgdown):
 *   - a register face is the face box of a coarsened fine box on its low or
 *     high side in one direction; CrseInit copies mult * flux onto it,
 *     FineAdd adds mult * flux of each of the ratio^2 fine faces, the
 *     transverse offsets looping outer (second transverse) to inner;
 *   - ClearInternalBorders zeroes a low face where it meets any box's high
 *     face (and vice versa), periodic images included;
 *   - Reflux visits the orientations x, y, z low, then x, y, z high, and
 *     updates the coarse cell just outside each face: s += -scale*f/vol
 *     below a low face, s += scale*f/vol above a high face;
 *   - average_down sums the ratio^3 fine cells (k, j, i offsets, i
 *     innermost) and multiplies by 1/ratio^3.
 */
#include <stdlib.h>
#include <string.h>

#include "orc.h"

typedef struct { int lo[3], hi[3]; } abox_t;

typedef struct {
    abox_t fb;
    int fine_box, dir, side;
    long off, n, koff;
} aface_t;

struct orc_flux_register {
    int ratio[3], crse_n[3], periodic, nfine;
    abox_t *fine;
    aface_t *faces;
    int nface;
    long total, ncell;
    real *reg;
    unsigned char *keep;
};

static long anpts(abox_t b)
{
    return (long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}

static int aisect(abox_t a, abox_t b, abox_t *o)
{
    for (int d = 0; d < 3; ++d) {
        o->lo[d] = a.lo[d] > b.lo[d] ? a.lo[d] : b.lo[d];
        o->hi[d] = a.hi[d] < b.hi[d] ? a.hi[d] : b.hi[d];
        if (o->lo[d] > o->hi[d]) return 0;
    }
    return 1;
}

/* offset of (i,j,k) in a Fortran-ordered box */
static long aidx(abox_t b, int i, int j, int k)
{
    return (long)(i - b.lo[0]) +
           (long)(b.hi[0] - b.lo[0] + 1) * ((j - b.lo[1]) + (long)(b.hi[1] - b.lo[1] + 1) * (k - b.lo[2]));
}

static abox_t mkbox(const int *lo, const int *hi)
{
    abox_t b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
    return b;
}

orc_flux_register_t *orc_flux_register_create(int nfine, const int *fine_boxes, const int *ratio,
                                              const int *crse_n, int bc)
{
    orc_flux_register_t *R = (orc_flux_register_t *)calloc(1, sizeof(orc_flux_register_t));
    R->periodic = bc == ORC_BC_PERIODIC;
    R->nfine = nfine;
    for (int d = 0; d < 3; ++d) { R->ratio[d] = ratio[d]; R->crse_n[d] = crse_n[d]; }
    R->fine = (abox_t *)malloc(sizeof(abox_t) * nfine);
    R->faces = (aface_t *)malloc(sizeof(aface_t) * 6 * nfine);
    for (int b = 0; b < nfine; ++b) R->fine[b] = mkbox(fine_boxes + 6 * b, fine_boxes + 6 * b + 3);
    long koff = 0;
    R->nface = 0;
    for (int b = 0; b < nfine; ++b) {
        abox_t c;
        for (int d = 0; d < 3; ++d) {
            c.lo[d] = R->fine[b].lo[d] / ratio[d];
            c.hi[d] = (R->fine[b].hi[d] + 1) / ratio[d] - 1;
        }
        for (int dir = 0; dir < 3; ++dir)
            for (int side = 0; side < 2; ++side) {
                aface_t *f = &R->faces[R->nface++];
                f->fb = c;
                f->fb.lo[dir] = f->fb.hi[dir] = side ? c.hi[dir] + 1 : c.lo[dir];
                f->fine_box = b;
                f->dir = dir;
                f->side = side;
                f->n = anpts(f->fb);
                f->koff = koff;
                f->off = (long)ORC_NVAR * koff;
                koff += f->n;
            }
    }
    R->ncell = koff;
    R->total = (long)ORC_NVAR * koff;
    R->reg = (real *)calloc((size_t)(R->total > 0 ? R->total : 1), sizeof(real));
    R->keep = (unsigned char *)malloc((size_t)(koff > 0 ? koff : 1));
    memset(R->keep, 1, (size_t)koff);
    for (int a = 0; a < R->nface; ++a) {
        const aface_t *f = &R->faces[a];
        const int d = f->dir, nd = crse_n[d];
        for (int g = 0; g < R->nface; ++g) {
            const aface_t *h = &R->faces[g];
            if (h->dir != d || h->side == f->side) continue;
            for (int shift = -1; shift <= 1; ++shift) {
                if (shift && !R->periodic) continue;
                abox_t hs = h->fb, is;
                hs.lo[d] += shift * nd;
                hs.hi[d] += shift * nd;
                if (!aisect(f->fb, hs, &is)) continue;
                for (int k = is.lo[2]; k <= is.hi[2]; ++k)
                    for (int j = is.lo[1]; j <= is.hi[1]; ++j)
                        for (int i = is.lo[0]; i <= is.hi[0]; ++i) R->keep[f->koff + aidx(f->fb, i, j, k)] = 0;
            }
        }
    }
    return R;
}

void orc_flux_register_destroy(orc_flux_register_t *R)
{
    if (!R) return;
    free(R->fine);
    free(R->faces);
    free(R->reg);
    free(R->keep);
    free(R);
}

long orc_flux_register_size(const orc_flux_register_t *R) { return R->total; }
real *orc_flux_register_data(orc_flux_register_t *R) { return R->reg; }

void orc_flux_register_set_val(orc_flux_register_t *R, real v)
{
    for (long t = 0; t < R->total; ++t) R->reg[t] = v;
}

void orc_flux_register_crse_init(orc_flux_register_t *R, const real *flux, const int *flux_lo,
                                 const int *flux_hi, int dir, real mult)
{
    const abox_t fbx = mkbox(flux_lo, flux_hi);
    const long fn = anpts(fbx);
    for (int a = 0; a < R->nface; ++a) {
        const aface_t *f = &R->faces[a];
        abox_t is;
        if (f->dir != dir || !aisect(f->fb, fbx, &is)) continue;
        for (int c = 0; c < ORC_NVAR; ++c)
            for (int k = is.lo[2]; k <= is.hi[2]; ++k)
                for (int j = is.lo[1]; j <= is.hi[1]; ++j)
                    for (int i = is.lo[0]; i <= is.hi[0]; ++i)
                        R->reg[f->off + aidx(f->fb, i, j, k) + f->n * c] =
                            flux[aidx(fbx, i, j, k) + fn * c] * mult;
    }
}

void orc_flux_register_fine_add(orc_flux_register_t *R, int fine_box, const real *flux, const int *flux_lo,
                                const int *flux_hi, int dir, real mult)
{
    const abox_t fbx = mkbox(flux_lo, flux_hi);
    const long fn = anpts(fbx);
    const int r0 = R->ratio[0], r1 = R->ratio[1], r2 = R->ratio[2];
#define FLX(ii, jj, kk, c) flux[aidx(fbx, ii, jj, kk) + fn * (c)]
    for (int a = 0; a < R->nface; ++a) {
        const aface_t *f = &R->faces[a];
        if (f->fine_box != fine_box || f->dir != dir) continue;
        for (int c = 0; c < ORC_NVAR; ++c)
            for (int k = f->fb.lo[2]; k <= f->fb.hi[2]; ++k)
                for (int j = f->fb.lo[1]; j <= f->fb.hi[1]; ++j)
                    for (int i = f->fb.lo[0]; i <= f->fb.hi[0]; ++i) {
                        real *r = &R->reg[f->off + aidx(f->fb, i, j, k) + f->n * c];
                        if (dir == 0) {
                            for (int koff = 0; koff < r2; ++koff)
                                for (int joff = 0; joff < r1; ++joff)
                                    *r = *r + mult * FLX(r0 * i, r1 * j + joff, r2 * k + koff, c);
                        } else if (dir == 1) {
                            for (int koff = 0; koff < r2; ++koff)
                                for (int ioff = 0; ioff < r0; ++ioff)
                                    *r = *r + mult * FLX(r0 * i + ioff, r1 * j, r2 * k + koff, c);
                        } else {
                            for (int joff = 0; joff < r1; ++joff)
                                for (int ioff = 0; ioff < r0; ++ioff)
                                    *r = *r + mult * FLX(r0 * i + ioff, r1 * j + joff, r2 * k, c);
                        }
                    }
    }
#undef FLX
}

void orc_flux_register_clear_internal_borders(orc_flux_register_t *R)
{
    for (int a = 0; a < R->nface; ++a) {
        const aface_t *f = &R->faces[a];
        for (long l = 0; l < f->n; ++l)
            if (!R->keep[f->koff + l])
                for (int c = 0; c < ORC_NVAR; ++c) R->reg[f->off + l + f->n * c] = 0.0;
    }
}

void orc_flux_register_reflux(orc_flux_register_t *R, real *state, const int *state_lo, const int *state_hi,
                              const int *lo, const int *hi, const real *vol, const int *vol_lo,
                              const int *vol_hi, real scale)
{
    const abox_t sb = mkbox(state_lo, state_hi), vb = mkbox(vol_lo, vol_hi), valid = mkbox(lo, hi);
    const long sn = anpts(sb);
    for (int side = 0; side < 2; ++side)
        for (int dir = 0; dir < 3; ++dir)
            for (int a = 0; a < R->nface; ++a) {
                const aface_t *f = &R->faces[a];
                if (f->dir != dir || f->side != side) continue;
                for (int k = f->fb.lo[2]; k <= f->fb.hi[2]; ++k)
                    for (int j = f->fb.lo[1]; j <= f->fb.hi[1]; ++j)
                        for (int i = f->fb.lo[0]; i <= f->fb.hi[0]; ++i) {
                            int c[3] = {i, j, k};
                            if (side == 0) c[dir] -= 1;
                            const int nd = R->crse_n[dir];
                            if (c[dir] < 0 || c[dir] >= nd) {
                                if (!R->periodic) continue;
                                c[dir] += c[dir] < 0 ? nd : -nd;
                            }
                            int in = 1;
                            for (int d = 0; d < 3; ++d)
                                in = in && c[d] >= valid.lo[d] && c[d] <= valid.hi[d];
                            if (!in) continue;
                            const long l = aidx(f->fb, i, j, k);
                            const long sc = aidx(sb, c[0], c[1], c[2]);
                            const real v = vol[aidx(vb, c[0], c[1], c[2])];
                            for (int q = 0; q < ORC_NVAR; ++q) {
                                const real fv = R->reg[f->off + l + f->n * q];
                                if (side == 0) state[sc + sn * q] = state[sc + sn * q] + -scale * fv / v;
                                else state[sc + sn * q] = state[sc + sn * q] + scale * fv / v;
                            }
                        }
            }
}

void orc_average_down(const real *fine, const int *fine_lo, const int *fine_hi, real *crse, const int *crse_lo,
                      const int *crse_hi, const int *lo, const int *hi, const int *ratio, int ncomp)
{
    const abox_t fb = mkbox(fine_lo, fine_hi), cb = mkbox(crse_lo, crse_hi);
    const long fn = anpts(fb), cn = anpts(cb);
    const real volfrac = 1.0 / (real)(ratio[0] * ratio[1] * ratio[2]);
    for (int n = 0; n < ncomp; ++n)
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const int ii = i * ratio[0], jj = j * ratio[1], kk = k * ratio[2];
                    real c = 0.;
                    for (int kref = 0; kref < ratio[2]; ++kref)
                        for (int jref = 0; jref < ratio[1]; ++jref)
                            for (int iref = 0; iref < ratio[0]; ++iref)
                                c = c + fine[aidx(fb, ii + iref, jj + jref, kk + kref) + fn * n];
                    crse[aidx(cb, i, j, k) + cn * n] = volfrac * c;
                }
}

/* the register's fine box that equals box b of a level (-1 if none) */
int orc_flux_register_find(const orc_flux_register_t *R, const int *lo, const int *hi)
{
    for (int b = 0; b < R->nfine; ++b) {
        int same = 1;
        for (int d = 0; d < 3; ++d) same = same && R->fine[b].lo[d] == lo[d] && R->fine[b].hi[d] == hi[d];
        if (same) return b;
    }
    return -1;
}

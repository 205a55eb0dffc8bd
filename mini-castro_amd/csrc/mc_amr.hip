// AMR coupling between a coarse and a fine level on gfx950: the flux
// register of a fine level and the average-down of its state.
//
// The reference advances a fine level with subcycled steps and keeps the
// conservation of the composite solution with an AMReX FluxRegister
// (Source/Castro.H:143, defined at Source/Castro.cpp:69):
//   * after each coarse step, CrseInit(fluxes[d], d, ..., -1) stores minus the
//     coarse fluxes (dt * F * area, the store_flux output) on the coarse
//     faces that bound the fine region (Source/Castro_advance.cpp:82-84);
//   * after each fine step, FineAdd(fluxes[d], d, ..., +1) adds the fine
//     fluxes summed over the ratio^2 fine faces of each coarse face
//     (Source/Castro_advance.cpp:86-88);
//   * Castro::reflux (Source/Castro.cpp:451-486) zeroes the register where
//     two fine boxes meet (ClearInternalBorders), adds the register divided
//     by the cell volume to the coarse cell just outside each face (minus on
//     a low face, plus on a high face) and clears the register;
//   * Castro::avgDown (Source/Castro.cpp:488-505) replaces every coarse cell
//     under the fine region by the mean of its ratio^3 fine cells
//     (amrex::average_down).
// AMReX is not part of the reference tree; these follow its published
// algorithm, with the loop orders of its 3-D kernels.
//
// Layout: the register holds, per fine box and per (direction, side), the
// face box of the coarsened fine box on that side, NVAR components in
// Fortran order; a byte mask per face cell marks the cells ClearInternalBorders
// zeroes.  Every kernel maps one thread to one face cell or coarse cell.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mc_amr.h"
#include "mc_common.h"
#include "mc_eos.h"
#include "mc_fused_impl.h"
#include "mc_internal.h"
#include "../../include/mc_castro.h"

namespace mc {

__global__ void k_reg_fill(double *p, long n, double v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__device__ __forceinline__ long face_local(const Box &fb, int i, int j, int k)
{
    const int nx = fb.hi[0] - fb.lo[0] + 1, ny = fb.hi[1] - fb.lo[1] + 1;
    return (long)(i - fb.lo[0]) + (long)nx * ((j - fb.lo[1]) + (long)ny * (k - fb.lo[2]));
}

// CrseInit (COPY): register = mult * flux on the face cells in `isect`
__global__ void k_reg_crse_init(Box fb, long n, double *reg, Arr flux, BoxIter isect, double mult)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(isect, t, i, j, k)) return;
    const long l = face_local(fb, i, j, k);
    for (int c = 0; c < NVAR; ++c) reg[l + n * c] = flux(i, j, k, c) * mult;
}

// FineAdd: register += mult * (each of the ratio^2 fine faces of a coarse face)
__global__ void k_reg_fine_add(Box fb, long n, double *reg, Arr flx, int dir, int r0, int r1, int r2, double mult)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(make_iter(fb), t, i, j, k)) return;
    const long l = face_local(fb, i, j, k);
    for (int c = 0; c < NVAR; ++c) {
        double r = reg[l + n * c];
        if (dir == 0) {
            const int ii = r0 * i;
            for (int koff = 0; koff < r2; ++koff)
                for (int joff = 0; joff < r1; ++joff) r += mult * flx(ii, r1 * j + joff, r2 * k + koff, c);
        } else if (dir == 1) {
            const int jj = r1 * j;
            for (int koff = 0; koff < r2; ++koff)
                for (int ioff = 0; ioff < r0; ++ioff) r += mult * flx(r0 * i + ioff, jj, r2 * k + koff, c);
        } else {
            const int kk = r2 * k;
            for (int joff = 0; joff < r1; ++joff)
                for (int ioff = 0; ioff < r0; ++ioff) r += mult * flx(r0 * i + ioff, r1 * j + joff, kk, c);
        }
        reg[l + n * c] = r;
    }
}

// FineAdd from the fused engine's final fluxes (one instance covering the
// fine level, whole-instance face arrays): component c of face (i,j,k) in
// direction dir is FN[dir] at eidx, and store_flux's dt * F * area is formed
// as k_store_flux forms it before the sum of k_reg_fine_add
__device__ __forceinline__ double fn_component(double *FN, const FInst &I, int c, idx_t e)
{
    if (c < UFS) return EF(FN, I, NFN, c, e);
    const int sp = c - UFS;
    const dbl2 v = SP2(species_pairs(FN, I, NFN, FN_X), sp >> 1, e);
    return (sp & 1) ? v.y : v.x;
}

__global__ void k_reg_fine_add_fn(Box fb, long n, double *reg, double *FN, const FInst *inst, int dir, int r0,
                                  int r1, int r2, double dt, double area, double mult)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(make_iter(fb), t, i, j, k)) return;
    const FInst &I = inst[0];
    const long l = face_local(fb, i, j, k);
    for (int c = 0; c < NVAR; ++c) {
        double r = reg[l + n * c];
        if (dir == 0) {
            const int ii = r0 * i;
            for (int koff = 0; koff < r2; ++koff)
                for (int joff = 0; joff < r1; ++joff)
                    r += mult * (dt * fn_component(FN, I, c, eidx(I, ii, r1 * j + joff, r2 * k + koff)) * area);
        } else if (dir == 1) {
            const int jj = r1 * j;
            for (int koff = 0; koff < r2; ++koff)
                for (int ioff = 0; ioff < r0; ++ioff)
                    r += mult * (dt * fn_component(FN, I, c, eidx(I, r0 * i + ioff, jj, r2 * k + koff)) * area);
        } else {
            const int kk = r2 * k;
            for (int joff = 0; joff < r1; ++joff)
                for (int ioff = 0; ioff < r0; ++ioff)
                    r += mult * (dt * fn_component(FN, I, c, eidx(I, r0 * i + ioff, r1 * j + joff, kk)) * area);
        }
        reg[l + n * c] = r;
    }
}

// ClearInternalBorders: zero the face cells another fine box shares
__global__ void k_reg_clear(long n, double *reg, const unsigned char *keep)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || keep[t]) return;
    for (int c = 0; c < NVAR; ++c) reg[t + n * c] = 0.0;
}

// Reflux of one register face into the coarse cell just outside it, if that
// cell (after the periodic wrap) lies in `valid`
__global__ void k_reg_reflux(Box fb, long n, const double *reg, int dir, int side, Arr s, Arr vol, double vconst,
                             Box valid, int n0, int n1, int n2, int periodic, double scale)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(make_iter(fb), t, i, j, k)) return;
    const long l = face_local(fb, i, j, k);
    int c[3] = {i, j, k};
    if (side == 0) c[dir] -= 1;  // a low face's outside cell lies below it
    const int nd = dir == 0 ? n0 : (dir == 1 ? n1 : n2);
    if (c[dir] < 0 || c[dir] >= nd) {
        if (!periodic) return;
        c[dir] += c[dir] < 0 ? nd : -nd;
    }
    for (int d = 0; d < 3; ++d)
        if (c[d] < valid.lo[d] || c[d] > valid.hi[d]) return;
    // the cell volume: the volume FAB, or one value for every cell (vol
    // without data)
    const double v = vol.p ? vol(c[0], c[1], c[2], 0) : vconst;
    for (int q = 0; q < NVAR; ++q) {
        const double f = reg[l + n * q];
        if (side == 0) s(c[0], c[1], c[2], q) += -scale * f / v;
        else s(c[0], c[1], c[2], q) += scale * f / v;
    }
}

// amrex::average_down on Cartesian cells: mean of the ratio^3 fine cells
__global__ void k_avgdown(Arr fine, Arr crse, BoxIter region, int r0, int r1, int r2, int ncomp, double volfrac)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(region, t, i, j, k)) return;
    const int ii = i * r0, jj = j * r1, kk = k * r2;
    for (int n = 0; n < ncomp; ++n) {
        double c = 0.;
        for (int kref = 0; kref < r2; ++kref)
            for (int jref = 0; jref < r1; ++jref)
                for (int iref = 0; iref < r0; ++iref) c += fine(ii + iref, jj + jref, kk + kref, n);
        crse(i, j, k, n) = volfrac * c;
    }
}

// stream order across the levels' streams (see mc_flux_register::last)
static int reg_begin(mc_flux_register *R, hipStream_t s)
{
    if (R->used && R->last_stream != s)
        return check_hip(hipStreamWaitEvent(s, R->last, 0), "flux register: wait for the other stream");
    return 0;
}
static int reg_end(mc_flux_register *R, hipStream_t s)
{
    R->used = true;
    R->last_stream = s;
    return check_hip(hipEventRecord(R->last, s), "flux register: record");
}

static bool inside(const Box &a, const Box &b)  // a within b
{
    for (int d = 0; d < 3; ++d)
        if (a.lo[d] < b.lo[d] || a.hi[d] > b.hi[d]) return false;
    return true;
}

static bool intersect(const Box &a, const Box &b, Box &o)
{
    for (int d = 0; d < 3; ++d) {
        o.lo[d] = std::max(a.lo[d], b.lo[d]);
        o.hi[d] = std::min(a.hi[d], b.hi[d]);
        if (o.lo[d] > o.hi[d]) return false;
    }
    return true;
}

int flux_register_fill(mc_flux_register *R, double v, hipStream_t s)
{
    if (reg_begin(R, s)) return -1;
    if (R->total > 0)
        hipLaunchKernelGGL(k_reg_fill, dim3(grid_for(R->total)), dim3(kBlock), 0, s, R->reg, R->total, v);
    if (check_hip(hipGetLastError(), "flux register fill")) return -1;
    return reg_end(R, s);
}

int flux_register_crse_init(mc_flux_register *R, const double *flux, const Box &fbox, int dir, double mult,
                            hipStream_t s)
{
    const Arr a = make_arr(flux, fbox);
    if (reg_begin(R, s)) return -1;
    for (const RegFace &f : R->faces) {
        Box is;
        if (f.dir != dir || !intersect(f.fb, fbox, is)) continue;
        const BoxIter it = make_iter(is);
        hipLaunchKernelGGL(k_reg_crse_init, dim3(grid_for(it.n)), dim3(kBlock), 0, s, f.fb, f.n, R->reg + f.off, a,
                           it, mult);
    }
    if (check_hip(hipGetLastError(), "flux register CrseInit")) return -1;
    return reg_end(R, s);
}

int flux_register_fine_add(mc_flux_register *R, int fine_box, const double *flux, const Box &fbox, int dir,
                           double mult, hipStream_t s)
{
    if (fine_box < 0 || fine_box >= (int)R->fine.size() || dir < 0 || dir > 2) {
        set_error("flux register FineAdd: bad fine box or direction");
        return -1;
    }
    // the fine faces on both boundaries of the box in dir must be in the FAB
    Box need = R->fine[fine_box];
    need.hi[dir] += 1;
    if (!inside(need, fbox)) {
        set_error("flux register FineAdd: the flux FAB does not hold the fine box's faces");
        return -1;
    }
    const Arr a = make_arr(flux, fbox);
    if (reg_begin(R, s)) return -1;
    for (const RegFace &f : R->faces) {
        if (f.fine_box != fine_box || f.dir != dir) continue;
        hipLaunchKernelGGL(k_reg_fine_add, dim3(grid_for(f.n)), dim3(kBlock), 0, s, f.fb, f.n, R->reg + f.off, a,
                           dir, R->ratio[0], R->ratio[1], R->ratio[2], mult);
    }
    if (check_hip(hipGetLastError(), "flux register FineAdd")) return -1;
    return reg_end(R, s);
}

int flux_register_fine_add_fused(mc_flux_register *R, const FusedLevel *F, double dt, const double *area,
                                 double mult, hipStream_t s)
{
    if (F->ninst != 1 || F->slab) {
        set_error("flux register FineAdd: the fused level must hold one instance with whole face arrays "
                  "(no z-slabs)");
        return -1;
    }
    for (int d = 0; d < 3; ++d)
        if (F->hinst[0].n[d] != R->crse_n[d] * R->ratio[d]) {
            set_error("flux register FineAdd: the fused level is not the register's refined domain");
            return -1;
        }
    if (reg_begin(R, s)) return -1;
    for (const RegFace &f : R->faces)
        hipLaunchKernelGGL(k_reg_fine_add_fn, dim3(grid_for(f.n)), dim3(kBlock), 0, s, f.fb, f.n, R->reg + f.off,
                           F->ctx.FN[f.dir], F->inst, f.dir, R->ratio[0], R->ratio[1], R->ratio[2], dt,
                           area[f.dir], mult);
    if (check_hip(hipGetLastError(), "flux register FineAdd (fused)")) return -1;
    return reg_end(R, s);
}

// CrseInit of every register face from the fused engine of the coarse
// level: register = mult * (dt * F * area), as store_flux and
// k_reg_crse_init form it
__global__ void k_reg_crse_init_fn(Box fb, long n, double *reg, double *FN, const FInst *inst, double dt,
                                   double area, double mult)
{
    int i, j, k;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!iter_cell(make_iter(fb), t, i, j, k)) return;
    const FInst &I = inst[0];
    const idx_t e = eidx(I, i, j, k);
    const long l = face_local(fb, i, j, k);
    for (int c = 0; c < NVAR; ++c) reg[l + n * c] = (dt * fn_component(FN, I, c, e) * area) * mult;
}

int flux_register_crse_init_fused(mc_flux_register *R, const FusedLevel *F, double dt, const double *area,
                                  double mult, hipStream_t s)
{
    if (F->ninst != 1 || F->slab) {
        set_error("flux register CrseInit: the fused level must hold one instance with whole face arrays "
                  "(no z-slabs)");
        return -1;
    }
    for (int d = 0; d < 3; ++d)
        if (F->hinst[0].n[d] != R->crse_n[d]) {
            set_error("flux register CrseInit: the fused level is not the register's coarse domain");
            return -1;
        }
    if (reg_begin(R, s)) return -1;
    for (const RegFace &f : R->faces)
        hipLaunchKernelGGL(k_reg_crse_init_fn, dim3(grid_for(f.n)), dim3(kBlock), 0, s, f.fb, f.n, R->reg + f.off,
                           F->ctx.FN[f.dir], F->inst, dt, area[f.dir], mult);
    if (check_hip(hipGetLastError(), "flux register CrseInit (fused)")) return -1;
    return reg_end(R, s);
}

int flux_register_clear_internal(mc_flux_register *R, hipStream_t s)
{
    long koff = 0;
    if (reg_begin(R, s)) return -1;
    for (const RegFace &f : R->faces) {
        hipLaunchKernelGGL(k_reg_clear, dim3(grid_for(f.n)), dim3(kBlock), 0, s, f.n, R->reg + f.off, R->keep + koff);
        koff += f.n;
    }
    if (check_hip(hipGetLastError(), "flux register ClearInternalBorders")) return -1;
    return reg_end(R, s);
}

int flux_register_reflux(mc_flux_register *R, double *state, const Box &sbox, const Box &valid, const double *vol,
                         const Box &vbox, double scale, hipStream_t s, double vconst)
{
    if (!inside(valid, sbox) || (vol && !inside(valid, vbox))) {
        set_error("flux register Reflux: the state or volume FAB does not hold the box");
        return -1;
    }
    const Arr S = make_arr(state, sbox), V = make_arr(vol, vol ? vbox : sbox);
    if (reg_begin(R, s)) return -1;
    // orientations in AMReX's order: x, y, z low faces, then x, y, z high faces
    for (int side = 0; side < 2; ++side)
        for (int dir = 0; dir < 3; ++dir)
            for (const RegFace &f : R->faces) {
                if (f.dir != dir || f.side != side) continue;
                hipLaunchKernelGGL(k_reg_reflux, dim3(grid_for(f.n)), dim3(kBlock), 0, s, f.fb, f.n, R->reg + f.off,
                                   dir, side, S, V, vconst, valid, R->crse_n[0], R->crse_n[1], R->crse_n[2],
                                   R->periodic, scale);
            }
    if (check_hip(hipGetLastError(), "flux register Reflux")) return -1;
    return reg_end(R, s);
}

int average_down(const double *fine, const Box &fbox, double *crse, const Box &cbox, const Box &region,
                 const int *ratio, int ncomp, hipStream_t s)
{
    Box need = region;
    for (int d = 0; d < 3; ++d) {
        need.lo[d] = region.lo[d] * ratio[d];
        need.hi[d] = region.hi[d] * ratio[d] + ratio[d] - 1;
    }
    if (!inside(region, cbox) || !inside(need, fbox) || ncomp < 1 || ncomp > NVAR) {
        set_error("average_down: the FABs do not hold the region or its fine cells");
        return -1;
    }
    const BoxIter it = make_iter(region);
    const double volfrac = 1.0 / (double)(ratio[0] * ratio[1] * ratio[2]);
    hipLaunchKernelGGL(k_avgdown, dim3(grid_for(it.n)), dim3(kBlock), 0, s, make_arr(fine, fbox),
                       make_arr(crse, cbox), it, ratio[0], ratio[1], ratio[2], ncomp, volfrac);
    return check_hip(hipGetLastError(), "average_down");
}

} // namespace mc

using namespace mc;

extern "C" {

mc_flux_register_t *mc_flux_register_create(int nfine, const int *fine_boxes, const int *ratio,
                                            const int *crse_n_cell, int bc)
{
    if (nfine <= 0 || !fine_boxes || !ratio || !crse_n_cell || (bc != MC_BC_PERIODIC && bc != MC_BC_OUTFLOW)) {
        set_error("mc_flux_register_create: bad arguments");
        return nullptr;
    }
    mc_flux_register *R = new mc_flux_register();
    R->periodic = bc == MC_BC_PERIODIC;
    for (int d = 0; d < 3; ++d) {
        R->ratio[d] = ratio[d];
        R->crse_n[d] = crse_n_cell[d];
        if (ratio[d] < 1 || crse_n_cell[d] < 1) {
            set_error("mc_flux_register_create: ratio and domain must be positive");
            delete R;
            return nullptr;
        }
    }
    // fine boxes: inside the refined domain, coarsenable by the ratio
    for (int b = 0; b < nfine; ++b) {
        Box f;
        for (int d = 0; d < 3; ++d) {
            f.lo[d] = fine_boxes[6 * b + d];
            f.hi[d] = fine_boxes[6 * b + 3 + d];
            if (f.lo[d] < 0 || f.hi[d] >= crse_n_cell[d] * ratio[d] || f.hi[d] < f.lo[d] || f.lo[d] % ratio[d] ||
                (f.hi[d] + 1) % ratio[d]) {
                set_error("mc_flux_register_create: a fine box is outside the refined domain or not coarsenable");
                delete R;
                return nullptr;
            }
        }
        R->fine.push_back(f);
    }
    long koff = 0;
    for (int b = 0; b < nfine; ++b) {
        Box c;
        for (int d = 0; d < 3; ++d) {
            c.lo[d] = R->fine[b].lo[d] / ratio[d];
            c.hi[d] = (R->fine[b].hi[d] + 1) / ratio[d] - 1;
        }
        for (int dir = 0; dir < 3; ++dir)
            for (int side = 0; side < 2; ++side) {
                RegFace f;
                f.fb = c;
                f.fb.lo[dir] = f.fb.hi[dir] = side ? c.hi[dir] + 1 : c.lo[dir];
                f.fine_box = b;
                f.dir = dir;
                f.side = side;
                f.n = box_npts(f.fb);
                f.off = (long)NVAR * koff;
                koff += f.n;
                R->faces.push_back(f);
            }
    }
    R->total = (long)NVAR * koff;
    // ClearInternalBorders: a low face cell shared with some box's high face
    // (directly or through the periodic wrap) is cleared, and vice versa
    std::vector<unsigned char> keep((size_t)koff, 1);
    long k0 = 0;
    for (const RegFace &f : R->faces) {
        const int d = f.dir, nd = crse_n_cell[d];
        for (const RegFace &g : R->faces) {
            if (g.dir != d || g.side == f.side) continue;
            for (int shift = -1; shift <= 1; ++shift) {
                if (shift && !R->periodic) continue;
                Box gs = g.fb;
                gs.lo[d] += shift * nd;
                gs.hi[d] += shift * nd;
                Box is;
                if (!intersect(f.fb, gs, is)) continue;
                for (int k = is.lo[2]; k <= is.hi[2]; ++k)
                    for (int j = is.lo[1]; j <= is.hi[1]; ++j)
                        for (int i = is.lo[0]; i <= is.hi[0]; ++i) {
                            const long l = (long)(i - f.fb.lo[0]) +
                                           (long)(f.fb.hi[0] - f.fb.lo[0] + 1) *
                                               ((j - f.fb.lo[1]) + (long)(f.fb.hi[1] - f.fb.lo[1] + 1) * (k - f.fb.lo[2]));
                            keep[k0 + l] = 0;
                        }
            }
        }
        k0 += f.n;
    }
    if (check_hip(hipEventCreateWithFlags(&R->last, hipEventDisableTiming), "hipEventCreate(flux register)") ||
        check_hip(hipMalloc(&R->reg, sizeof(double) * std::max(R->total, 1L)), "hipMalloc(flux register)") ||
        check_hip(hipMalloc(&R->keep, std::max<size_t>(keep.size(), 1)), "hipMalloc(flux register mask)") ||
        check_hip(hipMemcpy(R->keep, keep.data(), keep.size(), hipMemcpyHostToDevice), "flux register mask h2d")) {
        mc_flux_register_destroy(R);
        return nullptr;
    }
    if (flux_register_fill(R, 0.0, cur_stream()) || check_hip(hipStreamSynchronize(cur_stream()), "register sync")) {
        mc_flux_register_destroy(R);
        return nullptr;
    }
    return R;
}

void mc_flux_register_destroy(mc_flux_register_t *R)
{
    if (!R) return;
    (void)hipDeviceSynchronize();
    if (R->reg) (void)hipFree(R->reg);
    if (R->keep) (void)hipFree(R->keep);
    if (R->last) (void)hipEventDestroy(R->last);
    delete R;
}

long mc_flux_register_size(const mc_flux_register_t *R) { return R ? R->total : -1; }

int mc_flux_register_get(mc_flux_register_t *R, double *host)
{
    if (reg_begin(R, cur_stream())) return -1;
    if (check_hip(hipMemcpyAsync(host, R->reg, sizeof(double) * R->total, hipMemcpyDeviceToHost, cur_stream()),
                  "flux register d2h"))
        return -1;
    return check_hip(hipStreamSynchronize(cur_stream()), "flux register sync");
}

int mc_flux_register_set_val(mc_flux_register_t *R, double v) { return flux_register_fill(R, v, cur_stream()); }

int mc_flux_register_crse_init(mc_flux_register_t *R, const double *flux, const int *flux_lo, const int *flux_hi,
                               int dir, double mult)
{
    if (dir < 0 || dir > 2) {
        set_error("mc_flux_register_crse_init: direction must be 0, 1 or 2");
        return -1;
    }
    return flux_register_crse_init(R, flux, make_box(flux_lo, flux_hi), dir, mult, cur_stream());
}

int mc_flux_register_fine_add(mc_flux_register_t *R, int fine_box, const double *flux, const int *flux_lo,
                              const int *flux_hi, int dir, double mult)
{
    return flux_register_fine_add(R, fine_box, flux, make_box(flux_lo, flux_hi), dir, mult, cur_stream());
}

int mc_flux_register_clear_internal_borders(mc_flux_register_t *R)
{
    return flux_register_clear_internal(R, cur_stream());
}

int mc_flux_register_reflux(mc_flux_register_t *R, double *state, const int *state_lo, const int *state_hi,
                            const int *lo, const int *hi, const double *vol, const int *vol_lo, const int *vol_hi,
                            double scale)
{
    if (!vol) {
        set_error("mc_flux_register_reflux: no volume FAB");
        return -1;
    }
    return flux_register_reflux(R, state, make_box(state_lo, state_hi), make_box(lo, hi), vol,
                                make_box(vol_lo, vol_hi), scale, cur_stream());
}

int mc_average_down(const double *fine, const int *fine_lo, const int *fine_hi, double *crse, const int *crse_lo,
                    const int *crse_hi, const int *lo, const int *hi, const int *ratio, int ncomp)
{
    return average_down(fine, make_box(fine_lo, fine_hi), crse, make_box(crse_lo, crse_hi), make_box(lo, hi), ratio,
                        ncomp, cur_stream());
}

} // extern "C"

// Internal definitions shared by the fused engine (mc_fused.hip) and its
// discrete adjoint (mc_adjoint.hip): array layouts, launch plans and the
// per-face load/store helpers.
#pragma once

#include <vector>

#include "mc_common.h"
#include "mc_fused.h"
#include "mc_hydro.h"

namespace mc {

// the cleaned state U1 holds the NU1 components before the species in the
// cell layout, its species live in their own array of pairs (Ctx::U1X)
constexpr int NU1 = UFS;
// primitive components of the fused path; the species of q live in their
// own array of pairs (Ctx::QX)
enum { FQ_RHO = 0, FQ_U, FQ_V, FQ_W, FQ_P, FQ_RE, FQ_C, FQ_G, NFQ };
// The face arrays below store their components in interleaved pairs (EF,
// EF2): 2 * NSPEC_PAIRS component slots for NSPEC species, every pair
// starting at an even component and every component count even.
constexpr int NSPEC_PAIRS = (NSPEC + 1) / 2;
// traced / corrected interface states (velocities in absolute order)
enum { FS_RHO = 0, FS_U, FS_V, FS_W, FS_P, FS_RE, FS_X, NFS = FS_X + 2 * NSPEC_PAIRS };
// transverse-stage fluxes: conserved URHO..UEDEN, Godunov p, species,
// Godunov normal u (and a padding slot)
enum { FF_RHO = 0, FF_MX, FF_MY, FF_MZ, FF_E, FF_GP, FF_X, FF_GU = FF_X + 2 * NSPEC_PAIRS, NFF = FF_GU + 2 };
// final fluxes: conserved URHO..UTEMP, Godunov p, species, Godunov normal u
// (and a padding slot)
static_assert(UFS + NSPEC == NVAR, "the species are the last conserved components");
enum { FN_GP = UFS, FN_X, FN_GU = FN_X + 2 * NSPEC_PAIRS, NFN = FN_GU + 2 };
static_assert(NFS % 2 == 0 && NFF % 2 == 0 && NFN % 2 == 0, "face arrays hold whole pairs");
static_assert(FS_X % 2 == 0 && FF_X % 2 == 0 && FN_X % 2 == 0, "species pairs start at even components");
static_assert(FS_RHO == 0 && FS_U == 1 && FS_V == 2 && FS_W == 3 && FS_P == 4 && FS_RE == 5, "state pairs");
static_assert(FF_E == UEDEN && FF_GP == FF_E + 1 && FN_GP == UTEMP + 1, "flux pairs");

struct FInst {
    int n[3];
    int periodic;
    double dx[3];
    long offD, offE;  // cell offsets in the cell / extended index spaces
    long nD, nE;      // cell counts
    long offQ, nQ;    // species pairs of q: offset and cell count, both even
    int ext[3];       // n + 2 in x and y; in z the planes of one slab + 2
    // The face stages work on a z-slab of cells [zlo, zlo + zn) (the whole
    // instance unless the engine runs in slabs): the extended arrays hold the
    // slab's ring [zlo - 1, zlo + zn] in z, and the launch plans cover it.
    int zlo, zn;
};

// Brick tiling of a launch: blocks of bx*by*bz threads, one brick each.
struct Plan {
    const FInst *inst;
    const int *blk0;  // first block of each instance, ninst+1 entries
    int ninst;
    int bx, by, bz;
    int shift;        // first index: -shift
    int add;          // extent: n + add (cells [0,n): 0,0; faces [0,n]: 0,1; extended [-1,n]: 1,2)
    int zg;           // brick order: 0 raster, else groups of zg brick planes (plan_cell_at)
};

struct Ctx {
    const FInst *inst;
    double *U, *U1, *U1X, *Q, *QX, *FL, *DV;
    double *SM[3], *SP[3], *FD[3], *FT[6], *FN[3];
    double *dt_cur, *dt_est, *time;
    int *steps;
    EosTable tab;
};

__device__ __forceinline__ int find_inst(const Plan &p, int b)
{
    int lo = 0, hi = p.ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.blk0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Cell of logical block lb's thread in the plan's range; false if outside.
__device__ __forceinline__ bool plan_cell_at(const Plan &p, int lb, int &m, int &i, int &j, int &k)
{
    m = find_inst(p, lb);
    const FInst &I = p.inst[m];
    const int e0 = I.n[0] + p.add, e1 = I.n[1] + p.add, e2 = I.zn + p.add;
    const int nb0 = (e0 + p.bx - 1) / p.bx, nb1 = (e1 + p.by - 1) / p.by;
    const int b = lb - p.blk0[m];
    int bi = b % nb0, bj = (b / nb0) % nb1, bk = b / (nb0 * nb1);
    if (p.zg > 1) {
        // groups of zg brick planes: x fastest, then z within the group, then
        // y, so that a brick's z-neighbours run soon after it (the last
        // group may hold fewer planes)
        const int nb2 = (e2 + p.bz - 1) / p.bz;
        const int r = b / nb0, grp = r / (nb1 * p.zg), rr = r - grp * nb1 * p.zg;
        const int gz = min(p.zg, nb2 - grp * p.zg);
        bk = grp * p.zg + rr % gz;
        bj = rr / gz;
    }
    const int t = threadIdx.x;
    const int ti = t % p.bx, tj = (t / p.bx) % p.by, tk = t / (p.bx * p.by);
    const int li = bi * p.bx + ti, lj = bj * p.by + tj, lk = bk * p.bz + tk;
    i = li - p.shift;
    j = lj - p.shift;
    k = lk - p.shift + I.zlo;
    return li < e0 && lj < e1 && lk < e2;
}

__device__ __forceinline__ bool plan_cell(const Plan &p, int &m, int &i, int &j, int &k)
{
    return plan_cell_at(p, logical_block(), m, i, j, k);
}

// A stage with ny sub-launches (directions or pairs): either one block row
// per sub-launch (gridDim.y == ny), or interleaved at brick granularity
// (gridDim.y == 1, block b runs sub-launch b % ny on brick b / ny) so that
// the sub-launches of a brick run together on one XCD and share its L2.
__device__ __forceinline__ bool plan_cell_sub(const Plan &p, int ny, int &sub, int &m, int &i, int &j, int &k)
{
    int lb = logical_block();
    if (gridDim.y > 1) {
        sub = blockIdx.y;
    } else {
        sub = lb % ny;
        lb /= ny;
    }
    return plan_cell_at(p, lb, m, i, j, k);
}

// Bricks that overlap along direction d: consecutive bricks along d start
// bx-1 (by-1, bz-1) cells apart, so a brick's first layer along d (l = 0)
// is the last layer of the brick below it, or the ring cell -1 for the
// first brick.  Every cell of the range is a layer l >= 1 of exactly one
// brick except that ring cell.  The brick count along d is
// ceil((e - 1) / (b - 1)) for an extent of e cells (b >= 2).  Bricks are
// numbered x fastest, then y, then z, or in groups of zg brick planes.
__device__ __forceinline__ bool plan_cell_ovl(const Plan &p, int d, int lb, int &m, int &i, int &j, int &k, int &l)
{
    m = find_inst(p, lb);
    const FInst &I = p.inst[m];
    const int e0 = I.n[0] + p.add, e1 = I.n[1] + p.add, e2 = I.zn + p.add;
    const int s0 = p.bx - (d == 0), s1 = p.by - (d == 1), s2 = p.bz - (d == 2);
    const int nb0 = (e0 - (d == 0) + s0 - 1) / s0, nb1 = (e1 - (d == 1) + s1 - 1) / s1;
    const int b = lb - p.blk0[m];
    int bi = b % nb0, bj = (b / nb0) % nb1, bk = b / (nb0 * nb1);
    if (p.zg > 1) {  // groups of zg brick planes, as in plan_cell_at
        const int nb2 = (e2 - (d == 2) + s2 - 1) / s2;
        const int r = b / nb0, grp = r / (nb1 * p.zg), rr = r - grp * nb1 * p.zg;
        const int gz = min(p.zg, nb2 - grp * p.zg);
        bk = grp * p.zg + rr % gz;
        bj = rr / gz;
    }
    const int t = threadIdx.x;
    const int ti = t % p.bx, tj = (t / p.bx) % p.by, tk = t / (p.bx * p.by);
    const int li = bi * s0 + ti, lj = bj * s1 + tj, lk = bk * s2 + tk;
    l = d == 0 ? ti : (d == 1 ? tj : tk);
    i = li - p.shift;
    j = lj - p.shift;
    k = lk - p.shift + I.zlo;
    return li < e0 && lj < e1 && lk < e2;
}

// the three direction plans of an overlapping stage (K_tflux, K_trans1c,
// K_finalc)
// and their block counts
struct Plan3 {
    Plan p[3];
    int nb[3];
};

__device__ __forceinline__ int bmap(int i, int n, int periodic)
{
    return periodic ? ((i % n) + n) % n : min(max(i, 0), n - 1);
}

// Cell and face indices within one instance are 32-bit and so are the byte
// offsets of an instance's component arrays (fewer than 2^28 extended cells
// per instance, checked at creation, so even 16-byte pair offsets fit).  The array base of a component is
// uniform across a workgroup (one instance per workgroup), so every access is
// a scalar base plus a 32-bit per-lane byte offset.
typedef unsigned idx_t;

// Every array base CA / EA index is uniform across a wavefront: a workgroup
// belongs to one instance, and its direction / pair index comes from the
// block index (plan_cell_sub) or compile-time loops (a component chosen per
// lane goes through CA_lane instead).  Stating it keeps the base in scalar
// registers (one 32-bit vector offset per access instead of a 64-bit
// address).
__device__ __forceinline__ double *uniform_ptr(double *p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<double *>(((unsigned long long)hi << 32) | lo);
}

// a double in the global address space (loads and stores through it are
// global_load / global_store with the scalar base)
typedef __attribute__((address_space(1))) double gdouble;

__device__ __forceinline__ gdouble &at(double *base, idx_t idx)
{
    typedef __attribute__((address_space(1))) char gchar;
    gchar *g = (gchar *)uniform_ptr(base);
    return *(gdouble *)(g + idx * 8u);
}

// Face arrays (traced states SM/SP, fluxes F^d and F^{t|d}, final fluxes)
// hold their C components in interleaved pairs: components 2r and 2r+1
// share one region of 2*nE doubles, element idx of component c at double
// offset (c & ~1)*nE + 2*idx + (c & 1) of the instance's part.  One 16-byte
// access per lane moves both components of a pair (EF2), half the load and
// store instructions of separate components; EF reaches one component.  C,
// nE and the instance offsets are even, so the pairs are 16-byte aligned,
// and nE < 2^28 keeps the 16-byte lane offset in 32 bits.
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) dbl2 gdbl2;
__device__ __forceinline__ gdouble &EF(double *a, const FInst &I, int C, int c, idx_t idx)
{
    typedef __attribute__((address_space(1))) char gchar;
    gchar *g = (gchar *)uniform_ptr(a + ((long)C * I.offE + (long)(c & ~1) * I.nE + (c & 1)));
    return *(gdouble *)(g + idx * 16u);
}
// components c (even) and c+1
__device__ __forceinline__ gdbl2 &EF2(double *a, const FInst &I, int C, int c, idx_t idx)
{
    typedef __attribute__((address_space(1))) char gchar;
    gchar *g = (gchar *)uniform_ptr(a + ((long)C * I.offE + (long)c * I.nE));
    return *(gdbl2 *)(g + idx * 16u);
}

// The species of a face array (species block at the even component cX) as
// NSPEC_PAIRS pairs, and the species kept apart in pairs of a cell array
// (q, U1; count nQ).  The base and pair stride are taken once, before a
// kernel's species loop, so that the loop's stores (which may alias the
// instance table as far as the compiler knows) do not force the instance
// fields to be loaded again in every pass.
struct SpeciesPairs {
    double *base;
    long stride;  // doubles from one pair region to the next
};
__device__ __forceinline__ SpeciesPairs species_pairs(double *a, const FInst &I, int C, int cX)
{
    return SpeciesPairs{a + ((long)C * I.offE + (long)cX * I.nE), 2 * I.nE};
}
// the species of a cell array kept apart in pairs (q, U1), over the
// instance's cells
__device__ __forceinline__ SpeciesPairs cell_species(double *QX, const FInst &I)
{
    return SpeciesPairs{QX + 2L * NSPEC_PAIRS * I.offQ, 2 * I.nQ};
}
// pair p at element idx (idx < 2^28)
__device__ __forceinline__ gdbl2 &SP2(const SpeciesPairs &S, int p, idx_t idx)
{
    typedef __attribute__((address_space(1))) char gchar;
    gchar *g = (gchar *)uniform_ptr(S.base + p * S.stride);
    return *(gdbl2 *)(g + idx * 16u);
}

// Flux divergence of one conserved component (Source/Castro_hydro_nd.F90:
// 418-422): the six face fluxes times their areas, over the cell volume.
__device__ __forceinline__ double flux_div(double f0, double f0x, double f1, double f1y, double f2, double f2z,
                                           double a0, double a1, double a2, double volinv)
{
    return 0.0 + (f0 * a0 - f0x * a0 + f1 * a1 - f1y * a1 + f2 * a2 - f2z * a2) * volinv;
}

// index of a (boundary-mapped) cell in the cell space of instance I
__device__ __forceinline__ idx_t cidx(const FInst &I, int i, int j, int k)
{
    return (idx_t)bmap(i, I.n[0], I.periodic) +
           (idx_t)I.n[0] * ((idx_t)bmap(j, I.n[1], I.periodic) + (idx_t)I.n[1] * (idx_t)bmap(k, I.n[2], I.periodic));
}
// index of an extended cell / face (i, j in [-1, n], k in the slab's ring
// [zlo - 1, zlo + zn])
__device__ __forceinline__ idx_t eidx(const FInst &I, int i, int j, int k)
{
    return (idx_t)(i + 1) + (idx_t)I.ext[0] * ((idx_t)(j + 1) + (idx_t)I.ext[1] * (idx_t)(k - I.zlo + 1));
}
// first cell / face index of direction a the face stages of a slab cover,
// and the last cell: the instance's bounds in x and y, the slab's in z
__device__ __forceinline__ int lo_index(const FInst &I, int a) { return a == 2 ? I.zlo : 0; }
__device__ __forceinline__ int hi_cell(const FInst &I, int a) { return a == 2 ? I.zlo + I.zn - 1 : I.n[a] - 1; }
// component c of a cell array with C components
__device__ __forceinline__ gdouble &CA(double *a, const FInst &I, int C, int c, idx_t idx)
{
    return at(uniform_ptr(a + ((long)C * I.offD + (long)c * I.nD)), idx);
}
// component c chosen per lane (c not uniform across the wavefront)
__device__ __forceinline__ double CA_lane(const double *a, const FInst &I, int C, int c, idx_t idx)
{
    return a[(long)C * I.offD + (long)c * I.nD + idx];
}
__device__ __forceinline__ gdouble &EA(double *a, const FInst &I, int C, int c, idx_t idx)
{
    return at(uniform_ptr(a + ((long)C * I.offE + (long)c * I.nE)), idx);
}

// Hydro part (rho, velocities permuted to direction d, p, rhoe) of a state
__device__ __forceinline__ void load_fs_h(double *S, const FInst &I, idx_t e, int d, FaceH &f)
{
    const dbl2 ru = EF2(S, I, NFS, FS_RHO, e), vw = EF2(S, I, NFS, FS_V, e), pe = EF2(S, I, NFS, FS_P, e);
    f.rho = ru.x;
    f.un = d == 0 ? ru.y : (d == 1 ? vw.x : vw.y);  // FS_U + d
    f.ut = d == 0 ? vw.x : ru.y;                     // FS_V, else FS_U
    f.utt = d == 2 ? vw.x : vw.y;                    // FS_V at d = 2, else FS_W
    f.p = pe.x;
    f.rhoe = pe.y;
}

__device__ __forceinline__ void load_qin_h(double *S, const FInst &I, idx_t e, QInH &q)
{
    const dbl2 ru = EF2(S, I, NFS, FS_RHO, e), vw = EF2(S, I, NFS, FS_V, e), pe = EF2(S, I, NFS, FS_P, e);
    q.rho = ru.x;
    q.u = ru.y;
    q.v = vw.x;
    q.w = vw.y;
    q.p = pe.x;
    q.rhoe = pe.y;
    q.game = 0.0;
}

// The velocity components are passed by value: a conditional between two
// fields of one struct selects an address, which keeps the struct in
// scratch memory when the direction is a run-time value.
__device__ __forceinline__ FaceH face_of(int d, double rho, double u, double v, double w, double p, double rhoe)
{
    FaceH f;
    f.rho = rho;
    f.un = d == 0 ? u : (d == 1 ? v : w);
    f.ut = d == 0 ? v : u;
    f.utt = d == 2 ? v : w;
    f.p = p;
    f.rhoe = rhoe;
    return f;
}

__device__ __forceinline__ FaceH to_face(const QStateH &o, int d)
{
    return face_of(d, o.rho, o.u, o.v, o.w, o.p, o.rhoe);
}

// Non-species part of a transverse-stage flux; momenta in absolute order.
__device__ __forceinline__ void store_tflux_h(double *F, const FInst &I, idx_t e, int d, const RiemannH &o)
{
    // momentum FF_MX + d is the normal one; then FF_MY (d = 0) or FF_MX, then
    // FF_MY (d = 2) or FF_MZ
    const double mx = d == 0 ? o.f_mn : o.f_mt;
    const double my = d == 0 ? o.f_mt : (d == 1 ? o.f_mn : o.f_mtt);
    const double mz = d == 2 ? o.f_mn : o.f_mtt;
    EF2(F, I, NFF, FF_RHO, e) = dbl2{o.f_rho, mx};
    EF2(F, I, NFF, FF_MY, e) = dbl2{my, mz};
    EF2(F, I, NFF, FF_E, e) = dbl2{o.f_E, o.g_p};
    EF(F, I, NFF, FF_GU, e) = o.g_un;
}

// Riemann problem on face (i,j,k) of direction d (cells i-e_d and i)
__device__ __forceinline__ void face_riemann_h(const Ctx &x, const FInst &I, int d, int i, int j, int k,
                                               const FaceH &L, const FaceH &R, RiemannH &o)
{
    const idx_t cr = cidx(I, i, j, k);
    const idx_t cl = cidx(I, i - (d == 0), j - (d == 1), k - (d == 2));
    riemann_hydro(L, R, CA(x.Q, I, NFQ, FQ_C, cl), CA(x.Q, I, NFQ, FQ_C, cr), CA(x.Q, I, NFQ, FQ_G, cl),
                  CA(x.Q, I, NFQ, FQ_G, cr), o);
}

// URHO..UEDEN of a transverse-stage flux and its Godunov p, u
__device__ __forceinline__ void load_tflux_h(double *F, const FInst &I, idx_t e, double *v, double &gp, double &gu)
{
    const dbl2 a = EF2(F, I, NFF, FF_RHO, e), b = EF2(F, I, NFF, FF_MY, e), c = EF2(F, I, NFF, FF_E, e);
    v[URHO] = a.x;
    v[UMX] = a.y;
    v[UMY] = b.x;
    v[UMZ] = b.y;
    v[UEDEN] = c.x;
    gp = c.y;
    gu = EF(F, I, NFF, FF_GU, e);
}

static __constant__ int c_pair_t[6] = {1, 2, 0, 2, 0, 1};   // state (face) direction
static __constant__ int c_pair_d[6] = {0, 0, 1, 1, 2, 2};   // flux direction

// index of the F^{a|b} array
__device__ __forceinline__ int pair_index(int a, int b)
{
    // combos: (1,0) (2,0) (0,1) (2,1) (0,2) (1,2)
    return 2 * b + ((a == 2) || (b == 2 && a == 1) ? 1 : 0);
}

// box arena <-> instance-0 cell layout
struct ArenaBox {
    int lo[3], hi[3];
    long off;
};

// ------------------------------------------------------------------ host side
struct KernelTiling {
    int bx, by, bz;
    int zg = 0;  // brick order (Plan::zg)
};

struct FusedLevel {
    std::vector<FInst> hinst;
    int ninst = 0;
    FInst *inst = nullptr;
    long totD = 0, totE = 0, totQ = 0;
    Ctx ctx;
    double *arena_buf = nullptr;
    // plans: one block table per (tiling, ghost) pair
    struct PlanBuf {
        KernelTiling t;
        int shift, add, slab;  // slab -1: the whole instance (cell stages)
        int ovl;               // direction along which bricks overlap by one cell, -1 none
        int *blk0 = nullptr;
        int nblocks = 0;
    };
    std::vector<PlanBuf> plans;
    // per stage: est prim flat trace flux1 trans1 final update tflux
    // trans1c finalc; the last three are (W, L, X): bricks (W, L, 1) along
    // y, (W, 1, L) along z and (X, W*L/X, 1) along x, each overlapping its
    // neighbour along that direction by one cell
    KernelTiling tiling[FST_N];
    int variant = 0;  // FV_* flags
    // K_update without post_timestep's clean and estdt (a coarse level of a
    // flux register, whose post_timestep runs after the reflux)
    bool defer_post = false;
    // z-slabs of the face stages: planes per slab (0: the whole instance),
    // slab count, and one instance table per slab (zlo, zn set; slab 0 of a
    // whole-instance engine is `inst` itself)
    int slab = 0, nslab = 1;
    std::vector<FInst *> sinst;
    std::vector<std::vector<FInst>> hsinst;
    bool have_est = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double graph_stop = -2.0;
    const double *graph_tab = nullptr;  // EOS table the captured graph reads
    const unsigned long long *graph_fail = nullptr;  // and its Newton-failure counter
    // single-level import/export
    ArenaBox *abox = nullptr;
    long *abstart = nullptr;
    int nabox = 0;
    long arena_n = 0;
};

enum { ST_EST = FST_EST, ST_PRIM = FST_PRIM, ST_FLAT = FST_FLAT, ST_TRACE = FST_TRACE, ST_FLUX1 = FST_FLUX1,
       ST_TRANS1 = FST_TRANS1, ST_FINAL = FST_FINAL, ST_UPDATE = FST_UPDATE, ST_TFLUX = FST_TFLUX,
       ST_TRANS1C = FST_TRANS1C, ST_FINALC = FST_FINALC, NSTAGE = FST_N };
// launch plan of a stage under its current tiling (block table cached);
// slab >= 0 selects a z-slab of the face stages, -1 the whole instance
Plan get_plan(FusedLevel *F, int stage, int slab = 0);
int plan_blocks(FusedLevel *F, const Plan &p);
void fill(double *p, long n, double v, hipStream_t s);

} // namespace mc

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
// The passive species of the state and flux arrays are stored in pairs
// (SP2 below): 2 * NSPEC_PAIRS component slots for NSPEC species.
constexpr int NSPEC_PAIRS = (NSPEC + 1) / 2;
// traced / corrected interface states (velocities in absolute order)
enum { FS_RHO = 0, FS_U, FS_V, FS_W, FS_P, FS_RE, FS_X, NFS = FS_X + 2 * NSPEC_PAIRS };
// transverse-stage fluxes: conserved URHO..UEDEN, species, Godunov p and normal u
enum { FF_RHO = 0, FF_MX, FF_MY, FF_MZ, FF_E, FF_X, FF_GP = FF_X + 2 * NSPEC_PAIRS, FF_GU, NFF };
// final fluxes: NVAR conserved components (species in pairs), Godunov p and normal u
static_assert(UFS + NSPEC == NVAR, "the species are the last conserved components");
enum { FN_X = UFS, FN_GP = FN_X + 2 * NSPEC_PAIRS, FN_GU, NFN };

struct FInst {
    int n[3];
    int periodic;
    double dx[3];
    long offD, offE;  // cell offsets in the cell / extended index spaces
    long nD, nE;      // cell counts
    long offQ, nQ;    // species pairs of q: offset and cell count, both even
    int ext[3];       // n + 2
};

// Brick tiling of a launch: blocks of bx*by*bz threads, one brick each.
struct Plan {
    const FInst *inst;
    const int *blk0;  // first block of each instance, ninst+1 entries
    int ninst;
    int bx, by, bz;
    int shift;        // first index: -shift
    int add;          // extent: n + add (cells [0,n): 0,0; faces [0,n]: 0,1; extended [-1,n]: 1,2)
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
    const int e0 = I.n[0] + p.add, e1 = I.n[1] + p.add, e2 = I.n[2] + p.add;
    const int nb0 = (e0 + p.bx - 1) / p.bx, nb1 = (e1 + p.by - 1) / p.by;
    const int b = lb - p.blk0[m];
    const int bi = b % nb0, bj = (b / nb0) % nb1, bk = b / (nb0 * nb1);
    const int t = threadIdx.x;
    const int ti = t % p.bx, tj = (t / p.bx) % p.by, tk = t / (p.bx * p.by);
    const int li = bi * p.bx + ti, lj = bj * p.by + tj, lk = bk * p.bz + tk;
    i = li - p.shift;
    j = lj - p.shift;
    k = lk - p.shift;
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

__device__ __forceinline__ int bmap(int i, int n, int periodic)
{
    return periodic ? ((i % n) + n) % n : min(max(i, 0), n - 1);
}

// Cell and face indices within one instance are 32-bit and so are the byte
// offsets of an instance's component arrays (fewer than 2^29 extended cells
// per instance, checked at creation).  The array base of a component is
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

// The species of a state or flux array (C components per element, species
// block at component cX) in pairs: species 2p and 2p+1 share components
// cX+2p and cX+2p+1 as one interleaved region, element idx of species n at
// double offset (cX + 2*(n/2))*nE + 2*idx + n%2 of the instance's part.  One
// 16-byte access per lane moves both species of a pair, half the load and
// store instructions of separate components.  nE and the instance offsets
// are even, so the regions are 16-byte aligned.  The base and pair stride
// are taken once, before a kernel's species loop, so that the loop's stores
// (which may alias the instance table as far as the compiler knows) do not
// force the instance fields to be loaded again in every pass.
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) dbl2 gdbl2;
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
// pair p at element idx; the lane offset is 64-bit (it can pass 2^32 bytes)
__device__ __forceinline__ gdbl2 &SP2(const SpeciesPairs &S, int p, idx_t idx)
{
    typedef __attribute__((address_space(1))) char gchar;
    gchar *g = (gchar *)uniform_ptr(S.base + p * S.stride);
    return *(gdbl2 *)(g + (unsigned long long)idx * 16u);
}

// Flux divergence of one conserved component (Source/Castro_hydro_nd.F90:
// 382-391): the six face fluxes times their areas, over the cell volume.
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
// index of an extended cell / face (i,j,k in [-1, n])
__device__ __forceinline__ idx_t eidx(const FInst &I, int i, int j, int k)
{
    return (idx_t)(i + 1) + (idx_t)I.ext[0] * ((idx_t)(j + 1) + (idx_t)I.ext[1] * (idx_t)(k + 1));
}
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
    const int un = FS_U + d, ut = d == 0 ? FS_V : FS_U, utt = d == 2 ? FS_V : FS_W;
    f.rho = EA(S, I, NFS, FS_RHO, e);
    f.un = EA(S, I, NFS, un, e);
    f.ut = EA(S, I, NFS, ut, e);
    f.utt = EA(S, I, NFS, utt, e);
    f.p = EA(S, I, NFS, FS_P, e);
    f.rhoe = EA(S, I, NFS, FS_RE, e);
}

__device__ __forceinline__ void load_qin_h(double *S, const FInst &I, idx_t e, QInH &q)
{
    q.rho = EA(S, I, NFS, FS_RHO, e);
    q.u = EA(S, I, NFS, FS_U, e);
    q.v = EA(S, I, NFS, FS_V, e);
    q.w = EA(S, I, NFS, FS_W, e);
    q.p = EA(S, I, NFS, FS_P, e);
    q.rhoe = EA(S, I, NFS, FS_RE, e);
    q.game = 0.0;
}

__device__ __forceinline__ FaceH to_face(const QStateH &o, int d)
{
    FaceH f;
    f.rho = o.rho;
    f.un = d == 0 ? o.u : (d == 1 ? o.v : o.w);
    f.ut = d == 0 ? o.v : o.u;
    f.utt = d == 2 ? o.v : o.w;
    f.p = o.p;
    f.rhoe = o.rhoe;
    return f;
}

// Non-species part of a transverse-stage flux; momenta in absolute order.
__device__ __forceinline__ void store_tflux_h(double *F, const FInst &I, idx_t e, int d, const RiemannH &o)
{
    const int mn = FF_MX + d, mt = d == 0 ? FF_MY : FF_MX, mtt = d == 2 ? FF_MY : FF_MZ;
    EA(F, I, NFF, FF_RHO, e) = o.f_rho;
    EA(F, I, NFF, mn, e) = o.f_mn;
    EA(F, I, NFF, mt, e) = o.f_mt;
    EA(F, I, NFF, mtt, e) = o.f_mtt;
    EA(F, I, NFF, FF_E, e) = o.f_E;
    EA(F, I, NFF, FF_GP, e) = o.g_p;
    EA(F, I, NFF, FF_GU, e) = o.g_un;
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
#pragma unroll
    for (int n = 0; n <= UEDEN; ++n) v[n] = EA(F, I, NFF, FF_RHO + n, e);
    gp = EA(F, I, NFF, FF_GP, e);
    gu = EA(F, I, NFF, FF_GU, e);
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
        int shift, add;
        int *blk0 = nullptr;
        int nblocks = 0;
    };
    std::vector<PlanBuf> plans;
    KernelTiling tiling[8];  // per stage: est prim flat trace flux1 trans1 final update
    bool have_est = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double graph_stop = -2.0;
    // single-level import/export
    ArenaBox *abox = nullptr;
    long *abstart = nullptr;
    int nabox = 0;
    long arena_n = 0;
};

enum { ST_EST = FST_EST, ST_PRIM = FST_PRIM, ST_FLAT = FST_FLAT, ST_TRACE = FST_TRACE, ST_FLUX1 = FST_FLUX1,
       ST_TRANS1 = FST_TRANS1, ST_FINAL = FST_FINAL, ST_UPDATE = FST_UPDATE, NSTAGE = FST_N };
// launch plan of a stage under its current tiling (block table cached)
Plan get_plan(FusedLevel *F, int stage);
int plan_blocks(FusedLevel *F, const Plan &p);
void fill(double *p, long n, double v, hipStream_t s);

} // namespace mc

// Common definitions for the MI355X CTU hydro kernels.
//
// Component maps follow Source/Castro_nd.F90:14-57 (0-based here) and the
// FAB layout of the reference C-ABI: (ptr, lo[3], hi[3]) in Fortran order,
// i fastest and component slowest (Source/Castro.H:19).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mc {

constexpr int NSPEC = 13;
constexpr int NVAR = 20;
constexpr int QVAR = 21;
constexpr int NQAUX = 4;
constexpr int NGDNV = 6;

enum { URHO = 0, UMX, UMY, UMZ, UEDEN, UEINT, UTEMP, UFS };
enum { QRHO = 0, QU, QV, QW, QGAME, QPRES, QREINT, QTEMP, QFS };
enum { QGAMC = 0, QC, QDPDR, QDPDE };
enum { GDRHO = 0, GDU, GDV, GDW, GDPRES, GDGAME };

// Source/Castro_nd.F90:59-65
constexpr double small_dens = 1.0e-12;
constexpr double small_temp = 1.0e3;
constexpr double small_pres = 1.e-200;
constexpr double smallu = 1.e-12;
constexpr double small = 1.e-8;
constexpr double dual_energy_eta1 = 1.0;
constexpr double cfl = 0.5;

// amrex_constants_module values used on the path
constexpr double HALF = 0.5;
constexpr double FOURTH = 0.25;
constexpr double SIXTH = 1.0 / 6.0;
constexpr double TWO3RD = 2.0 / 3.0;
constexpr double FOUR3RD = 4.0 / 3.0;
constexpr double M_PI_D = 3.141592653589793238462643383279502884197;

// Source/network.F90:11-21; aion_inv = 1/aion evaluated in double.
__device__ __host__ inline double aion(int n)
{
    return n == 0 ? 4.0 : 12.0 + 4.0 * (n - 1);
}
__device__ __host__ inline double zion(int n) { return n == 0 ? 2.0 : 6.0 + 2.0 * (n - 1); }

// A box of cells or faces, inclusive bounds.
struct Box {
    int lo[3], hi[3];
};

__host__ __device__ inline long box_npts(const Box &b)
{
    return (long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
}

// Fortran-ordered view a(i,j,k,n).
struct Arr {
    double *p;
    int lo0, lo1, lo2;
    int nx, ny;
    long sxy, sn;

    __device__ __forceinline__ long idx(int i, int j, int k) const
    {
        return (long)(i - lo0) + (long)nx * (j - lo1) + sxy * (k - lo2);
    }
    __device__ __forceinline__ double &operator()(int i, int j, int k, int n) const
    {
        return p[idx(i, j, k) + sn * n];
    }
};

inline Arr make_arr(const double *p, const int *lo, const int *hi)
{
    Arr a;
    a.p = const_cast<double *>(p);
    a.lo0 = lo[0];
    a.lo1 = lo[1];
    a.lo2 = lo[2];
    a.nx = hi[0] - lo[0] + 1;
    a.ny = hi[1] - lo[1] + 1;
    a.sxy = (long)a.nx * a.ny;
    a.sn = a.sxy * (hi[2] - lo[2] + 1);
    return a;
}
inline Arr make_arr(const double *p, const Box &b) { return make_arr(p, b.lo, b.hi); }

inline Box make_box(const int *lo, const int *hi)
{
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = lo[d];
        b.hi[d] = hi[d];
    }
    return b;
}

// Launch shape for a flat loop over a box: i fastest.
// Logical block (brick) of this workgroup.  Consecutive workgroups are dealt
// round-robin to the 8 XCDs, each with its own L2; mapping workgroup b to
// brick start(b mod 8) + b/8 gives every XCD one contiguous run of bricks, so
// the neighbour cells and faces a stencil re-reads sit in the same L2.
__device__ __forceinline__ int logical_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

struct BoxIter {
    int lo0, lo1, lo2, nx, ny;
    long n;
};
__host__ __device__ inline BoxIter make_iter(const Box &b)
{
    BoxIter it;
    it.lo0 = b.lo[0];
    it.lo1 = b.lo[1];
    it.lo2 = b.lo[2];
    it.nx = b.hi[0] - b.lo[0] + 1;
    it.ny = b.hi[1] - b.lo[1] + 1;
    it.n = box_npts(b);
    return it;
}
__device__ __forceinline__ bool iter_cell(const BoxIter &it, long t, int &i, int &j, int &k)
{
    if (t >= it.n) return false;
    int tt = (int)(t % it.nx);
    long r = t / it.nx;
    i = it.lo0 + tt;
    j = it.lo1 + (int)(r % it.ny);
    k = it.lo2 + (int)(r / it.ny);
    return true;
}

constexpr int kBlock = 256;
inline unsigned grid_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Fortran SIGN(ONE, x): magnitude 1 carrying the sign bit of x.
__device__ __forceinline__ double fsign1(double x) { return copysign(1.0, x); }

} // namespace mc

// wd2_dev.hpp -- shared device-side definitions of the wolfd2 HIP path (gfx950).
//
// Layout on the device: every field is a double (or float) array of ldd*(ny+2)
// elements, element (i,j) at f[i + ldd*j] for i = 0..nx+1, j = 0..ny+1 -- the
// reference's (0:mnx,0:mny) column-major layout with the leading dimension
// rounded up to a multiple of 16 elements so that every grid row starts on a
// 128-byte boundary.  Fortran (1-based) indices are used throughout.
//
// FP64 kernels reproduce the reference's arithmetic bit for bit: the library is
// compiled with -ffp-contract=off and every expression keeps the reference's
// evaluation order (left to right, as written in momentum.f / pressure.f /
// utility.f / bound_cond.f).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wd2 {

// wolfd2.h:15-42
enum { RM_BLOCKG = 0, RM_INTERN = 1, RM_POROUS = 2 };
enum { BM_INTERN = 0, BM_WALL1 = 1, BM_WALL2 = 2, BM_INLET = 3, BM_OUTLT1 = 4, BM_OUTLT2 = 5 };
enum { WEST = 1, EAST = 2, SOUTH = 3, NORTH = 4 };
enum { VU = 1, VV = 2, VP = 3, VT = 4 };

// Region tables on the device (Fortran-ordered copies of the caller's tables).
//
// Boundary profiles (an extension: the reference holds one constant per
// region side and variable, dBCVal): poff[(ir-1) + mgri*((jr-1) + mgrj*((side-1)
// + 4*(var-1)))] for var = VU, VV, VP is -1 for a side that uses the constant,
// or the offset in pval of one value per grid line along the side -- indexed
// by j (0..ny+1) on the west and east sides, by i (0..nx+1) on the south and
// north sides.  bcp() returns the value for line k.
struct Regions {
    int nx, ny, ni, nj, mgri, mgrj;
    const int *brd, *rtype, *btp;
    const double *bcv, *poros, *porc1, *porc2;
    const int *poff;      // nullptr: no profiles
    const double *pval;

    __host__ __device__ int border(int ir, int jr, int side) const {
        return brd[(ir - 1) + mgri * ((jr - 1) + mgrj * (side - 1))];
    }
    __host__ __device__ int btype(int ir, int jr, int side) const {
        return btp[(ir - 1) + mgri * ((jr - 1) + mgrj * (side - 1))];
    }
    __host__ __device__ int type(int ir, int jr) const { return rtype[(ir - 1) + mgri * (jr - 1)]; }
    __host__ __device__ double bc(int ir, int jr, int side, int var) const {
        return bcv[(ir - 1) + mgri * ((jr - 1) + mgrj * ((side - 1) + 4 * (var - 1)))];
    }
    __host__ __device__ double tab(const double *t, int ir, int jr) const { return t[(ir - 1) + mgri * (jr - 1)]; }
    __host__ __device__ int prof(int ir, int jr, int side, int var) const {
        return poff ? poff[(ir - 1) + mgri * ((jr - 1) + mgrj * ((side - 1) + 4 * (var - 1)))] : -1;
    }
    // boundary value on grid line k of a side: the profile's, else the constant
    __host__ __device__ double bcp(int off, int ir, int jr, int side, int var, int k) const {
        return off >= 0 ? pval[off + k] : bc(ir, jr, side, var);
    }
};

// wolfd2.h:24-31: thermal region and boundary types.
enum { RT_NOSRCE = 0, RT_HEATGN = 1, RT_TEMPER = 2 };
enum { BT_INTERN = 0, BT_TEMPER = 1, BT_HTFLUX = 2 };

// Thermal region tables on the device (Fortran-ordered, extents of Regions):
// nTRgType(mgri,mgrj), nTemBdTp(mgri,mgrj,4), dTRgVal(mgri,mgrj),
// dHGSTval(mgri,mgrj).  Boundary temperatures are Regions::bc(.., VT).
struct Thermal {
    int mgri, mgrj;
    const int *trtype, *tbtp;
    const double *trval, *hgval;
    __host__ __device__ int rtype(int ir, int jr) const { return trtype[(ir - 1) + mgri * (jr - 1)]; }
    __host__ __device__ int btype(int ir, int jr, int side) const {
        return tbtp[(ir - 1) + mgri * ((jr - 1) + mgrj * (side - 1))];
    }
    __host__ __device__ double tval(int ir, int jr) const { return trval[(ir - 1) + mgri * (jr - 1)]; }
    __host__ __device__ double hg(int ir, int jr) const { return hgval[(ir - 1) + mgri * (jr - 1)]; }
};

// Metric order of grid.f:Grid (and of wd2_ctx_create's metrics argument),
// with lower-case aliases matching the Metrics member names.
enum {
    M_RAU, M_RBU, M_RBV, M_RGV, M_RAN, M_RBN, M_RGN, M_RAC, M_RBC, M_RGC, M_DJU, M_DJV, M_DJC, M_DJN,
    M_XEN, M_YEN, M_XZN, M_YZN, M_XEC, M_YEC, M_XZC, M_YZC, M_XEU, M_YEU, M_XZV, M_YZV, M_XZU, M_YZU,
    M_XEV, M_YEV, M_COUNT
};
enum {
    MI_rau, MI_rbu, MI_rbv, MI_rgv, MI_ran, MI_rbn, MI_rgn, MI_rac, MI_rbc, MI_rgc, MI_dju, MI_djv, MI_djc, MI_djn,
    MI_xen, MI_yen, MI_xzn, MI_yzn, MI_xec, MI_yec, MI_xzc, MI_yzc, MI_xeu, MI_yeu, MI_xzv, MI_yzv, MI_xzu, MI_yzu,
    MI_xev, MI_yev
};

// The constant part of a metric array: every entry (i,j) with i0 <= i <= i1
// and j0 <= j <= j1 holds c, bit for bit (found by scanning the array on the
// host), so a kernel reads c there instead of loading it -- the same value.
// On a uniform Cartesian grid every metric is one value over the whole
// grid, and the stencils then load only the fields.  Empty: i0 > i1.
struct MBox {
    double c;
    int i0, i1, j0, j1;
};

// Metric arrays used on the hot path (grid.f:Metric names) and their boxes.
struct Metrics {
    const double *ran, *rbn, *rgn, *rac, *rbc, *rgc, *rau, *rbu, *rbv, *rgv;
    const double *dju, *djv, *djc, *xen, *yen, *xzn, *yzn, *xec, *yec, *xzc, *yzc;
    const double *xeu, *yeu, *xzv, *yzv, *xzu, *yzu, *xev, *yev;
    MBox box[M_COUNT];
};

// A metric array with its box, as passed to kernels that take single arrays.
struct MArr {
    const double *p;
    MBox b;
};

__device__ inline double mget(const double *__restrict__ p, const MBox &b, int ld, int i, int j) {
    if (i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1) return b.c;
    return p[(size_t)i + (size_t)ld * (size_t)j];
}
__device__ inline double mget(const MArr &a, int ld, int i, int j) { return mget(a.p, a.b, ld, i, j); }
// m.f(i,j) of a Metrics m through its box
#define MGET(m, f, i, j) mget((m).f, (m).box[MI_##f], ld, (i), (j))

#define WD2_IX(i, j) ((size_t)(i) + (size_t)ld * (size_t)(j))

// Non-negative doubles order like their bit patterns read as unsigned integers.
__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ inline double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ inline double dmaxf(double a, double b) { return a > b ? a : b; }

// f(i,j), f(i+1,j) for even i (a 16-byte aligned pair: ld is a multiple of 16
// and every field starts 128-byte aligned), 0 for points outside the array
// (i outside 0..nx+1 or j outside 0..ny+1).
__device__ inline double2 load_pair(const double *__restrict__ f, int ld, int nx, int ny, int i, int j) {
    double2 v = make_double2(0.0, 0.0);
    if (j < 0 || j > ny + 1) return v;
    if (i >= 0 && i + 1 <= nx + 1) return *reinterpret_cast<const double2 *>(f + (size_t)i + (size_t)ld * j);
    if (i >= 0 && i <= nx + 1) v.x = f[(size_t)i + (size_t)ld * j];
    if (i + 1 >= 0 && i + 1 <= nx + 1) v.y = f[(size_t)i + 1 + (size_t)ld * j];
    return v;
}

// The same as one 16-byte load from a clamped address and two selects: the
// clamped pair lies inside the ld x (ny+2) field (ld is even and at least
// nx+2), so no branch is needed.
__device__ inline double2 load_pair_sel(const double *__restrict__ f, int ld, int nx, int ny, int i, int j) {
    const int ic = min(max(i, 0), ld - 2), jc = min(max(j, 0), ny + 1);
    const double2 v = *reinterpret_cast<const double2 *>(f + (size_t)ic + (size_t)ld * jc);
    const bool jin = j >= 0 && j <= ny + 1;
    return make_double2(jin && i >= 0 && i <= nx + 1 ? v.x : 0.0, jin && i + 1 >= 0 && i + 1 <= nx + 1 ? v.y : 0.0);
}

// Load that bypasses the (non-coherent) per-CU cache: used to read values that
// other workgroups of the same kernel published before an atomic ticket.
__device__ inline double ld_coherent(const double *p) {
    return bitsd(__hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Block-wide max of a double over NT threads (NT a multiple of 64, tid the
// flattened thread index); every thread gets the result.
template <int NT>
__device__ inline double block_max(double v, double *sh, int tid) {
    for (int off = 32; off > 0; off >>= 1) v = dmaxf(v, __shfl_xor(v, off));
    int lane = tid & 63, w = tid >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (tid == 0) {
        double m = sh[0];
        for (int k = 1; k < NT / 64; k++) m = dmaxf(m, sh[k]);
        sh[0] = m;
    }
    __syncthreads();
    double r = sh[0];
    __syncthreads();
    return r;
}

}  // namespace wd2

// k_momentum.hip -- the Douglas-Gunn ADI momentum step (momentum.f:199-1226).
//
// One kernel per reference stage, each a single pass over the grid:
//   ConvCoef (864-981)            face convective coefficients: cases 1,2,4,5, and
//                                 3,6 for the thermal energy equation
//   porous division (295-324)     cj*, c*s, c*n divided by the region porosity
//   DConvU/V + DDiffU/V           convection and full-tensor diffusion of the
//   (987-1109)                    star and n levels, four outputs per point
//   PorosCoef (1115-1226)         Dupuit-Forchheimer coefficients
//   assembly (350-384, 396-496,   upwinded tridiagonal rows + right-hand side
//   675-711, 723-821)             of the flattened chains, wall/blockage rows
//   scatter                       chain solution back into the field
// The chains themselves go to the parallel AltTridLU of k_trid.hip.  Region
// loops whose ranges overlap are resolved per point in the reference's
// region order, so the result is the reference's whatever the thread order.
//
// Every kernel is a template on T, the type of the temporaries and chains
// and of the arithmetic; the fields (u, v, metrics) stay FP64 and are read
// through cv<T>.  T = double is the reference's arithmetic, operation for
// operation (the casts are no-ops).  T = float is the mixed-precision path's
// momentum: FP32 stencils and ADI chains on FP64 fields, with the difference
// fn - fs of the right-hand side formed in FP64 before rounding.
#include <algorithm>
#include <type_traits>

#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int MX = 64, MY = 4;
static inline dim3 mgrid(int ni, int nj) { return dim3((ni + MX - 1) / MX, (nj + MY - 1) / MY); }
// XCD-aware order of the MX x MY tiles of an ni x nj range: workgroups go to
// the 8 XCDs round-robin, so workgroup b takes tile (b % 8) * per + b / 8 of
// the row-major tile grid, and each XCD works through a contiguous band of
// tile rows whose shared halo rows stay in its own L2.  xgrid: a 1-D grid of
// 8 * per workgroups; xtile: the workgroup's tile (false: none).
static inline dim3 xgrid(int ni, int nj) {
    const int nt = ((ni + MX - 1) / MX) * ((nj + MY - 1) / MY);
    return dim3(8 * ((nt + 7) / 8));
}
__device__ inline bool xtile(int ni, int nj, int &bx, int &by) {
    const int gx = (ni + MX - 1) / MX, nt = gx * ((nj + MY - 1) / MY), per = (nt + 7) / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    bx = t % gx;
    by = t / gx;
    return t < nt;
}

// ------------------------------------------------------------- ConvCoef
template <typename T, typename S>
__device__ inline T cv(S x) { return (T)x; }

// Field accessors of the stencil helpers: f(i, j) as T, read from memory
// (GAcc) or from a workgroup's LDS tile of the field (LAcc: the tile's
// origin (i0, j0) and row width w).  A tile holds the T values of the field
// (cv<T> of each element), so both give the same T for the same point.
template <typename T, typename E>
struct GAcc {
    const E *__restrict__ p;
    int ld;
    __device__ T operator()(int i, int j) const { return cv<T>(p[WD2_IX(i, j)]); }
};
template <typename T>
struct LAcc {
    const T *s;
    int i0, j0, w;
    __device__ T operator()(int i, int j) const { return s[(j - j0) * w + (i - i0)]; }
};
// A metric array read only at points inside its constant box (MBox): every
// read is the box's value, with no test and no load.  mpick<MK_CONST> gives
// one, mpick<MK_BOX> the array with its box (a load outside the box).
struct MConst {
    double c;
};
__device__ inline double mget(const MConst &a, int, int, int) { return a.c; }
enum { MK_BOX = 0, MK_CONST = 1 };
template <int MK>
__device__ inline std::conditional_t<MK == MK_CONST, MConst, MArr> mpick(const double *p, const MBox &b) {
    if constexpr (MK == MK_CONST)
        return MConst{b.c};
    else
        return MArr{p, b};
}
#define MPICK(m, f) mpick<MK>((m).f, (m).box[MI_##f])

// The centred face coefficients of ConvCoef cases 1 (cc_x1, cc_x2) and 2
// (cc_y1, cc_y2) at one point (momentum.f:901-925), shared by k_conv_coef and
// the fused right-hand-side kernels, which evaluate them on the fly: one
// expression, so the same bits either way.  The *a forms read u and v
// through accessors; the pointer forms read them from memory.
#define M(f, a, b) cv<T>(mget(f, ld, a, b))
template <typename T, typename AU, typename AV, typename MA>
__device__ inline T cc_x1a(int ld, int i, int j, T djac, const MA &xet, const MA &yet, const AU &U, const AV &V) {
    return (djac * M(yet, i, j) * (U(i, j) + U(i - 1, j)) - M(xet, i, j) * (V(i, j) + V(i, j - 1))) * T(0.5);
}
template <typename T, typename AU, typename AV, typename MA>
__device__ inline T cc_x2a(int ld, int i, int j, T djac, const MA &xzi, const MA &yzi, const AU &U, const AV &V) {
    return (M(xzi, i, j) * (V(i + 1, j) + V(i, j)) - djac * M(yzi, i, j) * (U(i, j + 1) + U(i, j))) * T(0.5);
}
template <typename T, typename AU, typename AV, typename MA>
__device__ inline T cc_y1a(int ld, int i, int j, T djac, const MA &xet, const MA &yet, const AU &U, const AV &V) {
    return (M(yet, i, j) * (U(i, j + 1) + U(i, j)) - djac * M(xet, i, j) * (V(i + 1, j) + V(i, j))) * T(0.5);
}
template <typename T, typename AU, typename AV, typename MA>
__device__ inline T cc_y2a(int ld, int i, int j, T djac, const MA &xzi, const MA &yzi, const AU &U, const AV &V) {
    return (djac * M(xzi, i, j) * (V(i, j) + V(i, j - 1)) - M(yzi, i, j) * (U(i, j) + U(i - 1, j))) * T(0.5);
}
// ConvCoef cases 4 and 5 (upwind Jacobian, momentum.f:927-961)
template <typename T, typename AU, typename AV, typename MA>
__device__ inline void cc_4a(int ld, int i, int j, T djac, const MA &xzi, const MA &xet, const MA &yzi,
                             const MA &yet, const AU &U, const AV &V, T &c1, T &c2) {
    const T dFour = T(4.0);
    c1 = djac * M(yet, i, j) * U(i, j) - M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour;
    c2 = M(xzi, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour - djac * M(yzi, i, j) * U(i, j);
}
template <typename T, typename AU, typename AV, typename MA>
__device__ inline void cc_5a(int ld, int i, int j, T djac, const MA &xzi, const MA &xet, const MA &yzi,
                             const MA &yet, const AU &U, const AV &V, T &c1, T &c2) {
    const T dFour = T(4.0);
    c1 = M(yet, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour - djac * M(xet, i, j) * V(i, j);
    c2 = djac * M(xzi, i, j) * V(i, j) - M(yzi, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour;
}
#undef M
template <typename T, typename E>
__device__ inline T cc_x1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return cc_x1a<T>(ld, i, j, djac, xet, yet, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld});
}
template <typename T, typename E>
__device__ inline T cc_x2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return cc_x2a<T>(ld, i, j, djac, xzi, yzi, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld});
}
template <typename T, typename E>
__device__ inline T cc_y1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return cc_y1a<T>(ld, i, j, djac, xet, yet, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld});
}
template <typename T, typename E>
__device__ inline T cc_y2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return cc_y2a<T>(ld, i, j, djac, xzi, yzi, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld});
}
template <typename T, typename E>
__device__ inline void cc_4(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const E *__restrict__ u, const E *__restrict__ v, T &c1, T &c2) {
    cc_4a<T>(ld, i, j, djac, xzi, xet, yzi, yet, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld}, c1, c2);
}
template <typename T, typename E>
__device__ inline void cc_5(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const E *__restrict__ u, const E *__restrict__ v, T &c1, T &c2) {
    cc_5a<T>(ld, i, j, djac, xzi, xet, yzi, yet, GAcc<T, E>{u, ld}, GAcc<T, E>{v, ld}, c1, c2);
}

template <int NCOMP, typename T>
__global__ __launch_bounds__(MX *MY) void k_conv_coef(int nx, int ny, int ld, double djac_, MArr xzi, MArr xet,
                                                      MArr yzi, MArr yet, const double *__restrict__ u,
                                                      const double *__restrict__ v, T *__restrict__ cc1,
                                                      T *__restrict__ cc2) {
    const int i = blockIdx.x * MX + threadIdx.x, j = blockIdx.y * MY + threadIdx.y;   // 0-based cover
    if (i > nx + 1 || j > ny + 1) return;
    const T dHalf = T(0.5), dFour = T(4.0), djac = cv<T>(djac_);
#define U(a, b) cv<T>(u[WD2_IX(a, b)])
#define V(a, b) cv<T>(v[WD2_IX(a, b)])
#define M(f, a, b) cv<T>(mget(f, ld, a, b))
    const size_t k = WD2_IX(i, j);
    if (NCOMP == 1) {
        if (i >= 1 && j >= 1) cc1[k] = cc_x1<T>(ld, i, j, djac, xet, yet, u, v);
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) cc2[k] = cc_x2<T>(ld, i, j, djac, xzi, yzi, u, v);
    } else if (NCOMP == 2) {
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) cc1[k] = cc_y1<T>(ld, i, j, djac, xet, yet, u, v);
        if (i >= 1 && j >= 1) cc2[k] = cc_y2<T>(ld, i, j, djac, xzi, yzi, u, v);
    } else if (NCOMP == 3) {   // thermal.f:110-112 (time-level-n face coefficients)
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
            cc1[k] = (M(yet, i, j) * U(i, j) -
                      M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour) * dHalf;
            cc2[k] = (M(xzi, i, j) * V(i, j) -
                      M(yzi, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour) * dHalf;
        }
    } else if (NCOMP == 6) {   // thermal.f:107-109 (cell-centre coefficients for the upwinded rows)
        if (i >= 1 && j >= 1) {
            cc1[k] = (M(yet, i, j) * (U(i, j) + U(i - 1, j)) - M(xet, i, j) * (V(i, j) + V(i, j - 1))) * dHalf;
            cc2[k] = (M(xzi, i, j) * (V(i, j) + V(i, j - 1)) - M(yzi, i, j) * (U(i, j) + U(i - 1, j))) * dHalf;
        }
    } else if (NCOMP == 4) {
        if (i <= nx && j >= 1 && j <= ny) cc_4<T>(ld, i, j, djac, xzi, xet, yzi, yet, u, v, cc1[k], cc2[k]);
    } else if (NCOMP == 5) {
        if (i >= 1 && i <= nx && j <= ny) cc_5<T>(ld, i, j, djac, xzi, xet, yzi, yet, u, v, cc1[k], cc2[k]);
    }
#undef U
#undef V
#undef M
}

template <typename T>
void launch_conv_coef(int nx, int ny, int ld, int ncomp, int njacob, const MArr &xzi, const MArr &xet,
                      const MArr &yzi, const MArr &yet, const double *u, const double *v, T *cc1, T *cc2,
                      hipStream_t s) {
    const double djac = njacob == 1 ? 2.0 : 1.0;
    const dim3 g = mgrid(nx + 2, ny + 2), b(MX, MY);
    switch (ncomp) {
    case 1: hipLaunchKernelGGL((k_conv_coef<1, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 2: hipLaunchKernelGGL((k_conv_coef<2, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 4: hipLaunchKernelGGL((k_conv_coef<4, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 5: hipLaunchKernelGGL((k_conv_coef<5, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 3: hipLaunchKernelGGL((k_conv_coef<3, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 6: hipLaunchKernelGGL((k_conv_coef<6, T>), g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    default: break;
    }
}
template void launch_conv_coef<double>(int, int, int, int, int, const MArr &, const MArr &, const MArr &,
                                       const MArr &, const double *, const double *, double *, double *, hipStream_t);
template void launch_conv_coef<float>(int, int, int, int, int, const MArr &, const MArr &, const MArr &, const MArr &,
                                      const double *, const double *, float *, float *, hipStream_t);

// --------------------------------------------------- porous-region division
template <typename T>
__global__ __launch_bounds__(MX *MY) void k_porous_div(Regions R, int ld, int comp, MomTempsT<T> t) {
    const int i = blockIdx.x * MX + threadIdx.x, j = blockIdx.y * MY + threadIdx.y;
    if (i > R.nx + 1 || j > R.ny + 1) return;
    const size_t k = WD2_IX(i, j);
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            if (R.type(ir, jr) != RM_POROUS) continue;
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            const bool in = comp == 1 ? (j >= jS + 1 && j <= jN && i >= iW && i <= iE)
                                      : (j >= jS && j <= jN && i >= iW + 1 && i <= iE);
            if (!in) continue;
            const T por = cv<T>(R.tab(R.poros, ir, jr));
            t.cj1[k] = t.cj1[k] / por;
            t.cj2[k] = t.cj2[k] / por;
            t.c1s[k] = t.c1s[k] / por;
            t.c2s[k] = t.c2s[k] / por;
            t.c1n[k] = t.c1n[k] / por;
            t.c2n[k] = t.c2n[k] / por;
        }
}

template <typename T>
void launch_porous_div(const Regions &R, int ld, int comp, const MomTempsT<T> &t, hipStream_t s) {
    hipLaunchKernelGGL(k_porous_div<T>, mgrid(R.nx + 2, R.ny + 2), dim3(MX, MY), 0, s, R, ld, comp, t);
}
template void launch_porous_div<double>(const Regions &, int, int, const MomTempsT<double> &, hipStream_t);
template void launch_porous_div<float>(const Regions &, int, int, const MomTempsT<float> &, hipStream_t);

// -------------------------------------------------------------- PorosCoef
// Each point takes the value of the last region (in loop order) whose write
// range covers it: zero for a non-porous region, the drag formula otherwise.
template <typename T, typename E>
__device__ inline T poros_val(int comp, int njacob, int ld, int i, int j, T porc1, T porc2,
                              const E *__restrict__ u, const E *__restrict__ v) {
    const T dFour = T(4.0);
    T unorm, unrm1 = T(0.0);
    if (comp == 1) {
        const T uu = cv<T>(u[WD2_IX(i, j)]);
        const T w = cv<T>(v[WD2_IX(i, j)]) + cv<T>(v[WD2_IX(i + 1, j)]) + cv<T>(v[WD2_IX(i, j - 1)]) +
                    cv<T>(v[WD2_IX(i + 1, j - 1)]) / dFour;
        unorm = sqrt(uu * uu + w * w);
        if (unorm > T(1.e-8)) unrm1 = (uu * uu) / unorm;
    } else {
        const T vv = cv<T>(v[WD2_IX(i, j)]);
        const T w = (cv<T>(u[WD2_IX(i - 1, j + 1)]) + cv<T>(u[WD2_IX(i, j + 1)]) + cv<T>(u[WD2_IX(i - 1, j)]) +
                     cv<T>(u[WD2_IX(i, j)])) / dFour;
        unorm = sqrt(w * w + vv * vv);
        if (unorm > T(1.e-8)) unrm1 = (vv * vv) / unorm;
    }
    if (njacob == 1) unorm = unrm1 + unorm;
    return porc1 + porc2 * unorm;
}

template <typename T>
__global__ __launch_bounds__(MX *MY) void k_poros_coef(Regions R, int ld, int comp, const double *__restrict__ us,
                                                       const double *__restrict__ vs, const double *__restrict__ un,
                                                       const double *__restrict__ vn, MomTempsT<T> t) {
    const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
    if (i > R.nx || j > R.ny) return;
    int wir = 0, wjr = 0, wpor = 0;
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            const bool por = R.type(ir, jr) == RM_POROUS;
            bool in;
            if (!por) in = j >= jS && j <= jN && i >= iW && i <= iE;
            else if (comp == 1) in = j >= jS + 1 && j <= jN && i >= iW && i <= iE;
            else in = j >= jS && j <= jN && i >= iW + 1 && i <= iE;
            if (in) { wir = ir; wjr = jr; wpor = por; }
        }
    if (wir == 0) return;
    const size_t k = WD2_IX(i, j);
    if (!wpor) {
        t.cpj[k] = T(0.0); t.cps[k] = T(0.0); t.cpn[k] = T(0.0);
        return;
    }
    const T c1 = cv<T>(R.tab(R.porc1, wir, wjr)), c2 = cv<T>(R.tab(R.porc2, wir, wjr));
    t.cpj[k] = poros_val<T>(comp, 1, ld, i, j, c1, c2, us, vs);
    t.cps[k] = poros_val<T>(comp, 0, ld, i, j, c1, c2, us, vs);
    t.cpn[k] = poros_val<T>(comp, 0, ld, i, j, c1, c2, un, vn);
}

template <typename T>
void launch_poros_coef(const Regions &R, int ld, int comp, const double *us, const double *vs, const double *un,
                       const double *vn, const MomTempsT<T> &t, hipStream_t s) {
    hipLaunchKernelGGL(k_poros_coef<T>, mgrid(R.nx, R.ny), dim3(MX, MY), 0, s, R, ld, comp, us, vs, un, vn, t);
}
template void launch_poros_coef<double>(const Regions &, int, int, const double *, const double *, const double *,
                                        const double *, const MomTempsT<double> &, hipStream_t);
template void launch_poros_coef<float>(const Regions &, int, int, const double *, const double *, const double *,
                                       const double *, const MomTempsT<float> &, hipStream_t);

// ---------------------------------------------------- convection/diffusion
// DConvU / DConvV (momentum.f:987-1030) at one point, from the four face
// coefficients around it: west/east c1 and south/north c2 (DConvU: c1(i,j),
// c1(i+1,j), c2(i,j-1), c2(i,j); DConvV: c1(i-1,j), c1(i,j), c2(i,j),
// c2(i,j+1)).
template <typename T, typename A>
__device__ inline T dconva(int i, int j, T c1w, T c1e, T c2s, T c2n, const A &U) {
    return -c2s * U(i, j - 1) - c1w * U(i - 1, j) + (c1e - c1w + c2n - c2s) * U(i, j) + c1e * U(i + 1, j) +
           c2n * U(i, j + 1);
}
template <typename T, typename E>
__device__ inline T dconv(int ld, int i, int j, T c1w, T c1e, T c2s, T c2n, const E *u) {
    return dconva<T>(i, j, c1w, c1e, c2s, c2n, GAcc<T, E>{u, ld});
}
template <typename T>
__device__ inline T dconv_u(int ld, int i, int j, const T *c1, const T *c2, const double *u) {
    return dconv<T>(ld, i, j, c1[WD2_IX(i, j)], c1[WD2_IX(i + 1, j)], c2[WD2_IX(i, j - 1)], c2[WD2_IX(i, j)], u);
}
template <typename T>
__device__ inline T dconv_v(int ld, int i, int j, const T *c1, const T *c2, const double *u) {
    return dconv<T>(ld, i, j, c1[WD2_IX(i - 1, j)], c1[WD2_IX(i, j)], c2[WD2_IX(i, j)], c2[WD2_IX(i, j + 1)], u);
}
#define F(f, a, b) cv<T>(mget(f, ld, a, b))
template <typename T, typename A, typename MA>
__device__ inline T ddiff_ua(int ld, int i, int j, const MA &ac, const MA &bc, const MA &bn, const MA &gn,
                             const A &U) {
    const T s1 = F(ac, i + 1, j) * (U(i + 1, j) - U(i, j)) - F(ac, i, j) * (U(i, j) - U(i - 1, j)) +
                 F(bc, i + 1, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                 F(bc, i, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const T s2 = F(bn, i, j) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                 F(bn, i, j - 1) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                 F(gn, i, j) * (U(i, j + 1) - U(i, j)) - F(gn, i, j - 1) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
}
template <typename T, typename A, typename MA>
__device__ inline T ddiff_va(int ld, int i, int j, const MA &an, const MA &bc, const MA &bn, const MA &gc,
                             const A &U) {
    const T s1 = F(an, i, j) * (U(i + 1, j) - U(i, j)) - F(an, i - 1, j) * (U(i, j) - U(i - 1, j)) +
                 F(bn, i, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                 F(bn, i - 1, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const T s2 = F(bc, i, j + 1) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                 F(bc, i, j) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                 F(gc, i, j + 1) * (U(i, j + 1) - U(i, j)) - F(gc, i, j) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
}
#undef F
template <typename T, typename E>
__device__ inline T ddiff_u(int ld, int i, int j, const MArr &ac, const MArr &bc, const MArr &bn, const MArr &gn,
                            const E *u) {
    return ddiff_ua<T>(ld, i, j, ac, bc, bn, gn, GAcc<T, E>{u, ld});
}
template <typename T, typename E>
__device__ inline T ddiff_v(int ld, int i, int j, const MArr &an, const MArr &bc, const MArr &bn, const MArr &gc,
                            const E *u) {
    return ddiff_va<T>(ld, i, j, an, bc, bn, gc, GAcc<T, E>{u, ld});
}

template <typename T>
__global__ __launch_bounds__(MX *MY) void k_conv_diff(int nx, int ny, int ld, int comp, Metrics m,
                                                      const double *__restrict__ fs, const double *__restrict__ fn,
                                                      MomTempsT<T> t) {
    if (comp == 1) {
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        t.cnvs[k] = dconv_u<T>(ld, i, j, t.c1s, t.c2s, fs);
        t.cnvn[k] = dconv_u<T>(ld, i, j, t.c1n, t.c2n, fn);
        t.difs[k] = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), fs);
        t.difn[k] = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), fn);
    } else {
        const int i = 2 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        t.cnvs[k] = dconv_v<T>(ld, i, j, t.c1s, t.c2s, fs);
        t.cnvn[k] = dconv_v<T>(ld, i, j, t.c1n, t.c2n, fn);
        t.difs[k] = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), fs);
        t.difn[k] = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), fn);
    }
}

template <typename T>
void launch_conv_diff(int nx, int ny, int ld, int comp, const Metrics &m, const double *fs, const double *fn,
                      const MomTempsT<T> &t, hipStream_t s) {
    const dim3 g = comp == 1 ? mgrid(nx, ny - 1) : mgrid(nx - 1, ny);
    hipLaunchKernelGGL(k_conv_diff<T>, g, dim3(MX, MY), 0, s, nx, ny, ld, comp, m, fs, fn, t);
}
template void launch_conv_diff<double>(int, int, int, int, const Metrics &, const double *, const double *,
                                       const MomTempsT<double> &, hipStream_t);
template void launch_conv_diff<float>(int, int, int, int, const Metrics &, const double *, const double *,
                                      const MomTempsT<float> &, hipStream_t);

// Fused right-hand-side terms for grids without porous regions.  The
// coefficients ConvCoef writes (cases 1 and 2 with njacob 0) are only read
// by DConvU / DConvV, so these kernels evaluate them where they are needed
// instead of storing them; each is the same expression as in k_conv_coef, so
// the bits are the same.  Every coefficient read lies in ConvCoef's write
// range (DConvU: c1 at i = 1..nx+1, c2 at j = 1..ny; DConvV: c1 at
// i = 1..nx, c2 at j = 1..ny+1).
//
// k_mom_n: the time-level-n terms cnvn and difn of one component, which
// depend only on un, vn and the metrics; the q-l iterations of one
// nAuxMomentum call reuse them.
template <int COMP, typename T, typename E = double>
__global__ __launch_bounds__(MX *MY) void k_mom_n(int nx, int ny, int ld, Metrics m, const E *__restrict__ un,
                                                  const E *__restrict__ vn, T *__restrict__ cnvn,
                                                  T *__restrict__ difn) {
    const T one = cv<T>(1.0);
    if (COMP == 1) {
        int bx, by;
        if (!xtile(nx, ny - 1, bx, by)) return;
        const int i = 1 + bx * MX + threadIdx.x, j = 2 + by * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const MArr xet = MARR(m, xec), yet = MARR(m, yec), xzi = MARR(m, xzn), yzi = MARR(m, yzn);
        const T c1w = cc_x1<T>(ld, i, j, one, xet, yet, un, vn), c1e = cc_x1<T>(ld, i + 1, j, one, xet, yet, un, vn);
        const T c2s = cc_x2<T>(ld, i, j - 1, one, xzi, yzi, un, vn), c2n = cc_x2<T>(ld, i, j, one, xzi, yzi, un, vn);
        const size_t k = WD2_IX(i, j);
        cnvn[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, un);
        difn[k] = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), un);
    } else {
        int bx, by;
        if (!xtile(nx - 1, ny, bx, by)) return;
        const int i = 2 + bx * MX + threadIdx.x, j = 1 + by * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const MArr xet = MARR(m, xen), yet = MARR(m, yen), xzi = MARR(m, xzc), yzi = MARR(m, yzc);
        const T c1w = cc_y1<T>(ld, i - 1, j, one, xet, yet, un, vn), c1e = cc_y1<T>(ld, i, j, one, xet, yet, un, vn);
        const T c2s = cc_y2<T>(ld, i, j, one, xzi, yzi, un, vn), c2n = cc_y2<T>(ld, i, j + 1, one, xzi, yzi, un, vn);
        const size_t k = WD2_IX(i, j);
        cnvn[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, vn);
        difn[k] = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), vn);
    }
}

template <typename T>
void launch_mom_n(int nx, int ny, int ld, int comp, const Metrics &m, const double *un, const double *vn, T *cnvn,
                  T *difn, hipStream_t s) {
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_n<1, T>), xgrid(nx, ny - 1), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn, difn);
    else
        hipLaunchKernelGGL((k_mom_n<2, T>), xgrid(nx - 1, ny), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn, difn);
}
template void launch_mom_n<double>(int, int, int, int, const Metrics &, const double *, const double *, double *,
                                   double *, hipStream_t);
template void launch_mom_n<float>(int, int, int, int, const Metrics &, const double *, const double *, float *,
                                  float *, hipStream_t);

// ------------------------------------------------------------ assembly
// comp 1 (x-momentum): rows (i=1..nx, j=2..ny), k = (j-2)*nx + i-1
// comp 2 (y-momentum): rows (i=2..nx, j=1..ny), k = (j-1)*(nx-1) + i-2
// split 1: first split step with its right-hand side; split 2: second split
// step (y-direction coefficients on the same x-ordered chain), right-hand
// side = first-split solution, wall and blockage identity rows.
//
// One chain row's coefficients (XMomentum 350-496 / YMomentum 675-821) from
// the upwind-Jacobian coefficient cj on the row's point (cj0) and its two
// neighbours along the split direction (cjm below, cjp above) and the two
// metric values of the split (mlo, mhi: rac(i,j), rac(i+1,j) / rgn(i,j-1),
// rgn(i,j) for comp 1; ran(i-1,j), ran(i,j) / rgc(i,j), rgc(i,j+1) for
// comp 2).  The two components sum the diagonal in different orders.
template <int COMP, typename T>
__device__ inline void asm_coef(T cjm, T cj0, T cjp, T mlo, T mhi, T rkj, T re1, T dk2, T cpj, T &a1, T &a2,
                                T &a3) {
    const T dZero = T(0.0), dOne = T(1.0);
    if (cj0 >= dZero) {
        a1 = rkj * (-cjm - re1 * mlo);
        a2 = COMP == 1 ? dOne + rkj * (cj0 + re1 * (mhi + mlo)) + dk2 * cpj
                       : rkj * (cj0 + re1 * (mhi + mlo)) + dk2 * cpj + dOne;
        a3 = rkj * (-re1 * mhi);
    } else {
        a1 = rkj * (-re1 * mlo);
        a2 = COMP == 1 ? dOne + rkj * (-cj0 + re1 * (mhi + mlo)) + dk2 * cpj
                       : rkj * (-cj0 + re1 * (mhi + mlo)) + dk2 * cpj + dOne;
        a3 = rkj * (cjp - re1 * mhi);
    }
}
// The first split's right-hand side (buoy: the y-momentum's buoyancy term).
template <int COMP, typename T, typename E>
__device__ inline T asm_rhs(E fsp, E fnp, T rkj, T re1, T dk2, T cnvs, T cnvn, T difs, T difn, T cps,
                            T cpn, T buoy) {
    const T r = cv<T>(fnp - fsp) + rkj * (-cnvs - cnvn) + rkj * re1 * (difs + difn) -
                dk2 * (cps * cv<T>(fsp) + cpn * cv<T>(fnp));
    return COMP == 1 ? r : r - buoy;
}
// Whether the second split's row at (i, j) is an identity row (a wall or
// inlet face on its split side, or a blocked region); a boundary type out of
// range sets err (2001 / 2003).
template <int COMP>
__device__ inline bool asm_ident(const Regions &R, int i, int j, int *err) {
    bool ident = false;
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            if (COMP == 1) {
                const bool jin = j >= jS + 1 && j <= jN;
                if (R.type(ir, jr) == RM_BLOCKG && jin && i >= iW && i <= iE) ident = true;
                const int tw = R.btype(ir, jr, WEST), te = R.btype(ir, jr, EAST);
                if ((tw == BM_WALL1 || tw == BM_WALL2 || tw == BM_INLET) && jin && i == iW) ident = true;
                if ((te == BM_WALL1 || te == BM_WALL2 || te == BM_INLET) && jin && i == iE) ident = true;
                if (tw < BM_INTERN || tw > BM_OUTLT2 || te < BM_INTERN || te > BM_OUTLT2) *err = 2001;
            } else {
                const bool iin = i >= iW + 1 && i <= iE;
                if (R.type(ir, jr) == RM_BLOCKG && iin && j >= jS && j <= jN) ident = true;
                const int ts = R.btype(ir, jr, SOUTH), tn = R.btype(ir, jr, NORTH);
                if ((ts == BM_WALL1 || ts == BM_WALL2 || ts == BM_INLET) && iin && j == jS) ident = true;
                if ((tn == BM_WALL1 || tn == BM_WALL2 || tn == BM_INLET) && iin && j == jN) ident = true;
                if (ts < BM_INTERN || ts > BM_OUTLT2 || tn < BM_INTERN || tn > BM_OUTLT2) *err = 2003;
            }
        }
    return ident;
}

template <typename T>
__global__ __launch_bounds__(MX *MY) void k_assemble(Regions R, int ld, int comp, int split, double dk_,
                                                     double dk2_, double re1_, double fr_, Metrics m, MomTempsT<T> t,
                                                     const double *__restrict__ fs, const double *__restrict__ fn,
                                                     const double *__restrict__ d, const double *__restrict__ dn,
                                                     const T *__restrict__ xprev, ChainT<T> ch, int *err,
                                                     int live) {
    const int nx = R.nx, ny = R.ny;
    int i, j;
    long k;
    if (comp == 1) {
        i = 1 + blockIdx.x * MX + threadIdx.x; j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        k = (long)(j - 2) * nx + (i - 1);
    } else {
        i = 2 + blockIdx.x * MX + threadIdx.x; j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        k = (long)(j - 1) * (nx - 1) + (i - 2);
    }
    const T dZero = T(0.0), dOne = T(1.0), dFour = T(4.0);
    const T dk = cv<T>(dk_), dk2 = cv<T>(dk2_), re1 = cv<T>(re1_), fr = cv<T>(fr_);
    const size_t p = WD2_IX(i, j);
#define TT(f, a, b) t.f[WD2_IX(a, b)]
#define MM(f, a, b) cv<T>(MGET(m, f, a, b))
    // Without porous regions the Dupuit-Forchheimer coefficients are the +0 the
    // arrays hold, and without a density field d and dn are +0: the values are
    // then used without loading them (the same arithmetic).
    const bool por = live & 1, den = live & 2;
    const T cpj = por ? t.cpj[p] : T(0), cps = por ? t.cps[p] : T(0), cpn = por ? t.cpn[p] : T(0);
    T a1, a2, a3;
    if (comp == 1) {
        const T rkj = dk2 * cv<T>(MGET(m, dju, i, j));
        if (split == 1) {
            asm_coef<1, T>(TT(cj1, i - 1, j), TT(cj1, i, j), TT(cj1, i + 1, j), MM(rac, i, j), MM(rac, i + 1, j), rkj,
                           re1, dk2, cpj, a1, a2, a3);
            ch.b[k] = asm_rhs<1, T>(fs[p], fn[p], rkj, re1, dk2, t.cnvs[p], t.cnvn[p], t.difs[p], t.difn[p], cps, cpn,
                                    dZero);
        } else {
            asm_coef<1, T>(TT(cj2, i, j - 1), TT(cj2, i, j), TT(cj2, i, j + 1), MM(rgn, i, j - 1), MM(rgn, i, j), rkj,
                           re1, dk2, cpj, a1, a2, a3);
            T bb = xprev[k];
            if (asm_ident<1>(R, i, j, err)) { a1 = dZero; a2 = dOne; a3 = dZero; bb = dZero; }
            ch.b[k] = bb;
        }
    } else {
        const T rkj = dk2 * cv<T>(MGET(m, djv, i, j));
        if (split == 1) {
            asm_coef<2, T>(TT(cj1, i - 1, j), TT(cj1, i, j), TT(cj1, i + 1, j), MM(ran, i - 1, j), MM(ran, i, j), rkj,
                           re1, dk2, cpj, a1, a2, a3);
            const T d1 = den ? cv<T>(d[WD2_IX(i, j + 1)]) : T(0), d0 = den ? cv<T>(d[p]) : T(0);
            const T n1 = den ? cv<T>(dn[WD2_IX(i, j + 1)]) : T(0), n0 = den ? cv<T>(dn[p]) : T(0);
            const T buoy = dk * (d1 + d0 + n1 + n0) / (dFour * fr);
            ch.b[k] = asm_rhs<2, T>(fs[p], fn[p], rkj, re1, dk2, t.cnvs[p], t.cnvn[p], t.difs[p], t.difn[p], cps, cpn,
                                    buoy);
        } else {
            asm_coef<2, T>(TT(cj2, i, j - 1), TT(cj2, i, j), TT(cj2, i, j + 1), MM(rgc, i, j), MM(rgc, i, j + 1), rkj,
                           re1, dk2, cpj, a1, a2, a3);
            T bb = xprev[k];
            if (asm_ident<2>(R, i, j, err)) { a1 = dZero; a2 = dOne; a3 = dZero; bb = dZero; }
            ch.b[k] = bb;
        }
    }
#undef TT
#undef MM
    ch.a[k] = a1;
    ch.d[k] = a2;
    ch.c[k] = a3;
}

template <typename T>
void launch_assemble(const Regions &R, int ld, int comp, int split, double dk, double re, double fr,
                     const Metrics &m, const MomTempsT<T> &t, const double *fs, const double *fn, const double *d,
                     const double *dn, const T *xprev, const ChainT<T> &ch, hipStream_t s, int *err, int live) {
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    const dim3 g = comp == 1 ? mgrid(R.nx, R.ny - 1) : mgrid(R.nx - 1, R.ny);
    hipLaunchKernelGGL(k_assemble<T>, g, dim3(MX, MY), 0, s, R, ld, comp, split, dk, dk2, re1, fr, m, t, fs, fn, d,
                       dn, xprev, ch, err, live);
}
template void launch_assemble<double>(const Regions &, int, int, int, double, double, double, const Metrics &,
                                      const MomTempsT<double> &, const double *, const double *, const double *,
                                      const double *, const double *, const ChainT<double> &, hipStream_t, int *,
                                      int);
template void launch_assemble<float>(const Regions &, int, int, int, double, double, double, const Metrics &,
                                     const MomTempsT<float> &, const double *, const double *, const double *,
                                     const double *, const float *, const ChainT<float> &, hipStream_t, int *, int);

// Fused assembly for grids without porous regions: the work of k_conv_coef,
// k_conv_diff (current iterate) and both k_assemble splits of one component
// in one pass over its rows.
//   - cj (ConvCoef case 4 / 5, njacob 1) is evaluated on the row's point and
//     its two neighbours along each split direction.  Where case 4 / 5 writes
//     nothing (cj1 at i = nx+1, cj2 at j = ny+1) the arrays hold +0, and so
//     does this kernel.
//   - cnvs and difs are evaluated at the row's point from the centred
//     coefficients around it, as k_mom_n does; cnvn and difn come from
//     k_mom_n.
//   - ch1 receives the first split's rows, ch2 the second split's
//     coefficients; zm flags the second split's identity rows, whose
//     right-hand side is +0 where the others take the first split's solution
//     (written into ch2.b by the first split's chain solve).
// Every value is the same expression as in k_conv_coef / k_conv_diff /
// k_assemble, so the chains are bit for bit those of the separate kernels.
//
// Two kernels share the rows.  A row whose metric reads (i-1..i+1,
// j-1..j+1) all lie inside their arrays' constant boxes takes the box values
// as constants -- the same values, with no per-read box test and no load --
// in k_mom_asm_cb; on a uniform grid that is every row but those next to the
// ghost points.  Those rows form a rectangle per component (found on the host
// from the boxes); k_mom_asm_gen takes the rest, up to four strips around it
// per component, or every row where no metric is constant.  k_mom_asm_cb
// first copies the us and vs values its rows' stencils read into LDS tiles
// as T (x-momentum: us on i-1..i+1, j-1..j+1 and vs on i-1..i+2, j-2..j+1
// around each row point; y-momentum the mirror image), one coalesced load per
// element in place of about 25 dependent loads per row.  Both kernels serve
// both components in one launch.

// One component's outputs: the time-level-n terms it reads, its two chains
// and its identity-row flags.
template <typename T>
struct AsmOut {
    const T *cnvn, *difn;
    ChainT<T> ch1, ch2;
    unsigned char *zm;
};

// One chain row: the fused assembly of the row at (i, j), with the us, vs
// stencils read through U, V and the metrics through mpick<MK>.
template <int COMP, int MK, typename T, typename E, typename AU, typename AV>
__device__ inline void mom_asm_row(const Regions &R, int ld, int i, int j, T dk, T dk2, T re1, T fr, const Metrics &m,
                                   const AsmOut<T> &o, const AU &U, const AV &V, const E *__restrict__ us,
                                   const E *__restrict__ vs, const E *__restrict__ un, const E *__restrict__ vn,
                                   const double *__restrict__ d, const double *__restrict__ dn, int *err, int live) {
    const int nx = R.nx, ny = R.ny;
    const long k = COMP == 1 ? (long)(j - 2) * nx + (i - 1) : (long)(j - 1) * (nx - 1) + (i - 2);
    const T dZero = T(0.0), dOne = T(1.0), dFour = T(4.0);
    const T one = cv<T>(1.0), two = cv<T>(2.0);
    const size_t p = WD2_IX(i, j);
    const ChainT<T> &ch1 = o.ch1, &ch2 = o.ch2;
#define MM(f, a, b) cv<T>(mget(MPICK(m, f), ld, a, b))
    T a1, a2, a3;
    if (COMP == 1) {
        const auto xzi = MPICK(m, xzu), xet = MPICK(m, xeu), yzi = MPICK(m, yzu), yet = MPICK(m, yeu);
        T c1m, c10, c1p = dZero, c2m, c20, c2p = dZero, junk;
        cc_4a<T>(ld, i - 1, j, two, xzi, xet, yzi, yet, U, V, c1m, junk);
        cc_4a<T>(ld, i, j, two, xzi, xet, yzi, yet, U, V, c10, c20);
        if (i + 1 <= nx) cc_4a<T>(ld, i + 1, j, two, xzi, xet, yzi, yet, U, V, c1p, junk);
        cc_4a<T>(ld, i, j - 1, two, xzi, xet, yzi, yet, U, V, junk, c2m);
        if (j + 1 <= ny) cc_4a<T>(ld, i, j + 1, two, xzi, xet, yzi, yet, U, V, junk, c2p);
        const auto cxe = MPICK(m, xec), cye = MPICK(m, yec), cxz = MPICK(m, xzn), cyz = MPICK(m, yzn);
        const T f1w = cc_x1a<T>(ld, i, j, one, cxe, cye, U, V), f1e = cc_x1a<T>(ld, i + 1, j, one, cxe, cye, U, V);
        const T f2s = cc_x2a<T>(ld, i, j - 1, one, cxz, cyz, U, V), f2n = cc_x2a<T>(ld, i, j, one, cxz, cyz, U, V);
        const T cnvs = dconva<T>(i, j, f1w, f1e, f2s, f2n, U);
        const T difs = ddiff_ua<T>(ld, i, j, MPICK(m, rac), MPICK(m, rbc), MPICK(m, rbn), MPICK(m, rgn), U);
        const T rkj = dk2 * MM(dju, i, j);
        asm_coef<1, T>(c1m, c10, c1p, MM(rac, i, j), MM(rac, i + 1, j), rkj, re1, dk2, dZero, a1, a2, a3);
        ch1.a[k] = a1; ch1.d[k] = a2; ch1.c[k] = a3;
        ch1.b[k] = asm_rhs<1, T>(us[p], un[p], rkj, re1, dk2, cnvs, o.cnvn[p], difs, o.difn[p], dZero, dZero, dZero);
        asm_coef<1, T>(c2m, c20, c2p, MM(rgn, i, j - 1), MM(rgn, i, j), rkj, re1, dk2, dZero, a1, a2, a3);
        const bool id = asm_ident<1>(R, i, j, err);
        if (id) { a1 = dZero; a2 = dOne; a3 = dZero; }
        ch2.a[k] = a1; ch2.d[k] = a2; ch2.c[k] = a3;
        o.zm[k] = id ? 1 : 0;
    } else {
        const auto xzi = MPICK(m, xzv), xet = MPICK(m, xev), yzi = MPICK(m, yzv), yet = MPICK(m, yev);
        T c1m, c10, c1p = dZero, c2m, c20, c2p = dZero, junk;
        cc_5a<T>(ld, i - 1, j, two, xzi, xet, yzi, yet, U, V, c1m, junk);
        cc_5a<T>(ld, i, j, two, xzi, xet, yzi, yet, U, V, c10, c20);
        if (i + 1 <= nx) cc_5a<T>(ld, i + 1, j, two, xzi, xet, yzi, yet, U, V, c1p, junk);
        cc_5a<T>(ld, i, j - 1, two, xzi, xet, yzi, yet, U, V, junk, c2m);
        if (j + 1 <= ny) cc_5a<T>(ld, i, j + 1, two, xzi, xet, yzi, yet, U, V, junk, c2p);
        const auto cxe = MPICK(m, xen), cye = MPICK(m, yen), cxz = MPICK(m, xzc), cyz = MPICK(m, yzc);
        const T f1w = cc_y1a<T>(ld, i - 1, j, one, cxe, cye, U, V), f1e = cc_y1a<T>(ld, i, j, one, cxe, cye, U, V);
        const T f2s = cc_y2a<T>(ld, i, j, one, cxz, cyz, U, V), f2n = cc_y2a<T>(ld, i, j + 1, one, cxz, cyz, U, V);
        const T cnvs = dconva<T>(i, j, f1w, f1e, f2s, f2n, V);
        const T difs = ddiff_va<T>(ld, i, j, MPICK(m, ran), MPICK(m, rbc), MPICK(m, rbn), MPICK(m, rgc), V);
        const T rkj = dk2 * MM(djv, i, j);
        asm_coef<2, T>(c1m, c10, c1p, MM(ran, i - 1, j), MM(ran, i, j), rkj, re1, dk2, dZero, a1, a2, a3);
        ch1.a[k] = a1; ch1.d[k] = a2; ch1.c[k] = a3;
        const bool den = live & 2;
        const T d1 = den ? cv<T>(d[WD2_IX(i, j + 1)]) : T(0), d0 = den ? cv<T>(d[p]) : T(0);
        const T n1 = den ? cv<T>(dn[WD2_IX(i, j + 1)]) : T(0), n0 = den ? cv<T>(dn[p]) : T(0);
        const T buoy = dk * (d1 + d0 + n1 + n0) / (dFour * fr);
        ch1.b[k] = asm_rhs<2, T>(vs[p], vn[p], rkj, re1, dk2, cnvs, o.cnvn[p], difs, o.difn[p], dZero, dZero, buoy);
        asm_coef<2, T>(c2m, c20, c2p, MM(rgc, i, j), MM(rgc, i, j + 1), rkj, re1, dk2, dZero, a1, a2, a3);
        const bool id = asm_ident<2>(R, i, j, err);
        if (id) { a1 = dZero; a2 = dOne; a3 = dZero; }
        ch2.a[k] = a1; ch2.d[k] = a2; ch2.c[k] = a3;
        o.zm[k] = id ? 1 : 0;
    }
#undef MM
}

// Rows i0..i1 x j0..j1 of one component (empty when i0 > i1 or j0 > j1).
struct AsmRect {
    int i0, i1, j0, j1;
};
__host__ __device__ inline int rect_tiles(const AsmRect &r) {
    return (r.i0 > r.i1 || r.j0 > r.j1) ? 0 : ((r.i1 - r.i0 + MX) / MX) * ((r.j1 - r.j0 + MY) / MY);
}

// The constant-box rows of one component: its LDS-tiled workgroup, after the
// us, vs tiles are staged, runs the constant-metric body on its rows inside
// the rectangle r.  The tiles of r are ordered for the XCDs as xtile does.
template <int COMP, typename T, typename E>
__device__ inline void mom_asm_cb_tile(int b, int nb, const AsmRect &r, const Regions &R, int ld, double dk_,
                                       double dk2_, double re1_, double fr_, const Metrics &m, const AsmOut<T> &o,
                                       const E *__restrict__ us, const E *__restrict__ vs,
                                       const E *__restrict__ un, const E *__restrict__ vn,
                                       const double *__restrict__ d, const double *__restrict__ dn, int *err, int live,
                                       unsigned char *tu_raw, unsigned char *tv_raw) {
    const int nx = R.nx, ny = R.ny;
    const int gx = (r.i1 - r.i0 + MX) / MX, nt = gx * ((r.j1 - r.j0 + MY) / MY), per = nb / 8;
    const int tile = (b % 8) * per + b / 8;
    if (tile >= nt) return;
    constexpr int NT = MX * MY;
    constexpr int WU = COMP == 1 ? MX + 2 : MX + 3, HU = COMP == 1 ? MY + 2 : MY + 3;
    constexpr int WV = COMP == 1 ? MX + 3 : MX + 2, HV = COMP == 1 ? MY + 3 : MY + 2;
    T *tu = reinterpret_cast<T *>(tu_raw), *tv = reinterpret_cast<T *>(tv_raw);
    // the tile's first row point (I0, J0) and the us, vs tiles around its rows
    const int I0 = r.i0 + (tile % gx) * MX, J0 = r.j0 + (tile / gx) * MY;
    const int ui0 = COMP == 1 ? I0 - 1 : I0 - 2, uj0 = J0 - 1;
    const int vi0 = I0 - 1, vj0 = COMP == 1 ? J0 - 2 : J0 - 1;
    {
        // Every tile point has i >= 0 and j >= 0; those past the field's last
        // column or row (nx + 1, ny + 1) are read by no row and hold 0.  The
        // loads use clamped indices, so all of a lane's loads are issued
        // before the first is waited for.
        const int tid = threadIdx.x + MX * threadIdx.y;
        E xu[(WU * HU + NT - 1) / NT], xv[(WV * HV + NT - 1) / NT];
#pragma unroll
        for (int q = 0; q < (WU * HU + NT - 1) / NT; q++) {
            const int e = min(tid + q * NT, WU * HU - 1);
            xu[q] = us[WD2_IX(min(ui0 + e % WU, nx + 1), min(uj0 + e / WU, ny + 1))];
        }
#pragma unroll
        for (int q = 0; q < (WV * HV + NT - 1) / NT; q++) {
            const int e = min(tid + q * NT, WV * HV - 1);
            xv[q] = vs[WD2_IX(min(vi0 + e % WV, nx + 1), min(vj0 + e / WV, ny + 1))];
        }
#pragma unroll
        for (int q = 0; q < (WU * HU + NT - 1) / NT; q++) {
            const int e = tid + q * NT;
            if (e < WU * HU) tu[e] = (ui0 + e % WU <= nx + 1 && uj0 + e / WU <= ny + 1) ? cv<T>(xu[q]) : T(0);
        }
#pragma unroll
        for (int q = 0; q < (WV * HV + NT - 1) / NT; q++) {
            const int e = tid + q * NT;
            if (e < WV * HV) tv[e] = (vi0 + e % WV <= nx + 1 && vj0 + e / WV <= ny + 1) ? cv<T>(xv[q]) : T(0);
        }
    }
    __syncthreads();
    const int i = I0 + threadIdx.x, j = J0 + threadIdx.y;
    if (i > r.i1 || j > r.j1) return;
    const LAcc<T> U{tu, ui0, uj0, WU}, V{tv, vi0, vj0, WV};
    mom_asm_row<COMP, MK_CONST, T, E>(R, ld, i, j, cv<T>(dk_), cv<T>(dk2_), cv<T>(re1_), cv<T>(fr_), m, o, U, V, us, vs,
                                  un, vn, d, dn, err, live);
}

// Workgroups [0, g1) take the x-momentum rectangle r1, the rest the
// y-momentum rectangle r2 (g1 and the total are multiples of 8, so each part
// keeps the XCD order).
template <typename T, typename E>
__global__ __launch_bounds__(MX *MY) void k_mom_asm_cb(Regions R, int ld, double dk_, double dk2_, double re1_,
                                                       double fr_, Metrics m, AsmOut<T> o1, AsmOut<T> o2,
                                                       const E *__restrict__ us, const E *__restrict__ vs,
                                                       const E *__restrict__ un, const E *__restrict__ vn,
                                                       const double *__restrict__ d, const double *__restrict__ dn,
                                                       int *err, int live, const int *stop, AsmRect r1, AsmRect r2,
                                                       int g1) {
    if (stop && *stop) return;
    // raw storage: T may be a type with a constructor (the tangent's Dual)
    constexpr int NU = (MX + 3) * (MY + 3) > (MX + 2) * (MY + 2) ? (MX + 3) * (MY + 3) : (MX + 2) * (MY + 2);
    __shared__ __align__(16) unsigned char tu_raw[NU * sizeof(T)], tv_raw[NU * sizeof(T)];
    const int b = (int)blockIdx.x;
    if (b < g1)
        mom_asm_cb_tile<1, T, E>(b, g1, r1, R, ld, dk_, dk2_, re1_, fr_, m, o1, us, vs, un, vn, d, dn, err, live,
                                 tu_raw, tv_raw);
    else
        mom_asm_cb_tile<2, T, E>(b - g1, (int)gridDim.x - g1, r2, R, ld, dk_, dk2_, re1_, fr_, m, o2, us, vs, un, vn, d,
                                 dn, err, live, tu_raw, tv_raw);
}

// The other rows: up to eight rectangles (at most four per component), one
// thread per row, reading us, vs and the metrics (through their boxes) from
// memory.
struct AsmGen {
    int n;
    int comp[8];
    AsmRect r[8];
    long start[9];   // first row of each rectangle in the launch; start[n] = the total
};
constexpr int AG_T = 256;
template <typename T, typename E>
__global__ __launch_bounds__(AG_T) void k_mom_asm_gen(Regions R, int ld, double dk_, double dk2_, double re1_,
                                                      double fr_, Metrics m, AsmOut<T> o1, AsmOut<T> o2,
                                                      const E *__restrict__ us, const E *__restrict__ vs,
                                                      const E *__restrict__ un, const E *__restrict__ vn,
                                                      const double *__restrict__ d, const double *__restrict__ dn,
                                                      int *err, int live, const int *stop, AsmGen g) {
    if (stop && *stop) return;
    const long q = (long)blockIdx.x * AG_T + threadIdx.x;
    if (q >= g.start[g.n]) return;
    int s = 0;
    while (q >= g.start[s + 1]) s++;
    const AsmRect &r = g.r[s];
    const int w = r.i1 - r.i0 + 1;
    const int i = r.i0 + (int)((q - g.start[s]) % w), j = r.j0 + (int)((q - g.start[s]) / w);
    const T dk = cv<T>(dk_), dk2 = cv<T>(dk2_), re1 = cv<T>(re1_), fr = cv<T>(fr_);
    const GAcc<T, E> U{us, ld}, V{vs, ld};
    if (g.comp[s] == 1)
        mom_asm_row<1, MK_BOX, T, E>(R, ld, i, j, dk, dk2, re1, fr, m, o1, U, V, us, vs, un, vn, d, dn, err, live);
    else
        mom_asm_row<2, MK_BOX, T, E>(R, ld, i, j, dk, dk2, re1, fr, m, o2, U, V, us, vs, un, vn, d, dn, err, live);
}

// The constant-box rectangle of component comp's rows (x-momentum: i =
// 1..nx, j = 2..ny; y-momentum: i = 2..nx, j = 1..ny): the rows whose metric
// reads i-1..i+1, j-1..j+1 lie inside the boxes of every metric the
// component reads.  The rest of the rows, as up to four strips, go to *g.
static AsmRect asm_rects(int nx, int ny, const Metrics &m, int comp, AsmGen &g) {
    static const int u1[13] = {MI_xzu, MI_xeu, MI_yzu, MI_yeu, MI_xec, MI_yec, MI_xzn,
                               MI_yzn, MI_rac, MI_rbc, MI_rbn, MI_rgn, MI_dju};
    static const int u2[13] = {MI_xzv, MI_xev, MI_yzv, MI_yev, MI_xen, MI_yen, MI_xzc,
                               MI_yzc, MI_ran, MI_rbc, MI_rbn, MI_rgc, MI_djv};
    const AsmRect all = comp == 1 ? AsmRect{1, nx, 2, ny} : AsmRect{2, nx, 1, ny};
    AsmRect c = all;
    for (int q = 0; q < 13; q++) {
        const MBox &b = m.box[comp == 1 ? u1[q] : u2[q]];
        c.i0 = std::max(c.i0, b.i0 + 1);
        c.i1 = std::min(c.i1, b.i1 - 1);
        c.j0 = std::max(c.j0, b.j0 + 1);
        c.j1 = std::min(c.j1, b.j1 - 1);
    }
    auto add = [&](int i0, int i1, int j0, int j1) {
        if (i0 > i1 || j0 > j1) return;
        g.comp[g.n] = comp;
        g.r[g.n] = AsmRect{i0, i1, j0, j1};
        g.start[g.n + 1] = g.start[g.n] + (long)(i1 - i0 + 1) * (j1 - j0 + 1);
        g.n++;
    };
    if (c.i0 > c.i1 || c.j0 > c.j1) {
        add(all.i0, all.i1, all.j0, all.j1);
        return AsmRect{1, 0, 1, 0};
    }
    add(all.i0, c.i0 - 1, all.j0, all.j1);   // west strip
    add(c.i1 + 1, all.i1, all.j0, all.j1);   // east strip
    add(c.i0, c.i1, all.j0, c.j0 - 1);       // south strip
    add(c.i0, c.i1, c.j1 + 1, all.j1);       // north strip
    return c;
}

// Both kernels for the components asked for (o1: x-momentum, o2:
// y-momentum; a null ch1.a skips a component).
template <typename T, typename E>
static void launch_asm(const Regions &R, int ld, double dk, double re, double fr, const Metrics &m,
                       const AsmOut<T> &o1, const AsmOut<T> &o2, const E *us, const E *vs, const E *un, const E *vn,
                       const double *d, const double *dn, hipStream_t s, int *err, int live, const int *stop) {
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    AsmGen g{};
    const AsmRect none{1, 0, 1, 0};
    const AsmRect r1 = o1.ch1.a ? asm_rects(R.nx, R.ny, m, 1, g) : none;
    const AsmRect r2 = o2.ch1.a ? asm_rects(R.nx, R.ny, m, 2, g) : none;
    const int g1 = 8 * ((rect_tiles(r1) + 7) / 8), g2 = 8 * ((rect_tiles(r2) + 7) / 8);
    if (g1 + g2 > 0)
        hipLaunchKernelGGL((k_mom_asm_cb<T, E>), dim3(g1 + g2), dim3(MX, MY), 0, s, R, ld, dk, dk2, re1, fr, m, o1, o2,
                           us, vs, un, vn, d, dn, err, live, stop, r1, r2, g1);
    if (g.n > 0) {
        const unsigned nb = (unsigned)((g.start[g.n] + AG_T - 1) / AG_T);
        hipLaunchKernelGGL((k_mom_asm_gen<T, E>), dim3(nb), dim3(AG_T), 0, s, R, ld, dk, dk2, re1, fr, m, o1, o2, us,
                           vs, un, vn, d, dn, err, live, stop, g);
    }
}

template <typename T>
void launch_mom_asm(const Regions &R, int ld, int comp, double dk, double re, double fr, const Metrics &m,
                    const MomTempsT<T> &t, const double *us, const double *vs, const double *un, const double *vn,
                    const double *d, const double *dn, const ChainT<T> &ch1, const ChainT<T> &ch2, unsigned char *zm,
                    hipStream_t s, int *err, int live, const int *stop) {
    const AsmOut<T> o{t.cnvn, t.difn, ch1, ch2, zm}, none{};
    launch_asm<T, double>(R, ld, dk, re, fr, m, comp == 1 ? o : none, comp == 1 ? none : o, us, vs, un, vn, d, dn, s,
                          err, live, stop);
}
template void launch_mom_asm<double>(const Regions &, int, int, double, double, double, const Metrics &,
                                     const MomTempsT<double> &, const double *, const double *, const double *,
                                     const double *, const double *, const double *, const ChainT<double> &,
                                     const ChainT<double> &, unsigned char *, hipStream_t, int *, int, const int *);
template void launch_mom_asm<float>(const Regions &, int, int, double, double, double, const Metrics &,
                                    const MomTempsT<float> &, const double *, const double *, const double *,
                                    const double *, const double *, const double *, const ChainT<float> &,
                                    const ChainT<float> &, unsigned char *, hipStream_t, int *, int, const int *);

template <typename T>
void launch_mom_asm_pair(const Regions &R, int ld, double dk, double re, double fr, const Metrics &m,
                         const MomTempsT<T> &t1, const MomTempsT<T> &t2, const double *us, const double *vs,
                         const double *un, const double *vn, const double *d, const double *dn,
                         const ChainT<T> &ch1x, const ChainT<T> &ch2x, unsigned char *zmx, const ChainT<T> &ch1y,
                         const ChainT<T> &ch2y, unsigned char *zmy, hipStream_t s, int *err, int live,
                         const int *stop) {
    const AsmOut<T> o1{t1.cnvn, t1.difn, ch1x, ch2x, zmx}, o2{t2.cnvn, t2.difn, ch1y, ch2y, zmy};
    launch_asm<T, double>(R, ld, dk, re, fr, m, o1, o2, us, vs, un, vn, d, dn, s, err, live, stop);
}
template void launch_mom_asm_pair<double>(const Regions &, int, double, double, double, const Metrics &,
                                          const MomTempsT<double> &, const MomTempsT<double> &, const double *,
                                          const double *, const double *, const double *, const double *,
                                          const double *, const ChainT<double> &, const ChainT<double> &,
                                          unsigned char *, const ChainT<double> &, const ChainT<double> &,
                                          unsigned char *, hipStream_t, int *, int, const int *);
template void launch_mom_asm_pair<float>(const Regions &, int, double, double, double, const Metrics &,
                                         const MomTempsT<float> &, const MomTempsT<float> &, const double *,
                                         const double *, const double *, const double *, const double *,
                                         const double *, const ChainT<float> &, const ChainT<float> &,
                                         unsigned char *, const ChainT<float> &, const ChainT<float> &,
                                         unsigned char *, hipStream_t, int *, int, const int *);

// The tangent-linear step's assembly: the same kernels on dual numbers
// (each field value with its directional derivative), so the chains' values
// are the FP64 kernels' bit for bit and their derivatives follow the same
// expressions by the chain rule.  No density field, no porous regions.
void launch_mom_n_tl(int nx, int ny, int ld, int comp, const Metrics &m, const Dual *un, const Dual *vn, Dual *cnvn,
                     Dual *difn, hipStream_t s) {
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_n<1, Dual, Dual>), xgrid(nx, ny - 1), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn,
                           difn);
    else
        hipLaunchKernelGGL((k_mom_n<2, Dual, Dual>), xgrid(nx - 1, ny), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn,
                           difn);
}
void launch_mom_asm_tl_pair(const Regions &R, int ld, double dk, double re, double fr, const Metrics &m,
                            const MomTempsT<Dual> &t1, const MomTempsT<Dual> &t2, const Dual *us, const Dual *vs,
                            const Dual *un, const Dual *vn, const ChainT<Dual> &ch1x, const ChainT<Dual> &ch2x,
                            unsigned char *zmx, const ChainT<Dual> &ch1y, const ChainT<Dual> &ch2y,
                            unsigned char *zmy, hipStream_t s, int *err) {
    const AsmOut<Dual> o1{t1.cnvn, t1.difn, ch1x, ch2x, zmx}, o2{t2.cnvn, t2.difn, ch1y, ch2y, zmy};
    launch_asm<Dual, Dual>(R, ld, dk, re, fr, m, o1, o2, us, vs, un, vn, nullptr, nullptr, s, err, 0, nullptr);
}

// ------------------------------------------------------------- scatter
template <typename T>
__global__ __launch_bounds__(MX *MY) void k_scatter(int nx, int ny, int ld, int comp, const T *__restrict__ x,
                                                    double *__restrict__ out) {
    if (comp == 1) {
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        out[WD2_IX(i, j)] = (double)x[(long)(j - 2) * nx + (i - 1)];
    } else {
        const int i = 2 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        out[WD2_IX(i, j)] = (double)x[(long)(j - 1) * (nx - 1) + (i - 2)];
    }
}

template <typename T>
void launch_scatter(int nx, int ny, int ld, int comp, const T *x, double *out, hipStream_t s) {
    const dim3 g = comp == 1 ? mgrid(nx, ny - 1) : mgrid(nx - 1, ny);
    hipLaunchKernelGGL(k_scatter<T>, g, dim3(MX, MY), 0, s, nx, ny, ld, comp, x, out);
}
template void launch_scatter<double>(int, int, int, int, const double *, double *, hipStream_t);
template void launch_scatter<float>(int, int, int, int, const float *, double *, hipStream_t);

// ------------------------------------------- q-l update and increment norms
// us += dus, vs += dvs on i=1..nx, j=1..ny (momentum.f:171-176) and
// DMaxNorm of dus -> val[0], of dvs -> val[1] (momentum.f:179-180), folded by
// a one-workgroup kernel.
__global__ __launch_bounds__(MX *MY) void k_update_norms(int nx, int ny, int ld, double *__restrict__ us,
                                                         double *__restrict__ vs, const double *__restrict__ dus,
                                                         const double *__restrict__ dvs, double *__restrict__ partials) {
    __shared__ double sh[MX * MY / 64];
    const int tid = threadIdx.x + MX * threadIdx.y;
    double m1 = 0.0, m2 = 0.0;
    const int i = 1 + blockIdx.x * MX + threadIdx.x;
    for (int j = 1 + blockIdx.y * MY + threadIdx.y; j <= ny; j += gridDim.y * MY) {
        if (i <= nx) {
            const size_t k = WD2_IX(i, j);
            const double du = dus[k], dv = dvs[k];
            us[k] = us[k] + du;
            vs[k] = vs[k] + dv;
            if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1) {
                m1 = dmaxf(m1, fabs(du));
                m2 = dmaxf(m2, fabs(dv));
            }
        }
    }
    m1 = block_max<MX * MY>(m1, sh, tid);
    m2 = block_max<MX * MY>(m2, sh, tid);
    const unsigned int bid = blockIdx.x + gridDim.x * blockIdx.y;
    if (tid == 0) { partials[2 * bid] = m1; partials[2 * bid + 1] = m2; }
}

__global__ __launch_bounds__(256) void k_update_norms_fold(int ld, const double *__restrict__ dus,
                                                           const double *__restrict__ dvs,
                                                           const double *__restrict__ partials, int nb, Scalars *sc) {
    __shared__ double sh[4];
    double r1 = 0.0, r2 = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) {
        r1 = dmaxf(r1, partials[2 * q]);
        r2 = dmaxf(r2, partials[2 * q + 1]);
    }
    r1 = block_max<256>(r1, sh, threadIdx.x);
    r2 = block_max<256>(r2, sh, threadIdx.x);
    if (threadIdx.x == 0) {
        sc->val[0] = dmaxf(fabs(dus[WD2_IX(5, 5)]), r1);
        sc->val[1] = dmaxf(fabs(dvs[WD2_IX(5, 5)]), r2);
    }
}

void launch_update_norms(int nx, int ny, int ld, double *us, double *vs, const double *dus, const double *dvs,
                         double *partials, Scalars *sc, hipStream_t s) {
    const int gx = (nx + MX - 1) / MX;
    const int gy = min((ny + MY - 1) / MY, max(1, 1024 / gx));
    hipLaunchKernelGGL(k_update_norms, dim3(gx, gy), dim3(MX, MY), 0, s, nx, ny, ld, us, vs, dus, dvs, partials);
    hipLaunchKernelGGL(k_update_norms_fold, dim3(1), dim3(256), 0, s, ld, dus, dvs, partials, gx * gy, sc);
}

// k_scatter (comp 1 from x1, comp 2 from x2) and k_update_norms per point:
// dus is taken from x1 where the X scatter writes it (j >= 2) and read
// elsewhere, dvs likewise (i >= 2), so every sum is the one the separate
// kernels give.  Of the increments taken from the chains only those at
// (5, 5) are stored: the fold reads the stored increments there, nothing
// else reads them during the q-l loop, and every point outside the chain
// rows keeps its stored value.  A wavefront covers one row j.
template <typename T>
__global__ __launch_bounds__(MX *MY) void k_scatter_update(int nx, int ny, int ld, const T *__restrict__ x1,
                                                           const T *__restrict__ x2, double *__restrict__ us,
                                                           double *__restrict__ vs, double *__restrict__ dus,
                                                           double *__restrict__ dvs, double *__restrict__ partials,
                                                           const int *stop) {
    if (*stop) return;
    __shared__ double sh[MX * MY / 64];
    const int tid = threadIdx.x + MX * threadIdx.y;
    double m1 = 0.0, m2 = 0.0;
    const int i = 1 + blockIdx.x * MX + threadIdx.x;
    for (int j = 1 + blockIdx.y * MY + threadIdx.y; j <= ny; j += gridDim.y * MY) {
        if (i <= nx) {
            const size_t k = WD2_IX(i, j);
            double du, dv;
            const bool at55 = i == 5 && j == 5;
            if (j >= 2) {
                du = (double)x1[(long)(j - 2) * nx + (i - 1)];
                if (at55) dus[k] = du;
            } else {
                du = dus[k];
            }
            if (i >= 2) {
                dv = (double)x2[(long)(j - 1) * (nx - 1) + (i - 2)];
                if (at55) dvs[k] = dv;
            } else {
                dv = dvs[k];
            }
            us[k] = us[k] + du;
            vs[k] = vs[k] + dv;
            if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1) {
                m1 = dmaxf(m1, fabs(du));
                m2 = dmaxf(m2, fabs(dv));
            }
        }
    }
    m1 = block_max<MX * MY>(m1, sh, tid);
    m2 = block_max<MX * MY>(m2, sh, tid);
    const unsigned int bid = blockIdx.x + gridDim.x * blockIdx.y;
    if (tid == 0) { partials[2 * bid] = m1; partials[2 * bid + 1] = m2; }
}

// k_update_norms_fold, then the q-l stopping test of iteration m
// (momentum.f:182-188: stop once max(DMaxNorm(dus), DMaxNorm(dvs)) <= qtol).
__global__ __launch_bounds__(256) void k_ql_fold(int ld, const double *__restrict__ dus,
                                                 const double *__restrict__ dvs, const double *__restrict__ partials,
                                                 int nb, Scalars *sc, double qtol, int m) {
    if (sc->ql_stop) return;
    __shared__ double sh[4];
    double r1 = 0.0, r2 = 0.0;
    for (int q = threadIdx.x; q < nb; q += 256) {
        r1 = dmaxf(r1, partials[2 * q]);
        r2 = dmaxf(r2, partials[2 * q + 1]);
    }
    r1 = block_max<256>(r1, sh, threadIdx.x);
    r2 = block_max<256>(r2, sh, threadIdx.x);
    if (threadIdx.x == 0) {
        const double v0 = dmaxf(fabs(dus[WD2_IX(5, 5)]), r1), v1 = dmaxf(fabs(dvs[WD2_IX(5, 5)]), r2);
        sc->val[0] = v0;
        sc->val[1] = v1;
        const double difmax = v0 > v1 ? v0 : v1;
        if (difmax <= qtol) sc->ql_stop = m;
    }
}

__global__ void k_ql_reset(Scalars *sc) { sc->ql_stop = 0; }
void launch_ql_reset(Scalars *sc, hipStream_t s) { hipLaunchKernelGGL(k_ql_reset, dim3(1), dim3(64), 0, s, sc); }

template <typename T>
void launch_scatter_update(int nx, int ny, int ld, const T *x1, const T *x2, double *us, double *vs, double *dus,
                           double *dvs, double *partials, Scalars *sc, double qtol, int m, hipStream_t s) {
    const int gx = (nx + MX - 1) / MX;
    const int gy = min((ny + MY - 1) / MY, max(1, 1024 / gx));
    hipLaunchKernelGGL(k_scatter_update<T>, dim3(gx, gy), dim3(MX, MY), 0, s, nx, ny, ld, x1, x2, us, vs, dus, dvs,
                       partials, &sc->ql_stop);
    hipLaunchKernelGGL(k_ql_fold, dim3(1), dim3(256), 0, s, ld, dus, dvs, partials, gx * gy, sc, qtol, m);
}
template void launch_scatter_update<double>(int, int, int, const double *, const double *, double *, double *,
                                            double *, double *, double *, Scalars *, double, int, hipStream_t);
template void launch_scatter_update<float>(int, int, int, const float *, const float *, double *, double *,
                                           double *, double *, double *, Scalars *, double, int, hipStream_t);

}  // namespace wd2

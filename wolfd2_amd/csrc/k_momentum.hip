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

// The centred face coefficients of ConvCoef cases 1 (cc_x1, cc_x2) and 2
// (cc_y1, cc_y2) at one point (momentum.f:901-925), shared by k_conv_coef and
// the fused right-hand-side kernels, which evaluate them on the fly: one
// expression, so the same bits either way.
#define U(a, b) cv<T>(u[WD2_IX(a, b)])
#define V(a, b) cv<T>(v[WD2_IX(a, b)])
#define M(f, a, b) cv<T>(mget(f, ld, a, b))
template <typename T, typename E>
__device__ inline T cc_x1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return (djac * M(yet, i, j) * (U(i, j) + U(i - 1, j)) - M(xet, i, j) * (V(i, j) + V(i, j - 1))) * T(0.5);
}
template <typename T, typename E>
__device__ inline T cc_x2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return (M(xzi, i, j) * (V(i + 1, j) + V(i, j)) - djac * M(yzi, i, j) * (U(i, j + 1) + U(i, j))) * T(0.5);
}
template <typename T, typename E>
__device__ inline T cc_y1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return (M(yet, i, j) * (U(i, j + 1) + U(i, j)) - djac * M(xet, i, j) * (V(i + 1, j) + V(i, j))) * T(0.5);
}
template <typename T, typename E>
__device__ inline T cc_y2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const E *__restrict__ u, const E *__restrict__ v) {
    return (djac * M(xzi, i, j) * (V(i, j) + V(i, j - 1)) - M(yzi, i, j) * (U(i, j) + U(i - 1, j))) * T(0.5);
}
// ConvCoef cases 4 and 5 (upwind Jacobian, momentum.f:927-961)
template <typename T, typename E>
__device__ inline void cc_4(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const E *__restrict__ u, const E *__restrict__ v, T &c1, T &c2) {
    const T dFour = T(4.0);
    c1 = djac * M(yet, i, j) * U(i, j) - M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour;
    c2 = M(xzi, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour - djac * M(yzi, i, j) * U(i, j);
}
template <typename T, typename E>
__device__ inline void cc_5(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const E *__restrict__ u, const E *__restrict__ v, T &c1, T &c2) {
    const T dFour = T(4.0);
    c1 = M(yet, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour - djac * M(xet, i, j) * V(i, j);
    c2 = djac * M(xzi, i, j) * V(i, j) - M(yzi, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour;
}
#undef U
#undef V
#undef M

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
#define U(a, b) cv<T>(u[WD2_IX(a, b)])
template <typename T, typename E>
__device__ inline T dconv(int ld, int i, int j, T c1w, T c1e, T c2s, T c2n, const E *u) {
    return -c2s * U(i, j - 1) - c1w * U(i - 1, j) + (c1e - c1w + c2n - c2s) * U(i, j) + c1e * U(i + 1, j) +
           c2n * U(i, j + 1);
}
template <typename T>
__device__ inline T dconv_u(int ld, int i, int j, const T *c1, const T *c2, const double *u) {
    return dconv<T>(ld, i, j, c1[WD2_IX(i, j)], c1[WD2_IX(i + 1, j)], c2[WD2_IX(i, j - 1)], c2[WD2_IX(i, j)], u);
}
template <typename T>
__device__ inline T dconv_v(int ld, int i, int j, const T *c1, const T *c2, const double *u) {
    return dconv<T>(ld, i, j, c1[WD2_IX(i - 1, j)], c1[WD2_IX(i, j)], c2[WD2_IX(i, j)], c2[WD2_IX(i, j + 1)], u);
}
template <typename T, typename E>
__device__ inline T ddiff_u(int ld, int i, int j, const MArr &ac, const MArr &bc, const MArr &bn, const MArr &gn,
                            const E *u) {
#define F(f, a, b) cv<T>(mget(f, ld, a, b))
    const T s1 = F(ac, i + 1, j) * (U(i + 1, j) - U(i, j)) - F(ac, i, j) * (U(i, j) - U(i - 1, j)) +
                 F(bc, i + 1, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                 F(bc, i, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const T s2 = F(bn, i, j) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                 F(bn, i, j - 1) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                 F(gn, i, j) * (U(i, j + 1) - U(i, j)) - F(gn, i, j - 1) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
}
template <typename T, typename E>
__device__ inline T ddiff_v(int ld, int i, int j, const MArr &an, const MArr &bc, const MArr &bn, const MArr &gc,
                            const E *u) {
    const T s1 = F(an, i, j) * (U(i + 1, j) - U(i, j)) - F(an, i - 1, j) * (U(i, j) - U(i - 1, j)) +
                 F(bn, i, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                 F(bn, i - 1, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const T s2 = F(bc, i, j + 1) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                 F(bc, i, j) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                 F(gc, i, j + 1) * (U(i, j + 1) - U(i, j)) - F(gc, i, j) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
#undef F
#undef U
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
template <int COMP, typename T, typename E = double>
__global__ __launch_bounds__(MX *MY) void k_mom_asm(Regions R, int ld, double dk_, double dk2_, double re1_,
                                                    double fr_, Metrics m, MomTempsT<T> t,
                                                    const E *__restrict__ us, const E *__restrict__ vs,
                                                    const E *__restrict__ un, const E *__restrict__ vn,
                                                    const double *__restrict__ d, const double *__restrict__ dn,
                                                    ChainT<T> ch1, ChainT<T> ch2, unsigned char *__restrict__ zm,
                                                    int *err, int live, const int *stop) {
    if (stop && *stop) return;
    const int nx = R.nx, ny = R.ny;
    int i, j;
    long k;
    int bx, by;
    if (COMP == 1) {
        if (!xtile(nx, ny - 1, bx, by)) return;
        i = 1 + bx * MX + threadIdx.x; j = 2 + by * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        k = (long)(j - 2) * nx + (i - 1);
    } else {
        if (!xtile(nx - 1, ny, bx, by)) return;
        i = 2 + bx * MX + threadIdx.x; j = 1 + by * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        k = (long)(j - 1) * (nx - 1) + (i - 2);
    }
    const T dZero = T(0.0), dOne = T(1.0), dFour = T(4.0);
    const T dk = cv<T>(dk_), dk2 = cv<T>(dk2_), re1 = cv<T>(re1_), fr = cv<T>(fr_);
    const T one = cv<T>(1.0), two = cv<T>(2.0);
    const size_t p = WD2_IX(i, j);
#define MM(f, a, b) cv<T>(MGET(m, f, a, b))
    T a1, a2, a3;
    if (COMP == 1) {
        const MArr xzi = MARR(m, xzu), xet = MARR(m, xeu), yzi = MARR(m, yzu), yet = MARR(m, yeu);
        T c1m, c10, c1p = dZero, c2m, c20, c2p = dZero, junk;
        cc_4<T>(ld, i - 1, j, two, xzi, xet, yzi, yet, us, vs, c1m, junk);
        cc_4<T>(ld, i, j, two, xzi, xet, yzi, yet, us, vs, c10, c20);
        if (i + 1 <= nx) cc_4<T>(ld, i + 1, j, two, xzi, xet, yzi, yet, us, vs, c1p, junk);
        cc_4<T>(ld, i, j - 1, two, xzi, xet, yzi, yet, us, vs, junk, c2m);
        if (j + 1 <= ny) cc_4<T>(ld, i, j + 1, two, xzi, xet, yzi, yet, us, vs, junk, c2p);
        const MArr cxe = MARR(m, xec), cye = MARR(m, yec), cxz = MARR(m, xzn), cyz = MARR(m, yzn);
        const T f1w = cc_x1<T>(ld, i, j, one, cxe, cye, us, vs), f1e = cc_x1<T>(ld, i + 1, j, one, cxe, cye, us, vs);
        const T f2s = cc_x2<T>(ld, i, j - 1, one, cxz, cyz, us, vs), f2n = cc_x2<T>(ld, i, j, one, cxz, cyz, us, vs);
        const T cnvs = dconv<T>(ld, i, j, f1w, f1e, f2s, f2n, us);
        const T difs = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), us);
        const T rkj = dk2 * cv<T>(MGET(m, dju, i, j));
        asm_coef<1, T>(c1m, c10, c1p, MM(rac, i, j), MM(rac, i + 1, j), rkj, re1, dk2, dZero, a1, a2, a3);
        ch1.a[k] = a1; ch1.d[k] = a2; ch1.c[k] = a3;
        ch1.b[k] = asm_rhs<1, T>(us[p], un[p], rkj, re1, dk2, cnvs, t.cnvn[p], difs, t.difn[p], dZero, dZero, dZero);
        asm_coef<1, T>(c2m, c20, c2p, MM(rgn, i, j - 1), MM(rgn, i, j), rkj, re1, dk2, dZero, a1, a2, a3);
        const bool id = asm_ident<1>(R, i, j, err);
        if (id) { a1 = dZero; a2 = dOne; a3 = dZero; }
        ch2.a[k] = a1; ch2.d[k] = a2; ch2.c[k] = a3;
        zm[k] = id ? 1 : 0;
    } else {
        const MArr xzi = MARR(m, xzv), xet = MARR(m, xev), yzi = MARR(m, yzv), yet = MARR(m, yev);
        T c1m, c10, c1p = dZero, c2m, c20, c2p = dZero, junk;
        cc_5<T>(ld, i - 1, j, two, xzi, xet, yzi, yet, us, vs, c1m, junk);
        cc_5<T>(ld, i, j, two, xzi, xet, yzi, yet, us, vs, c10, c20);
        if (i + 1 <= nx) cc_5<T>(ld, i + 1, j, two, xzi, xet, yzi, yet, us, vs, c1p, junk);
        cc_5<T>(ld, i, j - 1, two, xzi, xet, yzi, yet, us, vs, junk, c2m);
        if (j + 1 <= ny) cc_5<T>(ld, i, j + 1, two, xzi, xet, yzi, yet, us, vs, junk, c2p);
        const MArr cxe = MARR(m, xen), cye = MARR(m, yen), cxz = MARR(m, xzc), cyz = MARR(m, yzc);
        const T f1w = cc_y1<T>(ld, i - 1, j, one, cxe, cye, us, vs), f1e = cc_y1<T>(ld, i, j, one, cxe, cye, us, vs);
        const T f2s = cc_y2<T>(ld, i, j, one, cxz, cyz, us, vs), f2n = cc_y2<T>(ld, i, j + 1, one, cxz, cyz, us, vs);
        const T cnvs = dconv<T>(ld, i, j, f1w, f1e, f2s, f2n, vs);
        const T difs = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), vs);
        const T rkj = dk2 * cv<T>(MGET(m, djv, i, j));
        asm_coef<2, T>(c1m, c10, c1p, MM(ran, i - 1, j), MM(ran, i, j), rkj, re1, dk2, dZero, a1, a2, a3);
        ch1.a[k] = a1; ch1.d[k] = a2; ch1.c[k] = a3;
        const bool den = live & 2;
        const T d1 = den ? cv<T>(d[WD2_IX(i, j + 1)]) : T(0), d0 = den ? cv<T>(d[p]) : T(0);
        const T n1 = den ? cv<T>(dn[WD2_IX(i, j + 1)]) : T(0), n0 = den ? cv<T>(dn[p]) : T(0);
        const T buoy = dk * (d1 + d0 + n1 + n0) / (dFour * fr);
        ch1.b[k] = asm_rhs<2, T>(vs[p], vn[p], rkj, re1, dk2, cnvs, t.cnvn[p], difs, t.difn[p], dZero, dZero, buoy);
        asm_coef<2, T>(c2m, c20, c2p, MM(rgc, i, j), MM(rgc, i, j + 1), rkj, re1, dk2, dZero, a1, a2, a3);
        const bool id = asm_ident<2>(R, i, j, err);
        if (id) { a1 = dZero; a2 = dOne; a3 = dZero; }
        ch2.a[k] = a1; ch2.d[k] = a2; ch2.c[k] = a3;
        zm[k] = id ? 1 : 0;
    }
#undef MM
}

template <typename T>
void launch_mom_asm(const Regions &R, int ld, int comp, double dk, double re, double fr, const Metrics &m,
                    const MomTempsT<T> &t, const double *us, const double *vs, const double *un, const double *vn,
                    const double *d, const double *dn, const ChainT<T> &ch1, const ChainT<T> &ch2, unsigned char *zm,
                    hipStream_t s, int *err, int live, const int *stop) {
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_asm<1, T>), xgrid(R.nx, R.ny - 1), dim3(MX, MY), 0, s, R, ld, dk, dk2, re1, fr, m,
                           t, us, vs, un, vn, d, dn, ch1, ch2, zm, err, live, stop);
    else
        hipLaunchKernelGGL((k_mom_asm<2, T>), xgrid(R.nx - 1, R.ny), dim3(MX, MY), 0, s, R, ld, dk, dk2, re1, fr, m,
                           t, us, vs, un, vn, d, dn, ch1, ch2, zm, err, live, stop);
}
template void launch_mom_asm<double>(const Regions &, int, int, double, double, double, const Metrics &,
                                     const MomTempsT<double> &, const double *, const double *, const double *,
                                     const double *, const double *, const double *, const ChainT<double> &,
                                     const ChainT<double> &, unsigned char *, hipStream_t, int *, int, const int *);
template void launch_mom_asm<float>(const Regions &, int, int, double, double, double, const Metrics &,
                                    const MomTempsT<float> &, const double *, const double *, const double *,
                                    const double *, const double *, const double *, const ChainT<float> &,
                                    const ChainT<float> &, unsigned char *, hipStream_t, int *, int, const int *);

// The tangent-linear step's assembly: the same two kernels on dual numbers
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
void launch_mom_asm_tl(const Regions &R, int ld, int comp, double dk, double re, double fr, const Metrics &m,
                       const MomTempsT<Dual> &t, const Dual *us, const Dual *vs, const Dual *un, const Dual *vn,
                       const ChainT<Dual> &ch1, const ChainT<Dual> &ch2, unsigned char *zm, hipStream_t s, int *err) {
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_asm<1, Dual, Dual>), xgrid(R.nx, R.ny - 1), dim3(MX, MY), 0, s, R, ld, dk, dk2, re1,
                           fr, m, t, us, vs, un, vn, nullptr, nullptr, ch1, ch2, zm, err, 0, nullptr);
    else
        hipLaunchKernelGGL((k_mom_asm<2, Dual, Dual>), xgrid(R.nx - 1, R.ny), dim3(MX, MY), 0, s, R, ld, dk, dk2, re1,
                           fr, m, t, us, vs, un, vn, nullptr, nullptr, ch1, ch2, zm, err, 0, nullptr);
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
// dus is written where the X scatter writes it (j >= 2) and read elsewhere,
// dvs likewise (i >= 2), so every stored value and every sum is the one the
// separate kernels give.  A wavefront covers one row j.
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
            if (j >= 2) {
                du = (double)x1[(long)(j - 2) * nx + (i - 1)];
                dus[k] = du;
            } else {
                du = dus[k];
            }
            if (i >= 2) {
                dv = (double)x2[(long)(j - 1) * (nx - 1) + (i - 2)];
                dvs[k] = dv;
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

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

// ------------------------------------------------------------- ConvCoef
template <typename T>
__device__ inline T cv(double x) { return (T)x; }

// The centred face coefficients of ConvCoef cases 1 (cc_x1, cc_x2) and 2
// (cc_y1, cc_y2) at one point (momentum.f:901-925), shared by k_conv_coef and
// the fused right-hand-side kernels, which evaluate them on the fly: one
// expression, so the same bits either way.
#define U(a, b) cv<T>(u[WD2_IX(a, b)])
#define V(a, b) cv<T>(v[WD2_IX(a, b)])
#define M(f, a, b) cv<T>(mget(f, ld, a, b))
template <typename T>
__device__ inline T cc_x1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const double *__restrict__ u, const double *__restrict__ v) {
    return (djac * M(yet, i, j) * (U(i, j) + U(i - 1, j)) - M(xet, i, j) * (V(i, j) + V(i, j - 1))) * T(0.5);
}
template <typename T>
__device__ inline T cc_x2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const double *__restrict__ u, const double *__restrict__ v) {
    return (M(xzi, i, j) * (V(i + 1, j) + V(i, j)) - djac * M(yzi, i, j) * (U(i, j + 1) + U(i, j))) * T(0.5);
}
template <typename T>
__device__ inline T cc_y1(int ld, int i, int j, T djac, const MArr &xet, const MArr &yet,
                          const double *__restrict__ u, const double *__restrict__ v) {
    return (M(yet, i, j) * (U(i, j + 1) + U(i, j)) - djac * M(xet, i, j) * (V(i + 1, j) + V(i, j))) * T(0.5);
}
template <typename T>
__device__ inline T cc_y2(int ld, int i, int j, T djac, const MArr &xzi, const MArr &yzi,
                          const double *__restrict__ u, const double *__restrict__ v) {
    return (djac * M(xzi, i, j) * (V(i, j) + V(i, j - 1)) - M(yzi, i, j) * (U(i, j) + U(i - 1, j))) * T(0.5);
}
// ConvCoef cases 4 and 5 (upwind Jacobian, momentum.f:927-961)
template <typename T>
__device__ inline void cc_4(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const double *__restrict__ u, const double *__restrict__ v, T &c1, T &c2) {
    const T dFour = T(4.0);
    c1 = djac * M(yet, i, j) * U(i, j) - M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour;
    c2 = M(xzi, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour - djac * M(yzi, i, j) * U(i, j);
}
template <typename T>
__device__ inline void cc_5(int ld, int i, int j, T djac, const MArr &xzi, const MArr &xet, const MArr &yzi,
                            const MArr &yet, const double *__restrict__ u, const double *__restrict__ v, T &c1, T &c2) {
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
template <typename T>
__device__ inline T poros_val(int comp, int njacob, int ld, int i, int j, T porc1, T porc2,
                              const double *__restrict__ u, const double *__restrict__ v) {
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
template <typename T>
__device__ inline T dconv(int ld, int i, int j, T c1w, T c1e, T c2s, T c2n, const double *u) {
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
template <typename T>
__device__ inline T ddiff_u(int ld, int i, int j, const MArr &ac, const MArr &bc, const MArr &bn, const MArr &gn,
                            const double *u) {
#define F(f, a, b) cv<T>(mget(f, ld, a, b))
    const T s1 = F(ac, i + 1, j) * (U(i + 1, j) - U(i, j)) - F(ac, i, j) * (U(i, j) - U(i - 1, j)) +
                 F(bc, i + 1, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                 F(bc, i, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const T s2 = F(bn, i, j) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                 F(bn, i, j - 1) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                 F(gn, i, j) * (U(i, j + 1) - U(i, j)) - F(gn, i, j - 1) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
}
template <typename T>
__device__ inline T ddiff_v(int ld, int i, int j, const MArr &an, const MArr &bc, const MArr &bn, const MArr &gc,
                            const double *u) {
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
template <int COMP, typename T>
__global__ __launch_bounds__(MX *MY) void k_mom_n(int nx, int ny, int ld, Metrics m, const double *__restrict__ un,
                                                  const double *__restrict__ vn, T *__restrict__ cnvn,
                                                  T *__restrict__ difn) {
    const T one = cv<T>(1.0);
    if (COMP == 1) {
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const MArr xet = MARR(m, xec), yet = MARR(m, yec), xzi = MARR(m, xzn), yzi = MARR(m, yzn);
        const T c1w = cc_x1<T>(ld, i, j, one, xet, yet, un, vn), c1e = cc_x1<T>(ld, i + 1, j, one, xet, yet, un, vn);
        const T c2s = cc_x2<T>(ld, i, j - 1, one, xzi, yzi, un, vn), c2n = cc_x2<T>(ld, i, j, one, xzi, yzi, un, vn);
        const size_t k = WD2_IX(i, j);
        cnvn[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, un);
        difn[k] = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), un);
    } else {
        const int i = 2 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const MArr xet = MARR(m, xen), yet = MARR(m, yen), xzi = MARR(m, xzc), yzi = MARR(m, yzc);
        const T c1w = cc_y1<T>(ld, i - 1, j, one, xet, yet, un, vn), c1e = cc_y1<T>(ld, i, j, one, xet, yet, un, vn);
        const T c2s = cc_y2<T>(ld, i, j, one, xzi, yzi, un, vn), c2n = cc_y2<T>(ld, i, j + 1, one, xzi, yzi, un, vn);
        const size_t k = WD2_IX(i, j);
        cnvn[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, vn);
        difn[k] = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), vn);
    }
}

// k_mom_s: the upwind-Jacobian coefficients cj1, cj2 (ConvCoef case 4 / 5,
// njacob 1, on its write range) and the terms cnvs, difs of the current
// iterate, in one pass over us and vs.
template <int COMP, typename T>
__global__ __launch_bounds__(MX *MY) void k_mom_s(int nx, int ny, int ld, Metrics m, const double *__restrict__ us,
                                                  const double *__restrict__ vs, MomTempsT<T> t) {
    const T one = cv<T>(1.0), two = cv<T>(2.0);
    if (COMP == 1) {
        // cover i = 0..nx, j = 1..ny: case 4's range; DConvU / DDiffU on i >= 1, j >= 2
        const int i = blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        cc_4<T>(ld, i, j, two, MARR(m, xzu), MARR(m, xeu), MARR(m, yzu), MARR(m, yeu), us, vs, t.cj1[k], t.cj2[k]);
        if (i >= 1 && j >= 2) {
            const MArr xet = MARR(m, xec), yet = MARR(m, yec), xzi = MARR(m, xzn), yzi = MARR(m, yzn);
            const T c1w = cc_x1<T>(ld, i, j, one, xet, yet, us, vs), c1e = cc_x1<T>(ld, i + 1, j, one, xet, yet, us, vs);
            const T c2s = cc_x2<T>(ld, i, j - 1, one, xzi, yzi, us, vs), c2n = cc_x2<T>(ld, i, j, one, xzi, yzi, us, vs);
            t.cnvs[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, us);
            t.difs[k] = ddiff_u<T>(ld, i, j, MARR(m, rac), MARR(m, rbc), MARR(m, rbn), MARR(m, rgn), us);
        }
    } else {
        // cover i = 1..nx, j = 0..ny: case 5's range; DConvV / DDiffV on i >= 2, j >= 1
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        cc_5<T>(ld, i, j, two, MARR(m, xzv), MARR(m, xev), MARR(m, yzv), MARR(m, yev), us, vs, t.cj1[k], t.cj2[k]);
        if (i >= 2 && j >= 1) {
            const MArr xet = MARR(m, xen), yet = MARR(m, yen), xzi = MARR(m, xzc), yzi = MARR(m, yzc);
            const T c1w = cc_y1<T>(ld, i - 1, j, one, xet, yet, us, vs), c1e = cc_y1<T>(ld, i, j, one, xet, yet, us, vs);
            const T c2s = cc_y2<T>(ld, i, j, one, xzi, yzi, us, vs), c2n = cc_y2<T>(ld, i, j + 1, one, xzi, yzi, us, vs);
            t.cnvs[k] = dconv<T>(ld, i, j, c1w, c1e, c2s, c2n, vs);
            t.difs[k] = ddiff_v<T>(ld, i, j, MARR(m, ran), MARR(m, rbc), MARR(m, rbn), MARR(m, rgc), vs);
        }
    }
}

template <typename T>
void launch_mom_n(int nx, int ny, int ld, int comp, const Metrics &m, const double *un, const double *vn, T *cnvn,
                  T *difn, hipStream_t s) {
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_n<1, T>), mgrid(nx, ny - 1), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn, difn);
    else
        hipLaunchKernelGGL((k_mom_n<2, T>), mgrid(nx - 1, ny), dim3(MX, MY), 0, s, nx, ny, ld, m, un, vn, cnvn, difn);
}
template <typename T>
void launch_mom_s(int nx, int ny, int ld, int comp, const Metrics &m, const double *us, const double *vs,
                  const MomTempsT<T> &t, hipStream_t s) {
    if (comp == 1)
        hipLaunchKernelGGL((k_mom_s<1, T>), mgrid(nx + 1, ny), dim3(MX, MY), 0, s, nx, ny, ld, m, us, vs, t);
    else
        hipLaunchKernelGGL((k_mom_s<2, T>), mgrid(nx, ny + 1), dim3(MX, MY), 0, s, nx, ny, ld, m, us, vs, t);
}
template void launch_mom_n<double>(int, int, int, int, const Metrics &, const double *, const double *, double *,
                                   double *, hipStream_t);
template void launch_mom_n<float>(int, int, int, int, const Metrics &, const double *, const double *, float *,
                                  float *, hipStream_t);
template void launch_mom_s<double>(int, int, int, int, const Metrics &, const double *, const double *,
                                   const MomTempsT<double> &, hipStream_t);
template void launch_mom_s<float>(int, int, int, int, const Metrics &, const double *, const double *,
                                  const MomTempsT<float> &, hipStream_t);

// ------------------------------------------------------------ assembly
// comp 1 (x-momentum): rows (i=1..nx, j=2..ny), k = (j-2)*nx + i-1
// comp 2 (y-momentum): rows (i=2..nx, j=1..ny), k = (j-1)*(nx-1) + i-2
// split 1: first split step with its right-hand side; split 2: second split
// step (y-direction coefficients on the same x-ordered chain), right-hand
// side = first-split solution, wall and blockage identity rows.
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
            if (t.cj1[p] >= dZero) {
                a1 = rkj * (-TT(cj1, i - 1, j) - re1 * MM(rac, i, j));
                a2 = dOne + rkj * (TT(cj1, i, j) + re1 * (MM(rac, i + 1, j) + MM(rac, i, j))) + dk2 * cpj;
                a3 = rkj * (-re1 * MM(rac, i + 1, j));
            } else {
                a1 = rkj * (-re1 * MM(rac, i, j));
                a2 = dOne + rkj * (-TT(cj1, i, j) + re1 * (MM(rac, i + 1, j) + MM(rac, i, j))) + dk2 * cpj;
                a3 = rkj * (TT(cj1, i + 1, j) - re1 * MM(rac, i + 1, j));
            }
            ch.b[k] = cv<T>(fn[p] - fs[p]) + rkj * (-t.cnvs[p] - t.cnvn[p]) + rkj * re1 * (t.difs[p] + t.difn[p]) -
                      dk2 * (cps * cv<T>(fs[p]) + cpn * cv<T>(fn[p]));
        } else {
            if (t.cj2[p] >= dZero) {
                a1 = rkj * (-TT(cj2, i, j - 1) - re1 * MM(rgn, i, j - 1));
                a2 = dOne + rkj * (TT(cj2, i, j) + re1 * (MM(rgn, i, j) + MM(rgn, i, j - 1))) + dk2 * cpj;
                a3 = rkj * (-re1 * MM(rgn, i, j));
            } else {
                a1 = rkj * (-re1 * MM(rgn, i, j - 1));
                a2 = dOne + rkj * (-TT(cj2, i, j) + re1 * (MM(rgn, i, j) + MM(rgn, i, j - 1))) + dk2 * cpj;
                a3 = rkj * (TT(cj2, i, j + 1) - re1 * MM(rgn, i, j));
            }
            T bb = xprev[k];
            bool ident = false;
            for (int jr = 1; jr <= R.nj; jr++)
                for (int ir = 1; ir <= R.ni; ir++) {
                    const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
                    const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
                    const bool jin = j >= jS + 1 && j <= jN;
                    if (R.type(ir, jr) == RM_BLOCKG && jin && i >= iW && i <= iE) ident = true;
                    const int tw = R.btype(ir, jr, WEST), te = R.btype(ir, jr, EAST);
                    if ((tw == BM_WALL1 || tw == BM_WALL2 || tw == BM_INLET) && jin && i == iW) ident = true;
                    if ((te == BM_WALL1 || te == BM_WALL2 || te == BM_INLET) && jin && i == iE) ident = true;
                    if (tw < BM_INTERN || tw > BM_OUTLT2 || te < BM_INTERN || te > BM_OUTLT2) *err = 2001;
                }
            if (ident) { a1 = dZero; a2 = dOne; a3 = dZero; bb = dZero; }
            ch.b[k] = bb;
        }
    } else {
        const T rkj = dk2 * cv<T>(MGET(m, djv, i, j));
        if (split == 1) {
            if (t.cj1[p] >= dZero) {
                a1 = rkj * (-TT(cj1, i - 1, j) - re1 * MM(ran, i - 1, j));
                a2 = rkj * (TT(cj1, i, j) + re1 * (MM(ran, i, j) + MM(ran, i - 1, j))) + dk2 * cpj + dOne;
                a3 = rkj * (-re1 * MM(ran, i, j));
            } else {
                a1 = rkj * (-re1 * MM(ran, i - 1, j));
                a2 = rkj * (-TT(cj1, i, j) + re1 * (MM(ran, i, j) + MM(ran, i - 1, j))) + dk2 * cpj + dOne;
                a3 = rkj * (TT(cj1, i + 1, j) - re1 * MM(ran, i, j));
            }
            const T d1 = den ? cv<T>(d[WD2_IX(i, j + 1)]) : T(0), d0 = den ? cv<T>(d[p]) : T(0);
            const T n1 = den ? cv<T>(dn[WD2_IX(i, j + 1)]) : T(0), n0 = den ? cv<T>(dn[p]) : T(0);
            const T buoy = dk * (d1 + d0 + n1 + n0) / (dFour * fr);
            ch.b[k] = cv<T>(fn[p] - fs[p]) + rkj * (-t.cnvs[p] - t.cnvn[p]) + rkj * re1 * (t.difs[p] + t.difn[p]) -
                      dk2 * (cps * cv<T>(fs[p]) + cpn * cv<T>(fn[p])) - buoy;
        } else {
            if (t.cj2[p] >= dZero) {
                a1 = rkj * (-TT(cj2, i, j - 1) - re1 * MM(rgc, i, j));
                a2 = rkj * (TT(cj2, i, j) + re1 * (MM(rgc, i, j + 1) + MM(rgc, i, j))) + dk2 * cpj + dOne;
                a3 = rkj * (-re1 * MM(rgc, i, j + 1));
            } else {
                a1 = rkj * (-re1 * MM(rgc, i, j));
                a2 = rkj * (-TT(cj2, i, j) + re1 * (MM(rgc, i, j + 1) + MM(rgc, i, j))) + dk2 * cpj + dOne;
                a3 = rkj * (TT(cj2, i, j + 1) - re1 * MM(rgc, i, j + 1));
            }
            T bb = xprev[k];
            bool ident = false;
            for (int jr = 1; jr <= R.nj; jr++)
                for (int ir = 1; ir <= R.ni; ir++) {
                    const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
                    const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
                    const bool iin = i >= iW + 1 && i <= iE;
                    if (R.type(ir, jr) == RM_BLOCKG && iin && j >= jS && j <= jN) ident = true;
                    const int ts = R.btype(ir, jr, SOUTH), tn = R.btype(ir, jr, NORTH);
                    if ((ts == BM_WALL1 || ts == BM_WALL2 || ts == BM_INLET) && iin && j == jS) ident = true;
                    if ((tn == BM_WALL1 || tn == BM_WALL2 || tn == BM_INLET) && iin && j == jN) ident = true;
                    if (ts < BM_INTERN || ts > BM_OUTLT2 || tn < BM_INTERN || tn > BM_OUTLT2) *err = 2003;
                }
            if (ident) { a1 = dZero; a2 = dOne; a3 = dZero; bb = dZero; }
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

}  // namespace wd2

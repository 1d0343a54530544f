// k_momentum.hip -- the Douglas-Gunn ADI momentum step (momentum.f:199-1226).
//
// One kernel per reference stage, each a single pass over the grid:
//   ConvCoef (864-981)            face convective coefficients, cases 1,2,4,5
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
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int MX = 64, MY = 4;
static inline dim3 mgrid(int ni, int nj) { return dim3((ni + MX - 1) / MX, (nj + MY - 1) / MY); }

// ------------------------------------------------------------- ConvCoef
template <int NCOMP>
__global__ __launch_bounds__(MX *MY) void k_conv_coef(int nx, int ny, int ld, double djac,
                                                      const double *__restrict__ xzi, const double *__restrict__ xet,
                                                      const double *__restrict__ yzi, const double *__restrict__ yet,
                                                      const double *__restrict__ u, const double *__restrict__ v,
                                                      double *__restrict__ cc1, double *__restrict__ cc2) {
    const int i = blockIdx.x * MX + threadIdx.x, j = blockIdx.y * MY + threadIdx.y;   // 0-based cover
    if (i > nx + 1 || j > ny + 1) return;
    const double dHalf = 0.5, dFour = 4.0;
#define U(a, b) u[WD2_IX(a, b)]
#define V(a, b) v[WD2_IX(a, b)]
#define M(f, a, b) f[WD2_IX(a, b)]
    const size_t k = WD2_IX(i, j);
    if (NCOMP == 1) {
        if (i >= 1 && j >= 1)
            cc1[k] = (djac * M(yet, i, j) * (U(i, j) + U(i - 1, j)) - M(xet, i, j) * (V(i, j) + V(i, j - 1))) * dHalf;
        if (i >= 1 && i <= nx && j >= 1 && j <= ny)
            cc2[k] = (M(xzi, i, j) * (V(i + 1, j) + V(i, j)) - djac * M(yzi, i, j) * (U(i, j + 1) + U(i, j))) * dHalf;
    } else if (NCOMP == 2) {
        if (i >= 1 && i <= nx && j >= 1 && j <= ny)
            cc1[k] = (M(yet, i, j) * (U(i, j + 1) + U(i, j)) - djac * M(xet, i, j) * (V(i + 1, j) + V(i, j))) * dHalf;
        if (i >= 1 && j >= 1)
            cc2[k] = (djac * M(xzi, i, j) * (V(i, j) + V(i, j - 1)) - M(yzi, i, j) * (U(i, j) + U(i - 1, j))) * dHalf;
    } else if (NCOMP == 4) {
        if (i <= nx && j >= 1 && j <= ny) {
            cc1[k] = djac * M(yet, i, j) * U(i, j) -
                     M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour;
            cc2[k] = M(xzi, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / dFour -
                     djac * M(yzi, i, j) * U(i, j);
        }
    } else if (NCOMP == 5) {
        if (i >= 1 && i <= nx && j <= ny) {
            cc1[k] = M(yet, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour -
                     djac * M(xet, i, j) * V(i, j);
            cc2[k] = djac * M(xzi, i, j) * V(i, j) -
                     M(yzi, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / dFour;
        }
    }
#undef U
#undef V
#undef M
}

void launch_conv_coef(int nx, int ny, int ld, int ncomp, int njacob, const double *xzi, const double *xet,
                      const double *yzi, const double *yet, const double *u, const double *v, double *cc1,
                      double *cc2, hipStream_t s) {
    const double djac = njacob == 1 ? 2.0 : 1.0;
    const dim3 g = mgrid(nx + 2, ny + 2), b(MX, MY);
    switch (ncomp) {
    case 1: hipLaunchKernelGGL(k_conv_coef<1>, g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 2: hipLaunchKernelGGL(k_conv_coef<2>, g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 4: hipLaunchKernelGGL(k_conv_coef<4>, g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    case 5: hipLaunchKernelGGL(k_conv_coef<5>, g, b, 0, s, nx, ny, ld, djac, xzi, xet, yzi, yet, u, v, cc1, cc2); break;
    default: break;
    }
}

// --------------------------------------------------- porous-region division
__global__ __launch_bounds__(MX *MY) void k_porous_div(Regions R, int ld, int comp, MomTemps t) {
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
            const double por = R.tab(R.poros, ir, jr);
            t.cj1[k] = t.cj1[k] / por;
            t.cj2[k] = t.cj2[k] / por;
            t.c1s[k] = t.c1s[k] / por;
            t.c2s[k] = t.c2s[k] / por;
            t.c1n[k] = t.c1n[k] / por;
            t.c2n[k] = t.c2n[k] / por;
        }
}

void launch_porous_div(const Regions &R, int ld, int comp, const MomTemps &t, hipStream_t s) {
    hipLaunchKernelGGL(k_porous_div, mgrid(R.nx + 2, R.ny + 2), dim3(MX, MY), 0, s, R, ld, comp, t);
}

// -------------------------------------------------------------- PorosCoef
// Each point takes the value of the last region (in loop order) whose write
// range covers it: zero for a non-porous region, the drag formula otherwise.
__device__ inline double poros_val(int comp, int njacob, int ld, int i, int j, double porc1, double porc2,
                                   const double *__restrict__ u, const double *__restrict__ v) {
    const double dFour = 4.0;
    double unorm, unrm1 = 0.0;
    if (comp == 1) {
        const double uu = u[WD2_IX(i, j)];
        const double w = v[WD2_IX(i, j)] + v[WD2_IX(i + 1, j)] + v[WD2_IX(i, j - 1)] + v[WD2_IX(i + 1, j - 1)] / dFour;
        unorm = sqrt(uu * uu + w * w);
        if (unorm > 1.e-8) unrm1 = (uu * uu) / unorm;
    } else {
        const double vv = v[WD2_IX(i, j)];
        const double w = (u[WD2_IX(i - 1, j + 1)] + u[WD2_IX(i, j + 1)] + u[WD2_IX(i - 1, j)] + u[WD2_IX(i, j)]) / dFour;
        unorm = sqrt(w * w + vv * vv);
        if (unorm > 1.e-8) unrm1 = (vv * vv) / unorm;
    }
    if (njacob == 1) unorm = unrm1 + unorm;
    return porc1 + porc2 * unorm;
}

__global__ __launch_bounds__(MX *MY) void k_poros_coef(Regions R, int ld, int comp, const double *__restrict__ us,
                                                       const double *__restrict__ vs, const double *__restrict__ un,
                                                       const double *__restrict__ vn, MomTemps t) {
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
        t.cpj[k] = 0.0; t.cps[k] = 0.0; t.cpn[k] = 0.0;
        return;
    }
    const double c1 = R.tab(R.porc1, wir, wjr), c2 = R.tab(R.porc2, wir, wjr);
    t.cpj[k] = poros_val(comp, 1, ld, i, j, c1, c2, us, vs);
    t.cps[k] = poros_val(comp, 0, ld, i, j, c1, c2, us, vs);
    t.cpn[k] = poros_val(comp, 0, ld, i, j, c1, c2, un, vn);
}

void launch_poros_coef(const Regions &R, int ld, int comp, const double *us, const double *vs,
                       const double *un, const double *vn, const MomTemps &t, hipStream_t s) {
    hipLaunchKernelGGL(k_poros_coef, mgrid(R.nx, R.ny), dim3(MX, MY), 0, s, R, ld, comp, us, vs, un, vn, t);
}

// ---------------------------------------------------- convection/diffusion
__device__ inline double dconv_u(int ld, int i, int j, const double *c1, const double *c2, const double *u) {
#define C1(a, b) c1[WD2_IX(a, b)]
#define C2(a, b) c2[WD2_IX(a, b)]
#define U(a, b) u[WD2_IX(a, b)]
    return -C2(i, j - 1) * U(i, j - 1) - C1(i, j) * U(i - 1, j) +
           (C1(i + 1, j) - C1(i, j) + C2(i, j) - C2(i, j - 1)) * U(i, j) + C1(i + 1, j) * U(i + 1, j) +
           C2(i, j) * U(i, j + 1);
}
__device__ inline double dconv_v(int ld, int i, int j, const double *c1, const double *c2, const double *u) {
    return -C2(i, j) * U(i, j - 1) - C1(i - 1, j) * U(i - 1, j) +
           (C1(i, j) - C1(i - 1, j) + C2(i, j + 1) - C2(i, j)) * U(i, j) + C1(i, j) * U(i + 1, j) +
           C2(i, j + 1) * U(i, j + 1);
#undef C1
#undef C2
}
__device__ inline double ddiff_u(int ld, int i, int j, const double *ac, const double *bc, const double *bn,
                                 const double *gn, const double *u) {
#define F(f, a, b) f[WD2_IX(a, b)]
    const double s1 = F(ac, i + 1, j) * (U(i + 1, j) - U(i, j)) - F(ac, i, j) * (U(i, j) - U(i - 1, j)) +
                      F(bc, i + 1, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                      F(bc, i, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const double s2 = F(bn, i, j) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                      F(bn, i, j - 1) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                      F(gn, i, j) * (U(i, j + 1) - U(i, j)) - F(gn, i, j - 1) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
}
__device__ inline double ddiff_v(int ld, int i, int j, const double *an, const double *bc, const double *bn,
                                 const double *gc, const double *u) {
    const double s1 = F(an, i, j) * (U(i + 1, j) - U(i, j)) - F(an, i - 1, j) * (U(i, j) - U(i - 1, j)) +
                      F(bn, i, j) * (U(i + 1, j + 1) + U(i, j + 1) - U(i + 1, j - 1) - U(i, j - 1)) -
                      F(bn, i - 1, j) * (U(i, j + 1) + U(i - 1, j + 1) - U(i, j - 1) - U(i - 1, j - 1));
    const double s2 = F(bc, i, j + 1) * (U(i + 1, j + 1) + U(i + 1, j) - U(i - 1, j + 1) - U(i - 1, j)) -
                      F(bc, i, j) * (U(i + 1, j) + U(i + 1, j - 1) - U(i - 1, j) - U(i - 1, j - 1)) +
                      F(gc, i, j + 1) * (U(i, j + 1) - U(i, j)) - F(gc, i, j) * (U(i, j) - U(i, j - 1));
    return s1 + s2;
#undef F
#undef U
}

__global__ __launch_bounds__(MX *MY) void k_conv_diff(int nx, int ny, int ld, int comp, Metrics m,
                                                      const double *__restrict__ fs, const double *__restrict__ fn,
                                                      MomTemps t) {
    if (comp == 1) {
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        t.cnvs[k] = dconv_u(ld, i, j, t.c1s, t.c2s, fs);
        t.cnvn[k] = dconv_u(ld, i, j, t.c1n, t.c2n, fn);
        t.difs[k] = ddiff_u(ld, i, j, m.rac, m.rbc, m.rbn, m.rgn, fs);
        t.difn[k] = ddiff_u(ld, i, j, m.rac, m.rbc, m.rbn, m.rgn, fn);
    } else {
        const int i = 2 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        const size_t k = WD2_IX(i, j);
        t.cnvs[k] = dconv_v(ld, i, j, t.c1s, t.c2s, fs);
        t.cnvn[k] = dconv_v(ld, i, j, t.c1n, t.c2n, fn);
        t.difs[k] = ddiff_v(ld, i, j, m.ran, m.rbc, m.rbn, m.rgc, fs);
        t.difn[k] = ddiff_v(ld, i, j, m.ran, m.rbc, m.rbn, m.rgc, fn);
    }
}

void launch_conv_diff(int nx, int ny, int ld, int comp, const Metrics &m, const double *fs, const double *fn,
                      const MomTemps &t, hipStream_t s) {
    const dim3 g = comp == 1 ? mgrid(nx, ny - 1) : mgrid(nx - 1, ny);
    hipLaunchKernelGGL(k_conv_diff, g, dim3(MX, MY), 0, s, nx, ny, ld, comp, m, fs, fn, t);
}

// ------------------------------------------------------------ assembly
// comp 1 (x-momentum): rows (i=1..nx, j=2..ny), k = (j-2)*nx + i-1
// comp 2 (y-momentum): rows (i=2..nx, j=1..ny), k = (j-1)*(nx-1) + i-2
// split 1: first split step with its right-hand side; split 2: second split
// step (y-direction coefficients on the same x-ordered chain), right-hand
// side = first-split solution, wall and blockage identity rows.
__global__ __launch_bounds__(MX *MY) void k_assemble(Regions R, int ld, int comp, int split, double dk, double dk2,
                                                     double re1, double fr, Metrics m, MomTemps t,
                                                     const double *__restrict__ fs, const double *__restrict__ fn,
                                                     const double *__restrict__ d, const double *__restrict__ dn,
                                                     const double *__restrict__ xprev, Chain ch, int *err) {
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
    const double dZero = 0.0, dOne = 1.0, dFour = 4.0;
    const size_t p = WD2_IX(i, j);
#define T(f, a, b) t.f[WD2_IX(a, b)]
#define MM(f, a, b) m.f[WD2_IX(a, b)]
    double a1, a2, a3;
    if (comp == 1) {
        const double rkj = dk2 * m.dju[p];
        if (split == 1) {
            if (t.cj1[p] >= dZero) {
                a1 = rkj * (-T(cj1, i - 1, j) - re1 * MM(rac, i, j));
                a2 = dOne + rkj * (T(cj1, i, j) + re1 * (MM(rac, i + 1, j) + MM(rac, i, j))) + dk2 * t.cpj[p];
                a3 = rkj * (-re1 * MM(rac, i + 1, j));
            } else {
                a1 = rkj * (-re1 * MM(rac, i, j));
                a2 = dOne + rkj * (-T(cj1, i, j) + re1 * (MM(rac, i + 1, j) + MM(rac, i, j))) + dk2 * t.cpj[p];
                a3 = rkj * (T(cj1, i + 1, j) - re1 * MM(rac, i + 1, j));
            }
            ch.b[k] = fn[p] - fs[p] + rkj * (-t.cnvs[p] - t.cnvn[p]) + rkj * re1 * (t.difs[p] + t.difn[p]) -
                      dk2 * (t.cps[p] * fs[p] + t.cpn[p] * fn[p]);
        } else {
            if (t.cj2[p] >= dZero) {
                a1 = rkj * (-T(cj2, i, j - 1) - re1 * MM(rgn, i, j - 1));
                a2 = dOne + rkj * (T(cj2, i, j) + re1 * (MM(rgn, i, j) + MM(rgn, i, j - 1))) + dk2 * t.cpj[p];
                a3 = rkj * (-re1 * MM(rgn, i, j));
            } else {
                a1 = rkj * (-re1 * MM(rgn, i, j - 1));
                a2 = dOne + rkj * (-T(cj2, i, j) + re1 * (MM(rgn, i, j) + MM(rgn, i, j - 1))) + dk2 * t.cpj[p];
                a3 = rkj * (T(cj2, i, j + 1) - re1 * MM(rgn, i, j));
            }
            double bb = xprev[k];
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
        const double rkj = dk2 * m.djv[p];
        if (split == 1) {
            if (t.cj1[p] >= dZero) {
                a1 = rkj * (-T(cj1, i - 1, j) - re1 * MM(ran, i - 1, j));
                a2 = rkj * (T(cj1, i, j) + re1 * (MM(ran, i, j) + MM(ran, i - 1, j))) + dk2 * t.cpj[p] + dOne;
                a3 = rkj * (-re1 * MM(ran, i, j));
            } else {
                a1 = rkj * (-re1 * MM(ran, i - 1, j));
                a2 = rkj * (-T(cj1, i, j) + re1 * (MM(ran, i, j) + MM(ran, i - 1, j))) + dk2 * t.cpj[p] + dOne;
                a3 = rkj * (T(cj1, i + 1, j) - re1 * MM(ran, i, j));
            }
            const double buoy = dk * (d[WD2_IX(i, j + 1)] + d[p] + dn[WD2_IX(i, j + 1)] + dn[p]) / (dFour * fr);
            ch.b[k] = fn[p] - fs[p] + rkj * (-t.cnvs[p] - t.cnvn[p]) + rkj * re1 * (t.difs[p] + t.difn[p]) -
                      dk2 * (t.cps[p] * fs[p] + t.cpn[p] * fn[p]) - buoy;
        } else {
            if (t.cj2[p] >= dZero) {
                a1 = rkj * (-T(cj2, i, j - 1) - re1 * MM(rgc, i, j));
                a2 = rkj * (T(cj2, i, j) + re1 * (MM(rgc, i, j + 1) + MM(rgc, i, j))) + dk2 * t.cpj[p] + dOne;
                a3 = rkj * (-re1 * MM(rgc, i, j + 1));
            } else {
                a1 = rkj * (-re1 * MM(rgc, i, j));
                a2 = rkj * (-T(cj2, i, j) + re1 * (MM(rgc, i, j + 1) + MM(rgc, i, j))) + dk2 * t.cpj[p] + dOne;
                a3 = rkj * (T(cj2, i, j + 1) - re1 * MM(rgc, i, j + 1));
            }
            double bb = xprev[k];
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
#undef T
#undef MM
    ch.a[k] = a1;
    ch.d[k] = a2;
    ch.c[k] = a3;
}

void launch_assemble(const Regions &R, int ld, int comp, int split, double dk, double re, double fr,
                     const Metrics &m, const MomTemps &t, const double *fs, const double *fn, const double *d,
                     const double *dn, const double *xprev, const Chain &ch, hipStream_t s, int *err) {
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    const dim3 g = comp == 1 ? mgrid(R.nx, R.ny - 1) : mgrid(R.nx - 1, R.ny);
    hipLaunchKernelGGL(k_assemble, g, dim3(MX, MY), 0, s, R, ld, comp, split, dk, dk2, re1, fr, m, t, fs, fn, d, dn,
                       xprev, ch, err);
}

// ------------------------------------------------------------- scatter
__global__ __launch_bounds__(MX *MY) void k_scatter(int nx, int ny, int ld, int comp, const double *__restrict__ x,
                                                    double *__restrict__ out) {
    if (comp == 1) {
        const int i = 1 + blockIdx.x * MX + threadIdx.x, j = 2 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        out[WD2_IX(i, j)] = x[(long)(j - 2) * nx + (i - 1)];
    } else {
        const int i = 2 + blockIdx.x * MX + threadIdx.x, j = 1 + blockIdx.y * MY + threadIdx.y;
        if (i > nx || j > ny) return;
        out[WD2_IX(i, j)] = x[(long)(j - 1) * (nx - 1) + (i - 2)];
    }
}

void launch_scatter(int nx, int ny, int ld, int comp, const double *x, double *out, hipStream_t s) {
    const dim3 g = comp == 1 ? mgrid(nx, ny - 1) : mgrid(nx - 1, ny);
    hipLaunchKernelGGL(k_scatter, g, dim3(MX, MY), 0, s, nx, ny, ld, comp, x, out);
}

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

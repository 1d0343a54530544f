// k_thermal.hip -- the thermal energy equation and its helpers:
//   TempBoundCond (bound_cond.f:1030-1190)  temperature ghost and wall fills
//   ThermEnergy   (thermal.f:24-274)        two ADI splits, upwinded rows
//   EqState       (thermal.f:280-328)       ideal-gas density
//   Filter        (utility.f:31-247)        Shuman filter of u, v or t
//
// FP64, in the reference's operation order (the library is compiled with
// -ffp-contract=off), so every result equals the reference's bit for bit.
//
// ThermEnergy hands AltTridLU one chain of (nx-1)(ny-1) rows ordered i
// fastest, ind = (j-2)(nx-1) + i-1, for both splits: the second split's rows
// carry the j-direction coefficients in that i-fastest order
// (thermal.f:210-245).  The chain goes to the parallel AltTridLU of
// k_trid.hip; the assembly kernels write its rows and right-hand side in one
// pass over the grid.  Region loops are resolved per point in the reference's
// region order (heat source, fixed-temperature rows).
//
// Roofline (HBM): the assembly of the first split reads tn (9-point), the
// five metric arrays and four face-coefficient arrays and writes four chain
// arrays, about 18 values per point; the whole equation moves about 60 FP64
// values per point, a few percent of a momentum step.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int TX = 64, TY = 4;
static inline dim3 tgrid(int ni, int nj) { return dim3((ni + TX - 1) / TX, (nj + TY - 1) / TY); }
constexpr int TBC_THREADS = 1024;

#define PAR(var, lo, hi) for (int var = (lo) + (int)threadIdx.x; var <= (hi); var += TBC_THREADS)

// ---------------------------------------------------------- TempBoundCond
// One workgroup walks the regions in the reference's order; each line loop is
// spread over the workgroup and a barrier separates consecutive loops (a
// later side may read a corner an earlier one wrote).
__global__ __launch_bounds__(TBC_THREADS) void k_temp_bc(Regions R, Thermal TH, int ld, double *__restrict__ t,
                                                          int *__restrict__ err) {
    const double dTwo = 2.0;
#define T_(i, j) t[WD2_IX(i, j)]
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            if (TH.rtype(ir, jr) == RT_TEMPER) {
                const double tv = TH.tval(ir, jr);
                const int w = iE - iW, h = jN - jS;
                for (int k = threadIdx.x; k < w * h; k += TBC_THREADS) T_(iW + 1 + k % w, jS + 1 + k / w) = tv;
                __syncthreads();
                PAR(j, jS + 1, jN) T_(iW + 1, j) = dTwo * R.bc(ir, jr, WEST, VT) - T_(iW, j);
                __syncthreads();
                PAR(j, jS + 1, jN) T_(iE, j) = dTwo * R.bc(ir, jr, EAST, VT) - T_(iE + 1, j);
                __syncthreads();
                PAR(i, iW + 1, iE) T_(i, jS + 1) = dTwo * R.bc(ir, jr, SOUTH, VT) - T_(i, jS);
                __syncthreads();
                PAR(i, iW + 1, iE) T_(i, jN) = dTwo * R.bc(ir, jr, NORTH, VT) - T_(i, jN + 1);
                __syncthreads();
                continue;
            }
            for (int side = WEST; side <= NORTH; side++) {
                const int bt = TH.btype(ir, jr, side);
                const double b = R.bc(ir, jr, side, VT);
                if (bt == BT_TEMPER || bt == BT_HTFLUX) {
                    const bool fix = bt == BT_TEMPER;
                    // ghost g, first interior point n of each line of the side
                    if (side == WEST) {
                        PAR(j, jS + 1, jN) T_(iW, j) = fix ? dTwo * b - T_(iW + 1, j) : b + T_(iW + 1, j);
                    } else if (side == EAST) {
                        PAR(j, jS + 1, jN) T_(iE + 1, j) = fix ? dTwo * b - T_(iE, j) : b + T_(iE, j);
                    } else if (side == SOUTH) {
                        PAR(i, iW + 1, iE) T_(i, jS) = fix ? dTwo * b - T_(i, jS + 1) : b + T_(i, jS + 1);
                    } else {
                        PAR(i, iW + 1, iE) T_(i, jN + 1) = fix ? dTwo * b - T_(i, jN) : b + T_(i, jN);
                    }
                } else if (bt != BT_INTERN) {
                    if (threadIdx.x == 0) *err = 2000 + side;   // the reference STOPs
                }
                __syncthreads();
            }
        }
#undef T_
}

void launch_temp_bc(const Regions &R, const Thermal &TH, int ld, double *t, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_temp_bc, dim3(1), dim3(TBC_THREADS), 0, s, R, TH, ld, t, err);
}

// ------------------------------------------------------------ ThermEnergy
// heat source s(i,j) of thermal.f:120-147: the value of the last region whose
// interior holds the point
__device__ inline double heat_source(const Regions &R, const Thermal &TH, int i, int j) {
    double sv = 0.0;
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            if (i < R.border(ir, jr, WEST) + 1 || i > R.border(ir, jr, EAST)) continue;
            if (j < R.border(ir, jr, SOUTH) + 1 || j > R.border(ir, jr, NORTH)) continue;
            sv = TH.rtype(ir, jr) == RT_HEATGN ? TH.hg(ir, jr) : 0.0;
        }
    return sv;
}

// is (i,j) in the interior of a fixed-temperature region (thermal.f:223-245)
__device__ inline bool fixed_temp_point(const Regions &R, const Thermal &TH, int i, int j) {
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            if (TH.rtype(ir, jr) != RT_TEMPER) continue;
            if (i >= R.border(ir, jr, WEST) + 1 && i <= R.border(ir, jr, EAST) && j >= R.border(ir, jr, SOUTH) + 1 &&
                j <= R.border(ir, jr, NORTH))
                return true;
        }
    return false;
}

// First split (thermal.f:153-200): x-direction upwinded rows and the full
// explicit right-hand side.  cu1, cv1: ConvCoef case 6 of u, v; cun, cvn:
// case 3 of un, vn.
__global__ __launch_bounds__(TX *TY) void k_therm_x(Regions R, Thermal TH, int ld, double dk, double pe1,
                                                    Metrics m, const double *__restrict__ cu1,
                                                    const double *__restrict__ cun, const double *__restrict__ cvn,
                                                    const double *__restrict__ tn, double *__restrict__ A,
                                                    double *__restrict__ D, double *__restrict__ C,
                                                    double *__restrict__ B) {
    const int nx = R.nx, ny = R.ny;
    const int i = 2 + blockIdx.x * TX + threadIdx.x, j = 2 + blockIdx.y * TY + threadIdx.y;
    if (i > nx || j > ny) return;
    const long ind = (long)(j - 2) * (nx - 1) + (i - 2);
    const double dZero = 0.0, dOne = 1.0, dTwo = 2.0, dHalf = 0.5;
#define F(a, p, q) a[WD2_IX(p, q)]
#define G(f, p, q) MGET(m, f, p, q)   // a metric, through its constant box
    const double rkj = dk * G(djc, i, j) * dHalf;
    double a1, a2, a3;
    if (F(cu1, i, j) >= dZero) {
        a1 = rkj * (-F(cu1, i - 1, j) - pe1 * G(rau, i - 1, j));
        a2 = dOne + rkj * (F(cu1, i, j) + pe1 * (G(rau, i, j) + G(rau, i - 1, j)));
        a3 = rkj * (-pe1 * G(rau, i, j));
    } else {
        a1 = rkj * (-pe1 * G(rau, i - 1, j));
        a2 = dOne + rkj * (-F(cu1, i, j) + pe1 * (G(rau, i, j) + G(rau, i - 1, j)));
        a3 = rkj * (F(cu1, i + 1, j) - pe1 * G(rau, i, j));
    }
    const double c = -F(cvn, i, j - 1) * F(tn, i, j - 1) - F(cun, i - 1, j) * F(tn, i - 1, j) +
                     (F(cun, i, j) - F(cun, i - 1, j) + F(cvn, i, j) - F(cvn, i, j - 1)) * F(tn, i, j) +
                     F(cun, i, j) * F(tn, i + 1, j) + F(cvn, i, j) * F(tn, i, j + 1);
    const double d =
        G(rau, i, j) * (F(tn, i + 1, j) - F(tn, i, j)) - G(rau, i - 1, j) * (F(tn, i, j) - F(tn, i - 1, j)) +
        G(rbu, i, j) * (F(tn, i + 1, j + 1) + F(tn, i, j + 1) - F(tn, i + 1, j - 1) - F(tn, i, j - 1)) -
        G(rbu, i - 1, j) * (F(tn, i, j + 1) + F(tn, i - 1, j + 1) - F(tn, i, j - 1) - F(tn, i - 1, j - 1)) +
        G(rbv, i, j) * (F(tn, i + 1, j + 1) + F(tn, i + 1, j) - F(tn, i - 1, j + 1) - F(tn, i - 1, j)) -
        G(rbv, i, j - 1) * (F(tn, i + 1, j) + F(tn, i + 1, j - 1) - F(tn, i - 1, j) - F(tn, i - 1, j - 1)) +
        G(rgv, i, j) * (F(tn, i, j + 1) - F(tn, i, j)) - G(rgv, i, j - 1) * (F(tn, i, j) - F(tn, i, j - 1));
    A[ind] = a1;
    D[ind] = a2;
    C[ind] = a3;
    B[ind] = dk * heat_source(R, TH, i, j) * dHalf + dTwo * rkj * (-c + pe1 * d);
#undef F
#undef G
}

// Second split (thermal.f:210-252): y-direction upwinded rows in the same
// i-fastest order, right-hand side = the first split's solution x1; identity
// rows with zero right-hand side inside fixed-temperature regions.
__global__ __launch_bounds__(TX *TY) void k_therm_y(Regions R, Thermal TH, int ld, double dk, double pe1,
                                                    Metrics m, const double *__restrict__ cv1,
                                                    const double *__restrict__ x1, double *__restrict__ A,
                                                    double *__restrict__ D, double *__restrict__ C,
                                                    double *__restrict__ B) {
    const int nx = R.nx, ny = R.ny;
    const int i = 2 + blockIdx.x * TX + threadIdx.x, j = 2 + blockIdx.y * TY + threadIdx.y;
    if (i > nx || j > ny) return;
    const long ind = (long)(j - 2) * (nx - 1) + (i - 2);
    const double dZero = 0.0, dOne = 1.0, dHalf = 0.5;
#define F(a, p, q) a[WD2_IX(p, q)]
#define G(f, p, q) MGET(m, f, p, q)   // a metric, through its constant box
    if (fixed_temp_point(R, TH, i, j)) {
        A[ind] = dZero;
        D[ind] = dOne;
        C[ind] = dZero;
        B[ind] = dZero;
        return;
    }
    const double rkj = dk * G(djc, i, j) * dHalf;
    double a1, a2, a3;
    if (F(cv1, i, j) >= dZero) {
        a1 = rkj * (-F(cv1, i, j - 1) - pe1 * G(rgv, i, j - 1));
        a2 = dOne + rkj * (F(cv1, i, j) + pe1 * (G(rgv, i, j) + G(rgv, i, j - 1)));
        a3 = rkj * (-pe1 * G(rgv, i, j));
    } else {
        a1 = rkj * (-pe1 * G(rgv, i, j - 1));
        a2 = dOne + rkj * (-F(cv1, i, j) + pe1 * (G(rgv, i, j) + G(rgv, i, j - 1)));
        a3 = rkj * (F(cv1, i, j + 1) - pe1 * G(rgv, i, j));
    }
    A[ind] = a1;
    D[ind] = a2;
    C[ind] = a3;
    B[ind] = x1[ind];
#undef F
#undef G
}

// t(i,j) += x2(ind) (thermal.f:262-268)
__global__ __launch_bounds__(TX *TY) void k_therm_update(int nx, int ny, int ld, const double *__restrict__ x,
                                                         double *__restrict__ t) {
    const int i = 2 + blockIdx.x * TX + threadIdx.x, j = 2 + blockIdx.y * TY + threadIdx.y;
    if (i > nx || j > ny) return;
    t[WD2_IX(i, j)] = t[WD2_IX(i, j)] + x[(long)(j - 2) * (nx - 1) + (i - 2)];
}

void launch_therm_energy(const Regions &R, const Thermal &TH, int ld, double dk, double pe, const Metrics &m,
                         const double *un, const double *vn, const double *u, const double *v, const double *tn,
                         double *t, const ThermTemps &w, const ChainT<double> &ch, const TridWork &tw, int *err,
                         hipStream_t s) {
    const int nx = R.nx, ny = R.ny;
    const double pe1 = 1.0 / pe;
    launch_temp_bc(R, TH, ld, t, err, s);
    launch_conv_coef<double>(nx, ny, ld, 6, 0, MARR(m, xzc), MARR(m, xec), MARR(m, yzc), MARR(m, yec), u, v, w.cu1,
                             w.cv1, s);
    launch_conv_coef<double>(nx, ny, ld, 3, 0, MARR(m, xzv), MARR(m, xeu), MARR(m, yzv), MARR(m, yeu), un, vn, w.cun,
                             w.cvn, s);
    const long n = (long)(nx - 1) * (ny - 1);
    const dim3 g = tgrid(nx - 1, ny - 1), b(TX, TY);
    hipLaunchKernelGGL(k_therm_x, g, b, 0, s, R, TH, ld, dk, pe1, m, w.cu1, w.cun, w.cvn, tn, ch.a, ch.d, ch.c,
                       ch.b);
    launch_trid<double>(tw, n, n, ch.a, ch.d, ch.c, ch.b, ch.x, s);
    hipLaunchKernelGGL(k_therm_y, g, b, 0, s, R, TH, ld, dk, pe1, m, w.cv1, ch.x, ch.a, ch.d, ch.c, ch.b);
    launch_trid<double>(tw, n, n, ch.a, ch.d, ch.c, ch.b, ch.x, s);
    hipLaunchKernelGGL(k_therm_update, g, b, 0, s, nx, ny, ld, ch.x, t);
}

// ---------------------------------------------------------------- EqState
__global__ __launch_bounds__(TX *TY) void k_eq_state(int nx, int ny, int ld, double uref, double densref,
                                                     double tmax, double tref, double rconst,
                                                     const double *__restrict__ p, const double *__restrict__ t,
                                                     double *__restrict__ den) {
    const int i = 2 + blockIdx.x * TX + threadIdx.x, j = 2 + blockIdx.y * TY + threadIdx.y;
    if (i > nx || j > ny) return;
    const double dZero = 0.0, dOne = 1.0;
    const double pref = densref * rconst * tref;
    const size_t k = WD2_IX(i, j);
    const double c1 = p[k] * densref * (uref * uref) + pref;
    const double c2 = densref * rconst * (t[k] * (tmax - tref) + tref);
    double dv = c1 / c2 - dOne;
    if (fabs(dv) < 1.e-10) dv = dZero;
    den[k] = dv;
}

void launch_eq_state(int nx, int ny, int ld, double uref, double densref, double tmax, double tref, double rconst,
                     const double *p, const double *t, double *den, hipStream_t s) {
    hipLaunchKernelGGL(k_eq_state, tgrid(nx - 1, ny - 1), dim3(TX, TY), 0, s, nx, ny, ld, uref, densref, tmax,
                       tref, rconst, p, t, den);
}

// ----------------------------------------------------------------- Filter
// qh = qu, then the filtered value at every point some region's loops of
// utility.f:31-247 write (all of them read qu only, so the point's value does
// not depend on which region writes it); the caller copies qh back.
template <int COMP>
__global__ __launch_bounds__(TX *TY) void k_filter(Regions R, Thermal TH, int ld, double fp,
                                                   const double *__restrict__ qu, double *__restrict__ qh) {
    const int nx = R.nx, ny = R.ny;
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y;
    if (i > nx + 1 || j > ny + 1) return;
    bool hit = false;
    for (int jr = 1; jr <= R.nj && !hit; jr++)
        for (int ir = 1; ir <= R.ni && !hit; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            if (COMP == VU) {
                if (R.type(ir, jr) == RM_BLOCKG) continue;
                if (j < jS + 1 || j > jN) continue;
                if (i >= iW + 1 && i <= iE - 1) hit = true;
                if (i == iW && R.btype(ir, jr, WEST) == BM_OUTLT1) hit = true;
                const int te = R.btype(ir, jr, EAST);
                if (i == iE && (te == BM_INTERN || te == BM_OUTLT1)) hit = true;
            } else if (COMP == VV) {
                if (R.type(ir, jr) == RM_BLOCKG) continue;
                if (i < iW + 1 || i > iE) continue;
                if (j >= jS + 1 && j <= jN - 1) hit = true;
                if (j == jS && R.btype(ir, jr, SOUTH) == BM_OUTLT1) hit = true;
                const int tn = R.btype(ir, jr, NORTH);
                if (j == jN && (tn == BM_INTERN || tn == BM_OUTLT1)) hit = true;
            } else {
                // utility.f:209 compares the thermal region type with BT_TEMPER
                // (= RT_HEATGN): heat-generating regions are the ones skipped
                if (TH.rtype(ir, jr) == BT_TEMPER) continue;
                if (i >= iW + 1 && i <= iE && j >= jS + 1 && j <= jN) hit = true;
            }
        }
    const double dFour = 4.0;
#define Q(p, q) qu[WD2_IX(p, q)]
    qh[WD2_IX(i, j)] =
        hit ? (Q(i, j - 1) + Q(i - 1, j) + Q(i, j + 1) + Q(i + 1, j) + fp * Q(i, j)) / (fp + dFour) : Q(i, j);
#undef Q
}

void launch_filter(const Regions &R, const Thermal &TH, int ld, int ncomp, double fp, double *qu, double *qh,
                   hipStream_t s) {
    const dim3 g = tgrid(R.nx + 2, R.ny + 2), b(TX, TY);
    if (ncomp == VU)
        hipLaunchKernelGGL(k_filter<VU>, g, b, 0, s, R, TH, ld, fp, qu, qh);
    else if (ncomp == VV)
        hipLaunchKernelGGL(k_filter<VV>, g, b, 0, s, R, TH, ld, fp, qu, qh);
    else
        hipLaunchKernelGGL(k_filter<VT>, g, b, 0, s, R, TH, ld, fp, qu, qh);
    launch_copy(ld, 0, R.nx + 1, 0, R.ny + 1, qh, qu, s);
}

}  // namespace wd2

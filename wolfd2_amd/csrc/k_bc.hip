// k_bc.hip -- ghost-cell and wall fills (bound_cond.f:VelBoundCond 511-847,
// PresBoundCond 853-1024, VelOutflowBCs 1656-1874).
//
// These are O(nx+ny) line updates whose order matters: a later side reads
// values an earlier side wrote (corners), and the mass-conserving outlet
// formulas are first-order recurrences along the boundary.  One workgroup
// walks the regions and sides in the reference's order; each loop whose
// iterations are independent is spread over the workgroup, the recurrences
// run on one lane, and a barrier separates consecutive loops.
//
// Boundary values: the reference's dBCVal holds one constant per region side
// and variable.  A side may instead carry a profile, one value per grid line
// (Regions::bcp); every formula is the reference's with the constant replaced
// by the line's value, so a profile of constant values gives the same bits.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int BC_THREADS = 1024;

#define PAR(var, lo, hi) for (int var = (lo) + (int)threadIdx.x; var <= (hi); var += BC_THREADS)
#define SER(var, lo, hi) if (threadIdx.x == 0) for (int var = (lo); var <= (hi); var++)
#define SYNC() __syncthreads()

// VelBoundCond (outflow_only = 0) and VelOutflowBCs (outflow_only = 1).
template <typename T>
__global__ __launch_bounds__(BC_THREADS) void k_vel_bc(Regions R, int ld, T *__restrict__ u,
                                                       T *__restrict__ v, int outflow_only,
                                                       int *__restrict__ err, const int *stop) {
    if (stop && *stop) return;
    const T dTwo = 2, dThree = 3, dFour = 4, dFive = 5, dEight = 8;
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
#define U(i, j) u[WD2_IX(i, j)]
#define V(i, j) v[WD2_IX(i, j)]
            // ---- WEST
            {
                const int t = R.btype(ir, jr, WEST);
                const int ou = R.prof(ir, jr, WEST, VU), ov = R.prof(ir, jr, WEST, VV);
                auto bu = [&](int k) { return (T)R.bcp(ou, ir, jr, WEST, VU, k); };
                auto bv = [&](int k) { return (T)R.bcp(ov, ir, jr, WEST, VV, k); };
                if (!outflow_only && t == BM_WALL1) {
                    PAR(j, jS, jN) U(iW, j) = 0;
                    PAR(j, jS + 1, jN) V(iW, j) = dTwo * bv(j) - V(iW + 1, j);
                } else if (!outflow_only && t == BM_WALL2) {
                    PAR(j, jS, jN) U(iW, j) = 0;
                    PAR(j, jS + 1, jN) V(iW, j) = V(iW + 1, j);
                } else if (!outflow_only && t == BM_INLET) {
                    PAR(j, jS, jN) U(iW, j) = bu(j);
                    PAR(j, jS + 1, jN) V(iW, j) = dTwo * bv(j) - V(iW + 1, j);
                } else if (t == BM_OUTLT1) {
                    PAR(j, jS, jN) U(iW - 1, j) = bu(j) + U(iW, j);
                    PAR(j, jS + 1, jN) V(iW, j) = -V(iW + 1, j);
                } else if (t == BM_OUTLT2) {
                    if (!outflow_only) {
                        PAR(j, jS + 1, jN) U(iW, j) = U(iW + 1, j) + V(iW + 1, j) - V(iW + 1, j - 1);
                    } else {  // bound_cond.f:1727 has the opposite sign on the v terms
                        PAR(j, jS + 1, jN) U(iW, j) = U(iW + 1, j) - V(iW + 1, j) + V(iW + 1, j - 1);
                    }
                    SYNC();
                    SER(j, jS + 1, jN)
                    V(iW, j) = -V(iW, j - 1) + dFive * (V(iW + 1, j) - V(iW + 1, j - 1)) +
                               dEight * (U(iW + 1, j) - U(iW, j));
                } else if (t < BM_INTERN || t > BM_OUTLT2) {
                    if (threadIdx.x == 0) *err = 1000 + WEST;
                }
                SYNC();
            }
            // ---- EAST
            {
                const int t = R.btype(ir, jr, EAST);
                const int ou = R.prof(ir, jr, EAST, VU), ov = R.prof(ir, jr, EAST, VV);
                auto bu = [&](int k) { return (T)R.bcp(ou, ir, jr, EAST, VU, k); };
                auto bv = [&](int k) { return (T)R.bcp(ov, ir, jr, EAST, VV, k); };
                if (!outflow_only && t == BM_WALL1) {
                    PAR(j, jS, jN) U(iE, j) = 0;
                    PAR(j, jS + 1, jN) V(iE + 1, j) = dTwo * bv(j) - V(iE, j);
                } else if (!outflow_only && t == BM_WALL2) {
                    PAR(j, jS, jN) U(iE, j) = 0;
                    PAR(j, jS + 1, jN) V(iE + 1, j) = V(iE, j);
                } else if (!outflow_only && t == BM_INLET) {
                    PAR(j, jS, jN) U(iE, j) = bu(j);
                    PAR(j, jS + 1, jN) V(iE + 1, j) = dTwo * bv(j) - V(iE, j);
                } else if (t == BM_OUTLT1) {
                    PAR(j, jS, jN) U(iE + 1, j) = bu(j) + U(iE, j);
                    PAR(j, jS + 1, jN) V(iE + 1, j) = -V(iE, j);
                } else if (t == BM_OUTLT2) {
                    PAR(j, jS + 1, jN) U(iE, j) = U(iE - 1, j) - (V(iE, j) - V(iE, j - 1));
                    SYNC();
                    SER(j, jS + 1, jN - 1)
                    V(iE + 1, j) = V(iE + 1, j - 1) + dThree * (V(iE, j - 1) - V(iE, j)) -
                                   dFour * (U(iE, j) - U(iE - 1, j));
                } else if (t < BM_INTERN || t > BM_OUTLT2) {
                    if (threadIdx.x == 0) *err = 1000 + EAST;
                }
                SYNC();
            }
            // ---- SOUTH
            {
                const int t = R.btype(ir, jr, SOUTH);
                const int ou = R.prof(ir, jr, SOUTH, VU), ov = R.prof(ir, jr, SOUTH, VV);
                auto bu = [&](int k) { return (T)R.bcp(ou, ir, jr, SOUTH, VU, k); };
                auto bv = [&](int k) { return (T)R.bcp(ov, ir, jr, SOUTH, VV, k); };
                if (!outflow_only && t == BM_WALL1) {
                    PAR(i, iW + 1, iE) U(i, jS) = dTwo * bu(i) - U(i, jS + 1);
                    PAR(i, iW, iE) V(i, jS) = 0;
                } else if (!outflow_only && t == BM_WALL2) {
                    PAR(i, iW + 1, iE) U(i, jS) = U(i, jS + 1);
                    PAR(i, iW, iE) V(i, jS) = 0;
                } else if (!outflow_only && t == BM_INLET) {
                    PAR(i, iW + 1, iE) U(i, jS) = dTwo * bu(i) - U(i, jS + 1);
                    PAR(i, iW, iE) V(i, jS) = bv(i);
                } else if (t == BM_OUTLT1) {
                    PAR(i, iW + 1, iE) U(i, jS) = -U(i, jS + 1);
                    PAR(i, iW, iE) V(i, jS) = bv(i) + V(i, jS);
                } else if (t == BM_OUTLT2) {
                    PAR(i, iW + 1, iE) V(i, jS) = V(i, jS + 1) + (U(i, jS + 1) - U(i - 1, jS + 1));
                    SYNC();
                    SER(i, iW + 1, iE - 1)
                    U(i, jS) = U(i - 1, jS) + dThree * (U(i - 1, jS + 1) - U(i, jS + 1)) -
                               dFour * (V(i, jS + 1) - V(i, jS));
                } else if (t < BM_INTERN || t > BM_OUTLT2) {
                    if (threadIdx.x == 0) *err = 1000 + SOUTH;
                }
                SYNC();
            }
            // ---- NORTH
            {
                const int t = R.btype(ir, jr, NORTH);
                const int ou = R.prof(ir, jr, NORTH, VU), ov = R.prof(ir, jr, NORTH, VV);
                auto bu = [&](int k) { return (T)R.bcp(ou, ir, jr, NORTH, VU, k); };
                auto bv = [&](int k) { return (T)R.bcp(ov, ir, jr, NORTH, VV, k); };
                if (!outflow_only && t == BM_WALL1) {
                    PAR(i, iW + 1, iE) U(i, jN + 1) = dTwo * bu(i) - U(i, jN);
                    PAR(i, iW, iE) V(i, jN) = 0;
                } else if (!outflow_only && t == BM_WALL2) {
                    PAR(i, iW + 1, iE) U(i, jN + 1) = U(i, jN);
                    PAR(i, iW, iE) V(i, jN) = 0;
                } else if (!outflow_only && t == BM_INLET) {
                    PAR(i, iW + 1, iE) U(i, jN + 1) = dTwo * bu(i) - U(i, jN);
                    PAR(i, iW, iE) V(i, jN) = bv(i);
                } else if (t == BM_OUTLT1) {
                    PAR(i, iW + 1, iE) U(i, jN + 1) = -U(i, jN);
                    PAR(i, iW, iE) V(i, jN + 1) = bv(i) + V(i, jN);
                } else if (t == BM_OUTLT2) {
                    PAR(i, iW, iE) V(i, jN) = V(i, jN - 1) - (U(i, jN) - U(i - 1, jN));
                    SYNC();
                    SER(i, iW + 1, iE - 1)
                    U(i, jN + 1) = U(i - 1, jN + 1) + dThree * (U(i - 1, jN) - U(i, jN)) -
                                   dFour * (V(i, jN) - V(i, jN - 1));
                } else if (t < BM_INTERN || t > BM_OUTLT2) {
                    if (threadIdx.x == 0) *err = 1000 + NORTH;
                }
                SYNC();
            }
#undef U
#undef V
        }
}

// PresBoundCond.
template <typename T>
__global__ __launch_bounds__(BC_THREADS) void k_pres_bc(Regions R, int ld, T *__restrict__ p,
                                                        int *__restrict__ err) {
    const T dTwo = 2;
#define P(i, j) p[WD2_IX(i, j)]
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            // boundary value of the pressure on line k of a side
            auto bp = [&](int side, int k) { return (T)R.bcp(R.prof(ir, jr, side, VP), ir, jr, side, VP, k); };
            if (R.type(ir, jr) == RM_BLOCKG) {
                const int w = iE - iW, h = jN - jS;
                for (int k = threadIdx.x; k < w * h; k += BC_THREADS) P(iW + 1 + k % w, jS + 1 + k / w) = 0;
                SYNC();
                PAR(j, jS + 1, jN) P(iW + 1, j) = bp(WEST, j) + P(iW, j);
                SYNC();
                PAR(j, jS + 1, jN) P(iE, j) = bp(EAST, j) + P(iE + 1, j);
                SYNC();
                PAR(i, iW + 1, iE) P(i, jS + 1) = bp(SOUTH, i) + P(i, jS);
                SYNC();
                PAR(i, iW + 1, iE) P(i, jN) = bp(NORTH, i) + P(i, jN + 1);
                SYNC();
                continue;
            }
            int t = R.btype(ir, jr, WEST);
            if (t == BM_WALL1 || t == BM_WALL2 || t == BM_INLET) {
                PAR(j, jS + 1, jN) P(iW, j) = bp(WEST, j) + P(iW + 1, j);
            } else if (t == BM_OUTLT1 || t == BM_OUTLT2) {
                PAR(j, jS + 1, jN) P(iW, j) = dTwo * bp(WEST, j) - P(iW + 1, j);
            } else if (t != BM_INTERN && threadIdx.x == 0) {
                *err = 1000 + WEST;
            }
            SYNC();
            t = R.btype(ir, jr, EAST);
            if (t == BM_WALL1 || t == BM_WALL2 || t == BM_INLET) {
                PAR(j, jS + 1, jN) P(iE + 1, j) = bp(EAST, j) + P(iE, j);
            } else if (t == BM_OUTLT1 || t == BM_OUTLT2) {
                PAR(j, jS + 1, jN) P(iE + 1, j) = dTwo * bp(EAST, j) - P(iE, j);
            } else if (t != BM_INTERN && threadIdx.x == 0) {
                *err = 1000 + EAST;
            }
            SYNC();
            t = R.btype(ir, jr, SOUTH);
            if (t == BM_WALL1 || t == BM_WALL2 || t == BM_INLET) {
                PAR(i, iW + 1, iE) P(i, jS) = bp(SOUTH, i) + P(i, jS + 1);
            } else if (t == BM_OUTLT1 || t == BM_OUTLT2) {
                PAR(i, iW + 1, iE) P(i, jS) = dTwo * bp(SOUTH, i) - P(i, jS + 1);
            } else if (t != BM_INTERN && threadIdx.x == 0) {
                *err = 1000 + SOUTH;
            }
            SYNC();
            t = R.btype(ir, jr, NORTH);
            if (t == BM_WALL1 || t == BM_WALL2 || t == BM_INLET) {
                PAR(i, iW + 1, iE) P(i, jN + 1) = bp(NORTH, i) + P(i, jN);
            } else if (t == BM_OUTLT1 || t == BM_OUTLT2) {
                PAR(i, iW + 1, iE) P(i, jN + 1) = dTwo * bp(NORTH, i) - P(i, jN);
            } else if (t != BM_INTERN && threadIdx.x == 0) {
                *err = 1000 + NORTH;
            }
            SYNC();
        }
#undef P
}

template <typename T>
void launch_vel_bc(const Regions &R, int ld, T *u, T *v, bool outflow_only, int *err, hipStream_t s,
                   const int *stop) {
    hipLaunchKernelGGL(k_vel_bc<T>, dim3(1), dim3(BC_THREADS), 0, s, R, ld, u, v, outflow_only ? 1 : 0, err, stop);
}
template <typename T>
void launch_pres_bc(const Regions &R, int ld, T *p, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_pres_bc<T>, dim3(1), dim3(BC_THREADS), 0, s, R, ld, p, err);
}
template void launch_vel_bc<double>(const Regions &, int, double *, double *, bool, int *, hipStream_t,
                                    const int *);
template void launch_vel_bc<float>(const Regions &, int, float *, float *, bool, int *, hipStream_t, const int *);
template void launch_pres_bc<double>(const Regions &, int, double *, int *, hipStream_t);

}  // namespace wd2

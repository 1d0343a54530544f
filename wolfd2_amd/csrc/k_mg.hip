// k_mg.hip -- multigrid Poisson solver for the pressure equation, refined in
// FP64 (the tightly converged "tier B" solve; no counterpart in the reference,
// whose solvers are the six SOR / line-SOR variants of pressure.f).
//
// System: the 5-point operator of pressure.f:97-106 on the interior cells
// i = 2..nx, j = 2..ny, with the ghost pressures frozen at the values
// PresBoundCond set before the solve (main.f:797-803):
//     a1 p(i,j-1) + a2 p(i-1,j) + a3 p(i,j) + a4 p(i+1,j) + a5 p(i,j+1) = b(i,j)
//     a1 = rgv(i,j-1), a2 = rau(i-1,j), a4 = rau(i,j), a5 = rgv(i,j),
//     a3 = -a4 - a2 - a5 - a1.
// After Project, the divergence of an interior cell is dk (b - A p): the
// solve stops when max |dk (b - A p)| < sortol over the interior, the
// divergence the projection leaves (ghost-adjacent cells add the frozen-ghost
// lag of the reference's own algorithm).
//
// Defect correction: each cycle forms r = b - A p in FP64 on the fine grid,
// runs one V-cycle for A e = r in the arithmetic type T (float: the mixed
// path; double: the FP64 path), and adds e to p in FP64.
//
// V-cycle (cell-centred):
// * level 0 is the pressure grid, NX = nx-1 by NY = ny-1 cells; level l+1
//   has ceil(NX/2) x ceil(NY/2) cells, a coarse cell holding fine cells 2I,
//   2I+1 (and 2J, 2J+1) where they exist;
// * face coefficients: level 0 copies rau and rgv; a coarse face takes the
//   mean of the fine faces it covers; a boundary face (to the frozen ghost)
//   is scaled by s_l / s_(l-1) with s_l = 2^(l+1) / (2^l + 1), which keeps
//   the correction's zero at the fine ghost's centre on every level;
// * smoother: red-black Gauss-Seidel, NU sweeps before and after (colours
//   0, 1 down; 1, 0 up); restriction: the sum of a coarse cell's children;
//   prolongation: bilinear (9, 3, 3, 1) / 16 from the coarse cell and its
//   neighbours, zero outside the grid;
// * levels wider than 64 cells run as LDS tiles, one launch per level down
//   (smooth, residual, restrict) and one up (prolong, smooth); the levels of
//   at most 64 x 64 cells run in one workgroup (the bottom kernel, e and r in
//   LDS), down to a grid of at most 2 x 2 cells solved by 6 sweeps.
//
// The down kernel of level 0 also reduces max |r| per tile; every workgroup
// of the next kernel evaluates the stopping test from those maxima (every
// later kernel returns at entry once it has stopped the solve).
#include <cmath>

#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int MG_NU = 2;                       // smoothing sweeps before and after
constexpr int MG_OT = 32;                      // owned cells per tile side
constexpr int MG_HD = 2 * MG_NU + 2;           // down halo (2 NU + 1, even)
constexpr int MG_HU = 2 * MG_NU;               // up halo
constexpr int MG_TD = MG_OT + 2 * MG_HD;       // down tile side (44)
constexpr int MG_TU = MG_OT + 2 * MG_HU;       // up tile side (40)
constexpr int MG_TC = MG_TU / 2 + 2;           // coarse cells of the up tile (22)
constexpr int MG_NT1 = 1024;                   // threads of a coarser tile workgroup (few tiles: latency)
constexpr int MG_BSIDE = 64;                   // widest level the bottom kernel can hold
constexpr int MG_BT = 1024;                    // threads of the bottom workgroup
constexpr int MG_BCAP = 6400;                  // LDS cells per array (e, r) of the bottom levels
constexpr int MG_NCOARSE = 6;                  // sweeps on the coarsest grid (at most 2 x 2 cells)
constexpr int MG_NC = (MG_OT / 2) * (MG_OT / 2);   // coarse cells of a tile
static_assert(MG_HD % 2 == 0 && MG_HU % 2 == 0, "even halos keep the tile colours global");

// Level arrays (elements of T): cell (I, J), 0-based, at (I+1) + ld (J+1),
// with a zero ring (e, r, c); x-face west of cell I (I = 0..nx) at
// I + ld (J+1); y-face south of cell J (J = 0..ny) at (I+1) + ld J.
__device__ inline size_t mc(int ld, int I, int J) { return (size_t)(I + 1) + (size_t)ld * (J + 1); }
__device__ inline size_t mfx(int ld, int I, int J) { return (size_t)I + (size_t)ld * (J + 1); }
__device__ inline size_t mfy(int ld, int I, int J) { return (size_t)(I + 1) + (size_t)ld * J; }

// ------------------------------------------------------------ coefficients
// Level 0 faces from rau and rgv (field layout; uniform values when uni).
template <typename T>
__global__ void k_mg_coef0(int nx, int ny, int ld, const double *__restrict__ rau, const double *__restrict__ rgv,
                           int uni, double rau_c, double rgv_c, MgLevel L) {
    const int NX = nx - 1, NY = ny - 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    T *ax = (T *)L.ax, *ay = (T *)L.ay;
    if (I <= NX && J < NY) ax[mfx(L.ld, I, J)] = (T)(uni ? rau_c : rau[WD2_IX(I + 1, J + 2)]);
    if (I < NX && J <= NY) ay[mfy(L.ld, I, J)] = (T)(uni ? rgv_c : rgv[WD2_IX(I + 2, J + 1)]);
}

// Coarse faces of level c from level f: the mean of the fine faces a coarse
// face covers, boundary faces scaled by sb.
template <typename T>
__global__ void k_mg_coarsen(MgLevel F, MgLevel C, T sb) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
    const T *fx = (const T *)F.ax, *fy = (const T *)F.ay;
    T *cx = (T *)C.ax, *cy = (T *)C.ay;
    if (I <= C.nx && J < C.ny) {
        const int f = min(2 * I, F.nx);
        T v = fx[mfx(F.ld, f, 2 * J)];
        int n = 1;
        if (2 * J + 1 < F.ny) { v += fx[mfx(F.ld, f, 2 * J + 1)]; n = 2; }
        v = v / (T)n;
        if (I == 0 || I == C.nx) v *= sb;
        cx[mfx(C.ld, I, J)] = v;
    }
    if (I < C.nx && J <= C.ny) {
        const int f = min(2 * J, F.ny);
        T v = fy[mfy(F.ld, 2 * I, f)];
        int n = 1;
        if (2 * I + 1 < F.nx) { v += fy[mfy(F.ld, 2 * I + 1, f)]; n = 2; }
        v = v / (T)n;
        if (J == 0 || J == C.ny) v *= sb;
        cy[mfy(C.ld, I, J)] = v;
    }
}

// ------------------------------------------------------------ tile helpers
// One colour of a red-black Gauss-Seidel sweep on an LDS tile of side TW
// (E, R: TW x TW; AX: TW rows of TW + 1 x-faces; AY: TW + 1 rows of TW
// y-faces).  Updated: points off the tile's outer ring, inside the grid.
template <typename T, int TW, int NT>
__device__ inline void tile_colour(T *E, const T *R, const T *AX, const T *AY, int c, int gx0, int gy0, int NX,
                                   int NY) {
    constexpr int KW = TW / 2;
    for (int h = threadIdx.x; h < TW * KW; h += NT) {
        const int y = h / KW, x = 2 * (h % KW) + ((c ^ y) & 1);
        const int I = gx0 + x, J = gy0 + y;
        if (x < 1 || x > TW - 2 || y < 1 || y > TW - 2 || I < 0 || I >= NX || J < 0 || J >= NY) continue;
        const T aw = AX[y * (TW + 1) + x], ae = AX[y * (TW + 1) + x + 1];
        const T as = AY[y * TW + x], an = AY[(y + 1) * TW + x];
        const T a3 = -(aw + ae + as + an);
        const int q = y * TW + x;
        E[q] = (R[q] - (as * E[q - TW] + aw * E[q - 1] + ae * E[q + 1] + an * E[q + TW])) / a3;
    }
}

template <typename T, int TW>
__device__ inline T tile_res(const T *E, const T *R, const T *AX, const T *AY, int x, int y) {
    const T aw = AX[y * (TW + 1) + x], ae = AX[y * (TW + 1) + x + 1];
    const T as = AY[y * TW + x], an = AY[(y + 1) * TW + x];
    const int q = y * TW + x;
    const T a3 = -(aw + ae + as + an);
    return R[q] - (as * E[q - TW] + aw * E[q - 1] + a3 * E[q] + ae * E[q + 1] + an * E[q + TW]);
}

// The same two on a level whose faces all hold one x value ax and one y value
// ay (level 0 of a uniform Cartesian grid: rau and rgv are constants, and the
// faces to the frozen ghosts carry them too), with ia3 = 1 / a3: no face
// arrays, a multiply in place of the division.
template <typename T, int TW, int NT>
__device__ inline void tile_colour_u(T *E, const T *R, T ax, T ay, T ia3, int c, int gx0, int gy0, int NX, int NY) {
    constexpr int KW = TW / 2;
    for (int h = threadIdx.x; h < TW * KW; h += NT) {
        const int y = h / KW, x = 2 * (h % KW) + ((c ^ y) & 1);
        const int I = gx0 + x, J = gy0 + y;
        if (x < 1 || x > TW - 2 || y < 1 || y > TW - 2 || I < 0 || I >= NX || J < 0 || J >= NY) continue;
        const int q = y * TW + x;
        E[q] = (R[q] - (ay * (E[q - TW] + E[q + TW]) + ax * (E[q - 1] + E[q + 1]))) * ia3;
    }
}

template <typename T, int TW>
__device__ inline T tile_res_u(const T *E, const T *R, T ax, T ay, T a3, int x, int y) {
    const int q = y * TW + x;
    return R[q] - (ay * (E[q - TW] + E[q + TW]) + ax * (E[q - 1] + E[q + 1]) + a3 * E[q]);
}

// Faces of a tile of side TW at level L whose cell (0, 0) is (gx0, gy0).
template <typename T, int TW, int NT>
__device__ inline void tile_faces(const MgLevel &L, int gx0, int gy0, T *AX, T *AY) {
    const T *ax = (const T *)L.ax, *ay = (const T *)L.ay;
    for (int q = threadIdx.x; q < TW * (TW + 1); q += NT) {
        const int y = q / (TW + 1), x = q % (TW + 1);
        const int I = gx0 + x, J = gy0 + y;
        AX[q] = (I >= 0 && I <= L.nx && J >= 0 && J < L.ny) ? ax[mfx(L.ld, I, J)] : T(0);
    }
    for (int q = threadIdx.x; q < (TW + 1) * TW; q += NT) {
        const int y = q / TW, x = q % TW;
        const int I = gx0 + x, J = gy0 + y;
        AY[q] = (I >= 0 && I < L.nx && J >= 0 && J <= L.ny) ? ay[mfy(L.ld, I, J)] : T(0);
    }
}

__device__ inline bool mg_stopped(const Scalars *sc) { return sc->mg_done != 0; }

// The stopping test before V-cycle `cycle` (1-based), evaluated by every
// workgroup of the first kernel after the level-0 down kernel from that
// kernel's per-tile maxima of |r|: stop when max |dk r| < tol and at least
// minc cycles are applied (mg_nconv = cycles applied), or when `maxcyc`
// cycles are applied (mg_nconv = -1, not converged; a NaN never passes).
// All workgroups reach the same answer; workgroup 0 records it (val[7] =
// max |dk r|).  sh: NT / 64 doubles of LDS.
template <int NT>
__device__ inline bool mg_decide(const double *__restrict__ part, int nt, double dk, double tol, int cycle,
                                 int maxcyc, int minc, Scalars *sc, double *sh) {
    double m = 0.0;
    for (int k = threadIdx.x; k < nt; k += NT)
        if (!(part[k] <= m)) m = part[k];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        if (!(o <= m)) m = o;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    m = sh[0];
    for (int w = 1; w < NT / 64; w++)
        if (!(sh[w] <= m)) m = sh[w];
    __syncthreads();
    const double div = m * dk;
    const bool conv = div < tol && cycle - 1 >= minc, cap = !conv && cycle - 1 >= maxcyc;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->val[7] = div;
        sc->mg_cycles = cycle - 1;
        if (conv || cap) {
            sc->mg_done = 1;
            sc->mg_nconv = conv ? cycle - 1 : -1;
        }
    }
    return conv || cap;
}

// Arguments of the stopping test for the kernel that evaluates it.
struct MgDecide {
    const double *part;
    int nt, cycle, maxcyc;   // nt = 0: this kernel does not decide
    double dk, tol;
    int minc;                // cycles to apply before the tolerance may stop the solve
};

// XCD-aware tile order: consecutive workgroups go to different XCDs; each XCD
// gets a contiguous band of tiles.
__device__ inline int mg_tile(int ntiles) {
    const int per = (ntiles + 7) / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// ------------------------------------------------------------ down
// Level l: e = NU sweeps from 0 on A e = r, then r_(l+1) = restriction of
// r - A e.  L0: r = b - A p is formed here in FP64 from p (staged in LDS, the
// tile and one more ring) and b (ghosts frozen), stored in T for the up pass,
// and max |r| of the owned points goes to part[tile].  D.nt > 0: this launch
// first evaluates the stopping test of the level-0 kernel before it.
// UF: level 0 with uniform faces (rau_c, rgv_c), see tile_colour_u.
template <typename T, bool L0, int NT, bool STAGE, bool UF>
__global__ __launch_bounds__(NT) void k_mg_down(MgLevel L, MgLevel C, int ntx, int ntiles, int nx, int ny, int ld,
                                                const double *__restrict__ p, const double *__restrict__ b,
                                                const double *__restrict__ rau, const double *__restrict__ rgv,
                                                int uni, double rau_c, double rgv_c, double *__restrict__ part,
                                                MgDecide D, Scalars *sc) {
    if (mg_stopped(sc)) return;
    static_assert(L0 || !UF, "uniform faces are a level-0 shape");
    constexpr int TW = MG_TD, PW = TW + 2;
    __shared__ T E[TW * TW], R[TW * TW], AX[UF ? 1 : TW * (TW + 1)], AY[UF ? 1 : (TW + 1) * TW];
    __shared__ double PS[(L0 && STAGE) ? PW * PW : 1];
    __shared__ double red[NT / 64];
    if (D.nt > 0 && mg_decide<NT>(D.part, D.nt, D.dk, D.tol, D.cycle, D.maxcyc, D.minc, sc, red)) return;
    const int tile = mg_tile(ntiles);
    if (tile >= ntiles) return;
    const int I0 = (tile % ntx) * MG_OT, J0 = (tile / ntx) * MG_OT;
    const int gx0 = I0 - MG_HD, gy0 = J0 - MG_HD;
    const int NX = L.nx, NY = L.ny;
    T *e = (T *)L.e, *r = (T *)L.r;
    const T uax = (T)rau_c, uay = (T)rgv_c, ua3 = -(uax + uax + uay + uay), uia3 = T(1) / ua3;
    if (!UF) tile_faces<T, TW, NT>(L, gx0, gy0, AX, AY);
    if (L0 && STAGE) {
        // p on the tile and one more ring: field (i, j) = (I + 2, J + 2) for
        // I = gx0 - 1 .., 0 outside the field
        for (int q = threadIdx.x; q < PW * PW; q += NT) {
            const int i = gx0 + 1 + q % PW, j = gy0 + 1 + q / PW;
            PS[q] = (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 1) ? p[WD2_IX(i, j)] : 0.0;
        }
        __syncthreads();
    }
    double rmax = 0.0;
    for (int q = threadIdx.x; q < TW * TW; q += NT) {
        const int y = q / TW, x = q % TW;
        const int I = gx0 + x, J = gy0 + y;
        const bool in = I >= 0 && I < NX && J >= 0 && J < NY;
        E[q] = T(0);
        T rv = T(0);
        if (in) {
            if (L0) {
                // FP64 residual of the pressure system at (i, j) = (I+2, J+2)
                const int i = I + 2, j = J + 2;
                double a1, a2, a4, a5;
                if (uni) {
                    a1 = a5 = rgv_c;
                    a2 = a4 = rau_c;
                } else {
                    a1 = rgv[WD2_IX(i, j - 1)];
                    a2 = rau[WD2_IX(i - 1, j)];
                    a4 = rau[WD2_IX(i, j)];
                    a5 = rgv[WD2_IX(i, j)];
                }
                const double a3 = -a4 - a2 - a5 - a1;
                double res;
                if (STAGE) {
                    const int k = (y + 1) * PW + x + 1;
                    res = b[WD2_IX(i, j)] -
                          (a1 * PS[k - PW] + a2 * PS[k - 1] + a3 * PS[k] + a4 * PS[k + 1] + a5 * PS[k + PW]);
                } else {
                    res = b[WD2_IX(i, j)] - (a1 * p[WD2_IX(i, j - 1)] + a2 * p[WD2_IX(i - 1, j)] +
                                             a3 * p[WD2_IX(i, j)] + a4 * p[WD2_IX(i + 1, j)] + a5 * p[WD2_IX(i, j + 1)]);
                }
                rv = (T)res;
                const bool own = x >= MG_HD && x < MG_HD + MG_OT && y >= MG_HD && y < MG_HD + MG_OT;
                // a NaN residual wins the maximum (never converged)
                if (own && !(fabs(res) <= rmax)) rmax = fabs(res);
            } else {
                rv = r[mc(L.ld, I, J)];
            }
        }
        R[q] = rv;
    }
    if (L0) {
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(rmax, off);
            if (!(o <= rmax)) rmax = o;
        }
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rmax;
    }
    __syncthreads();
    if (L0 && threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < NT / 64; w++)
            if (!(red[w] <= m)) m = red[w];
        part[tile] = m;
    }
    for (int s = 0; s < MG_NU; s++)
        for (int c = 0; c < 2; c++) {
            if (UF)
                tile_colour_u<T, TW, NT>(E, R, uax, uay, uia3, c, gx0, gy0, NX, NY);
            else
                tile_colour<T, TW, NT>(E, R, AX, AY, c, gx0, gy0, NX, NY);
            __syncthreads();
        }
    // a coarse cell per thread: its children's residuals summed; e (and, on
    // level 0, r) of the owned children stored for the up pass
    T *rc = (T *)C.r;
    for (int q = threadIdx.x; q < MG_NC; q += NT) {
        const int cx = q % (MG_OT / 2), cy = q / (MG_OT / 2);
        const int Ic = I0 / 2 + cx, Jc = J0 / 2 + cy;
        T sum = T(0);
#pragma unroll
        for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++) {
                const int x = MG_HD + 2 * cx + dx, y = MG_HD + 2 * cy + dy;
                const int I = gx0 + x, J = gy0 + y;
                if (I < NX && J < NY) {
                    sum += UF ? tile_res_u<T, TW>(E, R, uax, uay, ua3, x, y) : tile_res<T, TW>(E, R, AX, AY, x, y);
                    e[mc(L.ld, I, J)] = E[y * TW + x];
                    if (L0) r[mc(L.ld, I, J)] = R[y * TW + x];
                }
            }
        if (Ic < C.nx && Jc < C.ny) rc[mc(C.ld, Ic, Jc)] = sum;
    }
}

// ------------------------------------------------------------ up
// Level l: c = e + prolongation of c_(l+1), then NU sweeps; c of the owned
// points is written (L0: p += c in FP64 instead).  The input e of the halos
// and the output c are separate arrays, so no tile reads what another writes.
// UF: level 0 with uniform faces ax, ay (see tile_colour_u).
template <typename T, bool L0, int NT, bool UF>
__global__ __launch_bounds__(NT) void k_mg_up(MgLevel L, MgLevel C, int ntx, int ntiles, int ld,
                                              double *__restrict__ p, const Scalars *sc, double rau_c, double rgv_c) {
    if (mg_stopped(sc)) return;
    static_assert(L0 || !UF, "uniform faces are a level-0 shape");
    constexpr int TW = MG_TU;
    __shared__ T E[TW * TW], R[TW * TW], AX[UF ? 1 : TW * (TW + 1)], AY[UF ? 1 : (TW + 1) * TW], EC[MG_TC * MG_TC];
    const int tile = mg_tile(ntiles);
    if (tile >= ntiles) return;
    const int I0 = (tile % ntx) * MG_OT, J0 = (tile / ntx) * MG_OT;
    const int gx0 = I0 - MG_HU, gy0 = J0 - MG_HU;
    const int cgx0 = gx0 / 2 - 1, cgy0 = gy0 / 2 - 1;   // gx0 is even
    const int NX = L.nx, NY = L.ny;
    const T *e = (const T *)L.e, *r = (const T *)L.r, *ec = (const T *)C.c;
    const T uax = (T)rau_c, uay = (T)rgv_c, uia3 = T(1) / -(uax + uax + uay + uay);
    if (!UF) tile_faces<T, TW, NT>(L, gx0, gy0, AX, AY);
    for (int q = threadIdx.x; q < MG_TC * MG_TC; q += NT) {
        const int Ic = cgx0 + q % MG_TC, Jc = cgy0 + q / MG_TC;
        EC[q] = (Ic >= 0 && Ic < C.nx && Jc >= 0 && Jc < C.ny) ? ec[mc(C.ld, Ic, Jc)] : T(0);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < TW * TW; q += NT) {
        const int y = q / TW, x = q % TW;
        const int I = gx0 + x, J = gy0 + y;
        T ev = T(0), rv = T(0);
        if (I >= 0 && I < NX && J >= 0 && J < NY) {
            const int ic = (I >> 1) - cgx0, jc = (J >> 1) - cgy0;
            const int dx = (I & 1) ? 1 : -1, dy = (J & 1) ? 1 : -1;
            const T pc = (T(9) * EC[jc * MG_TC + ic] + T(3) * EC[jc * MG_TC + ic + dx] +
                          T(3) * EC[(jc + dy) * MG_TC + ic] + EC[(jc + dy) * MG_TC + ic + dx]) *
                         T(0.0625);
            ev = e[mc(L.ld, I, J)] + pc;
            rv = r[mc(L.ld, I, J)];
        }
        E[q] = ev;
        R[q] = rv;
    }
    __syncthreads();
    for (int s = 0; s < MG_NU; s++)
        for (int c = 1; c >= 0; c--) {
            if (UF)
                tile_colour_u<T, TW, NT>(E, R, uax, uay, uia3, c, gx0, gy0, NX, NY);
            else
                tile_colour<T, TW, NT>(E, R, AX, AY, c, gx0, gy0, NX, NY);
            __syncthreads();
        }
    T *eo = (T *)L.c;
    for (int q = threadIdx.x; q < MG_OT * MG_OT; q += NT) {
        const int x = MG_HU + q % MG_OT, y = MG_HU + q / MG_OT;
        const int I = gx0 + x, J = gy0 + y;
        if (I >= NX || J >= NY) continue;
        if (L0)
            p[WD2_IX(I + 2, J + 2)] += (double)E[y * TW + x];
        else
            eo[mc(L.ld, I, J)] = E[y * TW + x];
    }
}

// ------------------------------------------------------------ bottom
// Levels lb..nlev-1 (each at most MG_BSIDE cells wide) in one workgroup: the
// V-cycle below level lb with the corrections and right-hand sides in LDS.
// The face coefficients are copied into LDS first where they fit
// (mg_bottom_faces_fit: compact rows of nx + 1 x-faces and nx y-faces per
// level), so the ~50 sweep phases between barriers read only LDS; otherwise
// they are read from the (small, cache-resident) level arrays.  Both give the
// same bits.  Level lb's r comes from the down kernel above; its correction
// goes to c for the up kernel above.  D.nt > 0: the stopping test is
// evaluated first.
struct MgBottom {
    MgLevel lev[MG_MAXLEV];
    int nlev;   // bottom levels
};

// LDS face values per direction: every bottom level of a 64-wide bottom in
// FP32, of a 32-wide one in FP64 (the FP64 corrections and right-hand sides
// take the rest of the LDS).
template <typename T>
constexpr int mg_bfcap() { return sizeof(T) == 4 ? 5632 : 1440; }

// Faces of one bottom level: the LDS copy (LF) or the level arrays.
template <typename T, bool LF>
struct BtFaces {
    const T *fx, *fy;   // LF: LDS rows of NX + 1 (x) and NX (y) values
    const T *ax, *ay;   // the level arrays
    int ld, NX;
    __device__ T x(int I, int J) const { return LF ? fx[I + (NX + 1) * J] : ax[mfx(ld, I, J)]; }
    __device__ T y(int I, int J) const { return LF ? fy[I + NX * J] : ay[mfy(ld, I, J)]; }
};

// Red-black Gauss-Seidel sweeps of level L (E, R in LDS with row stride
// nx + 2): nsweep sweeps of colours 0 then 1 (rev: 1 then 0), a sync after
// each colour, by S threads.  A thread's points (point q of a colour: row
// q / half, position q % half; q = thread + p S, at most P per thread and
// colour) keep their index, face
// coefficients, diagonal and right-hand side in registers across the sweeps
// (R does not change while a level is smoothed), so a phase reads only the
// four neighbours in E.  The update is the same expression on the same
// values as a point-by-point evaluation.
template <typename T, bool LF, int P, int S, typename Sync>
__device__ inline void bt_smooth(T *E, const T *R, const MgLevel &L, const BtFaces<T, LF> &F, int nsweep, bool rev,
                                 Sync sync) {
    const int NX = L.nx, NY = L.ny, W = NX + 2;
    const int half = (NX + 1) / 2;   // points of one colour per row, at most
    int kk[2][P];
    T cw[2][P], ce[2][P], cs[2][P], cn[2][P], ca[2][P], cr[2][P];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int p = 0; p < P; p++) {
            const int q = threadIdx.x + p * S;
            const int J = q / half, I = 2 * (q % half) + ((c ^ J) & 1);
            kk[c][p] = -1;
            if (q < half * NY && I < NX) {
                cw[c][p] = F.x(I, J);
                ce[c][p] = F.x(I + 1, J);
                cs[c][p] = F.y(I, J);
                cn[c][p] = F.y(I, J + 1);
                ca[c][p] = -(cw[c][p] + ce[c][p] + cs[c][p] + cn[c][p]);
                kk[c][p] = (I + 1) + W * (J + 1);
                cr[c][p] = R[kk[c][p]];
            }
        }
    for (int s = 0; s < nsweep; s++)
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            const int c = rev ? 1 - cc : cc;
#pragma unroll
            for (int p = 0; p < P; p++) {
                const int k = kk[c][p];
                if (k >= 0)
                    E[k] = (cr[c][p] - (cs[c][p] * E[k - W] + cw[c][p] * E[k - 1] + ce[c][p] * E[k + 1] +
                                        cn[c][p] * E[k + W])) /
                           ca[c][p];
            }
            sync();
        }
}

template <typename T, bool LF>
__global__ __launch_bounds__(MG_BT) void k_mg_bottom(MgBottom B, MgDecide D, Scalars *sc) {
    if (mg_stopped(sc)) return;
    constexpr int FC = LF ? mg_bfcap<T>() : 1;
    __shared__ T E[MG_BCAP], R[MG_BCAP];
    __shared__ T FX[FC], FY[FC];
    __shared__ double red[MG_BT / 64];
    if (D.nt > 0 && mg_decide<MG_BT>(D.part, D.nt, D.dk, D.tol, D.cycle, D.maxcyc, D.minc, sc, red)) return;
    // each level's offset in E and R, and in FX and FY (in LDS: indexed by
    // the level, they would otherwise go to scratch)
    __shared__ int off[MG_MAXLEV], ofx[MG_MAXLEV], ofy[MG_MAXLEV];
    if (threadIdx.x == 0) {
        int o = 0, ox = 0, oy = 0;
        for (int l = 0; l < B.nlev; l++) {
            const int NX = B.lev[l].nx, NY = B.lev[l].ny;
            off[l] = o;
            ofx[l] = ox;
            ofy[l] = oy;
            o += (NX + 2) * (NY + 2);
            ox += (NX + 1) * NY;
            oy += NX * (NY + 1);
        }
    }
    __syncthreads();
    auto faces = [&](int l) {
        const MgLevel &L = B.lev[l];
        return BtFaces<T, LF>{FX + ofx[l], FY + ofy[l], (const T *)L.ax, (const T *)L.ay, L.ld, L.nx};
    };
    // Set-up: every global load first (r of the first level, the faces),
    // then the LDS stores, so the loads' latencies overlap.  E is zeroed
    // everywhere, R everywhere but the first level's interior.
    constexpr int RB = (MG_BSIDE * MG_BSIDE + MG_BT - 1) / MG_BT;   // r values per thread
    constexpr int FB = LF ? (mg_bfcap<T>() + MG_BT - 1) / MG_BT : 1;   // face values per thread and direction
    const MgLevel &L0 = B.lev[0];
    const int n0 = L0.nx * L0.ny;
    T rv[RB], fxv[FB], fyv[FB];
#pragma unroll
    for (int it = 0; it < RB; it++) {
        const int q = threadIdx.x + it * MG_BT;
        if (q < n0) rv[it] = ((const T *)L0.r)[(size_t)(q % L0.nx + 1) + (size_t)L0.ld * (q / L0.nx + 1)];
    }
    const int nl = B.nlev - 1;
    const int nfx = ofx[nl] + (B.lev[nl].nx + 1) * B.lev[nl].ny, nfy = ofy[nl] + B.lev[nl].nx * (B.lev[nl].ny + 1);
    if (LF) {
#pragma unroll
        for (int it = 0; it < FB; it++) {
            const int q = threadIdx.x + it * MG_BT;
            if (q < nfx) {
                int l = 0;
                while (l < nl && q >= ofx[l + 1]) l++;
                const MgLevel &L = B.lev[l];
                const int u = q - ofx[l];
                fxv[it] = ((const T *)L.ax)[mfx(L.ld, u % (L.nx + 1), u / (L.nx + 1))];
            }
            if (q < nfy) {
                int l = 0;
                while (l < nl && q >= ofy[l + 1]) l++;
                const MgLevel &L = B.lev[l];
                const int u = q - ofy[l];
                fyv[it] = ((const T *)L.ay)[mfy(L.ld, u % L.nx, u / L.nx)];
            }
        }
    }
    for (int l = 0; l < B.nlev; l++) {
        const MgLevel &L = B.lev[l];
        const int W = L.nx + 2, n = W * (L.ny + 2);
        for (int q = threadIdx.x; q < n; q += MG_BT) {
            const int a = q % W, bb = q / W;   // padded position
            E[off[l] + q] = T(0);
            if (!(l == 0 && a >= 1 && a <= L.nx && bb >= 1 && bb <= L.ny)) R[off[l] + q] = T(0);
        }
    }
#pragma unroll
    for (int it = 0; it < RB; it++) {
        const int q = threadIdx.x + it * MG_BT;
        if (q < n0) R[off[0] + (q % L0.nx + 1) + (L0.nx + 2) * (q / L0.nx + 1)] = rv[it];
    }
    if (LF) {
#pragma unroll
        for (int it = 0; it < FB; it++) {
            const int q = threadIdx.x + it * MG_BT;
            if (q < nfx) FX[q] = fxv[it];
            if (q < nfy) FY[q] = fyv[it];
        }
    }
    __syncthreads();
    // The levels from lw on (at most 64 points of a colour: one per lane; the
    // levels shrink) run in wavefront 0 alone, whose phases need only a
    // wavefront sync; the wider ones use the whole workgroup between
    // barriers.  The operations and their order are the same either way.
    int lw = B.nlev;
    for (int l = B.nlev - 1; l >= 0 && ((B.lev[l].nx + 1) / 2) * B.lev[l].ny <= 64; l--) lw = l;
    const bool w0 = threadIdx.x < 64;
    auto sync = [](bool wave) {
        if (wave) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    };
    // a level's sweeps: in wavefront 0 (one point per lane and colour) or by the whole workgroup (at most MG_BSIDE^2 / 2
    // points per colour: two per thread)
    static_assert(MG_BSIDE * MG_BSIDE / 2 <= 2 * MG_BT, "two points per thread and colour");
    auto smooth = [&](T *El, const T *Rl, const MgLevel &L, const BtFaces<T, LF> &F, int nsweep, bool rev,
                      bool wave) {
        if (wave)
            bt_smooth<T, LF, 1, 64>(El, Rl, L, F, nsweep, rev, [&] { sync(true); });
        else
            bt_smooth<T, LF, 2, MG_BT>(El, Rl, L, F, nsweep, rev, [&] { sync(false); });
    };
    // pre-smoothing of level l and the restriction of its residual
    auto down = [&](int l, bool wave) {
        const int step = wave ? 64 : MG_BT;
        const MgLevel &L = B.lev[l];
        const BtFaces<T, LF> F = faces(l);
        const int NX = L.nx, NY = L.ny, W = NX + 2;
        T *El = E + off[l], *Rl = R + off[l];
        smooth(El, Rl, L, F, MG_NU, false, wave);
        const int NXc = B.lev[l + 1].nx, NYc = B.lev[l + 1].ny, Wc = NXc + 2;
        T *Rc = R + off[l + 1];
        for (int q = threadIdx.x; q < NXc * NYc; q += step) {
            const int Ic = q % NXc, Jc = q / NXc;
            T sum = T(0);
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const int I = 2 * Ic + dx, J = 2 * Jc + dy;
                    if (I >= NX || J >= NY) continue;
                    const T aw = F.x(I, J), ae = F.x(I + 1, J);
                    const T as = F.y(I, J), an = F.y(I, J + 1);
                    const T a3 = -(aw + ae + as + an);
                    const int k = (I + 1) + W * (J + 1);
                    sum += Rl[k] - (as * El[k - W] + aw * El[k - 1] + a3 * El[k] + ae * El[k + 1] + an * El[k + W]);
                }
            Rc[(Ic + 1) + Wc * (Jc + 1)] = sum;
        }
        sync(wave);
    };
    // the coarse correction of level l + 1 added to level l, post-smoothing
    auto up = [&](int l, bool wave) {
        const int step = wave ? 64 : MG_BT;
        const int NX = B.lev[l].nx, NY = B.lev[l].ny, W = NX + 2;
        const int NXc = B.lev[l + 1].nx, Wc = NXc + 2;
        T *El = E + off[l];
        const T *Ec = E + off[l + 1];
        for (int q = threadIdx.x; q < NX * NY; q += step) {
            const int I = q % NX, J = q / NX;
            const int ic = (I >> 1) + 1, jc = (J >> 1) + 1;   // padded coarse position
            const int dx = (I & 1) ? 1 : -1, dy = (J & 1) ? 1 : -1;
            // coarse neighbours beyond the grid read the zero ring
            const T pc = (T(9) * Ec[ic + Wc * jc] + T(3) * Ec[ic + dx + Wc * jc] + T(3) * Ec[ic + Wc * (jc + dy)] +
                          Ec[ic + dx + Wc * (jc + dy)]) *
                         T(0.0625);
            El[(I + 1) + W * (J + 1)] += pc;
        }
        sync(wave);
        smooth(El, R + off[l], B.lev[l], faces(l), MG_NU, true, wave);
    };
    // sweeps on the coarsest level
    auto coarsest = [&](bool wave) {
        const int l = B.nlev - 1;
        smooth(E + off[l], R + off[l], B.lev[l], faces(l), MG_NCOARSE, false, wave);
    };
    const int lb = lw < B.nlev ? lw : B.nlev - 1;   // levels below lb: whole workgroup
    for (int l = 0; l < lb; l++) down(l, false);
    if (lw < B.nlev) {
        if (w0) {
            for (int l = lw; l + 1 < B.nlev; l++) down(l, true);
            coarsest(true);
            for (int l = B.nlev - 2; l >= lw; l--) up(l, true);
        }
        __syncthreads();
    } else {
        coarsest(false);
    }
    for (int l = lb - 1; l >= 0; l--) up(l, false);
    const MgLevel &L = B.lev[0];
    T *e = (T *)L.c;
    const int W = L.nx + 2;
    for (int q = threadIdx.x; q < L.nx * L.ny; q += MG_BT) {
        const int I = q % L.nx, J = q / L.nx;
        e[(size_t)(I + 1) + (size_t)L.ld * (J + 1)] = E[(I + 1) + W * (J + 1)];
    }
}

// Whether the face coefficients of the bottom levels fit the kernel's LDS copy.
template <typename T>
static bool mg_bottom_faces_fit(const MgBottom &B) {
    int fx = 0, fy = 0;
    for (int l = 0; l < B.nlev; l++) {
        fx += (B.lev[l].nx + 1) * B.lev[l].ny;
        fy += B.lev[l].nx * (B.lev[l].ny + 1);
    }
    return fx <= mg_bfcap<T>() && fy <= mg_bfcap<T>();
}

template <typename T>
static void launch_mg_bottom(const MgBottom &B, const MgDecide &D, Scalars *sc, hipStream_t s) {
    if (mg_bottom_faces_fit<T>(B))
        hipLaunchKernelGGL((k_mg_bottom<T, true>), dim3(1), dim3(MG_BT), 0, s, B, D, sc);
    else
        hipLaunchKernelGGL((k_mg_bottom<T, false>), dim3(1), dim3(MG_BT), 0, s, B, D, sc);
}

// ------------------------------------------------------------ control
__global__ void k_mg_reset(Scalars *sc) {
    sc->mg_done = 0;
    sc->mg_nconv = -1;
    sc->mg_cycles = 0;
    sc->val[7] = 0.0;
}

// ------------------------------------------------------------ host side
int mg_tiles(int NX, int NY, int *ntx) {
    const int tx = (NX + MG_OT - 1) / MG_OT, ty = (NY + MG_OT - 1) / MG_OT;
    if (ntx) *ntx = tx;
    return tx * ty;
}
int mg_bottom_side() { return MG_BSIDE; }
int mg_bottom_cells() { return MG_BCAP; }
int mg_sweeps() { return MG_NU; }

template <typename T>
void launch_mg_coef(const MgHierarchy &H, int nx, int ny, int ld, const double *rau, const double *rgv,
                    const double *uni, hipStream_t s) {
    const MgLevel &L0 = H.lev[0];
    hipLaunchKernelGGL(k_mg_coef0<T>, dim3((L0.nx + 1 + 255) / 256, L0.ny + 1), dim3(256), 0, s, nx, ny, ld, rau,
                       rgv, uni ? 1 : 0, uni ? uni[0] : 0.0, uni ? uni[1] : 0.0, L0);
    for (int l = 1; l < H.nlev; l++) {
        const double sl = std::ldexp(1.0, l + 1) / (std::ldexp(1.0, l) + 1.0);
        const double sp = std::ldexp(1.0, l) / (std::ldexp(1.0, l - 1) + 1.0);
        const MgLevel &C = H.lev[l];
        hipLaunchKernelGGL(k_mg_coarsen<T>, dim3((C.nx + 1 + 255) / 256, C.ny + 1), dim3(256), 0, s, H.lev[l - 1], C,
                           (T)(sl / sp));
    }
}

template <typename T, int NT, bool STAGE>
static void down0_as(const MgHierarchy &H, int nx, int ny, int ld, const double *b, const double *p,
                     const double *rau, const double *rgv, const double *uni, double *part, Scalars *sc,
                     hipStream_t s) {
    int ntx = 0;
    const int nt = mg_tiles(H.lev[0].nx, H.lev[0].ny, &ntx);
    const MgDecide none{nullptr, 0, 0, 0, 0.0, 0.0};
    if (uni)
        hipLaunchKernelGGL((k_mg_down<T, true, NT, STAGE, true>), dim3(8 * ((nt + 7) / 8)), dim3(NT), 0, s, H.lev[0],
                           H.lev[1], ntx, nt, nx, ny, ld, p, b, rau, rgv, 1, uni[0], uni[1], part, none, sc);
    else
        hipLaunchKernelGGL((k_mg_down<T, true, NT, STAGE, false>), dim3(8 * ((nt + 7) / 8)), dim3(NT), 0, s,
                           H.lev[0], H.lev[1], ntx, nt, nx, ny, ld, p, b, rau, rgv, 0, 0.0, 0.0, part, none, sc);
}

// the level-0 down kernel in the shape the hierarchy's tuning selects
template <typename T>
static void down0(const MgHierarchy &H, int nx, int ny, int ld, const double *b, const double *p, const double *rau,
                  const double *rgv, const double *uni, double *part, Scalars *sc, hipStream_t s) {
    const bool big = H.tune.nt0 >= 512;
    if (H.tune.stage) {
        if (big)
            down0_as<T, 512, true>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
        else
            down0_as<T, 256, true>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
    } else {
        if (big)
            down0_as<T, 512, false>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
        else
            down0_as<T, 256, false>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
    }
}

template <typename T, int NT>
static void up0_as(const MgHierarchy &H, int ld, double *p, const double *uni, Scalars *sc, hipStream_t s) {
    int ntx = 0;
    const int nt = mg_tiles(H.lev[0].nx, H.lev[0].ny, &ntx);
    if (uni)
        hipLaunchKernelGGL((k_mg_up<T, true, NT, true>), dim3(8 * ((nt + 7) / 8)), dim3(NT), 0, s, H.lev[0], H.lev[1],
                           ntx, nt, ld, p, sc, uni[0], uni[1]);
    else
        hipLaunchKernelGGL((k_mg_up<T, true, NT, false>), dim3(8 * ((nt + 7) / 8)), dim3(NT), 0, s, H.lev[0],
                           H.lev[1], ntx, nt, ld, p, sc, 0.0, 0.0);
}

template <typename T>
static void up0(const MgHierarchy &H, int ld, double *p, const double *uni, Scalars *sc, hipStream_t s) {
    if (H.tune.nt0 >= 512)
        up0_as<T, 512>(H, ld, p, uni, sc, s);
    else
        up0_as<T, 256>(H, ld, p, uni, sc, s);
}

// One V-cycle of defect correction (cycle is 1-based), every kernel
// returning at entry once the solve has stopped; the stopping test of the
// cycle's level-0 residual runs at the start of the next kernel.
template <typename T>
void launch_mg_cycle(const MgHierarchy &H, int nx, int ny, int ld, const double *b, double *p, const double *rau,
                     const double *rgv, const double *uni, double dk, double tol, int cycle, int maxcyc,
                     double *part, Scalars *sc, hipStream_t s, hipEvent_t *ev_down, hipEvent_t *ev_up, int minc) {
    int ntx = 0;
    const int nt0 = mg_tiles(H.lev[0].nx, H.lev[0].ny, &ntx);
    if (ev_down) (void)hipEventRecord(ev_down[0], s);
    down0<T>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
    if (ev_down) (void)hipEventRecord(ev_down[1], s);
    const MgDecide D{part, nt0, cycle, maxcyc, dk, tol, minc};
    const MgDecide none{nullptr, 0, 0, 0, 0.0, 0.0};
    for (int l = 1; l < H.ntiled; l++) {
        const int nt = mg_tiles(H.lev[l].nx, H.lev[l].ny, &ntx);
        hipLaunchKernelGGL((k_mg_down<T, false, MG_NT1, false, false>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT1), 0, s, H.lev[l],
                           H.lev[l + 1], ntx, nt, nx, ny, ld, p, b, rau, rgv, 0, 0.0, 0.0, part, l == 1 ? D : none,
                           sc);
    }
    MgBottom B{};
    B.nlev = H.nlev - H.ntiled;
    for (int l = 0; l < B.nlev; l++) B.lev[l] = H.lev[H.ntiled + l];
    launch_mg_bottom<T>(B, H.ntiled == 1 ? D : none, sc, s);
    for (int l = H.ntiled - 1; l >= 1; l--) {
        const int nt = mg_tiles(H.lev[l].nx, H.lev[l].ny, &ntx);
        hipLaunchKernelGGL((k_mg_up<T, false, MG_NT1, false>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT1), 0, s, H.lev[l],
                           H.lev[l + 1], ntx, nt, ld, p, sc, 0.0, 0.0);
    }
    if (ev_up) (void)hipEventRecord(ev_up[0], s);
    up0<T>(H, ld, p, uni, sc, s);
    if (ev_up) (void)hipEventRecord(ev_up[1], s);
}

// `reps` back-to-back launches of the level-0 down (which = 2) or up
// (which = 3) kernel, or of the bottom kernel (4; it writes only the
// correction c of its first level), between one event pair, on the state of
// the last solve
// (the stop flag is cleared first and set again after; the up kernel adds
// the last correction to p each time, so the caller restores p).
template <typename T>
void launch_mg_kernel_bench(const MgHierarchy &H, int which, int reps, int nx, int ny, int ld, const double *b,
                            double *p, const double *rau, const double *rgv, const double *uni, double *part,
                            Scalars *sc, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    launch_mg_reset(sc, s);
    MgBottom B{};
    B.nlev = H.nlev - H.ntiled;
    for (int l = 0; l < B.nlev; l++) B.lev[l] = H.lev[H.ntiled + l];
    const MgDecide none{nullptr, 0, 0, 0, 0.0, 0.0};
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps; k++) {
        if (which == 2)
            down0<T>(H, nx, ny, ld, b, p, rau, rgv, uni, part, sc, s);
        else if (which == 3)
            up0<T>(H, ld, p, uni, sc, s);
        else
            launch_mg_bottom<T>(B, none, sc, s);
    }
    (void)hipEventRecord(e1, s);
}

void launch_mg_reset(Scalars *sc, hipStream_t s) { hipLaunchKernelGGL(k_mg_reset, dim3(1), dim3(1), 0, s, sc); }

template void launch_mg_coef<float>(const MgHierarchy &, int, int, int, const double *, const double *,
                                    const double *, hipStream_t);
template void launch_mg_coef<double>(const MgHierarchy &, int, int, int, const double *, const double *,
                                     const double *, hipStream_t);
template void launch_mg_cycle<float>(const MgHierarchy &, int, int, int, const double *, double *, const double *,
                                     const double *, const double *, double, double, int, int, double *, Scalars *,
                                     hipStream_t, hipEvent_t *, hipEvent_t *, int);
template void launch_mg_cycle<double>(const MgHierarchy &, int, int, int, const double *, double *, const double *,
                                      const double *, const double *, double, double, int, int, double *, Scalars *,
                                      hipStream_t, hipEvent_t *, hipEvent_t *, int);
template void launch_mg_kernel_bench<float>(const MgHierarchy &, int, int, int, int, int, const double *, double *,
                                            const double *, const double *, const double *, double *, Scalars *,
                                            hipStream_t, hipEvent_t, hipEvent_t);
template void launch_mg_kernel_bench<double>(const MgHierarchy &, int, int, int, int, int, const double *, double *,
                                             const double *, const double *, const double *, double *, Scalars *,
                                             hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace wd2

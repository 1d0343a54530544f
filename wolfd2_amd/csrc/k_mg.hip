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
// * levels wider than 32 cells run as LDS tiles, one launch per level down
//   (smooth, residual, restrict) and one up (prolong, smooth); the levels of
//   at most 32 x 32 cells run in one workgroup out of LDS (the bottom kernel),
//   down to a grid of at most 2 x 2 cells solved by 16 sweeps.
//
// The down kernel of level 0 also reduces max |r| per tile; the check kernel
// that follows decides (every later kernel returns at entry once it has).
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
constexpr int MG_NT = 256;                     // threads of a tile workgroup
constexpr int MG_BT = 1024;                    // threads of the bottom workgroup
constexpr int MG_BCAP = 2048;                  // LDS cells per array of the bottom levels
constexpr int MG_NCOARSE = 16;                 // sweeps on the coarsest grid
static_assert((MG_OT / 2) * (MG_OT / 2) == MG_NT, "one coarse cell per thread");
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

// XCD-aware tile order: consecutive workgroups go to different XCDs; each XCD
// gets a contiguous band of tiles.
__device__ inline int mg_tile(int ntiles) {
    const int per = (ntiles + 7) / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// ------------------------------------------------------------ down
// Level l: e = NU sweeps from 0 on A e = r, then r_(l+1) = restriction of
// r - A e.  L0: r = b - A p is formed here in FP64 from p and b (ghosts
// frozen), stored in T for the up pass, and max |r| of the owned points goes
// to part[tile].
template <typename T, bool L0>
__global__ __launch_bounds__(MG_NT) void k_mg_down(MgLevel L, MgLevel C, int ntx, int ntiles, int nx, int ny, int ld,
                                                   const double *__restrict__ p, const double *__restrict__ b,
                                                   const double *__restrict__ rau, const double *__restrict__ rgv,
                                                   int uni, double rau_c, double rgv_c, double *__restrict__ part,
                                                   const Scalars *sc) {
    if (mg_stopped(sc)) return;
    constexpr int TW = MG_TD;
    __shared__ T E[TW * TW], R[TW * TW], AX[TW * (TW + 1)], AY[(TW + 1) * TW];
    __shared__ double red[MG_NT / 64];
    const int tile = mg_tile(ntiles);
    if (tile >= ntiles) return;
    const int I0 = (tile % ntx) * MG_OT, J0 = (tile / ntx) * MG_OT;
    const int gx0 = I0 - MG_HD, gy0 = J0 - MG_HD;
    const int NX = L.nx, NY = L.ny;
    T *e = (T *)L.e, *r = (T *)L.r;
    tile_faces<T, TW, MG_NT>(L, gx0, gy0, AX, AY);
    double rmax = 0.0;
    for (int q = threadIdx.x; q < TW * TW; q += MG_NT) {
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
                const double res = b[WD2_IX(i, j)] - (a1 * p[WD2_IX(i, j - 1)] + a2 * p[WD2_IX(i - 1, j)] +
                                                      a3 * p[WD2_IX(i, j)] + a4 * p[WD2_IX(i + 1, j)] +
                                                      a5 * p[WD2_IX(i, j + 1)]);
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
        for (int w = 1; w < MG_NT / 64; w++)
            if (!(red[w] <= m)) m = red[w];
        part[tile] = m;
    }
    for (int s = 0; s < MG_NU; s++)
        for (int c = 0; c < 2; c++) {
            tile_colour<T, TW, MG_NT>(E, R, AX, AY, c, gx0, gy0, NX, NY);
            __syncthreads();
        }
    // one coarse cell per thread: its children's residuals summed; e (and,
    // on level 0, r) of the owned children stored for the up pass
    T *rc = (T *)C.r;
    const int cx = threadIdx.x % (MG_OT / 2), cy = threadIdx.x / (MG_OT / 2);
    const int Ic = I0 / 2 + cx, Jc = J0 / 2 + cy;
    T sum = T(0);
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const int x = MG_HD + 2 * cx + dx, y = MG_HD + 2 * cy + dy;
            const int I = gx0 + x, J = gy0 + y;
            if (I < NX && J < NY) {
                sum += tile_res<T, TW>(E, R, AX, AY, x, y);
                e[mc(L.ld, I, J)] = E[y * TW + x];
                if (L0) r[mc(L.ld, I, J)] = R[y * TW + x];
            }
        }
    if (Ic < C.nx && Jc < C.ny) rc[mc(C.ld, Ic, Jc)] = sum;
}

// ------------------------------------------------------------ up
// Level l: c = e + prolongation of c_(l+1), then NU sweeps; c of the owned
// points is written (L0: p += c in FP64 instead).  The input e of the halos
// and the output c are separate arrays, so no tile reads what another writes.
template <typename T, bool L0>
__global__ __launch_bounds__(MG_NT) void k_mg_up(MgLevel L, MgLevel C, int ntx, int ntiles, int ld,
                                                 double *__restrict__ p, const Scalars *sc) {
    if (mg_stopped(sc)) return;
    constexpr int TW = MG_TU;
    __shared__ T E[TW * TW], R[TW * TW], AX[TW * (TW + 1)], AY[(TW + 1) * TW], EC[MG_TC * MG_TC];
    const int tile = mg_tile(ntiles);
    if (tile >= ntiles) return;
    const int I0 = (tile % ntx) * MG_OT, J0 = (tile / ntx) * MG_OT;
    const int gx0 = I0 - MG_HU, gy0 = J0 - MG_HU;
    const int cgx0 = gx0 / 2 - 1, cgy0 = gy0 / 2 - 1;   // gx0 is even
    const int NX = L.nx, NY = L.ny;
    const T *e = (const T *)L.e, *r = (const T *)L.r, *ec = (const T *)C.c;
    tile_faces<T, TW, MG_NT>(L, gx0, gy0, AX, AY);
    for (int q = threadIdx.x; q < MG_TC * MG_TC; q += MG_NT) {
        const int Ic = cgx0 + q % MG_TC, Jc = cgy0 + q / MG_TC;
        EC[q] = (Ic >= 0 && Ic < C.nx && Jc >= 0 && Jc < C.ny) ? ec[mc(C.ld, Ic, Jc)] : T(0);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < TW * TW; q += MG_NT) {
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
            tile_colour<T, TW, MG_NT>(E, R, AX, AY, c, gx0, gy0, NX, NY);
            __syncthreads();
        }
    T *eo = (T *)L.c;
    for (int q = threadIdx.x; q < MG_OT * MG_OT; q += MG_NT) {
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
// Levels lb..nlev-1 (each at most 32 x 32 cells) in one workgroup: the
// V-cycle below level lb with everything in LDS.  Level lb's r comes from the
// down kernel above; its correction goes to c for the up kernel above.
struct MgBottom {
    MgLevel lev[MG_MAXLEV];
    int nlev;   // bottom levels
};

template <typename T>
__device__ inline void bt_colour(T *E, const T *R, const T *AX, const T *AY, int NX, int NY, int c) {
    const int W = NX + 2;
    for (int q = threadIdx.x; q < NX * NY; q += MG_BT) {
        const int I = q % NX, J = q / NX;
        if (((I + J) & 1) != c) continue;
        const T aw = AX[I + W * (J + 1)], ae = AX[I + 1 + W * (J + 1)];
        const T as = AY[I + 1 + W * J], an = AY[I + 1 + W * (J + 1)];
        const T a3 = -(aw + ae + as + an);
        const int k = (I + 1) + W * (J + 1);
        E[k] = (R[k] - (as * E[k - W] + aw * E[k - 1] + ae * E[k + 1] + an * E[k + W])) / a3;
    }
}

template <typename T>
__global__ __launch_bounds__(MG_BT) void k_mg_bottom(MgBottom B, const Scalars *sc) {
    if (mg_stopped(sc)) return;
    __shared__ T E[MG_BCAP], R[MG_BCAP], AX[MG_BCAP], AY[MG_BCAP];
    int off[MG_MAXLEV];
    int o = 0;
    for (int l = 0; l < B.nlev; l++) {
        off[l] = o;
        o += (B.lev[l].nx + 2) * (B.lev[l].ny + 2);
    }
    // load: faces of every level, r of the first; zero e (and the rings)
    for (int l = 0; l < B.nlev; l++) {
        const MgLevel &L = B.lev[l];
        const int W = L.nx + 2, n = W * (L.ny + 2);
        const T *ax = (const T *)L.ax, *ay = (const T *)L.ay, *r = (const T *)L.r;
        for (int q = threadIdx.x; q < n; q += MG_BT) {
            const int a = q % W, bb = q / W;   // padded position
            AX[off[l] + q] = (a <= L.nx && bb >= 1 && bb <= L.ny) ? ax[(size_t)a + (size_t)L.ld * bb] : T(0);
            AY[off[l] + q] = (a >= 1 && a <= L.nx && bb <= L.ny) ? ay[(size_t)a + (size_t)L.ld * bb] : T(0);
            E[off[l] + q] = T(0);
            R[off[l] + q] = (l == 0 && a >= 1 && a <= L.nx && bb >= 1 && bb <= L.ny) ? r[(size_t)a + (size_t)L.ld * bb]
                                                                                   : T(0);
        }
    }
    __syncthreads();
    // down
    for (int l = 0; l + 1 < B.nlev; l++) {
        const int NX = B.lev[l].nx, NY = B.lev[l].ny, W = NX + 2;
        T *El = E + off[l], *Rl = R + off[l], *Xl = AX + off[l], *Yl = AY + off[l];
        for (int s = 0; s < MG_NU; s++)
            for (int c = 0; c < 2; c++) {
                bt_colour(El, Rl, Xl, Yl, NX, NY, c);
                __syncthreads();
            }
        const int NXc = B.lev[l + 1].nx, NYc = B.lev[l + 1].ny, Wc = NXc + 2;
        T *Rc = R + off[l + 1];
        for (int q = threadIdx.x; q < NXc * NYc; q += MG_BT) {
            const int Ic = q % NXc, Jc = q / NXc;
            T sum = T(0);
            for (int dy = 0; dy < 2; dy++)
                for (int dx = 0; dx < 2; dx++) {
                    const int I = 2 * Ic + dx, J = 2 * Jc + dy;
                    if (I >= NX || J >= NY) continue;
                    const T aw = Xl[I + W * (J + 1)], ae = Xl[I + 1 + W * (J + 1)];
                    const T as = Yl[I + 1 + W * J], an = Yl[I + 1 + W * (J + 1)];
                    const T a3 = -(aw + ae + as + an);
                    const int k = (I + 1) + W * (J + 1);
                    sum += Rl[k] - (as * El[k - W] + aw * El[k - 1] + a3 * El[k] + ae * El[k + 1] + an * El[k + W]);
                }
            Rc[(Ic + 1) + Wc * (Jc + 1)] = sum;
        }
        __syncthreads();
    }
    {
        const int l = B.nlev - 1;
        const int NX = B.lev[l].nx, NY = B.lev[l].ny;
        for (int s = 0; s < MG_NCOARSE; s++)
            for (int c = 0; c < 2; c++) {
                bt_colour(E + off[l], R + off[l], AX + off[l], AY + off[l], NX, NY, c);
                __syncthreads();
            }
    }
    // up
    for (int l = B.nlev - 2; l >= 0; l--) {
        const int NX = B.lev[l].nx, NY = B.lev[l].ny, W = NX + 2;
        const int NXc = B.lev[l + 1].nx, NYc = B.lev[l + 1].ny, Wc = NXc + 2;
        T *El = E + off[l];
        const T *Ec = E + off[l + 1];
        for (int q = threadIdx.x; q < NX * NY; q += MG_BT) {
            const int I = q % NX, J = q / NX;
            const int ic = (I >> 1) + 1, jc = (J >> 1) + 1;   // padded coarse position
            const int dx = (I & 1) ? 1 : -1, dy = (J & 1) ? 1 : -1;
            // coarse neighbours beyond the grid read the zero ring
            const T pc = (T(9) * Ec[ic + Wc * jc] + T(3) * Ec[ic + dx + Wc * jc] + T(3) * Ec[ic + Wc * (jc + dy)] +
                          Ec[ic + dx + Wc * (jc + dy)]) *
                         T(0.0625);
            El[(I + 1) + W * (J + 1)] += pc;
        }
        __syncthreads();
        for (int s = 0; s < MG_NU; s++)
            for (int c = 1; c >= 0; c--) {
                bt_colour(El, R + off[l], AX + off[l], AY + off[l], NX, NY, c);
                __syncthreads();
            }
    }
    const MgLevel &L = B.lev[0];
    T *e = (T *)L.c;
    const int W = L.nx + 2;
    for (int q = threadIdx.x; q < L.nx * L.ny; q += MG_BT) {
        const int I = q % L.nx, J = q / L.nx;
        e[(size_t)(I + 1) + (size_t)L.ld * (J + 1)] = E[(I + 1) + W * (J + 1)];
    }
}

// ------------------------------------------------------------ control
__global__ void k_mg_reset(Scalars *sc) {
    sc->mg_done = 0;
    sc->mg_nconv = -1;
    sc->mg_cycles = 0;
    sc->val[7] = 0.0;
}

// Before V-cycle `cycle` (1-based): stop when max |dk r| < tol over the level-0
// tiles (nconv = cycles applied), or when `maxcyc` cycles are applied
// (nconv = -1: not converged).  val[7] receives max |dk r|.
__global__ __launch_bounds__(256) void k_mg_check(const double *__restrict__ part, int ntiles, double dk, double tol,
                                                  int cycle, int maxcyc, Scalars *sc) {
    if (sc->mg_done) return;
    __shared__ double red[4];
    double m = 0.0;
    for (int k = threadIdx.x; k < ntiles; k += 256)
        if (!(part[k] <= m)) m = part[k];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        if (!(o <= m)) m = o;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    m = red[0];
    for (int w = 1; w < 4; w++)
        if (!(red[w] <= m)) m = red[w];
    const double div = m * dk;
    sc->val[7] = div;
    sc->mg_cycles = cycle - 1;
    if (div < tol) {
        sc->mg_done = 1;
        sc->mg_nconv = cycle - 1;
    } else if (cycle - 1 >= maxcyc) {
        sc->mg_done = 1;
        sc->mg_nconv = -1;
    }
}

// ------------------------------------------------------------ host side
int mg_tiles(int NX, int NY, int *ntx) {
    const int tx = (NX + MG_OT - 1) / MG_OT, ty = (NY + MG_OT - 1) / MG_OT;
    if (ntx) *ntx = tx;
    return tx * ty;
}
int mg_bottom_side() { return MG_OT; }
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

// One V-cycle of defect correction (cycle is 1-based), every kernel
// returning at entry once the solve has stopped.
template <typename T>
void launch_mg_cycle(const MgHierarchy &H, int nx, int ny, int ld, const double *b, double *p, const double *rau,
                     const double *rgv, const double *uni, double dk, double tol, int cycle, int maxcyc,
                     double *part, Scalars *sc, hipStream_t s, hipEvent_t *ev_down, hipEvent_t *ev_up) {
    int ntx = 0;
    int nt = mg_tiles(H.lev[0].nx, H.lev[0].ny, &ntx);
    if (ev_down) (void)hipEventRecord(ev_down[0], s);
    hipLaunchKernelGGL((k_mg_down<T, true>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT), 0, s, H.lev[0], H.lev[1], ntx,
                       nt, nx, ny, ld, p, b, rau, rgv, uni ? 1 : 0, uni ? uni[0] : 0.0, uni ? uni[1] : 0.0, part, sc);
    if (ev_down) (void)hipEventRecord(ev_down[1], s);
    hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(256), 0, s, part, nt, dk, tol, cycle, maxcyc, sc);
    for (int l = 1; l < H.ntiled; l++) {
        nt = mg_tiles(H.lev[l].nx, H.lev[l].ny, &ntx);
        hipLaunchKernelGGL((k_mg_down<T, false>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT), 0, s, H.lev[l],
                           H.lev[l + 1], ntx, nt, nx, ny, ld, p, b, rau, rgv, 0, 0.0, 0.0, part, sc);
    }
    MgBottom B{};
    B.nlev = H.nlev - H.ntiled;
    for (int l = 0; l < B.nlev; l++) B.lev[l] = H.lev[H.ntiled + l];
    hipLaunchKernelGGL(k_mg_bottom<T>, dim3(1), dim3(MG_BT), 0, s, B, sc);
    for (int l = H.ntiled - 1; l >= 0; l--) {
        nt = mg_tiles(H.lev[l].nx, H.lev[l].ny, &ntx);
        if (l > 0)
            hipLaunchKernelGGL((k_mg_up<T, false>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT), 0, s, H.lev[l],
                               H.lev[l + 1], ntx, nt, ld, p, sc);
        else {
            if (ev_up) (void)hipEventRecord(ev_up[0], s);
            hipLaunchKernelGGL((k_mg_up<T, true>), dim3(8 * ((nt + 7) / 8)), dim3(MG_NT), 0, s, H.lev[0], H.lev[1],
                               ntx, nt, ld, p, sc);
            if (ev_up) (void)hipEventRecord(ev_up[1], s);
        }
    }
}

void launch_mg_reset(Scalars *sc, hipStream_t s) { hipLaunchKernelGGL(k_mg_reset, dim3(1), dim3(1), 0, s, sc); }

template void launch_mg_coef<float>(const MgHierarchy &, int, int, int, const double *, const double *,
                                    const double *, hipStream_t);
template void launch_mg_coef<double>(const MgHierarchy &, int, int, int, const double *, const double *,
                                     const double *, hipStream_t);
template void launch_mg_cycle<float>(const MgHierarchy &, int, int, int, const double *, double *, const double *,
                                     const double *, const double *, double, double, int, int, double *, Scalars *,
                                     hipStream_t, hipEvent_t *, hipEvent_t *);
template void launch_mg_cycle<double>(const MgHierarchy &, int, int, int, const double *, double *, const double *,
                                      const double *, const double *, double, double, int, int, double *, Scalars *,
                                      hipStream_t, hipEvent_t *, hipEvent_t *);

}  // namespace wd2

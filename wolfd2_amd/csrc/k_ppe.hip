// k_ppe.hip -- the pressure Poisson relaxations of pressure.f (Ppe 30-249).
//
//   SorRB / SorRBP (pressure.f:457-656), Cartesian grid without blockages:
//     k_sor_tiled runs KT sweeps per launch on LDS tiles (temporal blocking,
//     below), bit for bit the reference's iterates.
//   SorRB / SorRBP otherwise: two colour kernels per sweep.  Points of one
//     colour only read points of the other, so every point's arithmetic is
//     the reference's regardless of order; the unrelaxed |s| is max-reduced
//     through per-workgroup partials that the next sweep's kernel folds.
//   Sor (pressure.f:384-450): lexicographic Gauss-Seidel.  A point needs its
//     west and south neighbours of this sweep and its east and north neighbours
//     of the last one, so 64x64 tiles run in wavefront order (each workgroup
//     takes the next tile in row-major order from an atomic ticket and waits for
//     its west and south tiles), and inside a tile the points run along
//     anti-diagonals (a barrier between diagonals).
//   Slor / SlorRB / SlorRBP (pressure.f:673-1138): each j-line is an (nx-1)-row
//     AltTridLU; the lines of one colour are solved as one batch of chains.
//
// Sweep control stays on the device: the stopping test of a sweep (m > 1 and
// dif < sortol, pressure.f:443/534, or m = msorit) is evaluated on the device
// and raises sor_done; every later kernel of the solve returns at entry, so
// the host only looks at the flag every few dozen sweeps.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"
#include "wd2_sor.hpp"

namespace wd2 {

constexpr int RB_T = 256;

__device__ inline void sor_finish(Scalars *sc, double dif, double sortol, int msorit) {
    const int m = sc->sor_m + 1;
    sc->sor_m = m;
    sc->sor_dif = dif;
    if (m > 1 && dif < sortol) {
        sc->sor_done = 1;
        sc->sor_nconv = m;
    } else if (m >= msorit) {
        sc->sor_done = 1;
    }
}

__device__ inline bool last_arrival(unsigned int *counter, unsigned int nblocks, int tid) {
    __shared__ bool amlast;
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        const unsigned int t = atomicAdd(counter, 1u);
        amlast = (t == nblocks - 1);
        if (amlast) *counter = 0;
    }
    __syncthreads();
    return amlast;
}

// One point update of Sor / SorRB / SorRBP (pressure.f:425-431); returns |s|.
template <bool COEF>
__device__ inline double sor_point(int ld, int i, int j, const double *__restrict__ rau,
                                   const double *__restrict__ rgv, const PpeCoef &cf, const double *__restrict__ b,
                                   double sorrel, double *p) {
    const size_t k = WD2_IX(i, j);
    double a1, a2, a3, a4, a5;
    if (COEF) {
        a1 = cf.a1[k]; a2 = cf.a2[k]; a3 = cf.a3[k]; a4 = cf.a4[k]; a5 = cf.a5[k];
    } else {
        a1 = rgv[WD2_IX(i, j - 1)];
        a2 = rau[WD2_IX(i - 1, j)];
        a4 = rau[k];
        a5 = rgv[k];
        a3 = -a4 - a2 - a5 - a1;
    }
    double sum = b[k] - a1 * p[WD2_IX(i, j - 1)] - a2 * p[WD2_IX(i - 1, j)] - a4 * p[WD2_IX(i + 1, j)] -
                 a5 * p[WD2_IX(i, j + 1)];
    const double pij = p[k];
    sum = sum / a3 - pij;
    p[k] = pij + sorrel * sum;
    return fabs(sum);
}

// Max of n published partial maxima (written by earlier kernels), folded by
// one workgroup; every thread gets the result.
__device__ inline double fold_partials(const double *__restrict__ part, int n, double *sh, int tid) {
    double r = 0.0;
    for (int k = tid; k < n; k += RB_T) r = dmaxf(r, part[k]);
    return block_max<RB_T>(r, sh, tid);
}

// Sweep m of SorRB/SorRBP, one colour.  Before doing any work every
// workgroup evaluates the stopping test of sweep m-1 from that sweep's
// published partial maxima (pressure.f:534: m-1 > 1 and dif < sortol); all
// workgroups of both colour kernels reach the same decision, so the solve
// stops after exactly the sweep the reference stops after.  No device-scope
// fence is needed: partials cross workgroups only through kernel boundaries.
template <int COLOR, bool COEF>
__global__ __launch_bounds__(RB_T) void k_sor_rb(int nx, int ny, int ld, const double *__restrict__ rau,
                                                 const double *__restrict__ rgv, PpeCoef cf,
                                                 const double *__restrict__ b, double sorrel, double sortol, int m,
                                                 double *p, const double *__restrict__ prev_part, int nprev,
                                                 double *__restrict__ cur_part, Scalars *sc) {
    if (sc->sor_done) return;
    __shared__ double sh[RB_T / 64];
    const int tid = threadIdx.x;
    if (m - 1 > 1) {
        const double d = fold_partials(prev_part, nprev, sh, tid);
        if (d < sortol) {
            if (COLOR == 0 && blockIdx.x == 0 && blockIdx.y == 0 && tid == 0) {
                sc->sor_done = 1;
                sc->sor_nconv = m - 1;
                sc->sor_dif = d;
            }
            return;
        }
    }
    double mx = 0.0;
    for (int j = 2 + blockIdx.y; j <= ny; j += gridDim.y) {
        const int i0 = COLOR == 0 ? 2 + (j % 2) : 2 + ((j + 1) % 2);
        const int i = i0 + 2 * (blockIdx.x * RB_T + tid);
        if (i <= nx) mx = dmaxf(mx, sor_point<COEF>(ld, i, j, rau, rgv, cf, b, sorrel, p));
    }
    mx = block_max<RB_T>(mx, sh, tid);
    const unsigned int nb = gridDim.x * gridDim.y, bid = blockIdx.x + gridDim.x * blockIdx.y;
    if (tid == 0) cur_part[COLOR * nb + bid] = mx;
}

// Stopping test of the last launched sweep m (one workgroup).
__global__ __launch_bounds__(RB_T) void k_sor_check(const double *__restrict__ part, int n, int m, double sortol,
                                                    int msorit, Scalars *sc) {
    if (sc->sor_done) return;
    __shared__ double sh[RB_T / 64];
    const double d = fold_partials(part, n, sh, threadIdx.x);
    if (threadIdx.x == 0) {
        sc->sor_dif = d;
        if (m > 1 && d < sortol) {
            sc->sor_done = 1;
            sc->sor_nconv = m;
        } else if (m >= msorit) {
            sc->sor_done = 1;
        }
    }
}

static dim3 rb_grid(int nx, int ny) {
    const int per_row = (nx / 2 + RB_T - 1) / RB_T;
    const int rows = min(ny - 1, max(1, 512 / per_row));
    return dim3(per_row, rows);
}

int sor_rb_partials(int nx, int ny) {
    const dim3 g = rb_grid(nx, ny);
    return 2 * g.x * g.y;
}

void launch_sor_rb_sweep(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                         double sorrel, double sortol, int sweep, double *p, double *partials, Scalars *sc,
                         hipStream_t s, hipEvent_t *ev) {
    const dim3 g = rb_grid(nx, ny);
    const int np = 2 * g.x * g.y;
    const double *prev = partials + (size_t)((sweep - 1) & 1) * np;
    double *cur = partials + (size_t)(sweep & 1) * np;
    PpeCoef cf = coef ? *coef : PpeCoef{nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int color = 0; color < 2; color++) {
        if (ev) (void)hipEventRecord(ev[2 * color], s);
        if (coef) {
            if (color == 0)
                hipLaunchKernelGGL((k_sor_rb<0, true>), g, dim3(RB_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b, sorrel,
                                   sortol, sweep, p, prev, np, cur, sc);
            else
                hipLaunchKernelGGL((k_sor_rb<1, true>), g, dim3(RB_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b, sorrel,
                                   sortol, sweep, p, prev, np, cur, sc);
        } else {
            if (color == 0)
                hipLaunchKernelGGL((k_sor_rb<0, false>), g, dim3(RB_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b,
                                   sorrel, sortol, sweep, p, prev, np, cur, sc);
            else
                hipLaunchKernelGGL((k_sor_rb<1, false>), g, dim3(RB_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b,
                                   sorrel, sortol, sweep, p, prev, np, cur, sc);
        }
        if (ev) (void)hipEventRecord(ev[2 * color + 1], s);
    }
}

void launch_sor_rb_check(int nx, int ny, int sweep, double sortol, int msorit, const double *partials,
                         Scalars *sc, hipStream_t s) {
    const int np = sor_rb_partials(nx, ny);
    hipLaunchKernelGGL(k_sor_check, dim3(1), dim3(RB_T), 0, s, partials + (size_t)(sweep & 1) * np, np, sweep,
                       sortol, msorit, sc);
}

// --------------------------------------------- temporally blocked SorRB
// Cartesian grids without blockages: the right-hand side and the five
// coefficients do not change during the solve, so a workgroup can run KT
// sweeps on an LDS tile.  The tile holds TT x TT owned points plus a halo of
// TH = 2*KT cells: a red-black sweep moves the dependence front by two cells
// (black points read red neighbours, red points read the new black ones), so
// after KT sweeps every owned point has seen exactly the values it sees in
// the global sweep order, and its arithmetic is the reference's.  Points of
// the outermost ring are never updated; points outside the array are never
// read by an updated point.
//
// Only p lives in LDS (64 x 64 doubles, stored split by colour).  Each of
// the 512 lanes owns four horizontal pairs of points, with their right-hand
// side and coefficients (b, a1 = rgv(i,j-1), a2 = rau(i-1,j), a4 = rau(i,j),
// a5 = rgv(i,j)) in registers; in each half-sweep a lane updates the point of
// that colour of each of its pairs.
//
// The kernel reads p_in and writes the owned points to p_out (tiles read
// their neighbours' points as halo, so it cannot work in place).  The
// stopping test of sweep m is max|s| < sortol over the whole grid
// (pressure.f:534), which holds exactly when it holds on every tile: each tile
// publishes a bit mask of its sweeps whose owned-point maximum is below
// sortol, and the test of a launch is the AND of the masks.  Before any work,
// every workgroup evaluates the previous launch's test; all reach the same
// decision, so a solve that stopped is never advanced further.
constexpr int KT = 4, TH = 2 * KT, TW = 64, TT = TW - 2 * TH;   // 48 owned
constexpr int TP_T = 512;
constexpr int TP_PAIRS = TW * TW / 2 / TP_T;   // 4
constexpr int TK = TW / 2;                     // pairs per tile row

// RCP: every diagonal a3 of the launch is a power of two (the uniform-grid
// value, and -1 at points that are not updated), so x / a3 == x * (1 / a3)
// bit for bit (both are the correctly rounded value of the same real, and
// 1 / a3 is exact): the sweeps multiply by the reciprocal instead of dividing.
template <bool RCP>
__global__ __launch_bounds__(TP_T) void k_sor_tiled(int nx, int ny, int ld, const double *__restrict__ rau,
                                                    const double *__restrict__ rgv, const double *__restrict__ b,
                                                    double sorrel, int nsweep, const double *__restrict__ p_in,
                                                    double *__restrict__ p_out, unsigned *__restrict__ masks,
                                                    const unsigned *__restrict__ prev_masks, int prev_nsweep,
                                                    int prev_m0, int prev_launch, double sortol, int msorit,
                                                    int ntx, int ntiles, Scalars *sc, int force, int uni,
                                                    double rau_c, double rgv_c) {
    if (!force && sc->sor_done) return;
    // p split by colour: PC[c][y*TK + k] is the point of colour c ((i+j) & 1)
    // in pair k of tile row y, so that the neighbours a colour reads are
    // unit-stride across the lanes of a wavefront
    __shared__ double PC[2][TW * TK];
    __shared__ unsigned sh[TP_T / 64];
    if (!force && prev_nsweep > 0) {
        int conv = 0;
        const int k = tiled_stop(tiled_mask(prev_masks, ntiles, sh), prev_nsweep, prev_m0, msorit, &conv);
        if (k > 0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                sc->sor_done = 1;
                sc->sor_nconv = conv ? prev_m0 + k : 0;
                sc->sor_m = prev_m0 + k;
                sc->redo_launch = prev_launch;
                sc->redo_sweeps = k;
            }
            return;
        }
    }
    // XCD-aware tile order: consecutive workgroups go to different XCDs, so
    // give each XCD a contiguous band of tiles (their halos then share an L2).
    const int per = (ntiles + 7) / 8;
    const int tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int TI = tile % ntx, TJ = tile / ntx;
    const int i0 = 2 + TI * TT - TH, j0 = 2 + TJ * TT - TH;   // tile origin
    const int t = threadIdx.x;
    // Per pair r and colour c, everything the sweeps need for the pair's point
    // of that colour (h = its side, 0 left or 1 right): b and a1, a2, a4, a5
    // (+0, +0, +0, +0, 1 where the point is not updated, so a3 = -1 there),
    // a3 = -a4 - a2 - a5 - a1 (the reference's expression, evaluated once),
    // the LDS offsets of its west and east neighbours (q - 1 + h and q + h)
    // and of its south and north ones, clamped to the point's own slot where
    // they would leave the tile (the outer ring is never updated: what such a
    // point reads is discarded), and whether the point is updated / owned.
    double cb[TP_PAIRS][2], c1[TP_PAIRS][2], c2[TP_PAIRS][2], c3[TP_PAIRS][2], c4[TP_PAIRS][2], c5[TP_PAIRS][2];
    int ow[TP_PAIRS][2], oe[TP_PAIRS][2], os[TP_PAIRS], on[TP_PAIRS];
    bool up[TP_PAIRS][2];
    unsigned long long ownm[TP_PAIRS][2];
    int lcs[TP_PAIRS];
#define PX(r) (2 * ((t + TP_T * (r)) % TK))
#define PY(r) ((t + TP_T * (r)) / TK)
#define PB(r) (t + TP_T * (r))   // y*TK + k of the lane's pair r
#pragma unroll
    for (int r = 0; r < TP_PAIRS; r++) {
        const int x = PX(r), y = PY(r), kk = x / 2;
        const int i = i0 + x, j = j0 + y;   // i is even: the pair (i, i+1) is 16-byte aligned
        const double2 pv = load_pair_sel(p_in, ld, nx, ny, i, j);
        const int lc = (i + j) & 1;   // colour of the left point
        lcs[r] = lc;
        PC[lc][PB(r)] = pv.x;
        PC[lc ^ 1][PB(r)] = pv.y;
        const bool rowin = j >= 2 && j <= ny && y >= 1 && y <= TW - 2;
        const bool upd0 = rowin && i >= 2 && i <= nx && x >= 1;
        const bool upd1 = rowin && i + 1 >= 2 && i + 1 <= nx && x + 1 <= TW - 2;
        // coefficients as aligned pairs, read from clamped addresses and used
        // only where a point updates (i >= 2 and i + 1 <= nx + 1 there)
        const double2 bv = load_pair_sel(b, ld, nx, ny, i, j);
        double2 g0, g1, av;
        double aw;
        if (uni) {   // every rau / rgv entry an updated point reads has this value
            g0 = g1 = make_double2(rgv_c, rgv_c);
            av = make_double2(rau_c, rau_c);
            aw = rau_c;
        } else {
            g0 = load_pair_sel(rgv, ld, nx, ny, i, j - 1);
            g1 = load_pair_sel(rgv, ld, nx, ny, i, j);
            av = load_pair_sel(rau, ld, nx, ny, i, j);
            aw = rau[(size_t)min(max(i - 1, 0), ld - 1) + (size_t)ld * (size_t)min(max(j, 0), ny + 1)];
        }
        os[r] = y > 0 ? -TK : 0;
        on[r] = y < TW - 1 ? TK : 0;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int h = lc ^ c;
            const bool u = h ? upd1 : upd0;
            const bool own = u && x + h >= TH && x + h < TH + TT && y >= TH && y < TH + TT;
            const double bb = u ? (h ? bv.y : bv.x) : 0.0;
            const double a1 = u ? (h ? g0.y : g0.x) : 0.0;
            const double a2 = u ? (h ? av.x : aw) : 0.0;
            const double a4 = u ? (h ? av.y : av.x) : 0.0;
            const double a5 = u ? (h ? g1.y : g1.x) : 1.0;
            cb[r][c] = bb; c1[r][c] = a1; c2[r][c] = a2; c4[r][c] = a4; c5[r][c] = a5;
            c3[r][c] = -a4 - a2 - a5 - a1;
            if (RCP) c3[r][c] = 1.0 / c3[r][c];
            ow[r][c] = h ? 0 : (kk > 0 ? -1 : 0);
            oe[r][c] = h ? (kk < TK - 1 ? 1 : 0) : 0;
            up[r][c] = u;
            ownm[r][c] = __ballot(own);
        }
    }
    __syncthreads();
    // The sweeps are straight-line code: every point is evaluated and a point
    // that is not updated keeps its value by a select.  The stopping test
    // flags an owned point whose |s| is not below sortol (a NaN counts as not
    // converged), which is max |s| < sortol over the owned points otherwise;
    // it is gathered per wavefront as ballots.
    unsigned below = 0;   // bit s: the wavefront's owned max |s| of sweep s is < sortol
    for (int s = 0; s < nsweep; s++) {
        unsigned long long big = 0;
#pragma unroll
        for (int color = 0; color < 2; color++) {
            double *__restrict__ dst = PC[color];
            const double *__restrict__ src = PC[1 - color];
#pragma unroll
            for (int r = 0; r < TP_PAIRS; r++) {
                const int q = PB(r);
                const double pij = dst[q];
                // the reference's order: b - a1 pS - a2 pW - a4 pE - a5 pN
                double sum = cb[r][color] - c1[r][color] * src[q + os[r]] - c2[r][color] * src[q + ow[r][color]] -
                             c4[r][color] * src[q + oe[r][color]] - c5[r][color] * src[q + on[r]];
                sum = (RCP ? sum * c3[r][color] : sum / c3[r][color]) - pij;
                const double pn = pij + sorrel * sum;
                dst[q] = up[r][color] ? pn : pij;
                big |= __ballot(!(fabs(sum) < sortol)) & ownm[r][color];
            }
            __syncthreads();
        }
        if (big == 0) below |= 1u << s;
    }
#pragma unroll
    for (int r = 0; r < TP_PAIRS; r++) {
#pragma unroll
        for (int c = 0; c < 2; c++)
            if ((ownm[r][c] >> (t & 63)) & 1ull) {
                const int x = PX(r) + (lcs[r] ^ c), y = PY(r);
                p_out[WD2_IX(i0 + x, j0 + y)] = PC[c][PB(r)];
            }
    }
#undef PX
#undef PY
#undef PB
    // tile mask = AND over the wavefronts (below is the same in every lane of one)
    if ((t & 63) == 0) sh[t >> 6] = below;
    __syncthreads();
    if (t == 0) {
        unsigned r = ~0u;
        for (int w = 0; w < TP_T / 64; w++) r &= sh[w];
        masks[tile] = r;
    }
}

// Stopping test of the last launch of a batch (one workgroup).
__global__ __launch_bounds__(RB_T) void k_sor_tiled_check(const unsigned *__restrict__ masks, int ntiles,
                                                          int nsweep, int m0, int launch, int msorit, Scalars *sc) {
    if (sc->sor_done) return;
    __shared__ unsigned sh[RB_T / 64];
    int conv = 0;
    const int k = tiled_stop(tiled_mask(masks, ntiles, sh), nsweep, m0, msorit, &conv);
    if (threadIdx.x == 0 && k > 0) {
        sc->sor_done = 1;
        sc->sor_nconv = conv ? m0 + k : 0;
        sc->sor_m = m0 + k;
        sc->redo_launch = launch;
        sc->redo_sweeps = k;
    }
}

int sor_tiled_tiles(int nx, int ny, int *ntx) {
    const int tx = (nx - 1 + TT - 1) / TT, ty = (ny - 1 + TT - 1) / TT;
    if (ntx) *ntx = tx;
    return tx * ty;
}
int sor_tiled_sweeps() { return KT; }

void launch_sor_tiled(int nx, int ny, int ld, const Metrics &m, const double *b, double sorrel, int nsweep,
                      const double *p_in, double *p_out, unsigned *masks, const unsigned *prev_masks,
                      int prev_nsweep, int prev_m0, int prev_launch, double sortol, int msorit, Scalars *sc,
                      int force, const double *uni, hipStream_t s) {
    int ntx = 0;
    const int nt = sor_tiled_tiles(nx, ny, &ntx);
    const int grid = 8 * ((nt + 7) / 8);
    // the diagonal of an updated point on the uniform path, in the kernel's
    // order of operations (-a4 - a2 - a5 - a1); non-updated points have -1
    bool rcp = false;
    if (uni) {
        const double a3 = -uni[0] - uni[0] - uni[1] - uni[1];
        int e2 = 0;
        rcp = std::isfinite(a3) && a3 != 0.0 && std::frexp(a3, &e2) == -0.5;
    }
    if (rcp)
        hipLaunchKernelGGL(k_sor_tiled<true>, dim3(grid), dim3(TP_T), 0, s, nx, ny, ld, m.rau, m.rgv, b, sorrel, nsweep,
                           p_in, p_out, masks, prev_masks, prev_nsweep, prev_m0, prev_launch, sortol, msorit, ntx, nt,
                           sc, force, 1, uni[0], uni[1]);
    else
        hipLaunchKernelGGL(k_sor_tiled<false>, dim3(grid), dim3(TP_T), 0, s, nx, ny, ld, m.rau, m.rgv, b, sorrel,
                           nsweep, p_in, p_out, masks, prev_masks, prev_nsweep, prev_m0, prev_launch, sortol, msorit,
                           ntx, nt, sc, force, uni ? 1 : 0, uni ? uni[0] : 0.0, uni ? uni[1] : 0.0);
}

void launch_sor_tiled_check(int ntiles, int nsweep, int m0, int launch, int msorit, const unsigned *masks,
                            Scalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_sor_tiled_check, dim3(1), dim3(RB_T), 0, s, masks, ntiles, nsweep, m0, launch, msorit, sc);
}

__global__ void k_sor_reset(Scalars *sc) {
    sc->sor_done = 0;
    sc->sor_nconv = 0;
    sc->sor_m = 0;
    sc->sor_dif = 0.0;
    sc->val[6] = 0.0;
    sc->redo_launch = 0;
    sc->redo_sweeps = 0;
}
void launch_sor_reset(Scalars *sc, hipStream_t s) { hipLaunchKernelGGL(k_sor_reset, dim3(1), dim3(1), 0, s, sc); }

// ------------------------------------------------------- lexicographic Sor
constexpr int LX_T = 64;          // one wavefront per tile
constexpr int LX_TI = 64, LX_TJ = 64;

template <bool COEF>
__global__ __launch_bounds__(LX_T) void k_sor_lex(int nx, int ny, int ld, const double *__restrict__ rau,
                                                  const double *__restrict__ rgv, PpeCoef cf,
                                                  const double *__restrict__ b, double sorrel, double sortol,
                                                  int msorit, double *p, int *tile_flags, int sweep,
                                                  double *__restrict__ partials, Scalars *sc) {
    if (sc->sor_done) return;
    __shared__ int ticket;
    __shared__ double sh[1];
    const int t = threadIdx.x;
    const int ntx = (nx - 1 + LX_TI - 1) / LX_TI, nty = (ny - 1 + LX_TJ - 1) / LX_TJ;
    const unsigned int ntiles = ntx * nty;
    if (t == 0) ticket = (int)atomicAdd(&sc->counter[5], 1u);
    __syncthreads();
    const int tile = ticket;
    const int TI = tile % ntx, TJ = tile / ntx;
    // wait for the west and south tiles of this sweep
    if (t == 0) {
        if (TI > 0)
            while (__hip_atomic_load(&tile_flags[tile - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < sweep)
                __builtin_amdgcn_s_sleep(1);
        if (TJ > 0)
            while (__hip_atomic_load(&tile_flags[tile - ntx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < sweep)
                __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
    const int i0 = 2 + TI * LX_TI, j0 = 2 + TJ * LX_TJ;
    const int ni = min(LX_TI, nx - i0 + 1), nj = min(LX_TJ, ny - j0 + 1);
    // column i0 + t, rows j0 .. j0+nj-1, walked along anti-diagonals
    double mx = 0.0;
    for (int step = 0; step < nj + LX_T - 1; step++) {
        const int jj = step - t;
        if (t < ni && jj >= 0 && jj < nj) {
            const int i = i0 + t, j = j0 + jj;
            const size_t k = WD2_IX(i, j);
            double a1, a2, a3, a4, a5;
            if (COEF) {
                a1 = cf.a1[k]; a2 = cf.a2[k]; a3 = cf.a3[k]; a4 = cf.a4[k]; a5 = cf.a5[k];
            } else {
                a1 = rgv[WD2_IX(i, j - 1)]; a2 = rau[WD2_IX(i - 1, j)]; a4 = rau[k]; a5 = rgv[k];
                a3 = -a4 - a2 - a5 - a1;
            }
            // neighbours written during this sweep by other lanes or tiles are
            // read through the coherent path
            const double pS = ld_coherent(&p[WD2_IX(i, j - 1)]);
            const double pW = ld_coherent(&p[WD2_IX(i - 1, j)]);
            const double pE = p[WD2_IX(i + 1, j)], pN = p[WD2_IX(i, j + 1)];
            double sum = b[k] - a1 * pS - a2 * pW - a4 * pE - a5 * pN;
            const double pij = p[k];
            sum = sum / a3 - pij;
            const double pnew = pij + sorrel * sum;
            __hip_atomic_store((unsigned long long *)&p[k], dbits(pnew), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            mx = dmaxf(mx, fabs(sum));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    for (int off = 32; off > 0; off >>= 1) mx = dmaxf(mx, __shfl_xor(mx, off));
    if (t == 0) {
        partials[tile] = mx;
        __hip_atomic_store(&tile_flags[tile], sweep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (last_arrival(&sc->counter[6], ntiles, t)) {
        double r = 0.0;
        for (unsigned int k = t; k < ntiles; k += LX_T) r = dmaxf(r, ld_coherent(&partials[k]));
        for (int off = 32; off > 0; off >>= 1) r = dmaxf(r, __shfl_xor(r, off));
        if (t == 0) {
            sc->counter[5] = 0;
            sor_finish(sc, r, sortol, msorit);
        }
    }
    (void)sh;
}

// Lexicographic Sor, whole solve in one workgroup (grids of up to LR_MAXROWS
// rows).  Each lane owns one grid row j = 2 + lane and walks it from i = 2 to
// nx; the lanes of a wavefront run skewed by one column per lane, so at every
// step lane L updates the point whose south neighbour lane L-1 updated in the
// previous step (passed by a lane shuffle) and whose west neighbour is the
// lane's own previous point: exactly the reference's i-then-j order.  East,
// north and the point's old value are read from the previous sweep's field.
// Wavefront w starts LR_LAG steps after wavefront w-1 and takes the south
// values of its first row from a ring that w-1's last lane fills; with a
// workgroup barrier every LR_K steps a ring value is always at least one
// barrier old when it is read.  The stopping test (pressure.f:443) is applied
// in the kernel after every sweep, so the whole solve is one launch.
// Non-Cartesian grids re-form the right-hand side before every sweep
// (pressure.f:RhsPpe, as in the reference).
constexpr int LR_K = 16;                // steps between two barriers
constexpr int LR_LAG = 64 + LR_K;       // step lag between consecutive wavefronts
constexpr int LR_RING = 64;             // ring slots (> LR_K + 1)
constexpr int LR_D = 4;                 // steps of operand prefetch (divides LR_K)
constexpr int LR_MAXW = 8;
constexpr int LR_MAXROWS = 64 * LR_MAXW;

template <bool COEF>
__global__ __launch_bounds__(64 * LR_MAXW) void k_sor_lex_rows(int nx, int ny, int ld, const double *__restrict__ rau,
                                                              const double *__restrict__ rgv, PpeCoef cf,
                                                              const double *__restrict__ rhs,
                                                              const double *__restrict__ rbu,
                                                              const double *__restrict__ rbv, double *beff, int cart,
                                                              double sorrel, double sortol, int msorit, double *p,
                                                              Scalars *sc) {
    __shared__ double ring[LR_MAXW][LR_RING];
    __shared__ double red[LR_MAXW];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int nw = blockDim.x >> 6;
    const int j = 2 + t;
    const bool rowok = j <= ny;
    const int nc = nx - 1;                       // columns i = 2 .. nx
    // steps of a sweep, rounded up to whole blocks (the extra steps find no point)
    const int nsteps = (nc + LR_LAG * (nw - 1) + 63 + LR_K - 1) / LR_K * LR_K;
    const double *bb = cart ? rhs : beff;
    for (int m = 1;; m++) {
        if (!cart) {   // pressure.f:RhsPpe from the field at the start of the sweep
#define P(a, b) p[WD2_IX(a, b)]
            for (int q = t; q < (nx - 1) * (ny - 1); q += blockDim.x) {
                const int i = 2 + q % (nx - 1), jj = 2 + q / (nx - 1);
                beff[WD2_IX(i, jj)] =
                    rhs[WD2_IX(i, jj)] -
                    (rbu[WD2_IX(i, jj)] * (P(i + 1, jj + 1) + P(i, jj + 1) - P(i + 1, jj - 1) - P(i, jj - 1)) -
                     rbu[WD2_IX(i - 1, jj)] * (P(i, jj + 1) + P(i - 1, jj + 1) - P(i, jj - 1) - P(i - 1, jj - 1)) +
                     rbv[WD2_IX(i, jj)] * (P(i + 1, jj + 1) + P(i + 1, jj) - P(i - 1, jj + 1) - P(i - 1, jj)) -
                     rbv[WD2_IX(i, jj - 1)] * (P(i + 1, jj) + P(i + 1, jj - 1) - P(i - 1, jj) - P(i - 1, jj - 1)));
            }
#undef P
            __syncthreads();
        }
        // operands of the point LR_D steps ahead are loaded while the current
        // step runs (slot d holds column c + d of this lane's row)
        // (the point's old value is the previous column's east value, and on
        // the unit-coefficient path its a2 the previous column's a4)
        double oE[LR_D], oN[LR_D], oB[LR_D], o1[LR_D], o4[LR_D], o5[LR_D];
        double o2[COEF ? LR_D : 1], o3[COEF ? LR_D : 1];
        auto fetch = [&](int cc, int d) {
            if (rowok && cc >= 0 && cc < nc) {
                const int i = 2 + cc;
                const size_t k = WD2_IX(i, j);
                oE[d] = p[WD2_IX(i + 1, j)];
                oN[d] = p[WD2_IX(i, j + 1)];
                oB[d] = bb[k];
                if (COEF) {
                    o1[d] = cf.a1[k]; o2[d] = cf.a2[k]; o3[d] = cf.a3[k]; o4[d] = cf.a4[k]; o5[d] = cf.a5[k];
                } else {
                    o1[d] = rgv[WD2_IX(i, j - 1)];
                    o4[d] = rau[k];
                    o5[d] = rgv[k];
                }
            }
        };
        const int off = w * LR_LAG + lane;   // the lane's point at step st is column st - off
#pragma unroll
        for (int d = 0; d < LR_D; d++) fetch(d - off, d);
        double mx = 0.0, res = 0.0, prevE = 0.0, prevA4 = 0.0;
        for (int s0 = 0; s0 < nsteps; s0 += LR_K) {
#pragma unroll
            for (int q = 0; q < LR_K; q++) {
                const int st = s0 + q;
                const int d = q % LR_D;                   // slot of this step (LR_K is a multiple of LR_D)
                const int c = st - off;
                const double up = __shfl_up(res, 1);      // lane-1's point of the previous step
                if (rowok && c >= 0 && c < nc) {
                    const int i = 2 + c;
                    double pS;
                    if (lane > 0) pS = up;
                    else if (w > 0) pS = ring[w - 1][c % LR_RING];
                    else pS = p[WD2_IX(i, j - 1)];            // ghost row j = 1
                    const double pW = c > 0 ? res : p[WD2_IX(i - 1, j)];   // ghost column i = 1
                    const double pij = c > 0 ? prevE : p[WD2_IX(i, j)];
                    const double a1 = o1[d], a4 = o4[d], a5 = o5[d];
                    const double a2 = COEF ? o2[COEF ? d : 0] : (c > 0 ? prevA4 : rau[WD2_IX(i - 1, j)]);
                    const double a3 = COEF ? o3[COEF ? d : 0] : -a4 - a2 - a5 - a1;
                    double sum = oB[d] - a1 * pS - a2 * pW - a4 * oE[d] - a5 * oN[d];
                    sum = sum / a3 - pij;
                    res = pij + sorrel * sum;
                    p[WD2_IX(i, j)] = res;
                    mx = dmaxf(mx, fabs(sum));
                    if (lane == 63) ring[w][c % LR_RING] = res;
                }
                prevE = oE[d];        // the next column's old value and a2
                prevA4 = o4[d];
                fetch(c + LR_D, d);   // the slot is free again: column c + LR_D
            }
            __syncthreads();
        }
        // dif of the sweep and the stopping test (pressure.f:434-449)
        for (int off = 32; off > 0; off >>= 1) mx = dmaxf(mx, __shfl_xor(mx, off));
        if (lane == 0) red[w] = mx;
        __syncthreads();
        double dif = 0.0;
        for (int q = 0; q < nw; q++) dif = dmaxf(dif, red[q]);
        __syncthreads();
        const bool conv = m > 1 && dif < sortol;
        if (conv || m >= msorit) {
            if (t == 0) {
                sc->sor_m = m;
                sc->sor_dif = dif;
                sc->sor_nconv = conv ? m : 0;
                sc->sor_done = 1;
            }
            return;
        }
    }
}

int sor_lex_rows_max() { return LR_MAXROWS; }

void launch_sor_lex_rows(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *rhs,
                         double *beff, int cart, double sorrel, double sortol, int msorit, double *p, Scalars *sc,
                         hipStream_t s) {
    const int nw = (ny - 1 + 63) / 64;
    const PpeCoef cf = coef ? *coef : PpeCoef{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (coef)
        hipLaunchKernelGGL(k_sor_lex_rows<true>, dim3(1), dim3(64 * nw), 0, s, nx, ny, ld, m.rau, m.rgv, cf, rhs,
                           m.rbu, m.rbv, beff, cart, sorrel, sortol, msorit, p, sc);
    else
        hipLaunchKernelGGL(k_sor_lex_rows<false>, dim3(1), dim3(64 * nw), 0, s, nx, ny, ld, m.rau, m.rgv, cf, rhs,
                           m.rbu, m.rbv, beff, cart, sorrel, sortol, msorit, p, sc);
}

void launch_sor_lex_sweep(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                          double sorrel, double sortol, int msorit, double *p, int *tile_flags, int sweep,
                          double *partials, Scalars *sc, hipStream_t s) {
    const int ntx = (nx - 1 + LX_TI - 1) / LX_TI, nty = (ny - 1 + LX_TJ - 1) / LX_TJ;
    PpeCoef cf = coef ? *coef : PpeCoef{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (coef)
        hipLaunchKernelGGL(k_sor_lex<true>, dim3(ntx * nty), dim3(LX_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b, sorrel,
                           sortol, msorit, p, tile_flags, sweep, partials, sc);
    else
        hipLaunchKernelGGL(k_sor_lex<false>, dim3(ntx * nty), dim3(LX_T), 0, s, nx, ny, ld, m.rau, m.rgv, cf, b,
                           sorrel, sortol, msorit, p, tile_flags, sweep, partials, sc);
}

// ------------------------------------------------------------- line SOR
// Chains of the lines j = first_line + line_step*q (q < nlines): row r of line q
// is i = 2 + r (r < nx-1), with (a, d, c) = (a2, a3, a4) and
// b = rhs - (a1*p(i,j-1) + a5*p(i,j+1))   (pressure.f:746-771)
template <bool COEF>
__global__ __launch_bounds__(256) void k_slor_assemble(int nx, int ny, int ld, const double *__restrict__ rau,
                                                       const double *__restrict__ rgv, PpeCoef cf,
                                                       const double *__restrict__ b, const double *__restrict__ p,
                                                       int first_line, int line_step, int nlines,
                                                       double *__restrict__ ca, double *__restrict__ cd,
                                                       double *__restrict__ cc, double *__restrict__ cb,
                                                       const Scalars *sc) {
    if (sc->sor_done) return;
    const int L = nx - 1;
    const long r = (long)blockIdx.x * 256 + threadIdx.x;
    if (r >= (long)L * nlines) return;
    const int q = (int)(r / L), i = 2 + (int)(r % L);
    const int j = first_line + line_step * q;
    const size_t k = WD2_IX(i, j);
    double a1, a2, a3, a4, a5;
    if (COEF) {
        a1 = cf.a1[k]; a2 = cf.a2[k]; a3 = cf.a3[k]; a4 = cf.a4[k]; a5 = cf.a5[k];
    } else {
        a1 = rgv[WD2_IX(i, j - 1)]; a2 = rau[WD2_IX(i - 1, j)]; a4 = rau[k]; a5 = rgv[k];
        a3 = -a4 - a2 - a5 - a1;
    }
    ca[r] = a2;
    cd[r] = a3;
    cc[r] = a4;
    cb[r] = b[k] - (a1 * p[WD2_IX(i, j - 1)] + a5 * p[WD2_IX(i, j + 1)]);
}

void launch_slor_assemble(int nx, int ny, int ld, const Metrics &m, const PpeCoef *coef, const double *b,
                          const double *p, int first_line, int line_step, int nlines, double *ca, double *cd,
                          double *cc, double *cb, const Scalars *sc, hipStream_t s) {
    const long n = (long)(nx - 1) * nlines;
    PpeCoef cf = coef ? *coef : PpeCoef{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (coef)
        hipLaunchKernelGGL(k_slor_assemble<true>, dim3((n + 255) / 256), dim3(256), 0, s, nx, ny, ld, m.rau, m.rgv, cf,
                           b, p, first_line, line_step, nlines, ca, cd, cc, cb, sc);
    else
        hipLaunchKernelGGL(k_slor_assemble<false>, dim3((n + 255) / 256), dim3(256), 0, s, nx, ny, ld, m.rau, m.rgv,
                           cf, b, p, first_line, line_step, nlines, ca, cd, cc, cb, sc);
}

// Relaxation of the solved lines: p = plold + sorrel*(x - plold).  dif_mode 0:
// max |x - plold| (Slor, SlorRB); 1: max |plold - pnew| (SlorRBP, pressure.f:1122).
__global__ __launch_bounds__(256) void k_slor_relax(int nx, int ny, int ld, int first_line, int line_step,
                                                    int nlines, const double *__restrict__ x, double sorrel,
                                                    int dif_mode, double *__restrict__ p, Scalars *sc, int slot) {
    if (sc->sor_done) return;
    __shared__ double sh[4];
    const int L = nx - 1;
    const long r = (long)blockIdx.x * 256 + threadIdx.x;
    double mx = 0.0;
    if (r < (long)L * nlines) {
        const int q = (int)(r / L), i = 2 + (int)(r % L);
        const int j = first_line + line_step * q;
        const size_t k = WD2_IX(i, j);
        const double plold = p[k];
        const double sum = x[r] - plold;
        const double pnew = plold + sorrel * sum;
        p[k] = pnew;
        mx = dif_mode == 0 ? fabs(sum) : fabs(plold - pnew);
    }
    mx = block_max<256>(mx, sh, threadIdx.x);
    if (threadIdx.x == 0) atomicMax((unsigned long long *)&sc->val[slot], dbits(mx));
}

void launch_slor_relax(int nx, int ny, int ld, int first_line, int line_step, int nlines, const double *x,
                       double sorrel, int dif_mode, double *p, const double *pn, double *partials, Scalars *sc,
                       int slot, hipStream_t s) {
    const long n = (long)(nx - 1) * nlines;
    (void)pn;
    (void)partials;
    hipLaunchKernelGGL(k_slor_relax, dim3((n + 255) / 256), dim3(256), 0, s, nx, ny, ld, first_line, line_step,
                       nlines, x, sorrel, dif_mode, p, sc, slot);
}

__global__ void k_slor_finish(double sortol, int msorit, Scalars *sc) {
    if (sc->sor_done) return;
    const double dif = sc->val[6];
    sc->val[6] = 0.0;
    sor_finish(sc, dif, sortol, msorit);
}

void launch_slor_finish(int nparts, double sortol, int msorit, double *partials, Scalars *sc, hipStream_t s) {
    (void)nparts;
    (void)partials;
    hipLaunchKernelGGL(k_slor_finish, dim3(1), dim3(1), 0, s, sortol, msorit, sc);
}

void launch_slor_rbp_relax(int, int, int, int, double, const double *, double *, hipStream_t) {}

}  // namespace wd2

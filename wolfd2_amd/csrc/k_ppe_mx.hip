// k_ppe_mx.hip -- mixed-precision red-black SOR for the pressure Poisson
// equation: FP64 residual and pressure accumulation, FP32 relaxation sweeps.
//
// Each launch is one step of defect correction on an LDS tile.  The tile
// loads p, the right-hand side b and the coefficients in FP64 and forms the
// residual r = b - A p at its points in FP64 (A is the 5-point operator of
// pressure.f:97-106: a1 = rgv(i,j-1), a2 = rau(i-1,j), a4 = rau(i,j),
// a5 = rgv(i,j), a3 = -a4-a2-a5-a1).  It then runs KM sweeps of the
// reference's red-black SOR (pressure.f:SorRB 459-540: black points, then
// red, s = (r - sum a e_nb)/a3 - e, e += sorrel*s) on the correction e, in
// FP32, from e = 0, and writes p + e in FP64 for its owned points.
//
// SOR is an affine iteration, so p + e after k sweeps is the FP64 SOR iterate
// after k sweeps from p, up to the FP32 rounding of e; the unrelaxed update s
// of the correction sweep is the reference's s.  The ghosts of p stay frozen
// (e = 0 there), as in the reference.  The temporal blocking is that of the
// FP64 tiled kernel: halo TH = 2*KM cells (a red-black sweep moves the
// dependence front by two cells), points of the outermost tile ring hold
// e = 0, which is the FP64 kernel's stale ring, and the owned points see the
// global sweep order exactly.
//
// The stopping test (max|s| < sortol over the grid, m > 1, at most msorit
// sweeps) uses the per-tile bit masks of wd2_sor.hpp, evaluated by the next
// launch, so a solve that has stopped is never advanced.
//
// Roofline (HBM): per launch the tile reads p, b, rau, rgv (FP64) and writes
// p once; the KM sweeps run out of LDS and registers.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"
#include "wd2_sor.hpp"

namespace wd2 {

constexpr int KM = 8, MH = 2 * KM, MW = 96, MT = MW - 2 * MH;   // 64 owned per side
constexpr int MX_T = 768;                                         // 12 wavefronts
constexpr int MK = MW / 2;                                        // pairs per tile row
constexpr int PS = MK + 1, PR = 2 * PS;                           // LDS row of p: even, then odd columns
constexpr int SR = 6;                                             // rows of a lane's strip
constexpr int NB = MW / SR;                                       // strips per pair column
// Row stride of the colour-split correction: 6 strips of rows apart (the
// lanes k = 32..47 of one strip and k = 0..15 of the next share a half
// wavefront) the two halves fall on disjoint LDS banks: 6 * ES = 16 mod 32.
constexpr int ES = 56;
static_assert(MK * NB == MX_T, "one lane per (pair column, strip)");
static_assert(NB * SR == MW, "strips cover the tile");
static_assert((SR * ES) % 32 == 16 && ES >= MK, "conflict-free strip rows");

// Lane t owns pair column k = t % MK (points x = 2k, 2k+1) on the strip of
// rows y = SR*(t / MK) .. +SR-1.  Tile origins are even, so the colour
// (i + j) & 1 of the left point of row y is y & 1 = r & 1 (r the row within
// the strip): which point of a pair a colour updates is known at compile time.
// A point's four neighbours are of the other colour: its pair partner and,
// inside the strip, its north and south neighbours are in the lane's own
// registers; the other W/E neighbour (lane k -+ 1) and the N/S neighbours at
// the strip ends are read from LDS.  The FP64 residual reads p the same way.
//
// UNI: every rau and rgv entry the updated points read (rau on i = 1..nx,
// j = 2..ny; rgv on i = 2..nx, j = 1..ny) equals rau_c / rgv_c (a uniform
// Cartesian grid): the two values replace the loads (the same values, so the
// same arithmetic), and the sweep coefficients are the same at every point.
template <bool UNI>
__global__ __launch_bounds__(MX_T) void k_sor_tiled_mx(int nx, int ny, int ld, const double *__restrict__ rau,
                                                       const double *__restrict__ rgv, const double *__restrict__ b,
                                                       double sorrel, int nsweep,
                                                       const double *__restrict__ p_in, double *__restrict__ p_out,
                                                       unsigned *__restrict__ masks,
                                                       const unsigned *__restrict__ prev_masks, int prev_nsweep,
                                                       int prev_m0, int prev_launch, double sortol, int msorit,
                                                       int ntx, int ntiles, Scalars *sc, int force, double rau_c,
                                                       double rgv_c) {
    if (!force && sc->sor_done) return;
    // p by column parity within a row: column 2k + h of tile row y is
    // P[y*PR + h*PS + k], so the lanes of a row access consecutive doubles and
    // a lane's pair is one two-address access
    __shared__ double P[MW * PR];
    // the correction e split by colour: EC[c][y*ES + k] is the point of colour
    // c ((i+j) & 1) in pair k of tile row y
    __shared__ float EC[2][MW * ES];
    __shared__ unsigned sh[MX_T / 64];
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
    // XCD-aware tile order: consecutive workgroups go to different XCDs; each
    // XCD gets a contiguous band of tiles, whose halos then share its L2.
    const int per = (ntiles + 7) / 8;
    const int tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (tile >= ntiles) return;
    const int TI = tile % ntx, TJ = tile / ntx;
    const int i0 = 2 + TI * MT - MH, j0 = 2 + TJ * MT - MH;   // tile origin (even: i0 + x is 16-byte aligned)
    const int t = threadIdx.x;
    const int k = t % MK, y0 = SR * (t / MK), x = 2 * k;
    const int i = i0 + x;

    // which points of the strip are updated / owned (bit 2r + h)
    unsigned upd = 0, own = 0;
#pragma unroll
    for (int r = 0; r < SR; r++) {
        const int y = y0 + r, j = j0 + y;
        const bool rowin = j >= 2 && j <= ny && y >= 1 && y <= MW - 2;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const bool u = rowin && i + h >= 2 && i + h <= nx && x + h >= 1 && x + h <= MW - 2;
            const bool o = u && x + h >= MH && x + h < MH + MT && y >= MH && y < MH + MT;
            upd |= (u ? 1u : 0u) << (2 * r + h);
            own |= (o ? 1u : 0u) << (2 * r + h);
        }
    }

    // All loads are issued before the barrier, so their latencies overlap:
    // p and b of the strip (p into registers and the LDS tile; 0 outside the
    // array; b is read wherever it is, only updated points use it), unless
    // UNI the rau of the rows with points to update, and the rgv pairs of the
    // strip's rows and the row below (a row's a1 is the row below's a5).
    double2 pr[SR], bvr[SR], rav[SR], g[SR + 1];
    double ram[SR];
#pragma unroll
    for (int r = 0; r < SR; r++) {
        const int j = j0 + y0 + r;
        pr[r] = load_pair_sel(p_in, ld, nx, ny, i, j);
        bvr[r] = load_pair_sel(b, ld, nx, ny, i, j);
        ram[r] = rau_c;
        rav[r] = make_double2(rau_c, rau_c);
        if (!UNI && ((upd >> (2 * r)) & 3)) {
            // i >= 1 and i + 1 <= nx + 1 here: the pairs are inside the array
            ram[r] = rau[WD2_IX(i - 1, j)];
            rav[r] = *reinterpret_cast<const double2 *>(rau + WD2_IX(i, j));
        }
    }
#pragma unroll
    for (int q = 0; q <= SR; q++)
        g[q] = UNI ? make_double2(rgv_c, rgv_c) : load_pair(rgv, ld, nx, ny, i, j0 + y0 + q - 1);
#pragma unroll
    for (int r = 0; r < SR; r++) {
        const int c = (y0 + r) * PR + k;
        P[c] = pr[r].x;
        P[c + PS] = pr[r].y;
        EC[0][(y0 + r) * ES + k] = 0.f;
        EC[1][(y0 + r) * ES + k] = 0.f;
    }
    __syncthreads();
    // p at the strip's ends (the rows below and above), from LDS
    double2 pS = make_double2(0.0, 0.0), pN = make_double2(0.0, 0.0);
    if (y0 > 0) pS = make_double2(P[(y0 - 1) * PR + k], P[(y0 - 1) * PR + PS + k]);
    if (y0 + SR < MW) pN = make_double2(P[(y0 + SR) * PR + k], P[(y0 + SR) * PR + PS + k]);
    // the west neighbour of column x and the east neighbour of column x + 1,
    // kept inside the tile at its outer ring (those points are not updated)
    const int kw = k > 0 ? k - 1 : k, ke = k < MK - 1 ? k + 1 : k;

    // FP64 residual and the FP32 sweep coefficients, all scaled by 1/a3 in
    // FP32: the update is then four fused multiply-adds,
    // s = r/a3 - (a1/a3) e_S - (a2/a3) e_W - (a4/a3) e_E - (a5/a3) e_N - e.
    // These FP32 sweeps only produce the correction e; the iteration
    // converges to the fixed point of the FP64 residual.
    float cr[SR][2];
    float n1[UNI ? 1 : SR][2], n2[UNI ? 1 : SR][2], n4[UNI ? 1 : SR][2], n5[UNI ? 1 : SR][2];
    // UNI: every updated point has the same diagonal a3 (the same expression
    // of the same two values) and so the same FP32 reciprocal
    const double a3u = -rau_c - rau_c - rgv_c - rgv_c;
    const float i3u = 1.0f / (float)a3u;
    if (UNI) {
        n1[0][0] = n1[0][1] = n5[0][0] = n5[0][1] = -(float)rgv_c * i3u;
        n2[0][0] = n2[0][1] = n4[0][0] = n4[0][1] = -(float)rau_c * i3u;
    }
#pragma unroll
    for (int r = 0; r < SR; r++) {
        const int c0 = (y0 + r) * PR;
        const double ra[3] = {ram[r], rav[r].x, rav[r].y};
        const double2 ps = r > 0 ? pr[r - 1] : pS, pn = r < SR - 1 ? pr[r + 1] : pN;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            float rr = 0.f;
            if (UNI) {
                // straight-line: every point is evaluated and a point that is
                // not updated gets r = 0 by a select (what the clamped west and
                // east indices read at the ring is discarded)
                const double pw = h ? pr[r].x : P[c0 + PS + kw];
                const double pe = h ? P[c0 + ke] : pr[r].y;
                double res = fma(-rgv_c, h ? ps.y : ps.x, h ? bvr[r].y : bvr[r].x);
                res = fma(-rau_c, pw, res);
                res = fma(-a3u, h ? pr[r].y : pr[r].x, res);
                res = fma(-rau_c, pe, res);
                res = fma(-rgv_c, h ? pn.y : pn.x, res);
                rr = ((upd >> (2 * r + h)) & 1) ? (float)res * i3u : 0.f;
            } else if ((upd >> (2 * r + h)) & 1) {
                const double a1 = h ? g[r].y : g[r].x, a5 = h ? g[r + 1].y : g[r + 1].x;
                const double a2 = ra[h], a4 = ra[h + 1];
                const double a3 = UNI ? a3u : -a4 - a2 - a5 - a1;
                // an updated point has x >= 2 (h = 0) or x + 1 <= MW - 2 (h = 1)
                const double pw = h ? pr[r].x : P[c0 + PS + k - 1];
                const double pe = h ? P[c0 + k + 1] : pr[r].y;
                // the residual as one FMA chain (the defect only needs to be
                // accurate, not the reference's rounding)
                double res = fma(-a1, h ? ps.y : ps.x, h ? bvr[r].y : bvr[r].x);
                res = fma(-a2, pw, res);
                res = fma(-a3, h ? pr[r].y : pr[r].x, res);
                res = fma(-a4, pe, res);
                res = fma(-a5, h ? pn.y : pn.x, res);
                const float i3 = UNI ? i3u : 1.0f / (float)a3;
                rr = (float)res * i3;
                if (!UNI) {
                    n1[UNI ? 0 : r][h] = -(float)a1 * i3;
                    n2[UNI ? 0 : r][h] = -(float)a2 * i3;
                    n4[UNI ? 0 : r][h] = -(float)a4 * i3;
                    n5[UNI ? 0 : r][h] = -(float)a5 * i3;
                }
            } else if (!UNI) {
                n1[UNI ? 0 : r][h] = n2[UNI ? 0 : r][h] = n4[UNI ? 0 : r][h] = n5[UNI ? 0 : r][h] = 0.f;
            }
            cr[r][h] = rr;
        }
    }

    float e[SR][2];
#pragma unroll
    for (int r = 0; r < SR; r++) e[r][0] = e[r][1] = 0.f;
    const float w = (float)sorrel;
    const float tol = (float)sortol;
    // The sweeps are straight-line code: every point of the strip is
    // evaluated with its own relaxation factor wm, sorrel where the point is
    // updated and 0 elsewhere, so a point that is not updated keeps e = +0
    // (fmaf(0, s, +0) is +0 for finite s).  The neighbour indices that would
    // leave the tile at its outer ring are clamped to the lane's own slot, so
    // every read is of a written (finite) slot.  The stopping test flags an
    // owned point whose |s| is not below sortol (a NaN counts as not
    // converged, as in the FP64 kernel), which is max |s| >= sortol over the
    // owned points otherwise; it is gathered per wavefront in scalar lane
    // masks.
    float wm[SR][2];
    unsigned long long owm[SR][2];
#pragma unroll
    for (int r = 0; r < SR; r++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            wm[r][h] = ((upd >> (2 * r + h)) & 1) ? w : 0.f;
            owm[r][h] = __ballot((own >> (2 * r + h)) & 1);
        }
    const int qs0 = y0 > 0 ? -ES : 0, qn0 = y0 + SR < MW ? ES : 0;
    const int qw0 = k > 0 ? -1 : 0, qe0 = k < MK - 1 ? 1 : 0;
    unsigned below = 0;   // bit s: the wavefront's owned max |s| of sweep s is < sortol
    for (int s = 0; s < nsweep; s++) {
        unsigned long long big = 0;   // lanes of the wavefront with such a point
#pragma unroll
        for (int color = 0; color < 2; color++) {
            float *__restrict__ dst = EC[color];
            const float *__restrict__ src = EC[1 - color];
#pragma unroll
            for (int r = 0; r < SR; r++) {
                const int h = (r & 1) ^ color;   // the row's point of this colour
                const int rc = UNI ? 0 : r;
                const int q = (y0 + r) * ES + k;
                const float eS = r > 0 ? e[r - 1][h] : src[q + qs0];
                const float eN = r < SR - 1 ? e[r + 1][h] : src[q + qn0];
                const float eW = h ? e[r][0] : src[q + qw0];
                const float eE = h ? src[q + qe0] : e[r][1];
                float acc = fmaf(n1[rc][h], eS, cr[r][h]);
                acc = fmaf(n2[rc][h], eW, acc);
                acc = fmaf(n4[rc][h], eE, acc);
                acc = fmaf(n5[rc][h], eN, acc);
                const float sum = acc - e[r][h];
                e[r][h] = fmaf(wm[r][h], sum, e[r][h]);
                dst[q] = e[r][h];
                big |= __ballot(!(fabsf(sum) < tol)) & owm[r][h];
            }
            __syncthreads();
        }
        if (big == 0) below |= 1u << s;
    }

    // p + e on the owned points
#pragma unroll
    for (int r = 0; r < SR; r++)
#pragma unroll
        for (int h = 0; h < 2; h++)
            if ((own >> (2 * r + h)) & 1)
                p_out[WD2_IX(i + h, j0 + y0 + r)] = P[(y0 + r) * PR + h * PS + k] + (double)e[r][h];
    if ((t & 63) == 0) sh[t >> 6] = below;   // the same in every lane of the wavefront
    __syncthreads();
    if (t == 0) {
        unsigned r = ~0u;
        for (int q = 0; q < MX_T / 64; q++) r &= sh[q];
        masks[tile] = r;
    }
}

int sor_tiled_mx_tiles(int nx, int ny, int *ntx) {
    const int tx = (nx - 1 + MT - 1) / MT, ty = (ny - 1 + MT - 1) / MT;
    if (ntx) *ntx = tx;
    return tx * ty;
}
int sor_tiled_mx_sweeps() { return KM; }

void launch_sor_tiled_mx(int nx, int ny, int ld, const Metrics &m, const double *b, double sorrel, int nsweep,
                         const double *p_in, double *p_out, unsigned *masks, const unsigned *prev_masks,
                         int prev_nsweep, int prev_m0, int prev_launch, double sortol, int msorit, Scalars *sc,
                         int force, const double *uni, hipStream_t s) {
    int ntx = 0;
    const int nt = sor_tiled_mx_tiles(nx, ny, &ntx);
    const int grid = 8 * ((nt + 7) / 8);
    if (uni)
        hipLaunchKernelGGL(k_sor_tiled_mx<true>, dim3(grid), dim3(MX_T), 0, s, nx, ny, ld, m.rau, m.rgv, b, sorrel,
                           nsweep, p_in, p_out, masks, prev_masks, prev_nsweep, prev_m0, prev_launch, sortol, msorit,
                           ntx, nt, sc, force, uni[0], uni[1]);
    else
        hipLaunchKernelGGL(k_sor_tiled_mx<false>, dim3(grid), dim3(MX_T), 0, s, nx, ny, ld, m.rau, m.rgv, b, sorrel,
                           nsweep, p_in, p_out, masks, prev_masks, prev_nsweep, prev_m0, prev_launch, sortol, msorit,
                           ntx, nt, sc, force, 0.0, 0.0);
}

}  // namespace wd2

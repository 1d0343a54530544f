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
constexpr int MX_PAIRS = MW * MW / 2 / MX_T;                      // 6 pairs per lane
constexpr int MLD = MW + 1;                                       // padded LDS row of p
constexpr int MK = MW / 2;                                        // pairs per tile row
static_assert(MW * MW / 2 == MX_PAIRS * MX_T, "tile pairs must divide evenly");

__global__ __launch_bounds__(MX_T) void k_sor_tiled_mx(int nx, int ny, int ld, const double *__restrict__ rau,
                                                       const double *__restrict__ rgv, const double *__restrict__ b,
                                                       double sorrel, int nsweep,
                                                       const double *__restrict__ p_in, double *__restrict__ p_out,
                                                       unsigned *__restrict__ masks,
                                                       const unsigned *__restrict__ prev_masks, int prev_nsweep,
                                                       int prev_m0, int prev_launch, double sortol, int msorit,
                                                       int ntx, int ntiles, Scalars *sc, int force) {
    if (!force && sc->sor_done) return;
    __shared__ double P[MW * MLD];
    // the correction e split by colour: EC[c][y*MK + k] is the point of colour
    // c ((i+j) & 1) in pair k of tile row y, so that the neighbours a colour
    // reads are unit-stride across the lanes of a wavefront
    __shared__ float EC[2][MW * MK];
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
#define QX(r) (2 * ((t + MX_T * (r)) % MK))
#define QY(r) ((t + MX_T * (r)) / MK)
#define QB(r) (t + MX_T * (r))   // y*MK + k of the lane's pair r

    // p of the tile into LDS (FP64; 0 outside the array) as aligned pairs, e = 0
#pragma unroll
    for (int r = 0; r < MX_PAIRS; r++) {
        const int x = QX(r), y = QY(r);
        const int c = y * MLD + x;
        const double2 pv = load_pair(p_in, ld, nx, ny, i0 + x, j0 + y);
        P[c] = pv.x;
        P[c + 1] = pv.y;
        EC[0][QB(r)] = 0.f;
        EC[1][QB(r)] = 0.f;
    }
    __syncthreads();

    // FP64 residual and FP32 coefficients of the lane's points
    float cr[MX_PAIRS][2], c1[MX_PAIRS][2], c5[MX_PAIRS][2], ci3[MX_PAIRS][2];
    float cw[MX_PAIRS], cc[MX_PAIRS], ce[MX_PAIRS];   // rau at i-1, i, i+1 of the pair
    int flags[MX_PAIRS];
#pragma unroll
    for (int r = 0; r < MX_PAIRS; r++) {
        const int x = QX(r), y = QY(r);
        const int i = i0 + x, j = j0 + y;
        const int c0 = y * MLD + x;
        int f = 0;
        const bool rowin = j >= 2 && j <= ny && y >= 1 && y <= MW - 2;
        const bool upd0 = rowin && i >= 2 && i <= nx && x >= 1;
        const bool upd1 = rowin && i + 1 >= 2 && i + 1 <= nx && x + 1 <= MW - 2;
        double ra[3] = {0.0, 0.0, 0.0};
        double2 bv = make_double2(0.0, 0.0), g0 = bv, g1 = bv;
        if (upd0 || upd1) {
            // i >= 1 and i + 1 <= nx + 1 here: the pairs are inside the array
            ra[0] = rau[WD2_IX(i - 1, j)];
            const double2 av = *reinterpret_cast<const double2 *>(rau + WD2_IX(i, j));
            ra[1] = av.x;
            ra[2] = av.y;
            bv = *reinterpret_cast<const double2 *>(b + WD2_IX(i, j));
            g0 = *reinterpret_cast<const double2 *>(rgv + WD2_IX(i, j - 1));
            g1 = *reinterpret_cast<const double2 *>(rgv + WD2_IX(i, j));
        }
        cw[r] = (float)ra[0];
        cc[r] = (float)ra[1];
        ce[r] = (float)ra[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const bool upd = h ? upd1 : upd0;
            const bool own = upd && x + h >= MH && x + h < MH + MT && y >= MH && y < MH + MT;
            float rr = 0.f, a1f = 0.f, a5f = 0.f, i3f = 0.f;
            if (upd) {
                const double a1 = h ? g0.y : g0.x, a5 = h ? g1.y : g1.x;
                const double a2 = ra[h], a4 = ra[h + 1];
                const double a3 = -a4 - a2 - a5 - a1;
                const int c = c0 + h;
                const double res = (h ? bv.y : bv.x) - a1 * P[c - MLD] - a2 * P[c - 1] - a3 * P[c] -
                                   a4 * P[c + 1] - a5 * P[c + MLD];
                rr = (float)res;
                a1f = (float)a1;
                a5f = (float)a5;
                i3f = 1.0f / (float)a3;   // only scales the FP32 correction sweeps
            }
            cr[r][h] = rr; c1[r][h] = a1f; c5[r][h] = a5f; ci3[r][h] = i3f;
            f |= (upd ? 1 : 0) << h;
            f |= (own ? 1 : 0) << (2 + h);
        }
        f |= ((i + j) & 1) << 4;   // black iff (i + j) even
        flags[r] = f;
    }

    const float w = (float)sorrel;
    const float tol = (float)sortol;
    unsigned below = 0;   // bit s: this lane's owned max |s| of sweep s is < sortol
    for (int s = 0; s < nsweep; s++) {
        float mx = 0.f;
#pragma unroll
        for (int color = 0; color < 2; color++) {
            float *__restrict__ dst = EC[color];
            const float *__restrict__ src = EC[1 - color];
#pragma unroll
            for (int r = 0; r < MX_PAIRS; r++) {
                // h: the pair's point of this colour (0 left, 1 right); its west
                // neighbour is the other colour in pair k-1+h, its east one in k+h
                const int h = ((flags[r] >> 4) & 1) ^ color;
                if ((flags[r] >> h) & 1) {
                    const int q = QB(r);
                    // register selects (a runtime index into the arrays would go to scratch)
                    const float a1 = h ? c1[r][1] : c1[r][0], a5 = h ? c5[r][1] : c5[r][0];
                    const float a2 = h ? cc[r] : cw[r], a4 = h ? ce[r] : cc[r];
                    const float rr = h ? cr[r][1] : cr[r][0], i3 = h ? ci3[r][1] : ci3[r][0];
                    const float e = dst[q];
                    const float sum =
                        (rr - a1 * src[q - MK] - a2 * src[q - 1 + h] - a4 * src[q + h] - a5 * src[q + MK]) * i3 - e;
                    dst[q] = e + w * sum;
                    if ((flags[r] >> (2 + h)) & 1) mx = fmaxf(mx, fabsf(sum));
                }
            }
            __syncthreads();
        }
        if (mx < tol) below |= 1u << s;
    }

    // p + e on the owned points
#pragma unroll
    for (int r = 0; r < MX_PAIRS; r++) {
        const int x = QX(r), y = QY(r);
        const int lc = (flags[r] >> 4) & 1;   // colour of the pair's left point
#pragma unroll
        for (int h = 0; h < 2; h++)
            if ((flags[r] >> (2 + h)) & 1) {
                const int c = y * MLD + x + h;
                p_out[WD2_IX(i0 + x + h, j0 + y)] = P[c] + (double)EC[lc ^ h][QB(r)];
            }
    }
#undef QX
#undef QY
#undef QB
    for (int off = 32; off > 0; off >>= 1) below &= (unsigned)__shfl_xor((int)below, off);
    if ((t & 63) == 0) sh[t >> 6] = below;
    __syncthreads();
    if (t == 0) {
        unsigned r = ~0u;
        for (int k = 0; k < MX_T / 64; k++) r &= sh[k];
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
                         int force, hipStream_t s) {
    int ntx = 0;
    const int nt = sor_tiled_mx_tiles(nx, ny, &ntx);
    const int grid = 8 * ((nt + 7) / 8);
    hipLaunchKernelGGL(k_sor_tiled_mx, dim3(grid), dim3(MX_T), 0, s, nx, ny, ld, m.rau, m.rgv, b, sorrel, nsweep,
                       p_in, p_out, masks, prev_masks, prev_nsweep, prev_m0, prev_launch, sortol, msorit, ntx, nt,
                       sc, force);
}

}  // namespace wd2

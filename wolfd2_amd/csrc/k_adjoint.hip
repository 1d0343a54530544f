// k_adjoint.hip -- adjoint (transposed) kernels of the time step, FP64.
//
// The forward map differentiated is the step of main.f:690-972 on the
// multigrid path with a fixed number of q-l iterations (no stopping test)
// and a converged Poisson solve, on a one-region domain with no-slip walls
// (boundary profiles allowed), Cartesian metrics.  Every map below is the
// transpose of one forward piece; products that are bilinear in the fields
// (the convective fluxes, the upwinded chain coefficients) are transposed at
// the forward state, which the caller recomputes from checkpoints.
//
// Scatter-form transposes accumulate with FP64 atomics (the adjoint of a
// gather is a scatter); the sums are therefore not bit-reproducible, which
// does not matter to a gradient.
//
// Pieces and the forward lines they transpose:
// * chain solves (AltTridLU, momentum.f:1307-1339, with the first-row pivot
//   quirk H1: row 1's super-diagonal is divided by row 2's diagonal, i.e. the
//   matrix solved has S(1,2) = c1 d1 / d2): the transposed system is built
//   here and solved by the same parallel solver, whose own first-row quirk is
//   compensated; the coefficient adjoints are -mu_k x_{k-1}, -mu_k x_k,
//   -mu_k x_{k+1}, with the quirk's chain rule on rows 1 and 2;
// * the upwinded assembly of XMomentum (momentum.f:350-384, 396-428) and
//   YMomentum (675-754), the ConvCoef cases 1, 2, 4, 5 (864-981), DConvU/V
//   and DDiffU/V (987-1109);
// * VelBoundCond and PresBoundCond for no-slip walls (bound_cond.f:561-840,
//   853-1024), replayed in reverse order of assignment;
// * Project (utility.f:314-321, 384-391) and Divergence case 1
//   (pressure.f:287-299); the Poisson solve is self-adjoint on Cartesian grids
//   (a1 = rgv(i,j-1) is row (i,j-1)'s a5, a2 = rau(i-1,j) row (i-1,j)'s a4),
//   so its transpose is a solve with the same operator, and the frozen ghosts
//   receive minus their coupling coefficient times the solution.
#include <algorithm>

#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int AT = 256;

__device__ inline void aadd(double *p, double v) {
    if (v != 0.0) atomicAdd(p, v);
}

// ------------------------------------------------------------ chains
__device__ inline long apos(const Chains &cs, long k, long &L) {
    if (k < cs.n1) {
        L = cs.L1;
        return k % cs.L1;
    }
    L = cs.L2;
    return (k - cs.n1) % cs.L2;
}

// The transposed system of the one the chain solver solved for (A, D, C),
// in a form the same solver solves exactly: row k of S^T has sub-diagonal
// S(k-1,k) = C[k-1] (c1 d1 / d2 on a chain's first row), diagonal D[k] and
// super-diagonal S(k+1,k) = A[k+1]; on a chain's first row the super-diagonal
// is given as A[k+1] D[k+1] / D[k], which the solver's own first-row quirk
// turns back into A[k+1].
__global__ __launch_bounds__(AT) void k_adj_tchain(long n, Chains cs, const double *__restrict__ A,
                                                   const double *__restrict__ D, const double *__restrict__ C,
                                                   double *__restrict__ At, double *__restrict__ Dt,
                                                   double *__restrict__ Ct) {
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    long L;
    const long r = apos(cs, k, L);
    Dt[k] = D[k];
    double sub = 0.0, sup = 0.0;
    if (r > 0) sub = (r - 1 == 0) ? C[k - 1] * D[k - 1] / D[k] : C[k - 1];
    if (r < L - 1) sup = (r == 0) ? A[k + 1] * D[k + 1] / D[k] : A[k + 1];
    At[k] = sub;
    Ct[k] = sup;
}

// Coefficient adjoints -mu_k x_{k-1}, -mu_k x_k, -mu_k x_{k+1} of the
// matrix S of a solve S x = b (S with the chain solver's first-row quirk
// S(1,2) = c1 d1 / d2).  Z (optional): rows whose coefficients are constants
// (identity rows) get none.  XK (optional): x is taken as +0 on the rows not
// flagged in XK -- at a converged q-l fixed point the first split's x
// vanishes on every row but the second split's identity rows, and only
// those rows' columns of S carry a derivative.
__global__ __launch_bounds__(AT) void k_adj_coef(long n, Chains cs, const double *__restrict__ A,
                                                 const double *__restrict__ D, const double *__restrict__ C,
                                                 const double *__restrict__ mu, const double *__restrict__ x,
                                                 const unsigned char *__restrict__ Z,
                                                 const unsigned char *__restrict__ XK, double *__restrict__ Ab,
                                                 double *__restrict__ Db, double *__restrict__ Cb) {
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k >= n) return;
    long L;
    const long r = apos(cs, k, L);
    auto xv = [&](long q) { return (XK && !XK[q]) ? 0.0 : x[q]; };
    double ab = 0.0, db = 0.0, cb = 0.0;
    if (!(Z && Z[k])) {
        const double m = mu[k];
        db = -m * xv(k);
        if (r > 0) ab = -m * xv(k - 1);
        if (r < L - 1) {
            const double sb = -m * xv(k + 1);   // adjoint of the entry S(k, k+1)
            if (r == 0) {   // S(1,2) = c1 d1 / d2
                cb = sb * D[k] / D[k + 1];
                db += sb * C[k] / D[k + 1];
            } else {
                cb = sb;
            }
        }
    }
    // row 2's diagonal also enters S(1,2): d(c1 d1 / d2)/d d2 = -c1 d1 / d2^2
    if (r == 1 && !(Z && Z[k - 1])) db += mu[k - 1] * xv(k) * C[k - 1] * D[k - 1] / (D[k] * D[k]);
    Ab[k] = ab;
    Db[k] = db;
    Cb[k] = cb;
}

// lam in chain order from the adjoints of the increments (X rows i = 1..nx,
// j = 2..ny, then Y rows i = 2..nx, j = 1..ny).
// Rows flagged in Z (the second split's identity rows; null: none) get +0.
// Thread 0 also clears red[slot], red[slot + 1] (red null: none).
__global__ __launch_bounds__(AT) void k_adj_gather(int nx, int ny, int ld, const double *__restrict__ ub,
                                                   const double *__restrict__ vb, const unsigned char *__restrict__ Z,
                                                   double *__restrict__ lam, unsigned long long *red, int slot) {
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny;
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (red && k == 0) {
        red[slot] = 0ull;
        red[slot + 1] = 0ull;
    }
    if (Z && k < n1 + n2 && Z[k]) {
        lam[k] = 0.0;
        return;
    }
    if (k < n1) {
        const int i = 1 + (int)(k % nx), j = 2 + (int)(k / nx);
        lam[k] = ub[WD2_IX(i, j)];
    } else if (k < n1 + n2) {
        const long q = k - n1;
        const int i = 2 + (int)(q % (nx - 1)), j = 1 + (int)(q / (nx - 1));
        lam[k] = vb[WD2_IX(i, j)];
    }
}

__global__ __launch_bounds__(AT) void k_adj_zmask(long n, const unsigned char *__restrict__ Z,
                                                  const double *__restrict__ in, double *__restrict__ out) {
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k < n) out[k] = Z[k] ? 0.0 : in[k];
}

// ------------------------------------------------------------ assembly
// Forward values the transposes need, as the forward kernels form them.
#define F(a, b) f[WD2_IX(a, b)]
// ConvCoef case 4 (momentum.f:944-951), djac = 2: written for i = 0..nx, j = 1..ny
__device__ inline double cj1_x(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return 2.0 * m.yeu[WD2_IX(i, j)] * u[WD2_IX(i, j)] -
           m.xeu[WD2_IX(i, j)] * (v[WD2_IX(i + 1, j)] + v[WD2_IX(i, j)] + v[WD2_IX(i + 1, j - 1)] + v[WD2_IX(i, j - 1)]) /
               4.0;
}
__device__ inline double cj2_x(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return m.xzu[WD2_IX(i, j)] * (v[WD2_IX(i + 1, j)] + v[WD2_IX(i, j)] + v[WD2_IX(i + 1, j - 1)] + v[WD2_IX(i, j - 1)]) /
               4.0 -
           2.0 * m.yzu[WD2_IX(i, j)] * u[WD2_IX(i, j)];
}
// ConvCoef case 5 (momentum.f:953-961), djac = 2: written for i = 1..nx, j = 0..ny
__device__ inline double cj1_y(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return m.yev[WD2_IX(i, j)] * (u[WD2_IX(i, j + 1)] + u[WD2_IX(i - 1, j + 1)] + u[WD2_IX(i, j)] + u[WD2_IX(i - 1, j)]) /
               4.0 -
           2.0 * m.xev[WD2_IX(i, j)] * v[WD2_IX(i, j)];
}
__device__ inline double cj2_y(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return 2.0 * m.xzv[WD2_IX(i, j)] * v[WD2_IX(i, j)] -
           m.yzv[WD2_IX(i, j)] * (u[WD2_IX(i, j + 1)] + u[WD2_IX(i - 1, j + 1)] + u[WD2_IX(i, j)] + u[WD2_IX(i - 1, j)]) /
               4.0;
}
// ConvCoef case 1 (X, metrics xzn, xec, yzn, yec), djac = 1:
// c1 written i = 1..nx+1, j = 1..ny+1; c2 written i = 1..nx, j = 1..ny
__device__ inline double c1_x(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return (m.yec[WD2_IX(i, j)] * (u[WD2_IX(i, j)] + u[WD2_IX(i - 1, j)]) -
            m.xec[WD2_IX(i, j)] * (v[WD2_IX(i, j)] + v[WD2_IX(i, j - 1)])) *
           0.5;
}
__device__ inline double c2_x(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return (m.xzn[WD2_IX(i, j)] * (v[WD2_IX(i + 1, j)] + v[WD2_IX(i, j)]) -
            m.yzn[WD2_IX(i, j)] * (u[WD2_IX(i, j + 1)] + u[WD2_IX(i, j)])) *
           0.5;
}
// ConvCoef case 2 (Y, metrics xzc, xen, yzc, yen), djac = 1:
// c1 written i = 1..nx, j = 1..ny; c2 written i = 1..nx+1, j = 1..ny+1
__device__ inline double c1_y(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return (m.yen[WD2_IX(i, j)] * (u[WD2_IX(i, j + 1)] + u[WD2_IX(i, j)]) -
            m.xen[WD2_IX(i, j)] * (v[WD2_IX(i + 1, j)] + v[WD2_IX(i, j)])) *
           0.5;
}
__device__ inline double c2_y(const Metrics &m, int ld, const double *u, const double *v, int i, int j) {
    return (m.xzc[WD2_IX(i, j)] * (v[WD2_IX(i, j)] + v[WD2_IX(i, j - 1)]) -
            m.yzc[WD2_IX(i, j)] * (u[WD2_IX(i, j)] + u[WD2_IX(i - 1, j)])) *
           0.5;
}

// DDiffU transposed (momentum.f:1015-1045): d(i,j) += g spreads onto u.
__device__ inline void ddiff_u_t(const Metrics &m, int ld, double *ub, int i, int j, double g) {
    const double ac0 = m.rac[WD2_IX(i, j)], ac1 = m.rac[WD2_IX(i + 1, j)];
    const double bc0 = m.rbc[WD2_IX(i, j)], bc1 = m.rbc[WD2_IX(i + 1, j)];
    const double bn0 = m.rbn[WD2_IX(i, j)], bnm = m.rbn[WD2_IX(i, j - 1)];
    const double gn0 = m.rgn[WD2_IX(i, j)], gnm = m.rgn[WD2_IX(i, j - 1)];
    // s1
    aadd(&ub[WD2_IX(i + 1, j)], ac1 * g);
    aadd(&ub[WD2_IX(i, j)], -(ac1 + ac0) * g);
    aadd(&ub[WD2_IX(i - 1, j)], ac0 * g);
    aadd(&ub[WD2_IX(i + 1, j + 1)], bc1 * g);
    aadd(&ub[WD2_IX(i, j + 1)], (bc1 - bc0) * g);
    aadd(&ub[WD2_IX(i + 1, j - 1)], -bc1 * g);
    aadd(&ub[WD2_IX(i, j - 1)], (-bc1 + bc0) * g);
    aadd(&ub[WD2_IX(i - 1, j + 1)], -bc0 * g);
    aadd(&ub[WD2_IX(i - 1, j - 1)], bc0 * g);
    // s2
    aadd(&ub[WD2_IX(i + 1, j + 1)], bn0 * g);
    aadd(&ub[WD2_IX(i + 1, j)], (bn0 - bnm) * g);
    aadd(&ub[WD2_IX(i - 1, j + 1)], -bn0 * g);
    aadd(&ub[WD2_IX(i - 1, j)], (-bn0 + bnm) * g);
    aadd(&ub[WD2_IX(i + 1, j - 1)], -bnm * g);
    aadd(&ub[WD2_IX(i - 1, j - 1)], bnm * g);
    aadd(&ub[WD2_IX(i, j + 1)], gn0 * g);
    aadd(&ub[WD2_IX(i, j)], -(gn0 + gnm) * g);
    aadd(&ub[WD2_IX(i, j - 1)], gnm * g);
}

// DDiffV transposed (momentum.f:1079-1109).
__device__ inline void ddiff_v_t(const Metrics &m, int ld, double *vb, int i, int j, double g) {
    const double an0 = m.ran[WD2_IX(i, j)], anm = m.ran[WD2_IX(i - 1, j)];
    const double bn0 = m.rbn[WD2_IX(i, j)], bnm = m.rbn[WD2_IX(i - 1, j)];
    const double bc1 = m.rbc[WD2_IX(i, j + 1)], bc0 = m.rbc[WD2_IX(i, j)];
    const double gc1 = m.rgc[WD2_IX(i, j + 1)], gc0 = m.rgc[WD2_IX(i, j)];
    // s1
    aadd(&vb[WD2_IX(i + 1, j)], an0 * g);
    aadd(&vb[WD2_IX(i, j)], -(an0 + anm) * g);
    aadd(&vb[WD2_IX(i - 1, j)], anm * g);
    aadd(&vb[WD2_IX(i + 1, j + 1)], bn0 * g);
    aadd(&vb[WD2_IX(i, j + 1)], (bn0 - bnm) * g);
    aadd(&vb[WD2_IX(i + 1, j - 1)], -bn0 * g);
    aadd(&vb[WD2_IX(i, j - 1)], (-bn0 + bnm) * g);
    aadd(&vb[WD2_IX(i - 1, j + 1)], -bnm * g);
    aadd(&vb[WD2_IX(i - 1, j - 1)], bnm * g);
    // s2
    aadd(&vb[WD2_IX(i + 1, j + 1)], bc1 * g);
    aadd(&vb[WD2_IX(i + 1, j)], (bc1 - bc0) * g);
    aadd(&vb[WD2_IX(i - 1, j + 1)], -bc1 * g);
    aadd(&vb[WD2_IX(i - 1, j)], (-bc1 + bc0) * g);
    aadd(&vb[WD2_IX(i + 1, j - 1)], -bc0 * g);
    aadd(&vb[WD2_IX(i - 1, j - 1)], bc0 * g);
    aadd(&vb[WD2_IX(i, j + 1)], gc1 * g);
    aadd(&vb[WD2_IX(i, j)], -(gc1 + gc0) * g);
    aadd(&vb[WD2_IX(i, j - 1)], gc0 * g);
}

// XMomentum row (i, j) transposed: the adjoints of the first split's chain
// coefficients (a1b, d1b, c1b) and right-hand side (bb) go to the fields us,
// un through the assembly (momentum.f:350-384), DConvU and DDiffU, and to the
// ConvCoef outputs in A (which k_adj_cc takes on to us, vs, un, vn).  The
// second split's coefficients need no adjoint: at a converged q-l fixed point
// they only precondition.  with_n = 0 skips the time-level-n terms.  Without
// porous regions and buoyancy the porosity terms are +0.
__global__ __launch_bounds__(AT) void k_adj_asm_x(int nx, int ny, int ld, double dk, double re, Metrics m,
                                                  const double *__restrict__ us, const double *__restrict__ vs,
                                                  const double *__restrict__ un, const double *__restrict__ vn,
                                                  const double *__restrict__ a1b, const double *__restrict__ d1b,
                                                  const double *__restrict__ c1b, const double *__restrict__ bb,
                                                  int with_n, double *usb, double *unb, AdjCoefs A) {
    const long n1 = (long)nx * (ny - 1);
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k >= n1) return;
    const int i = 1 + (int)(k % nx), j = 2 + (int)(k / nx);
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    const double rkj = dk2 * m.dju[WD2_IX(i, j)];
    // split 1: upwinding on cj1(i,j); cj1(nx+1, j) is never written (+0)
    if (cj1_x(m, ld, us, vs, i, j) >= 0.0) {
        aadd(&A.cj1[WD2_IX(i - 1, j)], -rkj * a1b[k]);
        aadd(&A.cj1[WD2_IX(i, j)], rkj * d1b[k]);
    } else {
        aadd(&A.cj1[WD2_IX(i, j)], -rkj * d1b[k]);
        if (i + 1 <= nx) aadd(&A.cj1[WD2_IX(i + 1, j)], rkj * c1b[k]);
    }
    // right-hand side: b = un - us - rkj (cnvs + cnvn) + rkj re1 (difs + difn)
    const double b = bb[k];
    if (with_n) aadd(&unb[WD2_IX(i, j)], b);
    aadd(&usb[WD2_IX(i, j)], -b);
    const double gc = -rkj * b, gd = rkj * re1 * b;
    if (b == 0.0) return;
    // DConvU on (us; c1s, c2s) and (un; c1n, c2n)
    for (int lev = 0; lev < 1 + with_n; lev++) {
        const double *u = lev ? un : us, *v = lev ? vn : vs;
        double *ub = lev ? unb : usb, *c1 = lev ? A.c1n : A.c1s, *c2 = lev ? A.c2n : A.c2s;
        const double c1w = c1_x(m, ld, u, v, i, j), c1e = c1_x(m, ld, u, v, i + 1, j);
        const double c2s = c2_x(m, ld, u, v, i, j - 1), c2n = c2_x(m, ld, u, v, i, j);
        const double u0 = u[WD2_IX(i, j)];
        aadd(&ub[WD2_IX(i, j - 1)], -c2s * gc);
        aadd(&ub[WD2_IX(i - 1, j)], -c1w * gc);
        aadd(&ub[WD2_IX(i, j)], (c1e - c1w + c2n - c2s) * gc);
        aadd(&ub[WD2_IX(i + 1, j)], c1e * gc);
        aadd(&ub[WD2_IX(i, j + 1)], c2n * gc);
        aadd(&c1[WD2_IX(i, j)], (-u[WD2_IX(i - 1, j)] - u0) * gc);
        aadd(&c1[WD2_IX(i + 1, j)], (u0 + u[WD2_IX(i + 1, j)]) * gc);
        aadd(&c2[WD2_IX(i, j)], (u0 + u[WD2_IX(i, j + 1)]) * gc);
        aadd(&c2[WD2_IX(i, j - 1)], (-u[WD2_IX(i, j - 1)] - u0) * gc);
        ddiff_u_t(m, ld, ub, i, j, gd);
    }
}

// YMomentum row (i, j) transposed (momentum.f:675-754 with DConvV, DDiffV).
__global__ __launch_bounds__(AT) void k_adj_asm_y(int nx, int ny, int ld, double dk, double re, Metrics m,
                                                  const double *__restrict__ us, const double *__restrict__ vs,
                                                  const double *__restrict__ un, const double *__restrict__ vn,
                                                  const double *__restrict__ a1b, const double *__restrict__ d1b,
                                                  const double *__restrict__ c1b, const double *__restrict__ bb,
                                                  int with_n, double *vsb, double *vnb, AdjCoefs A) {
    const long n2 = (long)(nx - 1) * ny;
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k >= n2) return;
    const int i = 2 + (int)(k % (nx - 1)), j = 1 + (int)(k / (nx - 1));
    const double dk2 = dk * 0.5, re1 = 1.0 / re;
    const double rkj = dk2 * m.djv[WD2_IX(i, j)];
    // split 1: cj1(nx+1, j) is never written
    if (cj1_y(m, ld, us, vs, i, j) >= 0.0) {
        aadd(&A.cj1[WD2_IX(i - 1, j)], -rkj * a1b[k]);
        aadd(&A.cj1[WD2_IX(i, j)], rkj * d1b[k]);
    } else {
        aadd(&A.cj1[WD2_IX(i, j)], -rkj * d1b[k]);
        if (i + 1 <= nx) aadd(&A.cj1[WD2_IX(i + 1, j)], rkj * c1b[k]);
    }
    const double b = bb[k];
    if (with_n) aadd(&vnb[WD2_IX(i, j)], b);
    aadd(&vsb[WD2_IX(i, j)], -b);
    if (b == 0.0) return;
    const double gc = -rkj * b, gd = rkj * re1 * b;
    for (int lev = 0; lev < 1 + with_n; lev++) {
        const double *u = lev ? un : us, *v = lev ? vn : vs;
        double *vb = lev ? vnb : vsb, *c1 = lev ? A.c1n : A.c1s, *c2 = lev ? A.c2n : A.c2s;
        const double c1w = c1_y(m, ld, u, v, i - 1, j), c1e = c1_y(m, ld, u, v, i, j);
        const double c2s = c2_y(m, ld, u, v, i, j), c2n = c2_y(m, ld, u, v, i, j + 1);
        const double v0 = v[WD2_IX(i, j)];
        aadd(&vb[WD2_IX(i, j - 1)], -c2s * gc);
        aadd(&vb[WD2_IX(i - 1, j)], -c1w * gc);
        aadd(&vb[WD2_IX(i, j)], (c1e - c1w + c2n - c2s) * gc);
        aadd(&vb[WD2_IX(i + 1, j)], c1e * gc);
        aadd(&vb[WD2_IX(i, j + 1)], c2n * gc);
        aadd(&c1[WD2_IX(i, j)], (v0 + v[WD2_IX(i + 1, j)]) * gc);
        aadd(&c1[WD2_IX(i - 1, j)], (-v[WD2_IX(i - 1, j)] - v0) * gc);
        aadd(&c2[WD2_IX(i, j + 1)], (v0 + v[WD2_IX(i, j + 1)]) * gc);
        aadd(&c2[WD2_IX(i, j)], (-v[WD2_IX(i, j - 1)] - v0) * gc);
        ddiff_v_t(m, ld, vb, i, j, gd);
    }
}

// The ConvCoef outputs' adjoints of one component onto the fields (every
// point of the written ranges; entries outside them were never written and
// their adjoints are dropped).  Each thread clears the adjoints it reads, so
// A is all zero again for the next assembly transpose (with_n = 0: the
// time-level-n ones are zero and not read).
__global__ __launch_bounds__(AT) void k_adj_cc(int comp, int nx, int ny, int ld, Metrics m, AdjCoefs A,
                                               int with_n, double *usb, double *vsb, double *unb, double *vnb) {
    const int W = nx + 2;
    const long q = (long)blockIdx.x * AT + threadIdx.x;
    if (q >= (long)W * (ny + 2)) return;
    const int i = (int)(q % W), j = (int)(q / W);
    const size_t p = WD2_IX(i, j);
    const double gj1 = A.cj1[p], gj2 = A.cj2[p];
    const double gc[2][2] = {{A.c1s[p], A.c2s[p]}, {with_n ? A.c1n[p] : 0.0, with_n ? A.c2n[p] : 0.0}};
    A.cj1[p] = 0.0;
    A.cj2[p] = 0.0;
    A.c1s[p] = 0.0;
    A.c2s[p] = 0.0;
    if (with_n) {
        A.c1n[p] = 0.0;
        A.c2n[p] = 0.0;
    }
    if (comp == 1) {
        // case 4 (i = 0..nx, j = 1..ny): cj1 = 2 yeu u - xeu (4 v)/4, cj2 = xzu (4 v)/4 - 2 yzu u
        if (i <= nx && j >= 1 && j <= ny) {
            const double g1 = gj1, g2 = gj2;
            aadd(&usb[p], 2.0 * m.yeu[p] * g1 - 2.0 * m.yzu[p] * g2);
            const double gv = (-m.xeu[p] * g1 + m.xzu[p] * g2) / 4.0;
            if (gv != 0.0) {
                aadd(&vsb[WD2_IX(i + 1, j)], gv);
                aadd(&vsb[WD2_IX(i, j)], gv);
                aadd(&vsb[WD2_IX(i + 1, j - 1)], gv);
                aadd(&vsb[WD2_IX(i, j - 1)], gv);
            }
        }
        // case 1: c1 (i = 1..nx+1, j = 1..ny+1), c2 (i = 1..nx, j = 1..ny)
        for (int lev = 0; lev < 1 + with_n; lev++) {
            double *ub = lev ? unb : usb, *vb = lev ? vnb : vsb;
            const double g1 = gc[lev][0], g2 = gc[lev][1];
            if (i >= 1 && j >= 1) {
                aadd(&ub[WD2_IX(i, j)], 0.5 * m.yec[p] * g1);
                aadd(&ub[WD2_IX(i - 1, j)], 0.5 * m.yec[p] * g1);
                aadd(&vb[WD2_IX(i, j)], -0.5 * m.xec[p] * g1);
                aadd(&vb[WD2_IX(i, j - 1)], -0.5 * m.xec[p] * g1);
            }
            if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
                aadd(&vb[WD2_IX(i + 1, j)], 0.5 * m.xzn[p] * g2);
                aadd(&vb[WD2_IX(i, j)], 0.5 * m.xzn[p] * g2);
                aadd(&ub[WD2_IX(i, j + 1)], -0.5 * m.yzn[p] * g2);
                aadd(&ub[WD2_IX(i, j)], -0.5 * m.yzn[p] * g2);
            }
        }
    } else {
        // case 5 (i = 1..nx, j = 0..ny): cj1 = yev (4 u)/4 - 2 xev v, cj2 = 2 xzv v - yzv (4 u)/4
        if (i >= 1 && i <= nx && j <= ny) {
            const double g1 = gj1, g2 = gj2;
            aadd(&vsb[p], -2.0 * m.xev[p] * g1 + 2.0 * m.xzv[p] * g2);
            const double gu = (m.yev[p] * g1 - m.yzv[p] * g2) / 4.0;
            if (gu != 0.0) {
                aadd(&usb[WD2_IX(i, j + 1)], gu);
                aadd(&usb[WD2_IX(i - 1, j + 1)], gu);
                aadd(&usb[WD2_IX(i, j)], gu);
                aadd(&usb[WD2_IX(i - 1, j)], gu);
            }
        }
        // case 2: c1 (i = 1..nx, j = 1..ny), c2 (i = 1..nx+1, j = 1..ny+1)
        for (int lev = 0; lev < 1 + with_n; lev++) {
            double *ub = lev ? unb : usb, *vb = lev ? vnb : vsb;
            const double g1 = gc[lev][0], g2 = gc[lev][1];
            if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
                aadd(&ub[WD2_IX(i, j + 1)], 0.5 * m.yen[p] * g1);
                aadd(&ub[WD2_IX(i, j)], 0.5 * m.yen[p] * g1);
                aadd(&vb[WD2_IX(i + 1, j)], -0.5 * m.xen[p] * g1);
                aadd(&vb[WD2_IX(i, j)], -0.5 * m.xen[p] * g1);
            }
            if (i >= 1 && j >= 1) {
                aadd(&vb[WD2_IX(i, j)], 0.5 * m.xzc[p] * g2);
                aadd(&vb[WD2_IX(i, j - 1)], 0.5 * m.xzc[p] * g2);
                aadd(&ub[WD2_IX(i, j)], -0.5 * m.yzc[p] * g2);
                aadd(&ub[WD2_IX(i - 1, j)], -0.5 * m.yzc[p] * g2);
            }
        }
    }
}

// ------------------------------------------------------------ boundary fills
// VelBoundCond with no-slip walls on all four sides of one region covering
// the grid (iW = 1, iE = nx, jS = 1, jN = ny), replayed in reverse: sides N,
// S, E, W, and within a side the v assignments before the u ones.  The
// wall values (a profile value per line, else the constant) receive the
// adjoints of 2 U in bc_bar[side][var][line] (k = j on W/E, i on S/N).
__global__ __launch_bounds__(1024) void k_adj_vbc(Regions R, int ld, double *ub, double *vb, double *bcb) {
    const int nx = R.nx, ny = R.ny, t = threadIdx.x;
    const int W = (nx + 2 > ny + 2) ? nx + 2 : ny + 2;
    auto bc = [&](int side, int var, int k) -> double * { return bcb + ((side - 1) * 2 + (var - 1)) * (size_t)W + k; };
    // N: v(i, ny) = 0 (i = 1..nx); u(i, ny+1) = 2 U(i) - u(i, ny) (i = 2..nx)
    for (int i = 1 + t; i <= nx; i += 1024) vb[WD2_IX(i, ny)] = 0.0;
    __syncthreads();
    for (int i = 2 + t; i <= nx; i += 1024) {
        const double g = ub[WD2_IX(i, ny + 1)];
        ub[WD2_IX(i, ny)] -= g;
        *bc(NORTH, VU, i) += 2.0 * g;
        ub[WD2_IX(i, ny + 1)] = 0.0;
    }
    __syncthreads();
    // S: v(i, 1) = 0 (i = 1..nx); u(i, 1) = 2 U(i) - u(i, 2) (i = 2..nx)
    for (int i = 1 + t; i <= nx; i += 1024) vb[WD2_IX(i, 1)] = 0.0;
    __syncthreads();
    for (int i = 2 + t; i <= nx; i += 1024) {
        const double g = ub[WD2_IX(i, 1)];
        ub[WD2_IX(i, 2)] -= g;
        *bc(SOUTH, VU, i) += 2.0 * g;
        ub[WD2_IX(i, 1)] = 0.0;
    }
    __syncthreads();
    // E: v(nx+1, j) = 2 V(j) - v(nx, j) (j = 2..ny); u(nx, j) = 0 (j = 1..ny)
    for (int j = 2 + t; j <= ny; j += 1024) {
        const double g = vb[WD2_IX(nx + 1, j)];
        vb[WD2_IX(nx, j)] -= g;
        *bc(EAST, VV, j) += 2.0 * g;
        vb[WD2_IX(nx + 1, j)] = 0.0;
    }
    __syncthreads();
    for (int j = 1 + t; j <= ny; j += 1024) ub[WD2_IX(nx, j)] = 0.0;
    __syncthreads();
    // W: v(1, j) = 2 V(j) - v(2, j) (j = 2..ny); u(1, j) = 0 (j = 1..ny)
    for (int j = 2 + t; j <= ny; j += 1024) {
        const double g = vb[WD2_IX(1, j)];
        vb[WD2_IX(2, j)] -= g;
        *bc(WEST, VV, j) += 2.0 * g;
        vb[WD2_IX(1, j)] = 0.0;
    }
    __syncthreads();
    for (int j = 1 + t; j <= ny; j += 1024) ub[WD2_IX(1, j)] = 0.0;
}

// PresBoundCond with walls (p(ghost) = value + p(inner)) in reverse: N, S, E, W.
__global__ __launch_bounds__(1024) void k_adj_pbc(int nx, int ny, int ld, double *pb) {
    const int t = threadIdx.x;
    for (int i = 2 + t; i <= nx; i += 1024) {
        pb[WD2_IX(i, ny)] += pb[WD2_IX(i, ny + 1)];
        pb[WD2_IX(i, ny + 1)] = 0.0;
    }
    __syncthreads();
    for (int i = 2 + t; i <= nx; i += 1024) {
        pb[WD2_IX(i, 2)] += pb[WD2_IX(i, 1)];
        pb[WD2_IX(i, 1)] = 0.0;
    }
    __syncthreads();
    for (int j = 2 + t; j <= ny; j += 1024) {
        pb[WD2_IX(nx, j)] += pb[WD2_IX(nx + 1, j)];
        pb[WD2_IX(nx + 1, j)] = 0.0;
    }
    __syncthreads();
    for (int j = 2 + t; j <= ny; j += 1024) {
        pb[WD2_IX(2, j)] += pb[WD2_IX(1, j)];
        pb[WD2_IX(1, j)] = 0.0;
    }
}

// ------------------------------------------------------------ Project / Ppe
// Project transposed (utility.f:314-321, 384-391): u, v pass through; p
// receives minus the gradient operator's transpose of their adjoints.
__global__ __launch_bounds__(AT) void k_adj_project(int nx, int ny, int ld, double dk, Metrics m,
                                                    const double *__restrict__ ub, const double *__restrict__ vb,
                                                    double *pb) {
    const long q = (long)blockIdx.x * AT + threadIdx.x;
    if (q >= (long)(nx + 2) * (ny + 2)) return;
    const int i = (int)(q % (nx + 2)), j = (int)(q / (nx + 2));
    if (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny) {
        const double g = ub[WD2_IX(i, j)], djk = m.dju[WD2_IX(i, j)] * dk;
        const double gx = djk * m.yeu[WD2_IX(i, j)] * g, gy = djk * m.yzu[WD2_IX(i, j)] * g / 4.0;
        aadd(&pb[WD2_IX(i + 1, j)], -gx);
        aadd(&pb[WD2_IX(i, j)], gx);
        aadd(&pb[WD2_IX(i + 1, j + 1)], gy);
        aadd(&pb[WD2_IX(i, j + 1)], gy);
        aadd(&pb[WD2_IX(i + 1, j - 1)], -gy);
        aadd(&pb[WD2_IX(i, j - 1)], -gy);
    }
    if (i >= 2 && i <= nx && j >= 2 && j <= ny - 1) {
        const double g = vb[WD2_IX(i, j)], djk = m.djv[WD2_IX(i, j)] * dk;
        const double gx = djk * m.xev[WD2_IX(i, j)] * g / 4.0, gy = djk * m.xzv[WD2_IX(i, j)] * g;
        aadd(&pb[WD2_IX(i + 1, j + 1)], gx);
        aadd(&pb[WD2_IX(i + 1, j)], gx);
        aadd(&pb[WD2_IX(i - 1, j + 1)], -gx);
        aadd(&pb[WD2_IX(i - 1, j)], -gx);
        aadd(&pb[WD2_IX(i, j + 1)], -gy);
        aadd(&pb[WD2_IX(i, j)], gy);
    }
}

// Divergence case 1 (pressure.f:287-299) transposed, for g = lam / dk on
// the pressure cells i = 2..nx, j = 2..ny; and the frozen ghosts' share of
// the Poisson solve: pb(ghost) -= coupling * lam.
__global__ __launch_bounds__(AT) void k_adj_div(int nx, int ny, int ld, double dk, Metrics m,
                                                const double *__restrict__ lam, double *ub, double *vb, double *pb) {
    const long q = (long)blockIdx.x * AT + threadIdx.x;
    if (q >= (long)(nx - 1) * (ny - 1)) return;
    const int i = 2 + (int)(q % (nx - 1)), j = 2 + (int)(q / (nx - 1));
    const double l = lam[WD2_IX(i, j)], g = l / dk;
    if (i == 2) aadd(&pb[WD2_IX(1, j)], -m.rau[WD2_IX(1, j)] * l);
    if (i == nx) aadd(&pb[WD2_IX(nx + 1, j)], -m.rau[WD2_IX(nx, j)] * l);
    if (j == 2) aadd(&pb[WD2_IX(i, 1)], -m.rgv[WD2_IX(i, 1)] * l);
    if (j == ny) aadd(&pb[WD2_IX(i, ny + 1)], -m.rgv[WD2_IX(i, ny)] * l);
    if (g == 0.0) return;
    // ucij = yeu(i,j) u(i,j) - xeu(i,j) (v(i+1,j) + v(i,j) + v(i+1,j-1) + v(i,j-1)) / 4
    aadd(&ub[WD2_IX(i, j)], m.yeu[WD2_IX(i, j)] * g);
    double gv = -m.xeu[WD2_IX(i, j)] * g / 4.0;
    aadd(&vb[WD2_IX(i + 1, j)], gv);
    aadd(&vb[WD2_IX(i, j)], gv);
    aadd(&vb[WD2_IX(i + 1, j - 1)], gv);
    aadd(&vb[WD2_IX(i, j - 1)], gv);
    // - uci1j
    aadd(&ub[WD2_IX(i - 1, j)], -m.yeu[WD2_IX(i - 1, j)] * g);
    gv = m.xeu[WD2_IX(i - 1, j)] * g / 4.0;
    aadd(&vb[WD2_IX(i, j)], gv);
    aadd(&vb[WD2_IX(i - 1, j)], gv);
    aadd(&vb[WD2_IX(i, j - 1)], gv);
    aadd(&vb[WD2_IX(i - 1, j - 1)], gv);
    // vcij = xzv(i,j) v(i,j) - yzv(i,j) (u(i,j+1) + u(i-1,j+1) + u(i,j) + u(i-1,j)) / 4
    aadd(&vb[WD2_IX(i, j)], m.xzv[WD2_IX(i, j)] * g);
    double gu = -m.yzv[WD2_IX(i, j)] * g / 4.0;
    aadd(&ub[WD2_IX(i, j + 1)], gu);
    aadd(&ub[WD2_IX(i - 1, j + 1)], gu);
    aadd(&ub[WD2_IX(i, j)], gu);
    aadd(&ub[WD2_IX(i - 1, j)], gu);
    // - vcij1
    aadd(&vb[WD2_IX(i, j - 1)], -m.xzv[WD2_IX(i, j - 1)] * g);
    gu = m.yzv[WD2_IX(i, j - 1)] * g / 4.0;
    aadd(&ub[WD2_IX(i, j)], gu);
    aadd(&ub[WD2_IX(i - 1, j)], gu);
    aadd(&ub[WD2_IX(i, j - 1)], gu);
    aadd(&ub[WD2_IX(i - 1, j - 1)], gu);
}

// nu += z over n elements with red[slot] = max |z| and red[slot + 1] =
// max |nu| (the slots start at +0; a NaN wins).
__global__ __launch_bounds__(AT) void k_adj_update(long n, const double *__restrict__ z, double *__restrict__ nu,
                                                   unsigned long long *red, int slot) {
    __shared__ double sh[2][AT / 64];
    double mz = 0.0, mn = 0.0;
    for (long k = (long)blockIdx.x * AT + threadIdx.x; k < n; k += (long)gridDim.x * AT) {
        const double zk = z[k], v = nu[k] + zk;
        nu[k] = v;
        if (!(fabs(zk) <= mz)) mz = fabs(zk);
        if (!(fabs(v) <= mn)) mn = fabs(v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double a = __shfl_xor(mz, off), b = __shfl_xor(mn, off);
        if (!(a <= mz)) mz = a;
        if (!(b <= mn)) mn = b;
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = mz;
        sh[1][threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < AT / 64; w++) {
            if (!(sh[0][w] <= mz)) mz = sh[0][w];
            if (!(sh[1][w] <= mn)) mn = sh[1][w];
        }
        atomicMax(&red[slot], dbits(mz));
        atomicMax(&red[slot + 1], dbits(mn));
    }
}

// y += alpha x and y = 0 over n elements.
__global__ __launch_bounds__(AT) void k_adj_axpy(long n, double alpha, const double *__restrict__ x, double *y) {
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k < n) y[k] += alpha * x[k];
}
__global__ __launch_bounds__(AT) void k_adj_zero(long n, double *y) {
    const long k = (long)blockIdx.x * AT + threadIdx.x;
    if (k < n) y[k] = 0.0;
}

// The adjoint of a step's initial u (comp 1) or v (comp 2) from the
// converged q-l fixed point: out = fnb + fsb on the points the iterations
// leave at their initial value (us = un there), out = fnb on the points they
// update (u: i = 2..nx-1, j = 2..ny; v: i = 2..nx, j = 2..ny-1).
__global__ __launch_bounds__(AT) void k_adj_combine(int comp, int nx, int ny, int ld, const double *__restrict__ fsb,
                                                    const double *__restrict__ fnb, double *__restrict__ out) {
    const long q = (long)blockIdx.x * AT + threadIdx.x;
    if (q >= (long)(nx + 2) * (ny + 2)) return;
    const int i = (int)(q % (nx + 2)), j = (int)(q / (nx + 2));
    const bool act = comp == 1 ? (i >= 2 && i <= nx - 1 && j >= 2 && j <= ny)
                               : (i >= 2 && i <= nx && j >= 2 && j <= ny - 1);
    const size_t p = WD2_IX(i, j);
    out[p] = act ? fnb[p] : fnb[p] + fsb[p];
}

// red[slot] = max(red[slot], max |x|) over n elements (non-negative doubles
// order like their bit patterns; red must start at +0).  A NaN wins.
__global__ __launch_bounds__(AT) void k_adj_maxabs(long n, const double *__restrict__ x, unsigned long long *red,
                                                   int slot) {
    __shared__ double sh[AT / 64];
    double m = 0.0;
    for (long k = (long)blockIdx.x * AT + threadIdx.x; k < n; k += (long)gridDim.x * AT) {
        const double a = fabs(x[k]);
        if (!(a <= m)) m = a;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        if (!(o <= m)) m = o;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < AT / 64; w++)
            if (!(sh[w] <= m)) m = sh[w];
        atomicMax(&red[slot], dbits(m));
    }
}

// ------------------------------------------------------------ launchers
static inline dim3 ag(long n) { return dim3((unsigned)((n + AT - 1) / AT)); }

void launch_adj_tchain(long n, const Chains &cs, const double *A, const double *D, const double *C, double *At,
                       double *Dt, double *Ct, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_tchain, ag(n), dim3(AT), 0, s, n, cs, A, D, C, At, Dt, Ct);
}
void launch_adj_coef(long n, const Chains &cs, const double *A, const double *D, const double *C, const double *mu,
                     const double *x, const unsigned char *Z, const unsigned char *XK, double *Ab, double *Db,
                     double *Cb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_coef, ag(n), dim3(AT), 0, s, n, cs, A, D, C, mu, x, Z, XK, Ab, Db, Cb);
}
void launch_adj_gather(int nx, int ny, int ld, const double *ub, const double *vb, const unsigned char *Z,
                       double *lam, unsigned long long *red, int slot, hipStream_t s) {
    const long n = (long)nx * (ny - 1) + (long)(nx - 1) * ny;
    hipLaunchKernelGGL(k_adj_gather, ag(n), dim3(AT), 0, s, nx, ny, ld, ub, vb, Z, lam, red, slot);
}
void launch_adj_update(long n, const double *z, double *nu, unsigned long long *red, int slot, hipStream_t s) {
    const long nb = std::min<long>(1024, (n + AT - 1) / AT);
    hipLaunchKernelGGL(k_adj_update, dim3((unsigned)std::max<long>(nb, 1)), dim3(AT), 0, s, n, z, nu, red, slot);
}
void launch_adj_zmask(long n, const unsigned char *Z, const double *in, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_zmask, ag(n), dim3(AT), 0, s, n, Z, in, out);
}
void launch_adj_asm(int comp, int nx, int ny, int ld, double dk, double re, const Metrics &m, const double *us,
                    const double *vs, const double *un, const double *vn, const double *a1b, const double *d1b,
                    const double *c1b, const double *bb, int with_n, double *fsb, double *fnb, const AdjCoefs &A,
                    hipStream_t s) {
    if (comp == 1)
        hipLaunchKernelGGL(k_adj_asm_x, ag((long)nx * (ny - 1)), dim3(AT), 0, s, nx, ny, ld, dk, re, m, us, vs, un,
                           vn, a1b, d1b, c1b, bb, with_n, fsb, fnb, A);
    else
        hipLaunchKernelGGL(k_adj_asm_y, ag((long)(nx - 1) * ny), dim3(AT), 0, s, nx, ny, ld, dk, re, m, us, vs, un,
                           vn, a1b, d1b, c1b, bb, with_n, fsb, fnb, A);
}
void launch_adj_cc(int comp, int nx, int ny, int ld, const Metrics &m, const AdjCoefs &A, int with_n, double *usb,
                   double *vsb, double *unb, double *vnb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_cc, ag((long)(nx + 2) * (ny + 2)), dim3(AT), 0, s, comp, nx, ny, ld, m, A, with_n, usb,
                       vsb, unb, vnb);
}
void launch_adj_vbc(const Regions &R, int ld, double *ub, double *vb, double *bcb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_vbc, dim3(1), dim3(1024), 0, s, R, ld, ub, vb, bcb);
}
void launch_adj_pbc(int nx, int ny, int ld, double *pb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_pbc, dim3(1), dim3(1024), 0, s, nx, ny, ld, pb);
}
void launch_adj_project(int nx, int ny, int ld, double dk, const Metrics &m, const double *ub, const double *vb,
                        double *pb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_project, ag((long)(nx + 2) * (ny + 2)), dim3(AT), 0, s, nx, ny, ld, dk, m, ub, vb, pb);
}
void launch_adj_div(int nx, int ny, int ld, double dk, const Metrics &m, const double *lam, double *ub, double *vb,
                    double *pb, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_div, ag((long)(nx - 1) * (ny - 1)), dim3(AT), 0, s, nx, ny, ld, dk, m, lam, ub, vb, pb);
}
void launch_adj_axpy(long n, double alpha, const double *x, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_axpy, ag(n), dim3(AT), 0, s, n, alpha, x, y);
}
void launch_adj_zero(long n, double *y, hipStream_t s) { hipLaunchKernelGGL(k_adj_zero, ag(n), dim3(AT), 0, s, n, y); }
void launch_adj_combine(int comp, int nx, int ny, int ld, const double *fsb, const double *fnb, double *out,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_adj_combine, ag((long)(nx + 2) * (ny + 2)), dim3(AT), 0, s, comp, nx, ny, ld, fsb, fnb,
                       out);
}
void launch_adj_maxabs(long n, const double *x, unsigned long long *red, int slot, hipStream_t s) {
    const long nb = std::min<long>(1024, (n + AT - 1) / AT);
    hipLaunchKernelGGL(k_adj_maxabs, dim3((unsigned)std::max<long>(nb, 1)), dim3(AT), 0, s, n, x, red, slot);
}

}  // namespace wd2

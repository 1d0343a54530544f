// k_stencil.hip -- point stencils of the hot path: Project (utility.f:253-440),
// Divergence and RhsPpe (pressure.f:265-378), the PPE matrix (pressure.f:97-196),
// DMaxNorm / DiffMaxNorm (utility.f:446-507), and field copies.
//
// All are HBM-bound one-pass kernels over a 2-D grid of 64x4 threads (x along
// i, the contiguous direction).  Reductions write one partial per workgroup;
// a one-workgroup kernel folds the partials.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int SX = 64, SY = 4;

static inline dim3 grid2(int ni, int nj) { return dim3((ni + SX - 1) / SX, (nj + SY - 1) / SY); }

// ------------------------------------------------------------------ Project
__global__ __launch_bounds__(SX *SY) void k_project(Regions R, int ld, double dk, Metrics m,
                                                    const double *__restrict__ p, double *__restrict__ u,
                                                    double *__restrict__ v) {
    const int i = 1 + blockIdx.x * SX + threadIdx.x, j = 1 + blockIdx.y * SY + threadIdx.y;
    if (i > R.nx || j > R.ny) return;
    const double dFour = 4.0;
#define P(a, b) p[WD2_IX(a, b)]
    double uu = u[WD2_IX(i, j)], vv = v[WD2_IX(i, j)];
    bool wu = false, wv = false;
    auto upd_u = [&](void) {
        const double pzi = P(i + 1, j) - P(i, j);
        const double pet = (P(i + 1, j + 1) + P(i, j + 1) - P(i + 1, j - 1) - P(i, j - 1)) / dFour;
        const double djk = m.dju[WD2_IX(i, j)] * dk;
        uu = uu - djk * (m.yeu[WD2_IX(i, j)] * pzi - m.yzu[WD2_IX(i, j)] * pet);
        wu = true;
    };
    auto upd_v = [&](void) {
        const double pzi = (P(i + 1, j + 1) + P(i + 1, j) - P(i - 1, j + 1) - P(i - 1, j)) / dFour;
        const double pet = P(i, j + 1) - P(i, j);
        const double djk = m.djv[WD2_IX(i, j)] * dk;
        vv = vv - djk * (-m.xev[WD2_IX(i, j)] * pzi + m.xzv[WD2_IX(i, j)] * pet);
        wv = true;
    };
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            if (R.type(ir, jr) == RM_BLOCKG) continue;
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            const bool jin = j >= jS + 1 && j <= jN;
            if (jin && i >= iW + 1 && i <= iE - 1) upd_u();
            if (jin && i == iW && R.btype(ir, jr, WEST) == BM_OUTLT1) upd_u();
            if (jin && i == iE && (R.btype(ir, jr, EAST) == BM_INTERN || R.btype(ir, jr, EAST) == BM_OUTLT1))
                upd_u();
        }
    for (int jr = 1; jr <= R.nj; jr++)
        for (int ir = 1; ir <= R.ni; ir++) {
            if (R.type(ir, jr) == RM_BLOCKG) continue;
            const int iW = R.border(ir, jr, WEST), iE = R.border(ir, jr, EAST);
            const int jS = R.border(ir, jr, SOUTH), jN = R.border(ir, jr, NORTH);
            const bool iin = i >= iW + 1 && i <= iE;
            if (iin && j >= jS + 1 && j <= jN - 1) upd_v();
            if (iin && j == jS && R.btype(ir, jr, SOUTH) == BM_OUTLT1) upd_v();
            if (iin && j == jN && (R.btype(ir, jr, NORTH) == BM_INTERN || R.btype(ir, jr, NORTH) == BM_OUTLT1))
                upd_v();
        }
#undef P
    if (wu) u[WD2_IX(i, j)] = uu;
    if (wv) v[WD2_IX(i, j)] = vv;
}

void launch_project(const Regions &R, int ld, double dk, const Metrics &m, const double *p, double *u,
                    double *v, hipStream_t s) {
    hipLaunchKernelGGL(k_project, grid2(R.nx, R.ny), dim3(SX, SY), 0, s, R, ld, dk, m, p, u, v);
}

// --------------------------------------------------------------- Divergence
__global__ __launch_bounds__(SX *SY) void k_divergence(int nx, int ny, int ld, int nloc,
                                                       const double *__restrict__ xet, const double *__restrict__ yet,
                                                       const double *__restrict__ xzi, const double *__restrict__ yzi,
                                                       const double *__restrict__ u, const double *__restrict__ v,
                                                       double *__restrict__ div) {
    const int i = 1 + blockIdx.x * SX + threadIdx.x, j = 1 + blockIdx.y * SY + threadIdx.y;
    if (i > nx || j > ny) return;
#define U(a, b) u[WD2_IX(a, b)]
#define V(a, b) v[WD2_IX(a, b)]
#define M(f, a, b) f[WD2_IX(a, b)]
    if (nloc == 1) {
        const double ucij = M(yet, i, j) * U(i, j) -
                            M(xet, i, j) * (V(i + 1, j) + V(i, j) + V(i + 1, j - 1) + V(i, j - 1)) / 4.0;
        const double uci1j = M(yet, i - 1, j) * U(i - 1, j) -
                             M(xet, i - 1, j) * (V(i, j) + V(i - 1, j) + V(i, j - 1) + V(i - 1, j - 1)) / 4.0;
        const double vcij = M(xzi, i, j) * V(i, j) -
                            M(yzi, i, j) * (U(i, j + 1) + U(i - 1, j + 1) + U(i, j) + U(i - 1, j)) / 4.0;
        const double vcij1 = M(xzi, i, j - 1) * V(i, j - 1) -
                             M(yzi, i, j - 1) * (U(i, j) + U(i - 1, j) + U(i, j - 1) + U(i - 1, j - 1)) / 4.0;
        div[WD2_IX(i, j)] = ucij - uci1j + vcij - vcij1;
    } else {
        const double uci1j = M(yet, i + 1, j) * (U(i, j + 1) + U(i + 1, j + 1) + U(i, j) + U(i + 1, j)) / 4.0 -
                             M(xet, i + 1, j) * V(i + 1, j);
        const double ucij = M(yet, i, j) * (U(i - 1, j + 1) + U(i, j + 1) + U(i - 1, j) + U(i, j)) / 4.0 -
                            M(xet, i, j) * V(i, j);
        const double vcij1 = M(xzi, i, j + 1) * (V(i, j + 1) + V(i + 1, j + 1) + V(i, j) + V(i + 1, j)) / 4.0 -
                             M(yzi, i, j + 1) * U(i, j + 1);
        const double vcij = M(xzi, i, j) * (V(i, j) + V(i + 1, j) + V(i, j - 1) + V(i + 1, j - 1)) / 4.0 -
                            M(yzi, i, j) * U(i, j);
        div[WD2_IX(i, j)] = uci1j - ucij + vcij1 - vcij;
    }
#undef U
#undef V
#undef M
}

void launch_divergence(int nx, int ny, int ld, int nloc, const Metrics &m, const double *u, const double *v,
                       double *div, hipStream_t s) {
    const double *xet = nloc == 1 ? m.xeu : m.xev, *yet = nloc == 1 ? m.yeu : m.yev;
    const double *xzi = nloc == 1 ? m.xzv : m.xzu, *yzi = nloc == 1 ? m.yzv : m.yzu;
    hipLaunchKernelGGL(k_divergence, grid2(nx, ny), dim3(SX, SY), 0, s, nx, ny, ld, nloc, xet, yet, xzi, yzi, u, v,
                       div);
}

// ------------------------------------------------------------- PPE set-up
__global__ __launch_bounds__(SX *SY) void k_ppe_rhs(int nx, int ny, int ld, double dk,
                                                    const double *__restrict__ div,
                                                    const unsigned char *__restrict__ prow, double *__restrict__ rhs) {
    const int i = 2 + blockIdx.x * SX + threadIdx.x, j = 2 + blockIdx.y * SY + threadIdx.y;
    if (i > nx || j > ny) return;
    const double dv = (prow && prow[WD2_IX(i, j)]) ? 0.0 : div[WD2_IX(i, j)];
    rhs[WD2_IX(i, j)] = dv / dk;
}

void launch_ppe_rhs(int nx, int ny, int ld, double dk, const double *div, const unsigned char *prow,
                    double *rhs, hipStream_t s) {
    hipLaunchKernelGGL(k_ppe_rhs, grid2(nx - 1, ny - 1), dim3(SX, SY), 0, s, nx, ny, ld, dk, div, prow, rhs);
}

__global__ __launch_bounds__(SX *SY) void k_rhs_cross(int nx, int ny, int ld, const double *__restrict__ rbu,
                                                      const double *__restrict__ rbv, const double *__restrict__ rhs,
                                                      const double *__restrict__ p, double *__restrict__ beff) {
    const int i = 2 + blockIdx.x * SX + threadIdx.x, j = 2 + blockIdx.y * SY + threadIdx.y;
    if (i > nx || j > ny) return;
#define P(a, b) p[WD2_IX(a, b)]
    beff[WD2_IX(i, j)] =
        rhs[WD2_IX(i, j)] -
        (rbu[WD2_IX(i, j)] * (P(i + 1, j + 1) + P(i, j + 1) - P(i + 1, j - 1) - P(i, j - 1)) -
         rbu[WD2_IX(i - 1, j)] * (P(i, j + 1) + P(i - 1, j + 1) - P(i, j - 1) - P(i - 1, j - 1)) +
         rbv[WD2_IX(i, j)] * (P(i + 1, j + 1) + P(i + 1, j) - P(i - 1, j + 1) - P(i - 1, j)) -
         rbv[WD2_IX(i, j - 1)] * (P(i + 1, j) + P(i + 1, j - 1) - P(i - 1, j) - P(i - 1, j - 1)));
#undef P
}

void launch_rhs_cross(int nx, int ny, int ld, const Metrics &m, const double *rhs, const double *p,
                      double *beff, hipStream_t s) {
    hipLaunchKernelGGL(k_rhs_cross, grid2(nx - 1, ny - 1), dim3(SX, SY), 0, s, nx, ny, ld, m.rbu, m.rbv, rhs, p,
                       beff);
}

__global__ __launch_bounds__(SX *SY) void k_ppe_coef(int nx, int ny, int ld, const double *__restrict__ rau,
                                                     const double *__restrict__ rgv,
                                                     const unsigned char *__restrict__ prow, double *a1, double *a2,
                                                     double *a3, double *a4, double *a5) {
    const int i = 2 + blockIdx.x * SX + threadIdx.x, j = 2 + blockIdx.y * SY + threadIdx.y;
    if (i > nx || j > ny) return;
    const size_t k = WD2_IX(i, j);
    if (prow && prow[k]) {
        a1[k] = 0.0; a2[k] = 0.0; a3[k] = 1.0; a4[k] = 0.0; a5[k] = 0.0;
        return;
    }
    a1[k] = rgv[WD2_IX(i, j - 1)];
    a2[k] = rau[WD2_IX(i - 1, j)];
    a3[k] = -rau[k] - rau[WD2_IX(i - 1, j)] - rgv[k] - rgv[WD2_IX(i, j - 1)];
    a4[k] = rau[k];
    a5[k] = rgv[k];
}

void launch_ppe_coef(int nx, int ny, int ld, const Metrics &m, const unsigned char *prow, double *a1,
                     double *a2, double *a3, double *a4, double *a5, hipStream_t s) {
    hipLaunchKernelGGL(k_ppe_coef, grid2(nx - 1, ny - 1), dim3(SX, SY), 0, s, nx, ny, ld, m.rau, m.rgv, prow, a1, a2,
                       a3, a4, a5);
}

// ------------------------------------------------------------ max norms
// DMaxNorm (b == null): max(|a(5,5)|, max_{2..nx-1, 2..ny-1} |a|)
// DiffMaxNorm:          max(|a(2,2)-b(2,2)|, max_{2..nx-1, 2..ny-1} |a-b|)
// Workgroups publish partial maxima; a one-workgroup kernel folds them with
// the seed (the kernel boundary orders the two, no device-scope fence).
__global__ __launch_bounds__(SX *SY) void k_max_norm(int nx, int ny, int ld, const double *__restrict__ a,
                                                     const double *__restrict__ b, double *__restrict__ partials) {
    __shared__ double sh[SX * SY / 64];
    const int tid = threadIdx.x + SX * threadIdx.y;
    double mx = 0.0;
    for (int j = 2 + blockIdx.y * SY + threadIdx.y; j <= ny - 1; j += gridDim.y * SY) {
        const int i = 2 + blockIdx.x * SX + threadIdx.x;
        if (i <= nx - 1) mx = dmaxf(mx, b ? fabs(a[WD2_IX(i, j)] - b[WD2_IX(i, j)]) : fabs(a[WD2_IX(i, j)]));
    }
    mx = block_max<SX * SY>(mx, sh, tid);
    if (tid == 0) partials[blockIdx.x + gridDim.x * blockIdx.y] = mx;
}

__global__ __launch_bounds__(256) void k_max_norm_fold(int ld, const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ partials, int n, Scalars *sc,
                                                       int slot) {
    __shared__ double sh[4];
    double r = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) r = dmaxf(r, partials[k]);
    r = block_max<256>(r, sh, threadIdx.x);
    if (threadIdx.x == 0) {
        const double seed = b ? fabs(a[WD2_IX(2, 2)] - b[WD2_IX(2, 2)]) : fabs(a[WD2_IX(5, 5)]);
        sc->val[slot] = dmaxf(seed, r);
    }
}

void launch_max_norm(int nx, int ny, int ld, const double *a, const double *b, double *partials, Scalars *sc,
                     int slot, hipStream_t s) {
    const int gx = (max(nx - 2, 1) + SX - 1) / SX;
    const int gy = min((max(ny - 2, 1) + SY - 1) / SY, max(1, 1024 / gx));
    hipLaunchKernelGGL(k_max_norm, dim3(gx, gy), dim3(SX, SY), 0, s, nx, ny, ld, a, b, partials);
    hipLaunchKernelGGL(k_max_norm_fold, dim3(1), dim3(256), 0, s, ld, a, b, partials, gx * gy, sc, slot);
}

// -------------------------------------------------------------- copies
__global__ __launch_bounds__(SX *SY) void k_copy(int ld, int i0, int i1, int j0, int j1,
                                                 const double *__restrict__ src, double *__restrict__ dst) {
    const int i = i0 + blockIdx.x * SX + threadIdx.x, j = j0 + blockIdx.y * SY + threadIdx.y;
    if (i > i1 || j > j1) return;
    dst[WD2_IX(i, j)] = src[WD2_IX(i, j)];
}
__global__ __launch_bounds__(SX *SY) void k_fill(int ld, int i0, int i1, int j0, int j1, double val,
                                                 double *__restrict__ dst) {
    const int i = i0 + blockIdx.x * SX + threadIdx.x, j = j0 + blockIdx.y * SY + threadIdx.y;
    if (i > i1 || j > j1) return;
    dst[WD2_IX(i, j)] = val;
}

// Up to six fields copied in one pass (the time-level copies of a step).
constexpr int COPY_SET_MAX = 6;
struct CopySet {
    const double *src[COPY_SET_MAX];
    double *dst[COPY_SET_MAX];
    int n;
};
__global__ __launch_bounds__(SX *SY) void k_copy_set(int ld, int i0, int i1, int j0, int j1, CopySet cs) {
    const int i = i0 + blockIdx.x * SX + threadIdx.x, j = j0 + blockIdx.y * SY + threadIdx.y;
    if (i > i1 || j > j1) return;
    const size_t k = WD2_IX(i, j);
#pragma unroll
    for (int q = 0; q < COPY_SET_MAX; q++)
        if (q < cs.n) cs.dst[q][k] = cs.src[q][k];
}
void launch_copy_set(int ld, int i0, int i1, int j0, int j1, const double *const *src, double *const *dst, int n,
                     hipStream_t s) {
    CopySet cs{};
    cs.n = n < COPY_SET_MAX ? n : COPY_SET_MAX;
    for (int q = 0; q < cs.n; q++) {
        cs.src[q] = src[q];
        cs.dst[q] = dst[q];
    }
    hipLaunchKernelGGL(k_copy_set, grid2(i1 - i0 + 1, j1 - j0 + 1), dim3(SX, SY), 0, s, ld, i0, i1, j0, j1, cs);
}

void launch_copy(int ld, int i0, int i1, int j0, int j1, const double *src, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy, grid2(i1 - i0 + 1, j1 - j0 + 1), dim3(SX, SY), 0, s, ld, i0, i1, j0, j1, src, dst);
}
void launch_fill(int ld, int i0, int i1, int j0, int j1, double val, double *dst, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, grid2(i1 - i0 + 1, j1 - j0 + 1), dim3(SX, SY), 0, s, ld, i0, i1, j0, j1, val, dst);
}

// The multigrid solve's extrapolated initial guess on the interior cells
// (i = 2..nx, j = 2..ny) from the pressure p = p^n and the pressures at the
// start of the steps before, p1 = p^(n-1) and p2 = p^(n-2): linear, 2 p - p1,
// when p2 is null, else quadratic, 3 (p - p1) + p2.  The ghosts keep the
// values PresBoundCond set.
__global__ __launch_bounds__(SX *SY) void k_mg_extrap(int nx, int ny, int ld, double *__restrict__ p,
                                                      const double *__restrict__ p1, const double *__restrict__ p2) {
    const int i = 2 + blockIdx.x * SX + threadIdx.x, j = 2 + blockIdx.y * SY + threadIdx.y;
    if (i > nx || j > ny) return;
    const size_t k = WD2_IX(i, j);
    p[k] = p2 ? 3.0 * (p[k] - p1[k]) + p2[k] : 2.0 * p[k] - p1[k];
}
void launch_mg_extrap(int nx, int ny, int ld, double *p, const double *p1, const double *p2, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_extrap, grid2(nx - 1, ny - 1), dim3(SX, SY), 0, s, nx, ny, ld, p, p1, p2);
}

}  // namespace wd2

// k_tangent.hip -- tangent-linear (forward-mode) kernels of the time step,
// FP64.  No counterpart in the reference.
//
// The tangent-linear map is the derivative of the step the adjoint
// (k_adjoint.hip) transposes: main.f's step with each q-l loop and Poisson
// solve at its converged solution, on a one-region cavity with no-slip walls
// and the multigrid solver.  The linear pieces of the step (the boundary
// fills, the divergence and right-hand side, the Poisson solve with frozen
// ghosts, Project) are their own tangents and run as the forward kernels on
// the tangent fields.  The q-l loop's fixed point Pa S1(w)^-1 b(w) = 0 (Pa:
// rows that are not identity rows of the second split; S1 the first split's
// matrix with the chain solver's first-row quirk S(1,2) = c1 d1 / d2) has the
// tangent dw solving Pa S1^-1 (db - dS1 x1) = 0, x1 = S1^-1 b, where db and
// dS1 are the derivatives of the assembly along (dw, dwn).  The assembly's
// derivatives come from the forward kernels instantiated on dual numbers
// (k_momentum.hip); these kernels form r = db - dS1 x1 from them and add the
// iteration's increments, dw += S2^-1 Pa S1^-1 r, which is the linearisation
// of the forward q-l iteration and contracts as it does.
#include <algorithm>

#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int TT = 256;

__global__ __launch_bounds__(TT) void k_tl_pack(long n, const double *__restrict__ v, const double *__restrict__ d,
                                                Dual *__restrict__ out) {
    for (long k = (long)blockIdx.x * TT + threadIdx.x; k < n; k += (long)gridDim.x * TT) out[k] = Dual(v[k], d[k]);
}

// r_k = db_k - (dS1 x1)_k over a batch of chains (rows [0, n1) in chains of
// L1 rows, the rest of L2): the sub-diagonal term is absent on a chain's
// first row and the super-diagonal term on its last; on a first row (of a
// chain of more than one row) the super-diagonal is the quirk's c1 d1 / d2,
// whose derivative is (dc1 d1 + c1 dd1) / d2 - (c1 d1 / d2) dd2 / d2.
__global__ __launch_bounds__(TT) void k_tl_rhs(long n, Chains cs, const Dual *__restrict__ A,
                                               const Dual *__restrict__ D, const Dual *__restrict__ C,
                                               const Dual *__restrict__ B, const double *__restrict__ x1,
                                               double *__restrict__ r) {
    for (long k = (long)blockIdx.x * TT + threadIdx.x; k < n; k += (long)gridDim.x * TT) {
        const long Lk = k < cs.n1 ? cs.L1 : cs.L2;
        const long q = k < cs.n1 ? k % cs.L1 : (k - cs.n1) % cs.L2;
        double v = B[k].d - D[k].d * x1[k];
        if (q > 0) v -= A[k].d * x1[k - 1];
        if (q < Lk - 1) {
            double dc = C[k].d;
            if (q == 0) {
                const Dual c = C[k], d0 = D[k], d1 = D[k + 1];
                const double s = c.v * d0.v / d1.v;
                dc = (c.d * d0.v + c.v * d0.d) / d1.v - s * d1.d / d1.v;
            }
            v -= dc * x1[k + 1];
        }
        r[k] = v;
    }
}

// The chain increments z added to the tangent fields on the chain rows'
// points (u rows i = 1..nx, j = 2..ny, then v rows i = 2..nx, j = 1..ny), as
// the forward q-l update adds its increments.
__global__ __launch_bounds__(TT) void k_tl_scatter(int nx, int ny, int ld, const double *__restrict__ z,
                                                   double *__restrict__ du, double *__restrict__ dv) {
    const long n1 = (long)nx * (ny - 1), n2 = (long)(nx - 1) * ny;
    for (long k = (long)blockIdx.x * TT + threadIdx.x; k < n1 + n2; k += (long)gridDim.x * TT) {
        if (k < n1) {
            const int i = 1 + (int)(k % nx), j = 2 + (int)(k / nx);
            du[WD2_IX(i, j)] += z[k];
        } else {
            const long q = k - n1;
            const int i = 2 + (int)(q % (nx - 1)), j = 1 + (int)(q / (nx - 1));
            dv[WD2_IX(i, j)] += z[k];
        }
    }
}

static unsigned tl_blocks(long n) { return (unsigned)std::min<long>(std::max<long>((n + TT - 1) / TT, 1), 4096); }

void launch_tl_pack(long n, const double *v, const double *d, Dual *out, hipStream_t s) {
    hipLaunchKernelGGL(k_tl_pack, dim3(tl_blocks(n)), dim3(TT), 0, s, n, v, d, out);
}
void launch_tl_rhs(long n, const Chains &cs, const ChainT<Dual> &ch, const double *x1, double *r, hipStream_t s) {
    hipLaunchKernelGGL(k_tl_rhs, dim3(tl_blocks(n)), dim3(TT), 0, s, n, cs, ch.a, ch.d, ch.c, ch.b, x1, r);
}
void launch_tl_scatter(int nx, int ny, int ld, const double *z, double *du, double *dv, hipStream_t s) {
    const long n = (long)nx * (ny - 1) + (long)(nx - 1) * ny;
    hipLaunchKernelGGL(k_tl_scatter, dim3(tl_blocks(n)), dim3(TT), 0, s, nx, ny, ld, z, du, dv);
}

}  // namespace wd2

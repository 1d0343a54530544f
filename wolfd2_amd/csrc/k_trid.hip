// k_trid.hip -- bit-exact parallel AltTridLU (momentum.f:1307-1339).
//
// The reference solves each tridiagonal system with one sequential Thomas
// sweep: forward  d'_k = d_k - a_k c'_{k-1},  c'_k = c_k / d'_k,
//                 b'_k = (b_k - a_k b'_{k-1}) / d'_k,
// backward        x_k = b'_k - c'_k x_{k+1},
// with the first row's super-diagonal divided by the SECOND row's original
// diagonal (c'_1 = c_1 / d_2) and no c' for the last row.  The momentum
// chains are ~nx*ny rows long, so a sequential sweep is far too slow on a GPU.
//
// Both recurrences contract (the systems are diagonally dominant), so two
// runs of a recurrence started from different states agree bit for bit after
// a few dozen rows.  Each thread therefore runs its own 16-row segment from a
// guessed state (speculation), then re-runs it from its predecessor's end
// state until its values coincide bit for bit with the speculative ones
// (fix-up).  From the coalescence point on the two runs are identical, so the
// result equals the sequential sweep exactly.  Fix-ups are iterated inside a
// workgroup until no segment end changes; across workgroups a second kernel
// repairs each block's head from its predecessor's end state.  A segment that
// never coalesces raises a flag and a one-lane kernel recomputes the whole
// system sequentially, so the answer is exact in every case.
//
// A batch of independent chains of equal length L is solved at once: row k
// starts a chain when k % L == 0 and ends one when k % L == L-1.
#include "wd2_dev.hpp"
#include "wd2_launch.hpp"

namespace wd2 {

constexpr int TR_T = 256;   // threads per workgroup
constexpr int TR_S = 16;    // rows per thread
constexpr int TR_R = TR_T * TR_S;

__device__ inline bool same(double x, double y) { return dbits(x) == dbits(y); }

// One forward row.  kind: 0 interior, 1 chain start, 2 chain end.
__device__ inline void fwd_row(int kind, double a, double d, double c, double b, double dnext,
                               double cpv, double bpv, double &cpk, double &bpk) {
    if (kind == 1) {
        cpk = c / dnext;
        bpk = b / d;
    } else {
        const double dp = d - (a * cpv);
        cpk = (kind == 2) ? 0.0 : c / dp;
        bpk = (b - a * bpv) / dp;
    }
}

__device__ inline int row_kind(long k, long L) {
    const long r = k % L;
    return r == 0 ? 1 : (r == L - 1 ? 2 : 0);
}

__global__ __launch_bounds__(TR_T) void k_trid_fwd(long n, long L, const double *__restrict__ A,
                                                   const double *__restrict__ D, const double *__restrict__ Cc,
                                                   const double *__restrict__ B, double *__restrict__ CP,
                                                   double *__restrict__ BP, double *__restrict__ blk_c,
                                                   double *__restrict__ blk_b) {
    __shared__ double sc[TR_T], sb[TR_T];
    __shared__ int changed;
    const int t = threadIdx.x;
    const long seg0 = (long)blockIdx.x * TR_R + (long)t * TR_S;
    const int len = (int)max(0L, min((long)TR_S, n - seg0));
    double a[TR_S], d[TR_S], c[TR_S], b[TR_S], cp[TR_S], bp[TR_S];
#pragma unroll
    for (int s = 0; s < TR_S; s++) {
        if (s < len) {
            a[s] = A[seg0 + s]; d[s] = D[seg0 + s]; c[s] = Cc[seg0 + s]; b[s] = B[seg0 + s];
        } else {
            a[s] = 0; d[s] = 1; c[s] = 0; b[s] = 0;
        }
    }
    // the first-pivot quirk needs the original diagonal of the row after a chain start
#define DNEXT(s) (row_kind(seg0 + (s), L) == 1 ? D[seg0 + (s) + 1] : 1.0)
    // speculative run from the zero state
    double cpv = 0.0, bpv = 0.0;
#pragma unroll
    for (int s = 0; s < TR_S; s++) {
        if (s < len) {
            fwd_row(row_kind(seg0 + s, L), a[s], d[s], c[s], b[s], DNEXT(s), cpv, bpv, cp[s], bp[s]);
            cpv = cp[s]; bpv = bp[s];
        }
    }
    // fix-up rounds inside the workgroup
    const bool dep = len > 0 && t > 0 && row_kind(seg0, L) != 1;
    for (int round = 0; round <= TR_T; round++) {
        if (len > 0) { sc[t] = cp[len - 1]; sb[t] = bp[len - 1]; }
        if (t == 0) changed = 0;
        __syncthreads();
        if (dep) {
            double sc0 = sc[t - 1], sb0 = sb[t - 1];
            bool live = true;
#pragma unroll
            for (int s = 0; s < TR_S; s++) {
                if (live && s < len) {
                    double ck, bk;
                    fwd_row(row_kind(seg0 + s, L), a[s], d[s], c[s], b[s], DNEXT(s), sc0, sb0, ck, bk);
                    if (same(ck, cp[s]) && same(bk, bp[s])) {
                        live = false;
                    } else {
                        cp[s] = ck; bp[s] = bk; sc0 = ck; sb0 = bk;
                    }
                }
            }
            if (live) changed = 1;
        }
        __syncthreads();
        if (!changed) break;
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < TR_S; s++)
        if (s < len) { CP[seg0 + s] = cp[s]; BP[seg0 + s] = bp[s]; }
#undef DNEXT
    // block end state = state after the block's last valid row
    const long blk0 = (long)blockIdx.x * TR_R;
    const long last = min(blk0 + TR_R, n) - 1;
    if (last >= seg0 && last < seg0 + len) {
        blk_c[blockIdx.x] = cp[last - seg0];
        blk_b[blockIdx.x] = bp[last - seg0];
    }
}

// Repair each workgroup's head from its predecessor's end state.
__global__ void k_trid_fwd_fix(long n, long L, int nblk, const double *__restrict__ A,
                               const double *__restrict__ D, const double *__restrict__ Cc,
                               const double *__restrict__ B, double *__restrict__ CP, double *__restrict__ BP,
                               const double *__restrict__ blk_c, const double *__restrict__ blk_b,
                               int *__restrict__ fail) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk < 1 || blk >= nblk) return;
    const long k0 = (long)blk * TR_R, k1 = min(k0 + TR_R, n);
    if (row_kind(k0, L) == 1) return;
    double cpv = blk_c[blk - 1], bpv = blk_b[blk - 1];
    for (long k = k0; k < k1; k++) {
        const int kind = row_kind(k, L);
        const double dn = (kind == 1) ? D[k + 1] : 1.0;
        double ck, bk;
        fwd_row(kind, A[k], D[k], Cc[k], B[k], dn, cpv, bpv, ck, bk);
        if (same(ck, CP[k]) && same(bk, BP[k])) return;
        CP[k] = ck; BP[k] = bk; cpv = ck; bpv = bk;
    }
    atomicOr(fail, 1);
}

__global__ __launch_bounds__(TR_T) void k_trid_bwd(long n, long L, const double *__restrict__ CP,
                                                   const double *__restrict__ BP, double *__restrict__ X,
                                                   double *__restrict__ blk_x) {
    __shared__ double sx[TR_T];
    __shared__ int changed;
    const int t = threadIdx.x;
    const long seg0 = (long)blockIdx.x * TR_R + (long)t * TR_S;
    const int len = (int)max(0L, min((long)TR_S, n - seg0));
    double cp[TR_S], bp[TR_S], x[TR_S];
#pragma unroll
    for (int s = 0; s < TR_S; s++) {
        if (s < len) { cp[s] = CP[seg0 + s]; bp[s] = BP[seg0 + s]; } else { cp[s] = 0; bp[s] = 0; }
    }
    double xn = 0.0;
#pragma unroll
    for (int s = TR_S - 1; s >= 0; s--) {
        if (s < len) {
            x[s] = (row_kind(seg0 + s, L) == 2) ? bp[s] : bp[s] - (cp[s] * xn);
            xn = x[s];
        } else {
            x[s] = 0;
        }
    }
    const long lastrow = seg0 + len - 1;
    const bool dep = len > 0 && row_kind(lastrow, L) != 2 && lastrow < n - 1 && t < TR_T - 1 &&
                     seg0 + TR_S < n;
    for (int round = 0; round <= TR_T; round++) {
        if (len > 0) sx[t] = x[0];
        if (t == 0) changed = 0;
        __syncthreads();
        if (dep) {
            double xv = sx[t + 1];
            bool live = true;
#pragma unroll
            for (int s = TR_S - 1; s >= 0; s--) {
                if (live && s < len) {
                    const double xk = (row_kind(seg0 + s, L) == 2) ? bp[s] : bp[s] - (cp[s] * xv);
                    if (same(xk, x[s])) live = false; else { x[s] = xk; xv = xk; }
                }
            }
            if (live) changed = 1;
        }
        __syncthreads();
        if (!changed) break;
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < TR_S; s++)
        if (s < len) X[seg0 + s] = x[s];
    if (t == 0 && len > 0) blk_x[blockIdx.x] = x[0];
}

// Repair each workgroup's tail from its successor's first value.
__global__ void k_trid_bwd_fix(long n, long L, int nblk, const double *__restrict__ CP,
                               const double *__restrict__ BP, double *__restrict__ X,
                               const double *__restrict__ blk_x, int *__restrict__ fail) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nblk - 1) return;
    const long k0 = (long)blk * TR_R, k1 = min(k0 + TR_R, n);
    if (row_kind(k1 - 1, L) == 2) return;
    double xv = blk_x[blk + 1];
    for (long k = k1 - 1; k >= k0; k--) {
        const double xk = (row_kind(k, L) == 2) ? BP[k] : BP[k] - (CP[k] * xv);
        if (same(xk, X[k])) return;
        X[k] = xk;
        xv = xk;
    }
    atomicOr(fail, 1);
}

// Sequential AltTridLU of every chain, used only if a segment never coalesced.
__global__ void k_trid_fallback(long n, long L, const double *__restrict__ A, const double *__restrict__ D,
                                const double *__restrict__ Cc, const double *__restrict__ B,
                                double *__restrict__ CP, double *__restrict__ BP, double *__restrict__ X,
                                int *__restrict__ fail, int *__restrict__ nfallback) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (*fail == 0) return;
    for (long c0 = 0; c0 < n; c0 += L) {
        double cpv = 0, bpv = 0;
        for (long k = c0; k < c0 + L; k++) {
            const int kind = row_kind(k, L);
            fwd_row(kind, A[k], D[k], Cc[k], B[k], kind == 1 ? D[k + 1] : 1.0, cpv, bpv, CP[k], BP[k]);
            cpv = CP[k]; bpv = BP[k];
        }
        double xv = 0;
        for (long k = c0 + L - 1; k >= c0; k--) {
            X[k] = (row_kind(k, L) == 2) ? BP[k] : BP[k] - (CP[k] * xv);
            xv = X[k];
        }
    }
    *fail = 0;
    *nfallback += 1;
}

int trid_blocks(long n) { return (int)((n + TR_R - 1) / TR_R); }

void launch_trid(const TridWork &w, long n, long L, const double *A, const double *D, const double *C,
                 const double *B, double *X, hipStream_t s) {
    const int nb = trid_blocks(n);
    hipLaunchKernelGGL(k_trid_fwd, dim3(nb), dim3(TR_T), 0, s, n, L, A, D, C, B, w.cp, w.bp, w.blk_c, w.blk_b);
    if (nb > 1)
        hipLaunchKernelGGL(k_trid_fwd_fix, dim3((nb + 63) / 64), dim3(64), 0, s, n, L, nb, A, D, C, B, w.cp, w.bp,
                           w.blk_c, w.blk_b, w.fail);
    hipLaunchKernelGGL(k_trid_bwd, dim3(nb), dim3(TR_T), 0, s, n, L, w.cp, w.bp, X, w.blk_x);
    if (nb > 1)
        hipLaunchKernelGGL(k_trid_bwd_fix, dim3((nb + 63) / 64), dim3(64), 0, s, n, L, nb, w.cp, w.bp, X, w.blk_x,
                           w.fail);
    hipLaunchKernelGGL(k_trid_fallback, dim3(1), dim3(64), 0, s, n, L, A, D, C, B, w.cp, w.bp, X, w.fail,
                       w.nfallback);
}

}  // namespace wd2

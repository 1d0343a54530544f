// wd2_noalloc.hip -- AltTridLU on caller-owned device arrays without any
// allocation: the solver's scratch is one static device array per GPU.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The entry
// point takes the reference's AltTridLU arguments (momentum.f:1307-1339:
// a(3,n) rows of sub-diagonal, diagonal, super-diagonal, and b(n), overwritten
// by the solution), as device pointers, and queues its work on the caller's
// stream without synchronising.  The chain is split into the solver's
// structure-of-arrays layout in the static scratch, solved by the same
// kernels as wd2_alt_trid_lu (so with the same bits), and the solution is
// copied back into b.
//
// The scratch holds chains of up to WD2_TRID_NOALLOC_MAX_ROWS rows
// (about 15 MB); each GPU the library runs on keeps its own copy until the
// process ends.  Every call on one GPU shares that copy, so calls that overlap
// in time on one GPU (two streams, two host threads) race: serialise them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/wolfd2_amd.h"
#include "wd2_launch.hpp"

namespace wd2 {
int api_fail(int code, const char *msg);
}

using namespace wd2;

namespace {

constexpr long NA_ROWS = WD2_TRID_NOALLOC_MAX_ROWS;
constexpr long NA_PAD = (NA_ROWS + 15) / 16 * 16;   // rows per array (128-byte multiples)
constexpr int NA_ARRAYS = 7;                         // A, D, C, B, X, cp, bp
constexpr long NA_BLK = (NA_ROWS + 2047) / 2048 + 2;  // per-workgroup slots (2048 rows a workgroup)
constexpr long NA_DOUBLES = NA_ARRAYS * NA_PAD + 4 * NA_BLK + 2;

// The scratch: on the host this name is only a handle for hipGetSymbolAddress;
// the kernels receive the device address it returns.
static __device__ __align__(256) double scratch[NA_DOUBLES];

constexpr int NA_MAX_DEVICES = 64;
double *g_ptr[NA_MAX_DEVICES] = {};
std::mutex g_mu;

struct View {
    double *A, *D, *C, *B, *X;
    TridWork w;
};
View view(double *base) {
    View v;
    v.A = base;
    v.D = base + NA_PAD;
    v.C = base + 2 * NA_PAD;
    v.B = base + 3 * NA_PAD;
    v.X = base + 4 * NA_PAD;
    v.w.cp = base + 5 * NA_PAD;
    v.w.bp = base + 6 * NA_PAD;
    v.w.blk_c = base + NA_ARRAYS * NA_PAD;
    v.w.blk_b = v.w.blk_c + NA_BLK;
    v.w.blk_x = v.w.blk_b + NA_BLK;
    v.w.blk_xe = v.w.blk_x + NA_BLK;
    v.w.fail = reinterpret_cast<int *>(v.w.blk_xe + NA_BLK);
    v.w.nfallback = v.w.fail + 1;
    return v;
}

// The device address of this GPU's scratch, looked up once per device with
// that device current.
int scratch_ptr(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES)
        return api_fail(WD2_ERR_HIP, "wd2_alt_trid_lu_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return api_fail(WD2_ERR_HIP, "wd2_alt_trid_lu_noalloc: hipGetSymbolAddress of the scratch failed");
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return WD2_OK;
}

// a(3,n) -> A, D, C; b -> B; the solver's failure flag starts at zero
__global__ void k_na_split(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ A,
                           double *__restrict__ D, double *__restrict__ C, double *__restrict__ B, int *fail) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *fail = 0;
    if (k >= n) return;
    A[k] = a[3 * k];
    D[k] = a[3 * k + 1];
    C[k] = a[3 * k + 2];
    B[k] = b[k];
}

__global__ void k_na_copy(long n, const double *__restrict__ X, double *__restrict__ b) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = X[k];
}

}  // namespace

namespace wd2 {
// Sequential fallbacks of the no-allocation solver on the current device.
int noalloc_trid_fallbacks() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES || !g_ptr[dev]) return 0;
    int n = 0;
    if (hipMemcpy(&n, view(g_ptr[dev]).w.nfallback, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return n;
}
}  // namespace wd2

extern "C" int wd2_alt_trid_lu_noalloc(int n, const double *a, double *b, void *stream) {
    if (n < 2) return api_fail(WD2_ERR_ARG, "AltTridLU needs n >= 2");
    if ((long)n > NA_ROWS)
        return api_fail(WD2_ERR_ARG, "wd2_alt_trid_lu_noalloc: n exceeds WD2_TRID_NOALLOC_MAX_ROWS (the static scratch)");
    if (!a || !b) return api_fail(WD2_ERR_ARG, "wd2_alt_trid_lu_noalloc: null array");
    if (trid_blocks(n) + 2 > NA_BLK) return api_fail(WD2_ERR_ARG, "wd2_alt_trid_lu_noalloc: block slots exceeded");
    double *base = nullptr;
    int st = scratch_ptr(&base);
    if (st) return st;
    const View v = view(base);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int T = 256;
    const int nb = (int)((n + T - 1) / T);
    hipLaunchKernelGGL(k_na_split, dim3(nb), dim3(T), 0, s, (long)n, a, b, v.A, v.D, v.C, v.B, v.w.fail);
    launch_trid(v.w, (long)n, (long)n, (const double *)v.A, (const double *)v.D, (const double *)v.C,
                (const double *)v.B, v.X, s);
    hipLaunchKernelGGL(k_na_copy, dim3(nb), dim3(T), 0, s, (long)n, (const double *)v.X, b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_fail(WD2_ERR_HIP, hipGetErrorString(e));
    return WD2_OK;
}

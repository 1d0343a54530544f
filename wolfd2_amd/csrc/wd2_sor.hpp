// wd2_sor.hpp -- stopping test shared by the temporally blocked red-black
// SOR kernels (FP64 and mixed precision).
//
// The stopping test of sweep m is max|s| < sortol over the whole grid
// (pressure.f:534), which holds exactly when it holds on every tile: each tile
// publishes a bit mask of its sweeps whose owned-point maximum is below
// sortol, and the test of a launch is the AND of the masks.
#pragma once

#include "wd2_dev.hpp"

namespace wd2 {

// AND of the ntiles masks (every thread gets it).
__device__ inline unsigned tiled_mask(const unsigned *__restrict__ masks, int ntiles, unsigned *sh) {
    unsigned m = ~0u;
    for (int k = threadIdx.x; k < ntiles; k += blockDim.x) m &= masks[k];
    for (int off = 32; off > 0; off >>= 1) m &= (unsigned)__shfl_xor((int)m, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    unsigned r = ~0u;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r &= sh[w];
    __syncthreads();
    return r;
}

// 1-based sweep of a launch (nsweep sweeps after sweep m0) after which the
// solve stops, or 0; *conv tells whether it stopped converged.
__device__ inline int tiled_stop(unsigned mask, int nsweep, int m0, int msorit, int *conv) {
    for (int s = 0; s < nsweep; s++) {
        const int m = m0 + s + 1;
        const bool ok = m > 1 && ((mask >> s) & 1u);
        if (ok || m >= msorit) {
            *conv = ok ? 1 : 0;
            return s + 1;
        }
    }
    return 0;
}

}  // namespace wd2

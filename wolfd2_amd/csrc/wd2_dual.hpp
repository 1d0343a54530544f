// wd2_dual.hpp -- forward-mode dual numbers for the tangent-linear step.
//
// A Dual carries a value v and a directional derivative d.  Every operation
// computes v with the same IEEE operation the double type would use (so a
// kernel instantiated on Dual reproduces the FP64 kernel's values bit for
// bit) and d by the chain rule.  Comparisons look at v only: a branch taken
// on the values is the branch whose derivative is propagated.
#pragma once
#include <hip/hip_runtime.h>

namespace wd2 {

struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0.0), d(0.0) {}
    __host__ __device__ Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }

}  // namespace wd2

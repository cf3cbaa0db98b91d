// Exact modular products in double precision for primes below 2^51 (device).
// With |a| <= q (1 + eps) and 0 <= w < q (or a, w < 2^51, q < 2^51):
//   h = a w (rounded), e = fma(a, w, -h) = a w - h exactly,
//   Q = rint(h / q) within 1.25 of a w / q,  r = fma(-Q, q, h) + e = a w - Q q,
// every intermediate an integer below 2^53, so r is exact with |r| <= 1.25 q.
// fp_reduce(z) = z - rint(z / q) q maps an integer |z| < 2^53 to a
// representative of magnitude at most q/2 (1 + eps).  Files using these are
// compiled without floating-point contraction.
#pragma once
#include <hip/hip_runtime.h>

namespace nx {

__device__ __forceinline__ double fp_mulmod(double a, double w, double q, double qinv)
{
    const double h = a * w;
    const double e = __builtin_fma(a, w, -h);
    const double Q = __builtin_rint(h * qinv);
    return __builtin_fma(-Q, q, h) + e;
}

__device__ __forceinline__ double fp_reduce(double z, double q, double qinv)
{
    return __builtin_fma(-__builtin_rint(z * qinv), q, z);
}

// fp_reduce, then mapped to [0, q) as an integer
__device__ __forceinline__ uint64_t fp_final(double z, double q, double qinv)
{
    z = fp_reduce(z, q, qinv);
    return (uint64_t)(z < 0 ? z + q : z);
}

}  // namespace nx

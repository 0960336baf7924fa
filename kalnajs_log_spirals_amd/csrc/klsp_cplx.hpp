// klsp_cplx.hpp -- complex FP64 device helpers shared by the kernels of
// libklsp.so (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

namespace klsp {

// 1/(c + i d) by Smith's algorithm with the same range scaling as Julia's
// inv(::ComplexF64); the weights use it so c(omega) rounds like the reference.
__device__ __forceinline__ double2 d_inv(double c, double d)
{
    double s = 1.0;
    const double cd = fmax(fabs(c), fabs(d));
    const double eps = 2.220446049250313e-16;
    if (cd >= 0.5 * 1.7976931348623157e308) { c *= 0.5; d *= 0.5; s *= 0.5; }
    if (cd <= 2.2250738585072014e-308 * 2.0 / eps) {
        const double bs = 2.0 / (eps * eps);
        c *= bs; d *= bs; s *= bs;
    }
    double p, q;
    if (fabs(d) <= fabs(c)) {
        const double r = d / c, t = 1.0 / (c + d * r);
        p = t; q = -r * t;
    } else {
        const double r = c / d, t = 1.0 / (d + c * r);
        p = r * t; q = -t;
    }
    return make_double2(p * s, q * s);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// x - a b for complex values with four fused multiply-adds (the LU's
// multiply-subtract; zgetrf's BLAS kernels fuse the same way)
__device__ __forceinline__ double2 cfms(double2 x, double2 a, double2 b)
{
    return make_double2(fma(-a.x, b.x, fma(a.y, b.y, x.x)), fma(-a.x, b.y, fma(-a.y, b.x, x.y)));
}

// lane src's value in every lane (src wave-uniform)
__device__ __forceinline__ double bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

}  // namespace klsp

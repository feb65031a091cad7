// The two powers the path needs, correctly rounded in binary64 (round to
// nearest) up to a 2^-106 relative window around rounding midpoints.
//
// Both start from the correctly rounded sqrt / division and add one
// correction carried in double-double with error-free products (fma):
//   x^(3/2)  = x * sqrt(x),   sqrt(x) = s + (x - s*s)/(2 s)
//   x^(-1/2) = y (1 + r/2),   y = 1/sqrt(x), r = 1 - x y^2 (exact to 2^-106)
// so the result is RN(value + O(2^-106 value)).  They replace pow(x, 1.5)
// (calcRHS.f90:50-51, subrutinas.f90:196,400) and pow(x, -0.5)
// (subrutinas.f90:429,438-439): fewer instructions than a general pow, and
// closer to the exact value than either the device pow or glibc's.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd {

__device__ __forceinline__ double pow_3_2(double x) {
    double s = sqrt(x);
    double d = fma(-s, s, x);          // x - s^2, exact
    double sl = d / (2.0 * s);
    double p = x * s;
    double pe = fma(x, s, -p);         // x*s = p + pe, exact
    double r = p + (pe + x * sl);
    return (x == 0.0 || !(x < __builtin_huge_val())) ? x * s : r;
}

__device__ __forceinline__ double pow_m1_2(double x) {
    double y = 1.0 / sqrt(x);
    double p = y * y, pe = fma(y, y, -p);   // y^2 = p + pe
    double q = x * p, qe = fma(x, p, -q);   // x p = q + qe
    double r = ((1.0 - q) - qe) - x * pe;   // 1 - x y^2
    double z = y + y * (0.5 * r);
    // 0 -> inf, inf -> 0, NaN -> NaN as pow(x, -0.5) gives for x >= 0
    return (x == 0.0 || !(x < __builtin_huge_val())) ? y : z;
}

}  // namespace cfd

// FP64 sine and cosine for the network models' phase differences, in a form that keeps
// few registers live.
//
// The library sin/cos carry an argument reduction for huge |x| (a multi-word product
// with 2/pi) whose temporaries set the register count of every kernel that inlines them,
// although the network models never reach it: a fused vertex kernel with one inlined
// library sin per entry needs 106 VGPRs (4 waves per SIMD), 65 without.  Here:
//
//   * Cody-Waite reduction by pi/2: k = rint(x * 2/pi), r = ((x - k*P1) - k*P2) - k*P3,
//     where P1 and P2 hold the leading 33 significant bits each of pi/2 (so k*P1 and
//     k*P2 are exact for |k| < 2^20) and P3 the next 53 bits;
//   * on [-pi/4, pi/4] the Taylor polynomials of sin to x^15 and of cos to x^16, whose
//     truncation errors there are below 7e-17 and 3e-18 relative;
//   * |x| >= 2^20 * pi/2 and non-finite x go to the library function, out of line.
//
// The constants are exact rationals rounded to double: pi/2 split from its decimal
// expansion, and the coefficients (-1)^k / (2k+1)! and (-1)^k / (2k)!.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nd {

constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2_1 = 0x1.921fb54400000p+0;    // pi/2, leading 33 bits
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;   // the next 33 bits
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;   // the next 53 bits
constexpr double kReduceMax = 1647099.3291652855;   // 2^20 * pi/2

// sin(r) and cos(r) for |r| <= pi/4 (Horner in z = r^2)
__host__ __device__ __forceinline__ double sin_kernel(double r) {
    const double z = r * r;
    const double q = -0.16666666666666666 +
                     z * (0.008333333333333333 +
                          z * (-0.0001984126984126984 +
                               z * (2.7557319223985893e-06 +
                                    z * (-2.505210838544172e-08 + z * (1.6059043836821613e-10 +
                                                                       z * -7.647163731819816e-13)))));
    return r + (r * z) * q;
}

__host__ __device__ __forceinline__ double cos_kernel(double r) {
    const double z = r * r;
    const double q = 0.041666666666666664 +
                     z * (-0.001388888888888889 +
                          z * (2.48015873015873e-05 +
                               z * (-2.755731922398589e-07 +
                                    z * (2.08767569878681e-09 + z * (-1.1470745597729725e-11 +
                                                                     z * 4.779477332387385e-14)))));
    // 1 - z/2 + z^2 q, with the leading subtraction done first for accuracy near pi/4
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * q);
}

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __noinline__ double sin_far(double x) { return ::sin(x); }
__device__ __noinline__ double cos_far(double x) { return ::cos(x); }
#else
inline double sin_far(double x) { return std::sin(x); }
inline double cos_far(double x) { return std::cos(x); }
#endif

// quadrant q = k mod 4 of x = k*pi/2 + r: sin(x) = sin r, cos r, -sin r, -cos r
__host__ __device__ __forceinline__ double quadrant_value(int q, double r) {
    const double v = (q & 1) ? cos_kernel(r) : sin_kernel(r);
    return (q & 2) ? -v : v;
}

__host__ __device__ __forceinline__ double reduce_pio2(double x, int* q) {
    const double k = rint(x * kTwoOverPi);
    *q = static_cast<int>(k);
    return ((x - k * kPio2_1) - k * kPio2_2) - k * kPio2_3;
}

__host__ __device__ __forceinline__ double nd_sin(double x) {
    if (!(fabs(x) < kReduceMax)) return sin_far(x);
    int q;
    const double r = reduce_pio2(x, &q);
    return quadrant_value(q & 3, r);
}

__host__ __device__ __forceinline__ double nd_cos(double x) {
    if (!(fabs(x) < kReduceMax)) return cos_far(x);
    int q;
    const double r = reduce_pio2(x, &q);
    return quadrant_value((q + 1) & 3, r);
}

}  // namespace nd

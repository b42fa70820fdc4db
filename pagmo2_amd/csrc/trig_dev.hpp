// trig_dev.hpp -- cos and sin of moderate arguments for the separable kernels.
//
// Written from the mathematics: y = k*pi + r with k = rint(y / pi) and
// |r| <= pi/2 (two-constant reduction evaluated with fused multiply-adds, after
// Cody & Waite), cos(y) = (-1)^k cos(r), sin(y) = (-1)^k sin(r), and the
// Taylor series of cos(r) through r^22 and of sin(r) through r^23 in Horner
// form (truncation below 1e-19 on |r| <= pi/2).  Absolute error about 2e-16
// for |y| < 2^20; callers sum many such terms, so absolute accuracy is what
// their results depend on.  Constants: the double nearest pi, the double
// nearest its remainder, 1/pi, and (-1)^n / (2n)!, (-1)^n / (2n+1)! rounded.
#pragma once

#include <hip/hip_runtime.h>

namespace pgh
{
namespace dev
{

constexpr double kPiHi = 0x1.921fb54442d18p+1;
constexpr double kPiLo = 0x1.1a62633145c07p-53;
constexpr double kInvPi = 0x1.45f306dc9c883p-2;

__host__ __device__ __forceinline__ double reduce_pi(double y, double &k)
{
    k = rint(y * kInvPi);
    const double r = fma(-k, kPiHi, y);
    return fma(-k, kPiLo, r);
}

__host__ __device__ __forceinline__ double parity_sign(double k, double v)
{
    // k is an integer-valued double with |k| < 2^19 (|y| < 2^20)
    return (static_cast<int>(k) & 1) ? -v : v;
}

__host__ __device__ __forceinline__ double cos_reduced(double r)
{
    const double z = r * r;
    double p = -0x1.0ce396db7f853p-70;
    p = fma(p, z, 0x1.e542ba4020225p-62);
    p = fma(p, z, -0x1.6827863b97d97p-53);
    p = fma(p, z, 0x1.ae7f3e733b81fp-45);
    p = fma(p, z, -0x1.93974a8c07c9dp-37);
    p = fma(p, z, 0x1.1eed8eff8d898p-29);
    p = fma(p, z, -0x1.27e4fb7789f5cp-22);
    p = fma(p, z, 0x1.a01a01a01a01ap-16);
    p = fma(p, z, -0x1.6c16c16c16c17p-10);
    p = fma(p, z, 0x1.5555555555555p-5);
    p = fma(p, z, -0.5);
    return fma(p, z, 1.0);
}

__host__ __device__ __forceinline__ double sin_reduced(double r)
{
    const double z = r * r;
    double p = -0x1.761b41316381ap-75;
    p = fma(p, z, 0x1.71b8ef6dcf572p-66);
    p = fma(p, z, -0x1.2f49b46814157p-57);
    p = fma(p, z, 0x1.952c77030ad4ap-49);
    p = fma(p, z, -0x1.ae7f3e733b81fp-41);
    p = fma(p, z, 0x1.6124613a86d09p-33);
    p = fma(p, z, -0x1.ae64567f544e4p-26);
    p = fma(p, z, 0x1.71de3a556c734p-19);
    p = fma(p, z, -0x1.a01a01a01a01ap-13);
    p = fma(p, z, 0x1.1111111111111p-7);
    p = fma(p, z, -0x1.5555555555555p-3);
    return fma(p * z, r, r);
}

// cos(y), |y| < 2^20
__host__ __device__ __forceinline__ double cos_pi_reduced(double y)
{
    double k;
    const double r = reduce_pi(y, k);
    return parity_sign(k, cos_reduced(r));
}

// sin(y), |y| < 2^20
__host__ __device__ __forceinline__ double sin_pi_reduced(double y)
{
    double k;
    const double r = reduce_pi(y, k);
    return parity_sign(k, sin_reduced(r));
}

// Any argument: the reduction above when it is accurate, the math library's
// routine otherwise (|y| >= 2^20, inf, NaN).
__device__ __forceinline__ double cos_any(double y)
{
    return fabs(y) < 0x1p20 ? cos_pi_reduced(y) : cos(y);
}

__device__ __forceinline__ double sin_any(double y)
{
    return fabs(y) < 0x1p20 ? sin_pi_reduced(y) : sin(y);
}

} // namespace dev
} // namespace pgh

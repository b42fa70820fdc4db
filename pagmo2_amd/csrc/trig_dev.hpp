// trig_dev.hpp -- cos and sin of moderate arguments for the separable kernels.
//
// cos_2pi_small (the rastrigin/ackley streaming kernels) is described where it
// is defined.  The general routines are written from the mathematics: y = k*pi + r with k = rint(y / pi) and
// |r| <= pi/2 (two-constant reduction evaluated with fused multiply-adds, after
// Cody & Waite), cos(y) = (-1)^k cos(r), sin(y) = (-1)^k sin(r), and the
// Taylor series of cos(r) through r^22 and of sin(r) through r^23 in Horner
// form (truncation below 1e-19 on |r| <= pi/2).  Absolute error about 2e-16
// for |y| < 2^20; callers sum many such terms, so absolute accuracy is what
// their results depend on.  Constants: the double nearest pi, the double
// nearest its remainder, 1/pi, and (-1)^n / (2n)!, (-1)^n / (2n+1)! rounded.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

// S * cos(2 pi x) (S = 1 or 10) for moderate x (callers keep |x| < 2^10; the
// identity holds for |x| < 2^50).  2x = k + u with k = rint(2x) and |u| <= 1/2,
// u = 2x - k exact in binary64, so cos(2 pi x) = (-1)^k cos(pi u): no
// multiplication by pi before the reduction, hence no reduction error.  k is
// formed by adding 1.5 * 2^52 (rounding 2x to an integer, ties to even, in the
// low mantissa bits), whose lowest bit is then the parity of k.  cos(pi u) is
// the degree-8 polynomial in u^2 of smallest maximum error on [0, 1/4]
// (gen_cospi_coeffs.py: 3.9e-18 before rounding, about 3e-16 in binary64).
// Compared with the reference's cos(omega * x) the result differs by the
// rounding of omega * x there (at most |2 pi x| * 1.5e-16).
constexpr double kCosPi[9] = {0x1.0000000000000p+0,  -0x1.3bd3cc9be45dbp+2, 0x1.03c1f081b5993p+2,
                              -0x1.55d3c7e3bfc44p+0, 0x1.e1f506813c9d9p-3,  -0x1.a6d1efc968c10p-6,
                              0x1.f9d25488fb381p-10, -0x1.b69582d77d94dp-14, 0x1.16797132c8c4dp-18};
constexpr double kCosPi10[9] = {0x1.4000000000000p+3,  -0x1.8ac8bfc2dd752p+5, 0x1.44b26ca222ff8p+5,
                                -0x1.ab48b9dcafb55p+3, 0x1.2d392410c5e28p+1,  -0x1.084335dde178ap-2,
                                0x1.3c2374d59d030p-6,  -0x1.121d71c6ae7d0p-10, 0x1.5c17cd7f7af60p-15};

template <bool TIMES10>
__device__ __forceinline__ double cos_2pi_small(double x)
{
    constexpr double kShift = 0x1.8p52;
    const double *c = TIMES10 ? kCosPi10 : kCosPi;
    const double t = fma(x, 2.0, kShift);
    const double k = t - kShift;
    const double u = fma(x, 2.0, -k);
    const double z = u * u;
    double p = c[8];
#pragma unroll
    for (int i = 7; i >= 0; --i) p = fma(p, z, c[i]);
    // negate p when k is odd: adding bit 0 of t at bit 31 of p's high word
    // flips its sign bit (one shift-and-add)
    const uint32_t t_lo = static_cast<uint32_t>(__double2loint(t));
    const uint32_t p_hi = static_cast<uint32_t>(__double2hiint(p)) + (t_lo << 31);
    return __hiloint2double(static_cast<int>(p_hi), __double2loint(p));
}

// sin(2 pi x) for moderate x (|x| < 2^10), with the reduction of
// cos_2pi_small: sin(2 pi x) = (-1)^k sin(pi u), and sin(pi u) = u q(u^2)
// with q the degree-8 polynomial of smallest maximum error to
// sin(pi sqrt(z)) / sqrt(z) on [0, 1/4] (gen_cospi_coeffs.py: 6.5e-19 before
// rounding).  Compared with the reference's sin(omega * x) the result differs
// by the rounding of omega * x there (at most |2 pi x| * 1.5e-16).
constexpr double kSinPi[9] = {0x1.921fb54442d18p+1, -0x1.4abbce625be52p+2, 0x1.466bc6775aa7dp+1,
                              -0x1.32d2cce627c9ep-1, 0x1.50783485523f1p-4,  -0x1.e3074dfb5b971p-8,
                              0x1.e8f3677c2c58bp-12, -0x1.6f7ad23a648ddp-16, 0x1.9d46b0508cff6p-21};

__device__ __forceinline__ double sin_2pi_small(double x)
{
    constexpr double kShift = 0x1.8p52;
    const double t = fma(x, 2.0, kShift);
    const double k = t - kShift;
    const double u = fma(x, 2.0, -k);
    const double z = u * u;
    double q = kSinPi[8];
#pragma unroll
    for (int i = 7; i >= 0; --i) q = fma(q, z, kSinPi[i]);
    const double p = u * q;
    const uint32_t t_lo = static_cast<uint32_t>(__double2loint(t));
    const uint32_t p_hi = static_cast<uint32_t>(__double2hiint(p)) + (t_lo << 31);
    return __hiloint2double(static_cast<int>(p_hi), __double2loint(p));
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

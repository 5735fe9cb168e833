// FP64 cos / sin of 2 pi x and exp for the path-sum kernels, written for their argument
// ranges.
//
// The general library routines spend most of their instructions on cases the path sums
// never reach (Payne-Hanek reduction of huge radian arguments, overflow and special-value
// handling).  Here the phase arrives in revolutions, so the reduction to one period is
// exact (x - rint(x)), and the exponent is a finite number of moderate size.
//
//   cos2pi(x) = cos(2 pi x): w = x - rint(x) is exact, and cos(2 pi w) = sin(pi u) with
//       u = 1/2 - 2|w| in [-1/2, 1/2] (exact for 2|w| >= 1/4 by Sterbenz's lemma, within
//       2^-54 otherwise); sin(pi u) = u P(u^2), P of degree 8.
//   sin2pi(x) = cos(2 pi (x - rint(x) - 1/4)).  A caller that adds a table phase can fold
//       the -1/4 into the table and call cos2pi.
//   exp_fin(x) = e^x for finite x with |x| < 1e9: n = rint(x / ln 2) by the 1.5 * 2^52
//       rounding constant (the low word of x / ln 2 + 1.5 * 2^52 is n as an integer),
//       r = x - n ln 2 with ln 2 split in two parts (Cody and Waite's reduction),
//       |r| <= ln 2 / 2, e^r = Q(r) of degree 11, scaled by 2^n with ldexp (results below
//       the double range flush to zero, above it give inf).
//
// P and Q interpolate sin(pi u)/u (in u^2 on [0, 1/4]) and e^r (on [-ln 2/2, ln 2/2]) at
// Chebyshev nodes, solved in 60-digit decimal arithmetic by fp64math_coeffs.py in this
// folder; their interpolation errors are below 1e-18.  Kernels receive the coefficients
// as a by-value kernel argument (PolyCoefs): they then sit in scalar registers and every
// Horner step is one v_fma_f64 with a scalar operand (literal constants end up in vector
// registers, copied into the destination before each two-address v_fmac_f64).
// Against the host's long-double libm the functions are within 3e-16 absolute (cos,
// sin) and 1.5e-16 relative (exp), checked on dense random samples by the tests.  Neither handles NaN or inf inputs.
//
// Include after hip_runtime.h (the functions are __host__ __device__).
#pragma once

#include <cmath>
#include <cstdint>

namespace chm {

struct PolyCoefs {
    double p[9];    // sin(pi u) = u P(u^2), constant term first
    double q[12];   // e^r = Q(r), constant term first
};

constexpr PolyCoefs kPolyCoefs = {
    {3.141592653589793, -5.167712780049969, 2.5501640398773007, -0.5992645293189447,
     0.0821458865731029, -0.007370430505916944, 0.0004662998161898394,
     -2.1903497074626018e-05, 7.697826768224191e-07},
    {1.0, 1.0, 0.5000000000000019, 0.1666666666666668, 0.0416666666664881,
     0.008333333333319601, 0.0013888888952314775, 0.00019841269890047113,
     2.4801485482328494e-05, 2.755724091857897e-06, 2.763263963904103e-07,
     2.5110037605963777e-08}};

__host__ __device__ __forceinline__ double cos2pi(double x, const PolyCoefs &c) {
    const double w = x - rint(x);                // [-1/2, 1/2], exact
    const double u = fma(-2.0, fabs(w), 0.5);    // cos(2 pi w) = sin(pi u)
    const double z = u * u;
    double p = c.p[8];
#pragma unroll
    for (int i = 7; i >= 0; --i) p = fma(p, z, c.p[i]);
    return u * p;
}

__host__ __device__ __forceinline__ double sin2pi(double x, const PolyCoefs &c) {
    return cos2pi((x - rint(x)) - 0.25, c);
}

__host__ __device__ __forceinline__ void sincos2pi(double x, const PolyCoefs &c, double *s,
                                                   double *co) {
    const double f = x - rint(x);
    *co = cos2pi(f, c);
    *s = cos2pi(f - 0.25, c);
}

__host__ __device__ __forceinline__ double exp_fin(double x, const PolyCoefs &c) {
    constexpr double kRound = 6755399441055744.0;   // 1.5 * 2^52
    constexpr double kLn2Hi = 0.6931471805599453;
    constexpr double kLn2Lo = 2.3190468138462996e-17;
    // below -1100 the result is 0 anyway; the clamp keeps the integer part n in range
    // (an argument below about -1.5e9 would wrap it).  A comparison, not fmax, so that a
    // NaN argument stays NaN.
    x = (x < -1100.0) ? -1100.0 : x;
    const double t = fma(x, 1.4426950408889634, kRound);
    const double n = t - kRound;
    const int ni = (int)(uint32_t)__builtin_bit_cast(uint64_t, t);
    double r = fma(-n, kLn2Hi, x);
    r = fma(-n, kLn2Lo, r);
    double q = c.q[11];
#pragma unroll
    for (int i = 10; i >= 0; --i) q = fma(q, r, c.q[i]);
    return ldexp(q, ni);
}

}  // namespace chm

// Device-side FP64 exp / log with a fixed IEEE-754 operation sequence.
//
// The plain-FP64 form must reproduce the CPU restatement of XFluids bit for
// bit, including the chemistry step-size decisions that amplify one-ulp
// differences.  Vendor math libraries differ between host and device, so both
// sides use the same published methods:
//   exp: Cody & Waite reduction x = k ln2 + r with fused multiply-adds (ln2
//        split into the nearest double and a tail), Taylor series of exp(r) to
//        degree 13 by fused Horner steps, scaling by 2^k through the exponent
//        field in two factors.
//   log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
//        s = f/(2+f), log(1+f) = f - s (f - R), R = sum_{k=1..11} 2 s^{2k}/(2k+1).
// Accuracy: within 2 ulp of a correctly rounded result over the ranges used.
// All files including this header are compiled with -ffp-contract=off; the
// fused multiply-adds of exp are explicit fma calls on both sides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xf {

// a double constant materialised in a scalar register pair where it is used
__device__ __forceinline__ double sconst(double c)
{
    asm volatile("" : "+s"(c));
    return c;
}

__device__ __forceinline__ double dexp(double x)
{
    // the reduction runs on x clamped to [-746, 710] (NaN goes to -746), so k
    // stays in [-1076, 1024]; inputs outside the finite-result range are
    // replaced at the end
    const double xc = fmin(fmax(x, -746.0), 710.0);
    const double kf = (xc * 0x1.71547652b82fep+0 + 0x1.8p52) - 0x1.8p52;
    const int k = (int)kf;
    // ln2 = LN2_HI + LN2_LO with LN2_HI the double nearest ln2: the first fma
    // is exact (|result| < 1/2 and a multiple of 2^-54)
    double r = fma(-kf, 0x1.62e42fefa39efp-1, xc);
    r = fma(-kf, 0x1.abc9e3b39803fp-56, r);
    // the coefficients are forced into scalar registers right at their use:
    // the fma then reads them from there, where a literal in a vector
    // register per coefficient (hoisted out of loops) would cost two moves
    // and two vector registers each
    double p = 0x1.6124613a86d09p-33;
    p = fma(p, r, sconst(0x1.1eed8eff8d898p-29));
    p = fma(p, r, sconst(0x1.ae64567f544e4p-26));
    p = fma(p, r, sconst(0x1.27e4fb7789f5cp-22));
    p = fma(p, r, sconst(0x1.71de3a556c734p-19));
    p = fma(p, r, sconst(0x1.a01a01a01a01ap-16));
    p = fma(p, r, sconst(0x1.a01a01a01a01ap-13));
    p = fma(p, r, sconst(0x1.6c16c16c16c17p-10));
    p = fma(p, r, sconst(0x1.1111111111111p-7));
    p = fma(p, r, sconst(0x1.5555555555555p-5));
    p = fma(p, r, sconst(0x1.5555555555555p-3));
    p = fma(p, r, 0x1.0p-1);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    // 2^k as 2^k1 * 2^k2, both factors normal: the first product is exact and
    // the second rounds once, also for subnormal results
    const int k1 = k >> 1, k2 = k - k1;
    double e = (p * __longlong_as_double((long long)(k1 + 1023) << 52)) *
               __longlong_as_double((long long)(k2 + 1023) << 52);
    if (__builtin_expect(!(x >= -745.1332191019412 && x <= 709.782712893384), 0))
        e = x != x ? x : (x > 0.0 ? __longlong_as_double(0x7ff0000000000000LL) : 0.0);
    return e;
}

__device__ __forceinline__ double dlog(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __longlong_as_double(0x7ff8000000000000LL);
    if (x == 0.0) return -__longlong_as_double(0x7ff0000000000000LL);
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    if ((b >> 52) == 0x7ffULL) return x;
    int e = 0;
    if ((b >> 52) == 0ULL) {
        x = x * 0x1.0p54;
        b = (unsigned long long)__double_as_longlong(x);
        e = -54;
    }
    e += (int)(b >> 52) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.642c8590b2164p-4;
    R = 0x1.8618618618618p-4 + z * R;
    R = 0x1.af286bca1af28p-4 + z * R;
    R = 0x1.e1e1e1e1e1e1ep-4 + z * R;
    R = 0x1.1111111111111p-3 + z * R;
    R = 0x1.3b13b13b13b14p-3 + z * R;
    R = 0x1.745d1745d1746p-3 + z * R;
    R = 0x1.c71c71c71c71cp-3 + z * R;
    R = 0x1.2492492492492p-2 + z * R;
    R = 0x1.999999999999ap-2 + z * R;
    R = 0x1.5555555555555p-1 + z * R;
    R = z * R;
    const double t = f - s * (f - R);
    const double ed = (double)e;
    return ed * 0x1.62e42fee00000p-1 + (t + ed * 0x1.a39ef35793c76p-33);
}

// max/min with the reference's operand semantics on finite data
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

} // namespace xf

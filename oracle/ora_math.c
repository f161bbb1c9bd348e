/*
 * Deterministic FP64 exp/log for the oracle (test infrastructure only).
 *
 * The reference calls sycl::exp / sycl::log, whose results come from the SYCL
 * implementation's math library (AdaptiveCpp, unpinned) -- that boundary is
 * "parity unpinned".  To compare the CPU restatement with the GPU kernels
 * bit for bit, both sides evaluate exp and log with the same published method
 * and the same IEEE-754 operation sequence (no FMA contraction; the fused
 * multiply-adds of exp are explicit, correctly rounded fma calls):
 *
 *   exp: Cody & Waite argument reduction x = k*ln2 + r, |r| <= ln2/2, by two
 *        fma steps with ln2 split into the nearest double and a tail; exp(r)
 *        from its Taylor series to degree 13 by fused Horner steps; scaling
 *        by 2^k through the exponent bits, in two factors.
 *   log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact),
 *        s = f / (2 + f); log(1+f) = f - s*(f - R) with
 *        R = sum_{k>=1} 2 s^(2k) / (2k+1)  (the atanh series), to k = 11.
 *
 * Pinned by tests/test_oracle_math.py against libm (<= 2 ulp over wide
 * ranges, special values exact).
 */
#include "ora_internal.h"
#include <stdint.h>
#include <string.h>

static inline double bits2d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t d2bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

#define LN2_HI 0x1.62e42fee00000p-1
#define LN2_LO 0x1.a39ef35793c76p-33
#define INV_LN2 0x1.71547652b82fep+0
#define EXP_LN2_HI 0x1.62e42fefa39efp-1
#define EXP_LN2_LO 0x1.abc9e3b39803fp-56
#define RND_MAGIC 0x1.8p52
#define SQRT2 0x1.6a09e667f3bcdp+0

double ora_xexp(double x)
{
    /* reduction on x clamped to [-746, 710] (NaN goes to -746): k stays in
     * [-1076, 1024]; inputs outside the finite-result range are replaced at
     * the end */
    double xc = fmin(fmax(x, -746.0), 710.0);
    double kf = (xc * INV_LN2 + RND_MAGIC) - RND_MAGIC;   /* round to nearest, ties even */
    int k = (int)kf;
    double r = fma(-kf, EXP_LN2_HI, xc);                 /* exact: |r| < 1/2, a multiple of 2^-54 */
    r = fma(-kf, EXP_LN2_LO, r);
    double p = 0x1.6124613a86d09p-33;                     /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);                 /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);                 /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);                 /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);                 /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);                 /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);                 /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);                 /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);                  /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);                  /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);                  /* 1/3!  */
    p = fma(p, r, 0x1.0p-1);                              /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    /* 2^k = 2^k1 * 2^k2 with both factors normal: the first product is exact,
     * the second rounds once (also for subnormal results) */
    int k1 = k >> 1, k2 = k - k1;
    double e = (p * bits2d((uint64_t)(k1 + 1023) << 52)) * bits2d((uint64_t)(k2 + 1023) << 52);
    if (!(x >= -745.1332191019412 && x <= 709.782712893384))
        e = x != x ? x : (x > 0.0 ? bits2d(0x7ff0000000000000ULL) : 0.0);
    return e;
}

double ora_xlog(double x)
{
    if (x != x) return x;
    if (x < 0.0) return bits2d(0x7ff8000000000000ULL);
    if (x == 0.0) return -bits2d(0x7ff0000000000000ULL);
    uint64_t b = d2bits(x);
    if ((b >> 52) == 0x7ff) return x; /* +inf */
    int e = 0;
    if ((b >> 52) == 0) { /* subnormal: scale by 2^54 */
        x = x * 0x1.0p54;
        b = d2bits(x);
        e = -54;
    }
    e += (int)(b >> 52) - 1023;
    double m = bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0x1.642c8590b2164p-4;           /* 2/23 */
    R = 0x1.8618618618618p-4 + z * R;          /* 2/21 */
    R = 0x1.af286bca1af28p-4 + z * R;          /* 2/19 */
    R = 0x1.e1e1e1e1e1e1ep-4 + z * R;          /* 2/17 */
    R = 0x1.1111111111111p-3 + z * R;          /* 2/15 */
    R = 0x1.3b13b13b13b14p-3 + z * R;          /* 2/13 */
    R = 0x1.745d1745d1746p-3 + z * R;          /* 2/11 */
    R = 0x1.c71c71c71c71cp-3 + z * R;          /* 2/9  */
    R = 0x1.2492492492492p-2 + z * R;          /* 2/7  */
    R = 0x1.999999999999ap-2 + z * R;          /* 2/5  */
    R = 0x1.5555555555555p-1 + z * R;          /* 2/3  */
    R = z * R;
    double t = f - s * (f - R);
    double ed = (double)e;
    return ed * LN2_HI + (t + ed * LN2_LO);
}

double ora_exp(double x) { return ora_xexp(x); }
double ora_log(double x) { return ora_xlog(x); }

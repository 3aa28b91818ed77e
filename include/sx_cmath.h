/*
 * sx_cmath.h -- canonical double-precision log/exp/pow for decision-feeding values.
 *
 * The detection path feeds a few transcendental results into integer decisions:
 *   - the deblending sub-thresholds t_i = t0 * pow(peak/t0, i/N)
 *     (reference: src/cuda/cudadeblend.cu:199-217),
 *   - the ellipse correction abcor = ... / (2 pi log(t1t2) a b)
 *     (src/cuda/cudaanalyse.cu:1885-1894),
 *   - the CLEAN amplitude exponent pow(amp/dthresh, 1/CLEAN_PARAM)
 *     (src/cuda/cudaclean.cu:26-27),
 *   - the isophotal area thresholds thresh * pow(peak/thresh, i/8)
 *     (src/cuda/cudaanalyse.cu:2473-2478);
 * and the catalogue photometry uses them for its magnitudes (log10), the PHOTO-mode
 * exponentials and the star/galaxy network (cudaanalyse.cu:599-700, :2903-2921).
 * libm and the GPU math library round these differently in the last bit, so both the
 * CPU oracle and the HIP kernels use the functions below instead.  They are built only
 * from IEEE-754 +, -, *, / and exact bit manipulation, so any compiler that honours
 * IEEE double arithmetic without contraction (-ffp-contract=off) produces identical
 * bits on the host and on gfx950.
 *
 * Methods:
 *   log: argument reduction x = 2^k m, m in [sqrt(1/2), sqrt(2)); with f = m - 1 and
 *        s = f/(2+f), log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))) where
 *        R = sum_{j>=1} 2 s^{2j}/(2j+1) (the atanh series), truncated at j = 10
 *        (|s| <= 0.1716, truncation error < 1e-18 relative); k ln2 is added as a
 *        two-part constant (Cody & Waite reduction).
 *   exp: x = k ln2 + r with |r| <= ln2/2 (Cody & Waite two-part ln2), exp(r) by its
 *        Taylor series to degree 14 (truncation < 5e-18), then exact scaling by 2^k.
 *   pow: exp(y log x) for x > 0, with pow(x, 1) = x and pow(x, 0) = 1 exactly.
 * Accuracy (tested against glibc): a few ulp in double, far below the float rounding
 * the callers apply to the results.
 */
#ifndef SX_CMATH_H
#define SX_CMATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SX_HD __host__ __device__ inline
#else
#define SX_HD static inline
#endif

/* ln2 split so that k*SX_LN2_HI is exact for |k| < 2^20 */
#define SX_LN2_HI 0x1.62e42fee00000p-1
#define SX_LN2_LO 0x1.a39ef35793c76p-33

SX_HD uint64_t sx_d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
SX_HD double sx_u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

SX_HD double sx_log(double x)
{
    uint64_t u = sx_d2u(x);
    if (x != x || x < 0.0) return sx_u2d(0x7ff8000000000000ULL);       /* NaN */
    if (x == 0.0) return -sx_u2d(0x7ff0000000000000ULL);                /* -inf */
    if ((u >> 52) >= 0x7ff) return x;                                   /* +inf */
    int k = 0;
    if ((u >> 52) == 0) {                                               /* subnormal */
        x *= 0x1.0p54;
        k = -54;
        u = sx_d2u(x);
    }
    k += (int)(u >> 52) - 1023;
    double m = sx_u2d((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);   /* [1,2) */
    if (m > 1.4142135623730951) { m *= 0.5; k += 1; }                   /* [sqrt(1/2), sqrt(2)) */
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = z * (2.0 / 3.0 + z * (2.0 / 5.0 + z * (2.0 / 7.0 + z * (2.0 / 9.0 +
               z * (2.0 / 11.0 + z * (2.0 / 13.0 + z * (2.0 / 15.0 + z * (2.0 / 17.0 +
               z * (2.0 / 19.0 + z * (2.0 / 21.0))))))))));
    double hfsq = 0.5 * f * f;
    double dk = (double)k;
    return dk * SX_LN2_HI + (f - (hfsq - (s * (hfsq + R) + dk * SX_LN2_LO)));
}

SX_HD double sx_exp(double x)
{
    if (x != x) return x;
    if (x > 709.78) return sx_u2d(0x7ff0000000000000ULL);
    if (x < -745.2) return 0.0;
    double t = x * 1.4426950408889634;                 /* x / ln2 */
    double kd = t >= 0.0 ? (double)(int64_t)(t + 0.5) : -(double)(int64_t)(0.5 - t);
    int k = (int)kd;
    double r = (x - kd * SX_LN2_HI) - kd * SX_LN2_LO;
    double p = 1.0 / 87178291200.0;                    /* 1/14! */
    p = 1.0 / 6227020800.0 + r * p;                    /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    /* scale by 2^k in two exact steps so that results near the range ends stay exact */
    int k1 = k / 2, k2 = k - k1;
    double s1 = sx_u2d((uint64_t)(int64_t)(k1 + 1023) << 52);
    double s2 = sx_u2d((uint64_t)(int64_t)(k2 + 1023) << 52);
    return p * s1 * s2;
}

/* log10(x) = log(x) * (1/ln 10), the constant rounded to double; and its float form
   (the value rounded to float), for the magnitudes and the classifier inputs */
#define SX_INV_LN10 0.43429448190325182765
SX_HD double sx_log10(double x) { return sx_log(x) * SX_INV_LN10; }
SX_HD float sx_log10f(float x) { return (float)sx_log10((double)x); }

SX_HD double sx_pow(double x, double y)
{
    if (y == 1.0) return x;
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (!(x > 0.0)) {
        if (x == 0.0) return y > 0.0 ? 0.0 : sx_u2d(0x7ff0000000000000ULL);
        return sx_u2d(0x7ff8000000000000ULL);
    }
    return sx_exp(y * sx_log(x));
}

#endif

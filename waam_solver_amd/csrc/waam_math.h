/*
 * waam_math.h — single-precision exp() with a fixed, portable instruction sequence.
 *
 * The coupled step takes exp() of lattice-unit arguments in the arc heat source, the
 * arc-pressure and droplet-impact Gaussians and the surface-penetration weight
 * (reference: kernels.py:436, :489, :510, :3082, :1173). Cell flags are decided by
 * fp32 threshold tests on fields that these exponentials feed, so a GPU kernel and a
 * CPU implementation only agree bit for bit if they evaluate exp() identically. This
 * header is that one evaluation, compiled unchanged as plain C (gcc) and as HIP device
 * code (hipcc); every other operation on the path is a single IEEE-754 operation
 * (+ - * / sqrt, all correctly rounded on both targets when FMA contraction is off).
 *
 * Method (Cody & Waite, "Software Manual for the Elementary Functions", 1980):
 *   x = n*ln2 + r,   n = nearest integer to x/ln2,   |r| <= ln2/2 (+ rounding slack)
 *   ln2 is split as LN2_HI + LN2_LO with LN2_HI = 45426/65536 (15 significant bits), so
 *   n*LN2_HI is exact for |n| <= 2^8 and x - n*LN2_HI is exact (Sterbenz).
 *   e^r is the Taylor series sum_{k<=7} r^k/k! in Horner form (truncation error below
 *   4e-9 relative on |r| <= 0.35), and e^x = e^r * 2^n with the power of two built from
 *   exponent bits. Results that are subnormal are formed by two power-of-two scalings so
 *   that rounding happens exactly once.
 * Accuracy against a double-precision exp is checked in the CPU test suite.
 */
#ifndef WAAM_MATH_H
#define WAAM_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define WAAM_HD __host__ __device__ static inline
#else
#include <math.h>
#define WAAM_HD static inline
#endif

/* exp(x) is exactly 0 below this argument (exp(-103.98) < 2^-150). */
#define WAAM_EXP_ZERO_BELOW (-103.972084f)
/* exp(x) is +inf above this argument (exp(88.72284) > FLT_MAX). */
#define WAAM_EXP_INF_ABOVE (88.7228394f)

WAAM_HD float waam_f32_from_bits(uint32_t u)
{
    union { uint32_t u; float f; } v;
    v.u = u;
    return v.f;
}

/* 2^n for -126 <= n <= 127 (a normal float, exact). */
WAAM_HD float waam_pow2i(int n)
{
    return waam_f32_from_bits((uint32_t)(n + 127) << 23);
}

WAAM_HD float waam_expf(float x)
{
    if (!(x == x)) return x + x;                         /* NaN in, NaN out */
    if (x > WAAM_EXP_INF_ABOVE) return waam_f32_from_bits(0x7f800000u);
    if (x < WAAM_EXP_ZERO_BELOW) return 0.0f;

    const float LOG2E  = 1.44269502f;                     /* 1/ln2 rounded to fp32 */
    const float LN2_HI = 0.693145751953125f;              /* 45426 / 65536 */
    const float LN2_LO = 1.42860677e-06f;                 /* ln2 - LN2_HI */

    float nf = floorf(x * LOG2E + 0.5f);
    int n = (int)nf;
    float r = (x - nf * LN2_HI) - nf * LN2_LO;

    float p = 1.98412698e-4f;                             /* 1/7! */
    p = p * r + 1.38888889e-3f;                           /* 1/6! */
    p = p * r + 8.33333377e-3f;                           /* 1/5! */
    p = p * r + 4.16666679e-2f;                           /* 1/4! */
    p = p * r + 1.66666672e-1f;                           /* 1/3! */
    p = p * r + 0.5f;                                     /* 1/2! */
    p = p * r + 1.0f;
    p = p * r + 1.0f;

    if (n > 127) return (p * waam_pow2i(127)) * waam_pow2i(n - 127);
    if (n < -126) return (p * waam_pow2i(n + 64)) * waam_pow2i(-64);
    return p * waam_pow2i(n);
}

#endif /* WAAM_MATH_H */

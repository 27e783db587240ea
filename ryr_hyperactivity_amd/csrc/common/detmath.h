/*
 * Deterministic FP64 elementary functions shared by the HIP kernels and the
 * CPU oracle.  They use only IEEE-754 add/sub/mul/div, floor and integer bit
 * manipulation, evaluated with contraction off, so host (gcc, x86-64 SSE2)
 * and device (clang, gfx950) return identical bits for identical inputs.
 * That is what makes the RyR/LCC transition counts bit-exact between the GPU
 * and the oracle: every threshold the gating compares against (exp(-n p dt),
 * the V-dependent LCC rates, the Box-Muller factors) is computed here.
 *
 * Methods (textbook, written from the mathematics):
 *  exp : Cody-Waite reduction x = k ln2 + r with a 32-bit/remainder split of
 *        ln 2, |r| <= ln2/2, Horner-evaluated Taylor polynomial of degree 13
 *        (truncation < 5e-18 relative), scaling by 2^k through the exponent.
 *  log : x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *        log m = 2 atanh(s) by its odd series to s^23 (truncation < 1e-18).
 *  cos : k = nearest integer of x 2/pi, three-part Cody-Waite split of pi/2,
 *        Taylor sin/cos on |r| <= pi/4 selected by quadrant.
 *  pow : integral |y| <= 64 by right-to-left binary powering (so pow(x,2) is
 *        x*x and pow(x,3) is x*(x*x)); otherwise exp(y log x).
 * Accuracy against glibc is checked by the test suite (a few ulp).
 */
#ifndef RYR_DETMATH_H
#define RYR_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define RYR_DM __host__ __device__ inline
#else
#define RYR_DM static inline
#endif

/* RYR_STRICT_FP: no contraction of a*b+c (the exact FP64 path and the
 * oracle).  A unit defining RYR_CONTRACT_FAST (the FP32 production unit,
 * built with -ffp-contract=fast-honor-pragmas) leaves it empty, so the
 * compiler fuses there as it finds best.  RYR_EXPR_FP: fusion only inside one
 * expression, decided by the front end, so the same code inlined into two
 * kernels gives the same bits. */
#if defined(__clang__) && !defined(RYR_CONTRACT_FAST)
#define RYR_STRICT_FP _Pragma("clang fp contract(off)")
#else
#define RYR_STRICT_FP
#endif
#if defined(__clang__)
#define RYR_EXPR_FP _Pragma("clang fp contract(on)")
#else
#define RYR_EXPR_FP
#endif
/* RYR_EXACT_FP: no contraction in any unit.  The membrane step is FP64 in
 * both precisions and is formed in several places (the membrane workgroup,
 * the one-launch kernel's pending step and its per-wave V); without fusion
 * each of them rounds the same operations in the same order, so every
 * schedule gets the same bits in the FP32 unit too. */
#if defined(__clang__)
#define RYR_EXACT_FP _Pragma("clang fp contract(off)")
#else
#define RYR_EXACT_FP
#endif

RYR_DM double ryr_dm_from_bits(uint64_t u) { union { uint64_t u; double d; } c; c.u = u; return c.d; }
RYR_DM uint64_t ryr_dm_bits(double d) { union { uint64_t u; double d; } c; c.d = d; return c.u; }

/* 2^k for -1022 <= k <= 1023 */
RYR_DM double ryr_dm_pow2i(int k) { return ryr_dm_from_bits((uint64_t)(k + 1023) << 52); }

RYR_DM double ryr_exp(double x)
{
    RYR_STRICT_FP
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.1332191019412) return 0.0;
    const double ln2_hi = 0.6931471806019545;
    const double ln2_lo = -4.2009150726810846e-11;
    double kd = __builtin_floor(x * 1.4426950408889634 + 0.5);
    int k = (int)kd;
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k > 1023) return (p * 2.0) * ryr_dm_pow2i(k - 1);
    if (k < -1021) return (p * ryr_dm_pow2i(k + 1000)) * ryr_dm_pow2i(-1000);
    return p * ryr_dm_pow2i(k);
}

RYR_DM double ryr_log(double x)
{
    RYR_STRICT_FP
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int e = 0;
    uint64_t b = ryr_dm_bits(x);
    if (b < 0x0010000000000000ull) {          /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        b = ryr_dm_bits(x);
        e = -54;
    }
    e += (int)(b >> 52) - 1023;
    double m = ryr_dm_from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 0.043478260869565216;
    p = p * z + 0.047619047619047616;
    p = p * z + 0.05263157894736842;
    p = p * z + 0.058823529411764705;
    p = p * z + 0.06666666666666667;
    p = p * z + 0.07692307692307693;
    p = p * z + 0.09090909090909091;
    p = p * z + 0.1111111111111111;
    p = p * z + 0.14285714285714285;
    p = p * z + 0.2;
    p = p * z + 0.3333333333333333;
    double s2 = 2.0 * s;
    double lm = s2 + s2 * (z * p);
    double ed = (double)e;
    return ed * 0.6931471806019545 + (lm + ed * -4.2009150726810846e-11);
}

RYR_DM double ryr_cos(double x)
{
    RYR_STRICT_FP
    if (x != x) return x;
    double kd = __builtin_floor(x * 0.6366197723675814 + 0.5);
    double r = ((x - kd * 1.5707963267341256) - kd * 6.077100506303966e-11) - kd * 2.0222662487959506e-21;
    int q = ((int)(kd - 4.0 * __builtin_floor(kd * 0.25))) & 3;
    double z = r * r;
    double c = -1.5619206968586225e-16;
    c = c * z + 4.779477332387385e-14;
    c = c * z + -1.1470745597729725e-11;
    c = c * z + 2.08767569878681e-09;
    c = c * z + -2.755731922398589e-07;
    c = c * z + 2.48015873015873e-05;
    c = c * z + -0.001388888888888889;
    c = c * z + 0.041666666666666664;
    c = c * z + -0.5;
    c = c * z + 1.0;
    double s = -8.22063524662433e-18;
    s = s * z + 2.8114572543455206e-15;
    s = s * z + -7.647163731819816e-13;
    s = s * z + 1.6059043836821613e-10;
    s = s * z + -2.505210838544172e-08;
    s = s * z + 2.7557319223985893e-06;
    s = s * z + -0.0001984126984126984;
    s = s * z + 0.008333333333333333;
    s = s * z + -0.16666666666666666;
    s = r + r * (z * s);
    if (q == 0) return c;
    if (q == 1) return -s;
    if (q == 2) return -c;
    return s;
}

RYR_DM double ryr_pow(double x, double y)
{
    RYR_STRICT_FP
    if (y == __builtin_floor(y) && y <= 64.0 && y >= -64.0) {
        int n = (int)y;
        unsigned un = (unsigned)(n < 0 ? -n : n);
        double r = 1.0, base = x;
        while (un) {
            if (un & 1u) r = r * base;
            un >>= 1;
            if (un) base = base * base;
        }
        return n < 0 ? 1.0 / r : r;
    }
    if (x == 0.0) return y > 0.0 ? 0.0 : __builtin_inf();
    if (x < 0.0) return __builtin_nan("");
    return ryr_exp(y * ryr_log(x));
}

#endif /* RYR_DETMATH_H */

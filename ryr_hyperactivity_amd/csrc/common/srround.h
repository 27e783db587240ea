/*
 * Stochastic rounding of a double to float, and the counter-based random
 * bits that drive it.  Used by the FP32 production mode for the state fields
 * whose per-step increments are smaller than a float ulp of their value (the
 * NSR and JSR calcium, slow troponin and CSQN-bound calcium): each update is
 * formed in double and stored as one of the two floats around it, the upper
 * one with probability equal to the fraction of an ulp the value lies above
 * the lower one, so the stored value is an unbiased estimate of the double
 * one and increments far below an ulp still accumulate on average.
 *
 * The random bits are a function of (cell seed, step, field, CRU, sub-voxel)
 * only, so a cell's trajectory does not depend on the launch shape or on
 * which other cells share its batch.  The mixer is the 32-bit
 * xorshift-multiply finaliser "lowbias32" (C. Wellons, "Prospecting for Hash
 * Functions", 2018: constants 0x7feb352d, 0x846ca68b, shifts 16, 15, 16), a
 * bijection on 32-bit words.
 */
#ifndef RYR_SRROUND_H
#define RYR_SRROUND_H

#include <stdint.h>
#include "philox.h"

/* state fields rounded stochastically (the field word of the key) */
#define RYR_SR_CNSR 0u
#define RYR_SR_CATS 1u
#define RYR_SR_CJSR 2u
#define RYR_SR_CJSRB 3u

RYR_HD uint32_t ryr_mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* key of one field of one cell at one step (uniform over a launch's lanes) */
RYR_HD uint32_t ryr_sr_key(uint64_t seed, uint64_t step, uint32_t field)
{
    const uint32_t hi = ryr_mix32((uint32_t)(step >> 32) ^ (uint32_t)(seed >> 32) ^ (field * 0x9E3779B9u));
    return ryr_mix32(ryr_mix32((uint32_t)step ^ hi) ^ (uint32_t)seed);
}

/* 32 random bits for sub-voxel ii (0..7) of CRU n under a key */
RYR_HD uint32_t ryr_sr_bits(uint32_t key, uint32_t n, uint32_t ii)
{
    return ryr_mix32(((n << 3) | ii) ^ key);
}

/* v rounded to a neighbouring float: the 29 mantissa bits a float lacks
 * plus 29 random bits carry into the float's last bit with probability equal
 * to their fraction (the carry may move into the exponent, which is the next
 * binade's first float); the sum truncated to a float's mantissa converts
 * exactly.  Values below the float normal range round to nearest. */
RYR_HD float ryr_sr_f32(double v, uint32_t r)
{
#if defined(__HIPCC__)
    uint64_t b = (uint64_t)__builtin_bit_cast(long long, v);
#else
    union { double d; uint64_t u; } cv;
    cv.d = v;
    uint64_t b = cv.u;
#endif
    b += (uint64_t)(r >> 3);
    b &= ~(((uint64_t)1 << 29) - 1u);
#if defined(__HIPCC__)
    return (float)__builtin_bit_cast(double, (long long)b);
#else
    cv.u = b;
    return (float)cv.d;
#endif
}

#endif /* RYR_SRROUND_H */

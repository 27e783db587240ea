/*
 * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11), written from the
 * published description: two 32x32->64 multiplies per round, Weyl key bump,
 * ten rounds.  Shared by the HIP kernels and the CPU oracle so that both
 * consume the same uniforms.
 *
 * Stream layout used by the CRU update (replaces the per-CRU cuRAND XORWOW
 * state of cuda_Cell_Detailed.cu:187,1239-1240,1690,1695-1701):
 *   key     = 64-bit cell seed                 (k0 = low word, k1 = high word)
 *   counter = (block, CRU index, step lo, step hi)
 * One Philox block yields four 32-bit words = two 64-bit "slots".  Every draw
 * of the gating code (one uniform, or one 32-bit integer) consumes exactly one
 * slot, in program order, so the draw sequence of a CRU-step is addressable
 * without stored generator state.
 */
#ifndef RYR_PHILOX_H
#define RYR_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define RYR_HD __host__ __device__ inline
#else
#define RYR_HD static inline
#endif

#define RYR_PHILOX_M0 0xD2511F53u
#define RYR_PHILOX_M1 0xCD9E8D57u
#define RYR_PHILOX_W0 0x9E3779B9u
#define RYR_PHILOX_W1 0xBB67AE85u

typedef struct ryr_u32x4 { uint32_t v[4]; } ryr_u32x4;

RYR_HD ryr_u32x4 ryr_philox4x32_10(ryr_u32x4 ctr, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)RYR_PHILOX_M0 * (uint64_t)ctr.v[0];
        uint64_t p1 = (uint64_t)RYR_PHILOX_M1 * (uint64_t)ctr.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        ryr_u32x4 o;
        o.v[0] = hi1 ^ ctr.v[1] ^ k0;
        o.v[1] = lo1;
        o.v[2] = hi0 ^ ctr.v[3] ^ k1;
        o.v[3] = lo0;
        ctr = o;
        k0 += RYR_PHILOX_W0;
        k1 += RYR_PHILOX_W1;
    }
    return ctr;
}

/* 52 random bits -> double strictly inside (0,1): (x + 1/2) * 2^-52.
 * The extremes are 2^-53 and 1 - 2^-53, so log(u) and log(1-u) are finite. */
RYR_HD double ryr_slot_uniform(uint32_t a, uint32_t b)
{
    uint64_t x = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
    return ((double)x + 0.5) * 2.220446049250313080847e-16; /* 2^-52 */
}

/* Sequential draw stream of one CRU-step. */
typedef struct ryr_stream {
    uint32_t k0, k1;     /* seed */
    uint32_t cru;        /* CRU linear interior index */
    uint32_t step_lo, step_hi;
    uint32_t slot;       /* next slot to hand out */
    uint32_t w[4];       /* words of the current block */
} ryr_stream;

RYR_HD void ryr_stream_init(ryr_stream *s, uint64_t seed, uint32_t cru, uint64_t step)
{
    s->k0 = (uint32_t)seed;
    s->k1 = (uint32_t)(seed >> 32);
    s->cru = cru;
    s->step_lo = (uint32_t)step;
    s->step_hi = (uint32_t)(step >> 32);
    s->slot = 0;
    s->w[0] = s->w[1] = s->w[2] = s->w[3] = 0;
}

RYR_HD void ryr_stream_next_slot(ryr_stream *s, uint32_t *a, uint32_t *b)
{
    uint32_t sl = s->slot++;
    if ((sl & 1u) == 0u) {
        ryr_u32x4 c;
        c.v[0] = sl >> 1;
        c.v[1] = s->cru;
        c.v[2] = s->step_lo;
        c.v[3] = s->step_hi;
        ryr_u32x4 o = ryr_philox4x32_10(c, s->k0, s->k1);
        s->w[0] = o.v[0]; s->w[1] = o.v[1]; s->w[2] = o.v[2]; s->w[3] = o.v[3];
        *a = s->w[0]; *b = s->w[1];
    } else {
        *a = s->w[2]; *b = s->w[3];
    }
}

/* Draw addressing of the CRU update (one CRU-step; key = cell seed, counter =
 * (c0, CRU index, step lo, step hi); a block's words 0,1 form its first slot,
 * words 2,3 its second):
 *   - static slots 0..15, blocks c0 = 0..7 (slot s: block s >> 1, half s & 1):
 *     LCC channel l draws slot l; RyR transition t (0..7 in the order of the
 *     reference's ryrgating) draws its first uniform from slot 8 + t;
 *   - the further draws of transition t (the Knuth loop's uniforms, the
 *     Gaussian tries) are slots j = 0, 1, ... of its own substream, blocks
 *     c0 = ((t + 1) << RYR_SUB_SHIFT) | (j >> 1).
 * Each draw's address is a function of (seed, CRU, step, site, j) alone, so
 * the eight transitions are independent of each other's draw counts and can
 * be sampled in any order.  The LCC slots coincide with the first slots of
 * the sequential stream above. */
#define RYR_SUB_SHIFT 24
#define RYR_SLOT_RYR0 8

RYR_HD ryr_u32x4 ryr_block(uint64_t seed, uint32_t cru, uint64_t step, uint32_t c0)
{
    ryr_u32x4 c;
    c.v[0] = c0;
    c.v[1] = cru;
    c.v[2] = (uint32_t)step;
    c.v[3] = (uint32_t)(step >> 32);
    return ryr_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

RYR_HD uint32_t ryr_sub_block(int t, uint32_t j) { return ((uint32_t)(t + 1) << RYR_SUB_SHIFT) | (j >> 1); }

/* One uniform on (0,1) (stands where curand_uniform_double was called). */
RYR_HD double ryr_stream_uniform(ryr_stream *s)
{
    uint32_t a, b;
    ryr_stream_next_slot(s, &a, &b);
    return ryr_slot_uniform(a, b);
}

/* One 32-bit integer (stands where curand() was called). */
RYR_HD uint32_t ryr_stream_u32(ryr_stream *s)
{
    uint32_t a, b;
    ryr_stream_next_slot(s, &a, &b);
    return b;
}

#endif /* RYR_PHILOX_H */

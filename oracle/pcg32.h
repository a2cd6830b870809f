/*
 * PCG32 -- TEST INFRASTRUCTURE ONLY (shared by the oracle's C files).
 *
 * The XSH-RR 64/32 generator of M. E. O'Neill, "PCG: A Family of Simple
 * Fast Space-Efficient Statistically Good Algorithms for Random Number
 * Generation" (2014), as published in the pcg-c-basic sources
 * (pcg32_random_r, pcg32_srandom_r).  The render jobs key one stream per
 * pixel: initstate = a per-(seed, frame, sample) value, initseq = the pixel
 * index.  Pinned by the published pcg32-demo output (tests/test_oracle.py).
 */
#ifndef ORC_PCG32_H
#define ORC_PCG32_H

#include <stdint.h>

typedef struct {
    uint64_t state, inc;
} orc_pcg32;

static inline uint32_t orc_pcg_next(orc_pcg32 *r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static inline orc_pcg32 orc_pcg_seed(uint64_t initstate, uint64_t initseq)
{
    orc_pcg32 r;
    r.state = 0u;
    r.inc = (initseq << 1u) | 1u;
    orc_pcg_next(&r);
    r.state += initstate;
    orc_pcg_next(&r);
    return r;
}

/* The per-(seed, frame, sample) initstate of the render jobs' streams. */
static inline uint64_t orc_sample_seed(uint64_t seed, uint32_t frame, uint32_t sample)
{
    return seed + 0x9E3779B97F4A7C15ull * ((uint64_t)frame * 65536ull + sample + 1ull);
}

#endif

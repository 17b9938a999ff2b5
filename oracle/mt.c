/* Mersenne Twister as the reference's vendored mtwist 1.x uses it (test infrastructure).
 *
 * Follows thirdparty/mtwist.c:372-406 (mts_seed32: Knuth 69069 LCG fill from the
 * seed, 0 -> 4357, then an immediate refresh), thirdparty/mtwist.c:640-796
 * (mts_refresh: the MT19937 recurrence, N = 624, M = 397, A = 0x9908b0df) and
 * thirdparty/mtwist.h:623-634 (mt_drand: tempered word * 2^-32).
 *
 * mtwist stores its state reversed and hands words out as statevec[--stateptr];
 * with mt[i] = statevec[623 - i] that is the textbook order mt[0], mt[1], ...
 * This file keeps the textbook order (Matsumoto & Nishimura 1998, genrand with
 * the 1998 sgenrand seeding); the golden vectors dumped from the reference's own
 * mtwist.c pin the equivalence.
 */
#include "oracle.h"

#define MT_N 624
#define MT_M 397
#define MT_A 0x9908b0dfu
#define MT_UPPER 0x80000000u
#define MT_LOWER 0x7fffffffu

static void refresh(orc_mt *s)
{
    uint32_t *mt = s->mt;
    int k;
    for (k = 0; k < MT_N; ++k) {
        uint32_t y = (mt[k] & MT_UPPER) | (mt[(k + 1) % MT_N] & MT_LOWER);
        mt[k] = mt[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
    }
    s->idx = 0;
}

void orc_mt_seed(orc_mt *s, uint32_t seed)
{
    int i;
    if (seed == 0) seed = 4357u;
    s->mt[0] = seed;
    for (i = 1; i < MT_N; ++i) s->mt[i] = 69069u * s->mt[i - 1];
    refresh(s);
}

uint32_t orc_mt_u32(orc_mt *s)
{
    uint32_t y;
    if (s->idx >= MT_N) refresh(s);
    y = s->mt[s->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double orc_mt_drand(orc_mt *s)
{
    /* exact: a 32-bit integer times 2^-32 */
    return (double)orc_mt_u32(s) * (1.0 / 4294967296.0);
}

void orc_mt_fill(uint32_t seed, uint32_t *out, size_t n)
{
    orc_mt s;
    size_t i;
    orc_mt_seed(&s, seed);
    for (i = 0; i < n; ++i) out[i] = orc_mt_u32(&s);
}

/* Philox-4x32 and the uniform-float convention of the reference's random stream.
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle. Only tests/, the smoke check and bench.py's CPU-baseline leg
 * use it, as the checker; the product never links it.
 *
 * Round function and key schedule: Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11),
 * Philox-4x32 with multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl increments 0x9E3779B9 / 0xBB67AE85; the
 * reference reaches it through Triton's randint4x (triton/language/random.py:12-42, 88-110), key
 * (seed low word, seed high word), counter (offset low word, offset high word, 0, 0).
 * Uniform: u = fp32(x < 0 ? ~x : x) * fp32(4.6566127342e-10) with x the output as int32
 * (triton/language/random.py:126-143; SCK/bitwise.py:117-121).
 */
#ifndef SCK_ORACLE_PHILOX_H
#define SCK_ORACLE_PHILOX_H

#include <stdint.h>

static inline void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1, int rounds) {
    for (int r = 0; r < rounds; ++r) {
        uint64_t p0 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
        uint64_t p2 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
        uint32_t n0 = (uint32_t)(p0 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p2 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p0;
        c[3] = (uint32_t)p2;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

/* The four raw outputs of randint4x(seed, off) for a non-negative seed < 2^31 and a 64-bit offset. */
static inline void philox_raw4(int64_t seed, uint64_t off, uint32_t out[4], int rounds) {
    out[0] = (uint32_t)off;
    out[1] = (uint32_t)(off >> 32);
    out[2] = 0;
    out[3] = 0;
    philox4x32(out, (uint32_t)((uint64_t)seed & 0xFFFFFFFFu), (uint32_t)(((uint64_t)seed >> 32) & 0xFFFFFFFFu),
               rounds);
}

static inline float philox_u01(uint32_t x) {
    int32_t v = (int32_t)x;
    if (v < 0) v = ~v;
    return (float)v * 4.6566127342e-10f;
}

/* Per-sweep key of every kernel: (base_seed * 1000003 + sweep * 7919 + 12345) mod 2147483629
 * (SCK/generator.py:203-205 of the emitted source; SCK/bitwise.py:145-147). C remainder semantics. */
static inline int64_t sweep_seed(int64_t base_seed, int64_t sweep) {
    return (base_seed * 1000003LL + sweep * 7919LL + 12345LL) % 2147483629LL;
}

#endif

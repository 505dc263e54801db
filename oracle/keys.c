/* ORACLE -- test infrastructure only.
 * SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014; constants as published by Vigna) and the key
 * mapping of the heap-sort workloads:
 *   u32 key = (uint32_t)(((x >> 32) * 0xFFFFFFFF) >> 32), in [0, 2^32-2]
 *   u64 key = (u32 key << 32) | index
 * The reference draws unseeded rand() keys (test/heap_test.cu:92,111-116);
 * a seeded generator makes runs reproducible. */
#include "oracle.h"

uint64_t orc_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_keys_u32(uint64_t seed, int64_t n, uint32_t* out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t x = orc_splitmix64(&s);
        out[i] = (uint32_t)(((x >> 32) * 0xFFFFFFFFull) >> 32);
    }
}

void orc_keys_u64(uint64_t seed, int64_t n, uint64_t* out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t x = orc_splitmix64(&s);
        const uint32_t k = (uint32_t)(((x >> 32) * 0xFFFFFFFFull) >> 32);
        out[i] = ((uint64_t)k << 32) | (uint64_t)(uint32_t)i;
    }
}

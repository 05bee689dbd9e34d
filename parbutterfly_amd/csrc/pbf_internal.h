// Internal helpers shared by the host and device translation units of libparbutterfly_amd.
#pragma once
#include <stdint.h>

// SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014), used as a counter-based hash: draw k of stream s under seed.
static inline uint64_t pbf_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t pbf_rng(uint64_t seed, uint64_t stream, uint64_t k) {
  return pbf_mix64(pbf_mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + k * 0x9E3779B97F4A7C15ull);
}
// Uniform double in [0, 1) from the top 53 bits.
static inline double pbf_u01(uint64_t r) { return double(r >> 11) * (1.0 / 9007199254740992.0); }
// Uniform integer in [0, n) by multiply-shift (Lemire).
static inline uint64_t pbf_bounded(uint64_t r, uint64_t n) {
  return (uint64_t)(((unsigned __int128)r * n) >> 64);
}

// wedgeF of cmpWedgeCounts (butterfly_utils.h:1061-1069): C(d,2) of a degree d, with the
// product d * (d - 1) taken in the 32-bit degree type as there, i.e. modulo 2^32 (it wraps
// for d > 65536) before the halving.  The side choice compares 64-bit sums of these terms, so
// it is the reference's choice on every graph.
static inline int64_t pbf_side_term(uint32_t d) { return (int64_t)((uint32_t)(d * (d - 1u)) / 2u); }

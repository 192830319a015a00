// Deterministic pseudo-random numbers for the workload generators.
//
// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) seeds and mixes; xoshiro256** (Blackman, Vigna,
// "Scrambled linear pseudorandom number generators", 2018) draws streams.
// Only integer arithmetic decides a draw, so a seed reproduces the same graph
// on any x86-64 host.
#pragma once
#include <cstdint>

namespace stc {

static inline uint64_t splitmix64_next(uint64_t &state) {
  uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Stateless mix of one word (one SplitMix64 step from `x`).
static inline uint64_t mix64(uint64_t x) {
  uint64_t s = x;
  return splitmix64_next(s);
}

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix64_next(st);
  }
  static inline uint64_t rotl(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  inline uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Uniform integer in [0, bound) for bound < 2^32 (multiply-high mapping).
  inline uint32_t below(uint32_t bound) {
    return (uint32_t)(((next() >> 32) * (uint64_t)bound) >> 32);
  }
  // 53-bit uniform in [0, 1).
  inline double unit() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Integer threshold for "probability p" on a 53-bit draw.
static inline uint64_t prob_threshold53(double p) {
  if (p <= 0.0) return 0;
  if (p >= 1.0) return 1ULL << 53;
  return (uint64_t)(p * 9007199254740992.0);
}

// Sign of the undirected edge {a,b}: -1 with the probability encoded in
// `thr53`, a pure function of (seed, min, max).
static inline int8_t edge_sign(uint64_t seed, uint32_t a, uint32_t b,
                               uint64_t thr53) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  uint64_t h = mix64(seed ^ mix64(((uint64_t)lo << 32) | hi));
  return ((h >> 11) < thr53) ? (int8_t)-1 : (int8_t)1;
}

}  // namespace stc

// Pseudo-random generators used by the synthetic graph generators.
//
// SplitMix64: Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014 (the 64-bit finaliser with constants
// 0x9E3779B97F4A7C15 / 0xBF58476D1CE4E5B9 / 0x94D049BB133111EB as published
// by Vigna in the xoshiro reference page).
// xoshiro256**: Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 2021.  State is seeded from SplitMix64 as the
// authors recommend.
//
// The clique hash also uses splitmix64_mix(), so the device code includes
// this header (all functions are host+device).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define RDMCE_HD __host__ __device__
#else
#define RDMCE_HD
#endif

namespace rdmce {

// One SplitMix64 output for input state x (state advanced by the golden gamma).
RDMCE_HD inline uint64_t splitmix64_mix(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t r = splitmix64_mix(s);
    s += 0x9E3779B97F4A7C15ull;
    return r;
  }
};

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    SplitMix64 sm(seed);
    for (auto &w : s) w = sm.next();
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
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
  // Uniform double in [0,1) from the top 53 bits.
  double next_double() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Uniform integer in [0, bound) by Lemire's multiply-shift with rejection.
  uint64_t next_below(uint64_t bound) {
    if (bound <= 1) return 0;
    for (;;) {
      __uint128_t m = (__uint128_t)next() * bound;
      uint64_t lo = (uint64_t)m;
      if (lo >= bound || lo >= (uint64_t)(-bound) % bound) return (uint64_t)(m >> 64);
    }
  }
};

// Stream seed for (seed, stream index): independent generators for parallel
// work whose output does not depend on the thread count.
inline uint64_t stream_seed(uint64_t seed, uint64_t stream) {
  return splitmix64_mix(seed ^ splitmix64_mix(stream + 0x632BE59BD9B4E019ull));
}

}  // namespace rdmce

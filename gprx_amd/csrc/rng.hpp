// SplitMix64-seeded Xoshiro256++ (src/rng.rs:15-87), host side: the SVGP
// epoch shuffle and the optimizer's restart draws.
#pragma once

#include <cstdint>

namespace gprx {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) {
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      v = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Uniform in [0, mx] without modulo bias (Lemire's widening multiply).
  uint64_t up_to(uint64_t mx) {
    const uint64_t range = mx + 1;
    if (range == 0) return next();
    const uint64_t threshold = (0 - range) % range;
    for (;;) {
      const unsigned __int128 wide = (unsigned __int128)next() * range;
      if ((uint64_t)wide >= threshold) return (uint64_t)(wide >> 64);
    }
  }
  // (next >> 11) 2^-53 in [0, 1)
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // unit() clamped into [2^-53, 1 - 2^-53] (src/rng.rs:71-81)
  double open_unit() {
    const double u = unit();
    const double eps = 1.0 / 9007199254740992.0;
    return u <= eps ? eps : (u >= 1.0 - eps ? 1.0 - eps : u);
  }
};

}  // namespace gprx

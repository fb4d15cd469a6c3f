// gprx's SeededRng (src/rng.rs:15-87) for drawing the synthetic workloads at
// full size: SplitMix64 seeding, Xoshiro256++ output and state update,
// unit = (next >> 11) 2^-53, and the Box-Muller normal on open units.  The
// caller owns the four-word state, so a stream continues across calls in the
// reference's draw order.
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace {

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

inline uint64_t next(uint64_t* s) {
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

constexpr double kUnit = 1.0 / 9007199254740992.0;  // 2^-53

inline double unit(uint64_t* s) { return (double)(next(s) >> 11) * kUnit; }

inline double open_unit(uint64_t* s) {
  const double u = unit(s);
  return u <= kUnit ? kUnit : (u >= 1.0 - kUnit ? 1.0 - kUnit : u);
}

}  // namespace

extern "C" {

void gprx_rng_seed(uint64_t seed, uint64_t* s) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) {
    x += 0x9e3779b97f4a7c15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s[i] = z ^ (z >> 31);
  }
}

void gprx_rng_units(uint64_t* s, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = unit(s);
}

void gprx_rng_normals(uint64_t* s, double* out, size_t n) {
  const double two_pi = 6.283185307179586;
  for (size_t i = 0; i < n; ++i) {
    double u1 = open_unit(s);
    if (u1 < 2.2250738585072014e-308) u1 = 2.2250738585072014e-308;
    const double u2 = open_unit(s);
    out[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
  }
}

}  // extern "C"

// hc_rng.h - counter-based uniform generator for the vent heat-flux noise.
//
// The reference draws the per-cell noise of SOURCE_terms from the Fortran
// runtime's random_number stream (source80.inc:40-42), which cannot be
// reproduced on a device.  Callers that need that stream pass the draws in
// (hc_source_terms); the native loop uses this stateless generator instead:
// draw r of patch cell n at step m is a SplitMix64 hash of (seed, m, n, r),
// identical on host and device.
#pragma once
#include <stdint.h>
#ifdef __HIPCC__
#define HC_HD __host__ __device__
#else
#define HC_HD
#endif

namespace hc {
HC_HD inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in [0,1) with 53 random bits
HC_HD inline double hash_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = splitmix64(seed ^ splitmix64(a ^ splitmix64(b ^ splitmix64(c))));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
}  // namespace hc

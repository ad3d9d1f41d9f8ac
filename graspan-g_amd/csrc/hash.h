// Per-edge hash used by the closure checksum (host and device).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define GG_HD __host__ __device__
#else
#define GG_HD
#endif

namespace gg {

// triple_hash(s, d, l) = hash_mix(s * kHashS ^ d * kHashD ^ l * kHashL)
// (64-bit products; hash_mix is the splitmix64 finalizer).
constexpr uint64_t kHashS = 0x9E3779B97F4A7C15ull, kHashD = 0xC2B2AE3D27D4EB4Full, kHashL = 0x165667B19E3779F9ull;

GG_HD inline uint64_t hash_mix(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

GG_HD inline uint64_t triple_hash(uint32_t s, uint32_t d, uint32_t l) {
  return hash_mix(uint64_t(s) * kHashS ^ uint64_t(d) * kHashD ^ uint64_t(l) * kHashL);
}

}  // namespace gg

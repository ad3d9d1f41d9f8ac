// Per-edge hash used by the closure checksum (host and device).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define GG_HD __host__ __device__
#else
#define GG_HD
#endif

namespace gg {

GG_HD inline uint64_t triple_hash(uint32_t s, uint32_t d, uint32_t l) {
  uint64_t x = uint64_t(s) * 0x9E3779B97F4A7C15ull ^ uint64_t(d) * 0xC2B2AE3D27D4EB4Full ^
               uint64_t(l) * 0x165667B19E3779F9ull;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

}  // namespace gg

// SplitMix64 stream shared by the generators and the seeded cost perturbation.
#pragma once
#include <stdint.h>

namespace pds {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next_u64() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // U[0,1) from the top 53 bits.
  double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace pds

// rng.hpp -- SplitMix64 (Steele, Lea, Flood 2014), the generator behind every
// seeded input of this build (states, circuits, measurement shots).
#pragma once
#include <cstdint>

namespace sq {

struct SplitMix64 {
    uint64_t x;
    uint64_t next() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }   // [0, 1)
};

}  // namespace sq

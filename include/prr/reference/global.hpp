// global.hpp — the reference's global.hpp names (global.hpp:9-50) for scene
// scripts: EPSILON, kInfinity, clamp and get_random_float (a thread-local
// SplitMix64, Steele et al. 2014, seeded from std::random_device; top 24 bits).
#pragma once
#include <cstdint>
#include <limits>
#include <random>

#include "../prr.hpp"

inline constexpr const float EPSILON = 1e-3f;
const float kInfinity = std::numeric_limits<float>::max();

inline float clamp(const float &lo, const float &hi, const float &v) { return std::max(lo, std::min(hi, v)); }

inline float get_random_float() {
    thread_local uint64_t state = [] {
        std::random_device rd;
        return (uint64_t) rd() << 32 | rd();
    }();
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= (z >> 31);
    return (float) (z >> 40) * 0x1p-24f;
}

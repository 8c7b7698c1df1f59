// ref_seed.h — force-included (g++ -include) into every reference translation
// unit that oracle/Makefile compiles for the golden-vector harness
// (oracle/_ref/ref_harness). TEST INFRASTRUCTURE ONLY.
//
// The reference seeds its thread-local SplitMix64 from std::random_device on
// the first draw of each thread (global.hpp:35-41), which makes its outputs
// irreproducible. This header keeps the reference sources untouched and
// swaps in a seedable device: every call returns the 32-bit value v set by the
// harness, so a fresh thread starts from state (v << 32) | v regardless of
// the evaluation order of the two rd() calls. In counter mode each call
// returns a fresh value (used for statistical images only).
#pragma once
#include <atomic>
#include <cstdint>
#include <random>

namespace prr_refseed {
inline std::atomic<uint32_t> &value() {
    static std::atomic<uint32_t> v{0x12345678u};
    return v;
}
inline std::atomic<int> &counter_mode() {
    static std::atomic<int> m{0};
    return m;
}
inline std::atomic<uint32_t> &counter() {
    static std::atomic<uint32_t> c{1u};
    return c;
}
}  // namespace prr_refseed

namespace std {
struct prr_seeded_device {
    typedef unsigned int result_type;
    unsigned int operator()() {
        if (prr_refseed::counter_mode().load()) {
            uint32_t c = prr_refseed::counter().fetch_add(1u);
            return c * 2654435761u ^ 0x9E3779B9u;
        }
        return prr_refseed::value().load();
    }
};
}  // namespace std

#define random_device prr_seeded_device

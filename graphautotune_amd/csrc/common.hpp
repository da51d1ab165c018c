// Internal helpers shared by the host-side C++ of libgat.so.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gat.h"

namespace gat {

using VIndex = int32_t;  // include/core/types.hpp:4
using EIndex = int64_t;  // include/core/types.hpp:5

// Thread-local error message behind gat_last_error().
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *last_error();

// Status carrier used inside the library; converted to a return code at the
// C boundary.
struct Status {
    int code = GAT_OK;
    bool ok() const { return code == GAT_OK; }
};

#define GAT_FAIL(code, ...)      \
    do {                         \
        ::gat::set_error(__VA_ARGS__); \
        return (code);           \
    } while (0)

// splitmix64 + xoshiro256** (Blackman & Vigna, public domain algorithms).
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &w : s) w = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1) with 53 random bits
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, n) (Lemire's nearly-divisionless method)
    uint64_t below(uint64_t n) {
        __uint128_t m = (__uint128_t)next() * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            uint64_t t = (0 - n) % n;
            while (l < t) {
                m = (__uint128_t)next() * n;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    // Jump function: equivalent to 2^128 calls to next(); gives independent
    // streams for parallel generation.
    void jump() {
        static const uint64_t J[] = {0x180ec6d33cfd0abaull, 0xd5a61266f0c9392cull,
                                     0xa9582618e03fc9aaull, 0x39abdc4529b1661cull};
        uint64_t t[4] = {0, 0, 0, 0};
        for (uint64_t j : J)
            for (int b = 0; b < 64; b++) {
                if (j & (1ull << b))
                    for (int k = 0; k < 4; k++) t[k] ^= s[k];
                next();
            }
        for (int k = 0; k < 4; k++) s[k] = t[k];
    }
};

}  // namespace gat

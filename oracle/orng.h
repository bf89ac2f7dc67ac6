// ORACLE -- test infrastructure only (see ovec.h).
//
// The shared counter-based generator that replaces the reference's per-thread
// std::mt19937 + uniform_real_distribution<float>(1e-6, 1-1e-6)
// (src/component.cpp:5-10).  Philox-4x32-10 as defined in Salmon et al.,
// SC'11 ("Parallel random numbers: as easy as 1, 2, 3"), checked against the
// Random123 known-answer vectors in the tests.
//
// Stream definition (identical on the GPU): key = (seed, tag); draw n of the
// stream (pixel p, sample s) is word (n & 3) of philox(ctr = {p, s, n >> 2, 0}).
// Word -> float: ((w >> 8) * 2^-24) mapped affinely onto [1e-6, 1-1e-6].
#pragma once
#include <cstdint>

namespace ref {

inline void philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum : uint32_t {
    TAG_SCENE = 0x5C3E0000u,
    TAG_RAYS = 0x5C3E1000u,
    TAG_DIRECT = 0x0D1EC700u,
    TAG_INDIRECT = 0x1D1EC700u
};

// Drop-in for the reference's Generator (operator() returns the next float).
struct Generator {
    uint32_t key[2];
    uint32_t pix, smp, n;
    Generator(uint32_t seed, uint32_t tag, uint32_t pixel, uint32_t sample) : pix(pixel), smp(sample), n(0) {
        key[0] = seed;
        key[1] = tag;
    }
    float operator()() {
        uint32_t c[4] = {pix, smp, n >> 2, 0u}, o[4];
        philox4x32_10(c, key, o);
        uint32_t w = o[n & 3u];
        ++n;
        const float a = 1e-6f, b = 1.0f - 1e-6f;
        float u = float(w >> 8) * (1.0f / 16777216.0f);
        return a + (b - a) * u;
    }
};

}  // namespace ref

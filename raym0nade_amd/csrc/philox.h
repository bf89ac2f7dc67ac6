// Counter-based random numbers shared by the host scene builders and the
// gfx950 kernels: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), written from the paper's definition.
//
// The reference draws every random number from one std::mt19937 per CPU thread
// (src/component.cpp:5-10, seeded with the thread index at src/render.cpp:614-617),
// so its sample stream depends on scheduling.  Here a draw is a pure function of
// (key, counter): key = (seed, stream tag), counter = (pixel, sample, block, lane
// group), which makes every path replayable on any device.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RM_HD __host__ __device__ __forceinline__
#else
#define RM_HD inline
#endif

namespace rm {

struct U4 { uint32_t x, y, z, w; };

RM_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
    return uint32_t((uint64_t(a) * uint64_t(b)) >> 32);
}

// Ten Philox rounds; multipliers and Weyl key increments from the SC'11 paper.
RM_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
        uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// 24 high bits -> [0, 1).
RM_HD float u01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// The reference's Generator is uniform_real_distribution<float>(1e-6, 1-1e-6)
// (src/component.cpp:5-6); the same open range is kept.
RM_HD float gen_range(uint32_t x) {
    const float a = 1e-6f, b = 1.0f - 1e-6f;
    return a + (b - a) * u01(x);
}

// Stream tags (second key word together with the user seed in the first).
enum : uint32_t {
    TAG_SCENE = 0x5C3E0000u,   // procedural geometry / materials
    TAG_RAYS = 0x5C3E1000u,    // synthetic ray batches
    TAG_DIRECT = 0x0D1EC700u,  // primary-vertex light samples, counter.y = sample index
    TAG_INDIRECT = 0x1D1EC700u // indirect paths, counter.y = sample index
};

// Sequential draws from one (key, pixel, sample) stream: draw n comes from
// block n>>2, lane n&3.
struct Stream {
    uint32_t k0, k1, pix, smp, n;
    RM_HD float next() {
        U4 r = philox4x32_10(U4{pix, smp, n >> 2, 0u}, k0, k1);
        uint32_t l = n & 3u;
        ++n;
        uint32_t w = l == 0 ? r.x : (l == 1 ? r.y : (l == 2 ? r.z : r.w));
        return gen_range(w);
    }
};

}  // namespace rm

// The reference's 8-bit textures (ImageData, include/material.h:12-24;
// src/material.cpp:26-148) for host and gfx950 code: bytes in the texture's
// channel layout, integer mip levels, and samples of bytes / 255 read with the
// sampler's own stride (4 for vec4 lookups, 3 for vec3 ones), whatever the
// texture's channel count -- as the reference reads them.  A read past a
// level's last byte (undefined in the reference: a 4-byte read of 3-channel
// data) reads 0.
#pragma once
#include <math.h>
#include <stdint.h>
#include "philox.h"
#include "rm_layout.h"
#include "rm_vec.h"

namespace rm {

// get_basic (src/material.cpp:26-46): bw bytes at `index`, / 255.
RM_HD inline void tex8_texel(const uint8_t *lv, int64_t n, int64_t index, int bw, float out[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int64_t i = index + c;
        out[c] = (c < bw && i < n) ? static_cast<float>(lv[i]) / 255.0f : 0.0f;
    }
}

// get_bilinear (src/material.cpp:48-79): wrap-around bilinear, stride bw.
RM_HD inline void tex8_bilinear(const uint8_t *lv, int64_t n, int w, int h, float u, float v, int bw, float out[4]) {
    float x = u * float(w), y = v * float(h);
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float dx = x - float(x0), dy = y - float(y0);
    if (x0 < 0 || x0 >= w) { x0 %= w; if (x0 < 0) x0 += w; }
    if (y0 < 0 || y0 >= h) { y0 %= h; if (y0 < 0) y0 += h; }
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    float c00[4], c01[4], c10[4], c11[4];
    tex8_texel(lv, n, ((int64_t)y0 * w + x0) * bw, bw, c00);
    tex8_texel(lv, n, ((int64_t)y0 * w + x1) * bw, bw, c01);
    tex8_texel(lv, n, ((int64_t)y1 * w + x0) * bw, bw, c10);
    tex8_texel(lv, n, ((int64_t)y1 * w + x1) * bw, bw, c11);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float a = c00[c] * (1.0f - dx) + c01[c] * dx;
        float b = c10[c] * (1.0f - dx) + c11[c] * dx;
        out[c] = a * (1.0f - dy) + b * dy;
    }
}

// ImageData::get<Vec> (src/material.cpp:81-94) with the sampler stride bw (3
// for vec3, 4 for vec4): v flipped, depth clamped to [0, levels - 1], the
// two nearest levels blended.
RM_HD inline void tex8_get(const uint8_t *pool, const TexRec &T, float u, float v, float depth, int bw, float out[4]) {
    v = 1.0f - v;
    depth = gmax(gmin(depth, float(T.levels - 1)), 0.0f);  // std::max(std::min(depth, levels - 1), 0)
    int level = (int)depth;
    int next = level + 1 < T.levels - 1 ? level + 1 : T.levels - 1;
    float blend = depth - float(level);
    float a[4], b[4];
    tex8_bilinear(pool + T.offset[level], T.level_bytes[level], T.width >> level, T.height >> level, u, v, bw, a);
    tex8_bilinear(pool + T.offset[next], T.level_bytes[next], T.width >> next, T.height >> next, u, v, bw, b);
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = a[c] * (1.0f - blend) + b[c] * blend;
}

}  // namespace rm

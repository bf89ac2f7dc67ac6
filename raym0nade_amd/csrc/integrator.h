// Pieces of renderPixel (src/render.cpp:439-501) shared by the wavefront
// integrator and its adjoint: render arguments as the kernels read them, the
// frame's buffers, the camera ray with its differentials and the red nudge of
// near-grey primary albedo.
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/raym0nade_hip.h"
#include "shade.h"

namespace rm {

struct RArgs {
    V3 pos, dir, up, right;
    float acc, exposure, p_direct;
    int w, h, spp, max_depth, rr_min, K;
    uint32_t seed;
};

// RenderArgs as the kernels read them; max_depth <= 0 means 16 and
// rr_min_depth <= 0 means 1.  K (paths per pixel per round) is left to the caller.
inline RArgs rargs_from(const rm_render_args *a) {
    RArgs A{};
    A.pos = V3{a->position[0], a->position[1], a->position[2]};
    A.dir = V3{a->direction[0], a->direction[1], a->direction[2]};
    A.up = V3{a->up[0], a->up[1], a->up[2]};
    A.right = V3{a->right[0], a->right[1], a->right[2]};
    A.acc = a->accuracy;
    A.exposure = a->exposure;
    A.p_direct = a->p_direct;
    A.w = a->width;
    A.h = a->height;
    A.spp = a->spp;
    A.max_depth = a->max_depth > 0 ? a->max_depth : 16;
    A.rr_min = a->rr_min_depth > 0 ? a->rr_min_depth : 1;
    A.seed = a->seed;
    A.K = 1;
    return A;
}

struct FrameView {
    float *pos, *shapeN, *surfN, *emis, *base, *Dd, *Ds, *Id, *Is;
    float *mat;     // per pixel (metallic, roughness, opacity) of the primary hit
    int32_t *drop;  // per pixel: name (s * 8 + k) of the dropped indirect light sample, or -1
    float4 *rec;    // null, or per (pixel, indirect sample): sum dd | sum ds | sum w L
    int32_t rec_stride;
    unsigned long long *stats;  // RM_STAT_COUNT counters of the render (rm_frame_get_stats)
};

__device__ __forceinline__ void st3(float *p, int64_t i, V3 v) {
    p[i * 3] = v.x;
    p[i * 3 + 1] = v.y;
    p[i * 3 + 2] = v.z;
}

// renderPixel's camera ray and initRayDiff (src/render.cpp:439-477)
__device__ __forceinline__ void camera_ray(const RArgs &A, int x, int y, V3 &Dir, Diff &diff) {
    float rx = float(x) - float(A.w) / 2.0f, ry = float(y) - float(A.h) / 2.0f;
    V3 d = A.dir + A.acc * (rx * A.right + ry * A.up);
    diff.dPdx = v3(0.0f);
    diff.dPdy = v3(0.0f);
    V3 ddx = A.acc * A.right, ddy = A.acc * A.up;
    float dd = dot(d, d), dx = dot(d, ddx), dy = dot(d, ddy);
    diff.dDdx = (dd * ddx - d * dx) / (sqrtf(dd) * dd);
    diff.dDdy = (dd * ddy - d * dy) / (sqrtf(dd) * dd);
    Dir = normalize(d);
}

// Red nudge of a near-grey or near-black primary albedo (src/render.cpp:492-495);
// the reference compares against double literals and adds 4e-2 in double.
__device__ __forceinline__ void nudge_base(V3 &b) {
    float c0 = dot(b, W_RGB());
    if ((double)c0 < 2e-2 || (c0 < 0.8f && (double)length(b / c0 - v3(1.0f)) < 2e-2))
        b.x = (float)((double)b.x + 4e-2);
}

// sampleCount (src/render.cpp:126): light samples of a terminating vertex at
// depth d, 0 1 2 2 3 3 3 4 4 4 4 5 5 5 5 5 6 for d = 0..16.
__device__ __forceinline__ int sample_count(int d) {
    int k = 0;
    while ((k + 1) * k / 2 + 1 <= d) ++k;
    return k;
}

}  // namespace rm

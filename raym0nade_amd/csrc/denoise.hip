// Denoise behind rm_spatial_clamp and rm_filter: Photo::spatialClamp
// (src/image.cpp:30-82) and Photo::filter (src/image.cpp:84-213) over the
// frame's four demodulated radiance buffers (Dd, Ds, Id, Is as float4 rgb +
// Var, one 16-byte row per pixel).
//
// Layout and passes (all one lane per pixel, blocks renumbered so each XCD
// owns one contiguous band of rows, which keeps every stencil's neighbour
// rows in that XCD's L2):
//   spatialClamp  k_clamp_rows: horizontal 7-tap luminance blur of all four
//                 buffers into a scratch plane; k_clamp_cols: vertical blur
//                 and the in-place scaling.  Two kernels because a pixel's
//                 clamp must see its neighbours' unclamped luminance.
//   filter        k_filter_var: 3x3 variance estimate into the ping-pong
//                 buffer; k_atrous x5 (strides 1, 2, 4, 8, 16) alternating
//                 between the ping-pong buffer and the frame, so the fifth
//                 pass lands in the frame.  One pass serves all four buffers:
//                 the G-buffer terms of a tap (normal power, depth slope,
//                 material distance) are computed once and shared, only the
//                 radiance-distance term is per buffer.
// Every arithmetic step keeps the reference's operation order.  spatialClamp
// and filterVar round as the reference does.  In the a-trous taps the square
// roots, divisions and exponentials use the hardware's approximate
// instructions (v_sqrt, v_rsq, v_rcp, v_exp: about 1 ulp each, where the
// exactly rounded sequences cost ten times as many instructions): the tap
// weights move by about 1e-6 relative, and the filter is compared with the
// reference within 1e-4 relative.
#include <hip/hip_runtime.h>
#include "rm_frame.h"
#include "rm_vec.h"

namespace rm {

constexpr int DB = 256;

struct Quad {
    float4 *b[4];  // Dd, Ds, Id, Is
};

struct GView {
    const float *pos, *shapeN, *surfN, *mat;
};

__device__ __forceinline__ V3 rgb(float4 q) { return V3{q.x, q.y, q.z}; }
__device__ __forceinline__ float lum(float4 q) { return dot(rgb(q), V3{0.3f, 0.6f, 0.1f}); }
__device__ __forceinline__ V3 ld3g(const float *p, uint32_t i) {
    return V3{p[(size_t)i * 3], p[(size_t)i * 3 + 1], p[(size_t)i * 3 + 2]};
}

// pow(c, 1024) for c in [0, 1] (src/image.cpp:120) as ten squarings in fp64:
// the relative error before the final rounding is below 2^-43, so the result
// is the correctly rounded fp32 power except within 2^-43 of a rounding
// midpoint -- what a correctly rounded host powf returns -- and it costs a
// third of the device powf.
__device__ __forceinline__ float pow1024(float c) {
    double d = (double)c;
#pragma unroll
    for (int k = 0; k < 10; ++k) d = d * d;
    return (float)d;
}

// Blocks of one launch cover 4 buffers x nb blocks of pixels; returns the
// buffer and the pixel, or false past the end.
__device__ __forceinline__ bool quad_pixel(uint32_t nb, uint32_t n, int &b, uint32_t &i) {
    uint32_t L = xcd_block(blockIdx.x, gridDim.x);
    b = (int)(L / nb);
    i = (L % nb) * DB + threadIdx.x;
    return i < n;
}

// src/image.cpp:42-52
__global__ __launch_bounds__(DB) void k_clamp_rows(Quad R, int w, uint32_t n, uint32_t nb, float *__restrict__ b1) {
    int b;
    uint32_t i;
    if (!quad_pixel(nb, n, b, i)) return;
    const float h[7] = {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f};
    const float4 *r = R.b[b];
    int x = (int)(i % (uint32_t)w);
    float sum = 0.0f;
#pragma unroll
    for (int dx = -3; dx <= 3; ++dx) {
        int nx = x + dx;
        if (nx >= 0 && nx < w) sum += lum(r[(int64_t)i + dx]) * h[dx + 3];
    }
    b1[(size_t)b * n + i] = sum;
}

// src/image.cpp:54-74
__global__ __launch_bounds__(DB) void k_clamp_cols(Quad R, int w, int hgt, uint32_t n, uint32_t nb,
                                                   const float *__restrict__ b1) {
    int b;
    uint32_t i;
    if (!quad_pixel(nb, n, b, i)) return;
    const float h[7] = {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f};
    const float *c = b1 + (size_t)b * n;
    int y = (int)(i / (uint32_t)w);
    float sum = 0.0f;
#pragma unroll
    for (int dy = -3; dy <= 3; ++dy) {
        int ny = y + dy;
        if (ny >= 0 && ny < hgt) sum += c[(int64_t)i + (int64_t)dy * w] * h[dy + 3];
    }
    float4 q = R.b[b][i];
    float l = lum(q);
    float blurred = sum - h[3] * h[3] * l;
    if (l > 36.0f * blurred) {
        float s = blurred / (l + 1e-4f) / (1.0f - h[3] * h[3]);
        R.b[b][i] = make_float4(q.x * s, q.y * s, q.z * s, q.w);
    }
}

// src/image.cpp:84-107: Var <- weighted Var + E[L.L] - E[L].E[L] over 3x3.
__global__ __launch_bounds__(DB) void k_filter_var(Quad in, Quad out, int w, int hgt, uint32_t n, uint32_t nb) {
    int b;
    uint32_t i;
    if (!quad_pixel(nb, n, b, i)) return;
    const float h[3] = {0.25f, 0.5f, 0.25f};
    const float4 *r = in.b[b];
    int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
    V3 E = v3(0.0f);
    float E2 = 0.0f, V = 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            int nx = x + dx, ny = y + dy;
            if (nx < 0 || nx >= w || ny < 0 || ny >= hgt) continue;
            float4 q = r[(int64_t)i + (int64_t)dy * w + dx];
            float wt = h[dx + 1] * h[dy + 1];
            E += rgb(q) * wt;
            E2 += dot(rgb(q), rgb(q)) * wt;
            V += q.w * wt;
        }
    float4 c = r[i];
    out.b[b][i] = make_float4(c.x, c.y, c.z, V + E2 - dot(E, E));
}

// Approximate square root, reciprocal square root, reciprocal and e^x of the
// a-trous taps (about 1 ulp; e^x as 2^(x log2 e)).
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// src/image.cpp:109-189 for the four buffers at once.  Buffers 1 and 3 (Ds,
// Is) are the specular ones.
// The taps' sums may use fused multiply-adds (one rounding fewer per term,
// inside the same tolerance as the approximate instructions).
__global__ __launch_bounds__(DB) void k_atrous(GView G, Quad in, Quad out, int w, int hgt, uint32_t n, int step) {
#pragma clang fp contract(fast)
    uint32_t i = xcd_block(blockIdx.x, gridDim.x) * DB + threadIdx.x;
    if (i >= n) return;
    const V3 pp = ld3g(G.pos, i);
    if (!any_finite(pp)) {  // :169-170, the pixel keeps the buffers' zero
#pragma unroll
        for (int b = 0; b < 4; ++b) out.b[b][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float h[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    const V3 sn = ld3g(G.shapeN, i), un = ld3g(G.surfN, i);
    const float mp = G.mat[(size_t)i * 3], rp = G.mat[(size_t)i * 3 + 1], op = G.mat[(size_t)i * 3 + 2];
    const float rw = (rp < 0.04f) ? 0.04f : rp;  // std::max(Gp.roughness, eps_r)
    float4 Lp[4];
    float inv_den[4], ws[4], va[4];
    V3 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        Lp[b] = in.b[b][i];
        inv_den[b] = 1.0f / (1.0f * sqrtf(Lp[b].w) + 1e-2f);  // 1 / (sigma_l * sqrt(Lp.Var) + eps)
        ws[b] = 0.0f;
        va[b] = 0.0f;
        acc[b] = v3(0.0f);
    }
    const int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int ny = y + dy * step;
        if (ny < 0 || ny >= hgt) continue;
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int nx = x + dx * step;
            if (nx < 0 || nx >= w) continue;
            const uint32_t q = (uint32_t)ny * (uint32_t)w + (uint32_t)nx;
            const float base = h[dx + 2] * h[dy + 2];
            float4 Lq[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) Lq[b] = in.b[b][q];
            float wt[4] = {base, base, base, base};
            if (dx != 0 || dy != 0) {
                // getWeight (:109-156); g stays 0 on every early return
                float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                const V3 pq = ld3g(G.pos, q);
                if (any_finite(pq)) {
                    const float c = dot(un, ld3g(G.surfN, q));
                    const float pw = pow1024((0.0f < c) ? c : 0.0f);
                    if (!(pw < 1e-6f)) {
                        const V3 dd = pq - pp;
                        const V3 dir = dd * frsq(dot(dd, dd));
                        const float s = fabsf(dot(sn, dir));
                        const float t = s * frcp(fsqrt(1.0f - s * s) + 1e-4f);
                        const float mq = G.mat[(size_t)q * 3], oq = G.mat[(size_t)q * 3 + 2];
                        const V3 mv{mp - mq, 0.04f - 0.04f, op - oq};  // specular is 0.04 everywhere
                        const float md = fsqrt(dot(mv, mv));
                        const float k1 = 0.0f + -t / 1.0f;
                        // k = (k1 - rd) - md <= k1 - md for rd >= 0 (rounding is monotonic): when
                        // k1 - md is already below the cut every buffer's weight is 0.
                        if (!(k1 + -md / 1.0f < -7.5f)) {
#pragma unroll
                            for (int b = 0; b < 4; ++b) {
                                const V3 dl = rgb(Lp[b]) - rgb(Lq[b]);
                                const float rd = fsqrt(dot(dl, dl)) * inv_den[b];
                                float k = k1;
                                k += -rd;
                                k += -md / 1.0f;
                                if (k < -7.5f) continue;
                                float wgt = pw * fexp(k);
                                if (b & 1) wgt *= rw;
                                g[b] = isfinite(wgt) ? wgt : 0.0f;
                            }
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < 4; ++b) wt[b] = base * g[b];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                ws[b] += wt[b];
                acc[b] += rgb(Lq[b]) * wt[b];
                va[b] += Lq[b].w * wt[b] * wt[b];
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        V3 r = acc[b] / ws[b];
        out.b[b][i] = make_float4(r.x, r.y, r.z, va[b] / (ws[b] * ws[b]));
    }
}

}  // namespace rm

using namespace rm;

namespace {

Quad frame_quad(rm_frame *f) {
    Quad q;
    for (int b = 0; b < 4; ++b) q.b[b] = f->buf[RM_BUF_DD + b].as<float4>();
    return q;
}

// The ping-pong half: four planes of float4 per pixel (allocated by
// denoise_prepare).
Quad filt_quad(rm_frame *f) {
    const size_t n = (size_t)f->w * f->h;
    Quad q;
    for (int b = 0; b < 4; ++b) q.b[b] = f->filt.as<float4>() + (size_t)b * n;
    return q;
}

}  // namespace

int32_t rm::denoise_prepare(rm_frame *f) {
    const size_t need = (size_t)f->w * f->h * 4 * sizeof(float4);
    return f->filt.bytes < need ? f->filt.alloc(need) : RM_OK;
}

int32_t rm::clamp_launch(rm_frame *f, hipStream_t st) {
    if (f->filt.bytes < (size_t)f->w * f->h * 4 * sizeof(float4))
        return fail(RM_ERR_INVALID, "spatial clamp: work buffer not prepared");
    const uint32_t n = (uint32_t)((size_t)f->w * f->h), nb = (n + DB - 1) / DB;
    Quad R = frame_quad(f);
    float *b1 = reinterpret_cast<float *>(filt_quad(f).b[0]);  // 4 planes of n floats
    hipLaunchKernelGGL(k_clamp_rows, dim3(4 * nb), dim3(DB), 0, st, R, f->w, n, nb, b1);
    hipLaunchKernelGGL(k_clamp_cols, dim3(4 * nb), dim3(DB), 0, st, R, f->w, f->h, n, nb, (const float *)b1);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

int32_t rm::filter_launch(rm_frame *f, hipStream_t st) {
    if (f->filt.bytes < (size_t)f->w * f->h * 4 * sizeof(float4))
        return fail(RM_ERR_INVALID, "filter: work buffer not prepared");
    const uint32_t n = (uint32_t)((size_t)f->w * f->h), nb = (n + DB - 1) / DB;
    Quad A = frame_quad(f), B = filt_quad(f);
    GView G{f->buf[RM_BUF_POSITION].as<float>(), f->buf[RM_BUF_SHAPE_NORMAL].as<float>(),
            f->buf[RM_BUF_SURFACE_NORMAL].as<float>(), f->buf[RM_BUF_MATERIAL].as<float>()};
    hipLaunchKernelGGL(k_filter_var, dim3(4 * nb), dim3(DB), 0, st, A, B, f->w, f->h, n, nb);
    // five passes: B -> A, A -> B, B -> A, A -> B, B -> A
    for (int pass = 0, step = 1; pass < 5; ++pass, step *= 2) {
        const Quad &src = (pass & 1) ? A : B;
        const Quad &dst = (pass & 1) ? B : A;
        hipLaunchKernelGGL(k_atrous, dim3(nb), dim3(DB), 0, st, G, src, dst, f->w, f->h, n, step);
    }
    RM_HIP(hipGetLastError());
    return RM_OK;
}

extern "C" int32_t rm_spatial_clamp(rm_frame *f, void *stream) {
    if (!f) return fail(RM_ERR_INVALID, "rm_spatial_clamp: frame is NULL");
    int32_t rc = denoise_prepare(f);
    return rc ? rc : clamp_launch(f, (hipStream_t)stream);
}

extern "C" int32_t rm_filter(rm_frame *f, void *stream) {
    if (!f) return fail(RM_ERR_INVALID, "rm_filter: frame is NULL");
    int32_t rc = denoise_prepare(f);
    return rc ? rc : filter_launch(f, (hipStream_t)stream);
}

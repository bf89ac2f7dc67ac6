// Post-processing behind rm_postprocess: Photo::shade (src/image.cpp:215-246),
// Photo::depthFeildBlur (:285-356, dof.hip), Photo::bloom (:248-283),
// Photo::gammaCorrection (:454-468) and Photo::FXAA (:358-452), one lane per
// pixel.  Without bloom and depth of field, shade and the tone curve are
// fused in one pass; with them the linear image goes to a frame buffer
// first.  FXAA reads the tone-mapped image from a second buffer because
// every output pixel looks at up to 12 neighbours along the edge normal.
#include <hip/hip_runtime.h>
#include "rm_frame.h"
#include "rm_vec.h"

namespace rm {

constexpr int PB = 256;
enum : int {
    O_BaseColor = 1, O_Emission = 2, O_Direct = 4, O_Indirect = 8, O_Diffuse = 16, O_Specular = 32,
    O_ShapeNormal = 64, O_SurfaceNormal = 128, O_Bloom = 256, O_FXAA = 512, O_DoF = 1024,
    O_Linear = RM_SHADE_LINEAR,  // no gammaCorrection (not a Photo::ShadeOption)
    O_Clamp = RM_POST_SPATIAL_CLAMP, O_Filter = RM_POST_FILTER
};

__device__ __forceinline__ V3 ld3p(const float *p, int64_t i) { return V3{p[i * 3], p[i * 3 + 1], p[i * 3 + 2]}; }
__device__ __forceinline__ V3 ld4p(const float *p, int64_t i) { return V3{p[i * 4], p[i * 4 + 1], p[i * 4 + 2]}; }

struct PostIn {
    const float *shapeN, *surfN, *emis, *base, *Dd, *Ds, *Id, *Is;
};

// fp32 tanh and pow evaluated in fp64 and rounded once: the correctly
// rounded fp32 result except within ~2^-50 of a rounding midpoint, which is
// what glibc's tanhf / powf return nearly always.
__device__ __forceinline__ float tanh_f(float x) { return (float)tanh((double)x); }
__device__ __forceinline__ float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }

// gammaCorrection (src/image.cpp:456-468): max(0), tanh luminance soft clip
// above 0.75, min(1), ^(1/2.2)
__device__ __forceinline__ V3 gamma(V3 px) {
    px = vmax(px, v3(0.0f));
    float c = dot(px, V3{0.3f, 0.6f, 0.1f});
    if (c > 0.75f) px = px / c * (tanh_f(3.0f * (c - 0.75f)) / 3.0f + 0.75f);
    px = vmin(px, v3(1.0f));
    const float e = 1.0f / 2.2f;
    return V3{pow_f(px.x, e), pow_f(px.y, e), pow_f(px.z, e)};
}

__global__ __launch_bounds__(PB) void k_shade_gamma(PostIn I, int64_t n, int opt, float exposure, float *out) {
    int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= n) return;
    V3 px;
    if (opt & O_ShapeNormal) {
        px = (ld3p(I.shapeN, i) + v3(1.0f)) / 2.0f;
    } else if (opt & O_SurfaceNormal) {
        px = (ld3p(I.surfN, i) + v3(1.0f)) / 2.0f;
    } else {
        V3 rd = v3(0.0f), rs = v3(0.0f);
        if (opt & O_Direct) {
            if (opt & O_Diffuse) rd += ld4p(I.Dd, i);
            if (opt & O_Specular) rs += ld4p(I.Ds, i);
        }
        if (opt & O_Indirect) {
            if (opt & O_Diffuse) rd += ld4p(I.Id, i);
            if (opt & O_Specular) rs += ld4p(I.Is, i);
        }
        if (!(opt & (O_Direct | O_Indirect))) rd = v3(1.0f);
        V3 dc = (opt & O_BaseColor) ? ld3p(I.base, i) : v3(1.0f);
        px = dc * rd + rs;
        if (opt & O_Emission) px += ld3p(I.emis, i) * exposure;
    }
    if (!(opt & O_Linear)) px = gamma(px);
    out[i * 3] = px.x;
    out[i * 3 + 1] = px.y;
    out[i * 3 + 2] = px.z;
}

// gammaCorrection of a linear image (or a plain copy with apply = 0).
__global__ __launch_bounds__(PB) void k_gamma(const float *__restrict__ in, int64_t n, int apply, float *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= n) return;
    V3 px = ld3p(in, i);
    if (apply) px = gamma(px);
    out[i * 3] = px.x;
    out[i * 3 + 1] = px.y;
    out[i * 3 + 2] = px.z;
}

// bloom, first half (src/image.cpp:254-262): the compressed highlights,
// Clum >= 1 only (the level starts at zero elsewhere).
__global__ __launch_bounds__(PB) void k_bloom_level0(const float *__restrict__ img, int64_t n, float *__restrict__ lvl) {
    int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= n) return;
    V3 p = ld3p(img, i);
    float c = dot(p, V3{0.3f, 0.6f, 0.1f});
    V3 b = v3(0.0f);
    if (!(c < 1.0f)) {
        b = p / pow_f(c, 0.65f);
        b = V3{gmax(b.x - 1.0f, 0.0f), gmax(b.y - 1.0f, 0.0f), gmax(b.z - 1.0f, 0.0f)};
    }
    lvl[i * 3] = b.x;
    lvl[i * 3 + 1] = b.y;
    lvl[i * 3 + 2] = b.z;
}

// bloom, one level (src/image.cpp:264-282): the previous level blurred by the
// 5x5 B3-spline at stride `step` (zero padding, no renormalisation), and the
// image gains a sixth of it.  Blocks are renumbered per XCD band of rows.
__global__ __launch_bounds__(PB) void k_bloom_pass(const float *__restrict__ prev, int w, int h, int step,
                                                  float *__restrict__ next, float *__restrict__ img) {
    uint32_t i = xcd_block(blockIdx.x, gridDim.x) * PB + threadIdx.x;
    if (i >= (uint32_t)w * (uint32_t)h) return;
    const float hk[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
    V3 col = v3(0.0f);
#pragma unroll
    for (int ky = -2; ky <= 2; ++ky) {
        int ny = y + ky * step;
        if (ny < 0 || ny >= h) continue;
#pragma unroll
        for (int kx = -2; kx <= 2; ++kx) {
            int nx = x + kx * step;
            if (nx < 0 || nx >= w) continue;
            col += ld3p(prev, (int64_t)ny * w + nx) * hk[ky + 2] * hk[kx + 2];
        }
    }
    next[(size_t)i * 3] = col.x;
    next[(size_t)i * 3 + 1] = col.y;
    next[(size_t)i * 3 + 2] = col.z;
    V3 p = ld3p(img, i) + col / 6.0f;
    img[(size_t)i * 3] = p.x;
    img[(size_t)i * 3 + 1] = p.y;
    img[(size_t)i * 3 + 2] = p.z;
}

__device__ __forceinline__ float luma(const float *d, int64_t i) {
    return dot(V3{d[i * 3], d[i * 3 + 1], d[i * 3 + 2]}, V3{0.3f, 0.6f, 0.1f});
}

__global__ __launch_bounds__(PB) void k_fxaa(const float *__restrict__ d, int w, int h, float *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    int x = (int)(i % w), y = (int)(i / w);
    const float EDGE_MIN = 0.0312f, EDGE_MAX = 0.125f, SUBPIX = 0.75f;
    float M = luma(d, i);
    float N = y > 0 ? luma(d, i - w) : M;
    float S = y < h - 1 ? luma(d, i + w) : M;
    float E = x < w - 1 ? luma(d, i + 1) : M;
    float Wl = x > 0 ? luma(d, i - 1) : M;
    // std::min / std::max over an initializer list keep the first extreme
    float mn = N, mx = N;
    if (S < mn) mn = S;
    if (E < mn) mn = E;
    if (Wl < mn) mn = Wl;
    if (mx < S) mx = S;
    if (mx < E) mx = E;
    if (mx < Wl) mx = Wl;
    float range = mx - mn;
    float thr = EDGE_MIN < mx * EDGE_MAX ? mx * EDGE_MAX : EDGE_MIN;  // std::max(EDGE_MIN, mx*EDGE_MAX)
    V3 c0{d[i * 3], d[i * 3 + 1], d[i * 3 + 2]};
    if (range < thr) {
        out[i * 3] = c0.x;
        out[i * 3 + 1] = c0.y;
        out[i * 3 + 2] = c0.z;
        return;
    }
    float NW = (y > 0 && x > 0) ? luma(d, i - w - 1) : M;
    float NE = (y > 0 && x < w - 1) ? luma(d, i - w + 1) : M;
    float SW = (y < h - 1 && x > 0) ? luma(d, i + w - 1) : M;
    float SE = (y < h - 1 && x < w - 1) ? luma(d, i + w + 1) : M;
    float eh = fabsf((NW + Wl + SW) - (NE + E + SE)) * (1.0f / 3.0f);
    float ev = fabsf((NW + N + NE) - (SW + S + SE)) * (1.0f / 3.0f);
    bool horiz = eh >= ev;
    float step = horiz ? 1.0f / float(w) : 1.0f / float(h);
    float gs = (horiz ? eh : ev) / range;
    gs = gs < -2.0f ? -2.0f : (gs > 2.0f ? 2.0f : gs);  // std::clamp
    float u = float(x) / float(w), v = float(y) / float(h);
    V3 fin = c0;
    float best = 0.0f;
    for (int k = 0; k < 12; ++k) {
        float o = gs * step * float(k + 1);
        float su = u + (horiz ? 0.0f : o), sv = v + (horiz ? o : 0.0f);
        if (su < 0.0f || su > 1.0f || sv < 0.0f || sv > 1.0f) continue;
        int sx = (int)(su * float(w)), sy = (int)(sv * float(h));
        sx = sx < 0 ? 0 : (sx > w - 1 ? w - 1 : sx);
        sy = sy < 0 ? 0 : (sy > h - 1 ? h - 1 : sy);
        int64_t j = (int64_t)sy * w + sx;
        float dl = fabsf(luma(d, j) - M);
        if (dl > best) {
            best = dl;
            fin = V3{d[j * 3], d[j * 3 + 1], d[j * 3 + 2]};
        }
    }
    float sub = (fabsf(N + S - 2.0f * M) * 2.0f + fabsf(E + Wl - 2.0f * M)) * 0.25f;
    sub = 1.0f < sub ? 1.0f : sub;  // std::min(sub, 1)
    float t = sub * SUBPIX;
    V3 r = c0 * (1.0f - t) + fin * t;
    out[i * 3] = r.x;
    out[i * 3 + 1] = r.y;
    out[i * 3 + 2] = r.z;
}

}  // namespace rm

using namespace rm;

// Allocation half of rm_postprocess (kept out of any graph capture).
int32_t rm::post_prepare(rm_frame *f, int opt, bool in_batch) {
    if (!f) return fail(RM_ERR_INVALID, "rm_postprocess: frame is NULL");
    if ((opt & O_DoF) && in_batch)
        return fail(RM_ERR_UNSUPPORTED, "depth of field in a batch's post options: it sizes its work from the image "
                                        "(a host wait); run rm_postprocess on the frame after the batch");
    const size_t plane = (size_t)f->w * f->h * 12;
    int32_t rc;
    if ((opt & O_FXAA) && f->post.bytes < plane && (rc = f->post.alloc(plane))) return rc;
    if ((opt & (O_Bloom | O_DoF)) && f->lin.bytes < 2 * plane && (rc = f->lin.alloc(2 * plane))) return rc;
    if ((opt & O_Bloom) && f->lvl.bytes < 2 * plane && (rc = f->lvl.alloc(2 * plane))) return rc;
    if ((opt & (O_Clamp | O_Filter)) && (rc = denoise_prepare(f))) return rc;
    return RM_OK;
}

// Launch half: only kernel launches on `st` (capturable), except depth of
// field (dof.hip), which waits for its tile-list sizes.
int32_t rm::post_launch(rm_frame *f, int opt, float exposure, hipStream_t st) {
    const int64_t n = (int64_t)f->w * f->h;
    PostIn I;
    I.shapeN = f->buf[RM_BUF_SHAPE_NORMAL].as<float>();
    I.surfN = f->buf[RM_BUF_SURFACE_NORMAL].as<float>();
    I.emis = f->buf[RM_BUF_EMISSION].as<float>();
    I.base = f->buf[RM_BUF_BASE_COLOR].as<float>();
    I.Dd = f->buf[RM_BUF_DD].as<float>();
    I.Ds = f->buf[RM_BUF_DS].as<float>();
    I.Id = f->buf[RM_BUF_ID].as<float>();
    I.Is = f->buf[RM_BUF_IS].as<float>();
    float *pixels = f->buf[RM_BUF_PIXELS].as<float>();
    const unsigned g = (unsigned)((n + PB - 1) / PB);
    int32_t drc;
    if ((opt & O_Clamp) && (drc = clamp_launch(f, st))) return drc;
    if ((opt & O_Filter) && (drc = filter_launch(f, st))) return drc;
    if ((opt & O_FXAA) && f->post.bytes < (size_t)n * 12) return fail(RM_ERR_INVALID, "rm_postprocess: FXAA buffer not prepared");
    float *tone_out = (opt & O_FXAA) ? f->post.as<float>() : pixels;  // where the tone-mapped image goes
    if (!(opt & (O_Bloom | O_DoF))) {
        hipLaunchKernelGGL(k_shade_gamma, dim3(g), dim3(PB), 0, st, I, n, (int)opt, exposure, tone_out);
    } else {
        if (f->lin.bytes < (size_t)n * 24 || ((opt & O_Bloom) && f->lvl.bytes < (size_t)n * 24))
            return fail(RM_ERR_INVALID, "rm_postprocess: bloom / depth-of-field buffers not prepared");
        float *img = f->lin.as<float>(), *img2 = img + n * 3;
        hipLaunchKernelGGL(k_shade_gamma, dim3(g), dim3(PB), 0, st, I, n, (int)(opt | O_Linear), exposure, img);
        if (opt & O_DoF) {
            int32_t rc = dof_launch(f, img, img2, st);
            if (rc) return rc;
            std::swap(img, img2);
        }
        if (opt & O_Bloom) {
            float *la = f->lvl.as<float>(), *lb = la + n * 3;
            hipLaunchKernelGGL(k_bloom_level0, dim3(g), dim3(PB), 0, st, (const float *)img, n, la);
            for (int step = 1; step <= 16; step *= 2) {
                hipLaunchKernelGGL(k_bloom_pass, dim3(g), dim3(PB), 0, st, (const float *)la, f->w, f->h, step, lb, img);
                std::swap(la, lb);
            }
        }
        hipLaunchKernelGGL(k_gamma, dim3(g), dim3(PB), 0, st, (const float *)img, n, (opt & O_Linear) ? 0 : 1, tone_out);
    }
    if (opt & O_FXAA)
        hipLaunchKernelGGL(k_fxaa, dim3(g), dim3(PB), 0, st, (const float *)f->post.as<float>(), f->w, f->h, pixels);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

extern "C" int32_t rm_postprocess(rm_frame *f, int32_t opt, float exposure, void *stream) {
    int32_t rc = post_prepare(f, opt);
    if (rc) return rc;
    return post_launch(f, opt, exposure, (hipStream_t)stream);
}

extern "C" int32_t rm_frame_set_dof(rm_frame *f, float focus, float coc, const float camera_position[3]) {
    if (!f || !camera_position) return fail(RM_ERR_INVALID, "rm_frame_set_dof: NULL argument");
    f->dof_focus = focus;
    f->dof_coc = coc;
    for (int k = 0; k < 3; ++k) f->dof_cam[k] = camera_position[k];
    return RM_OK;
}

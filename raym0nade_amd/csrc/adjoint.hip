// Path-replay adjoint behind rm_adjoint: gradients of a loss on the linear
// Full image (Photo::shade, src/image.cpp:215-246, without gammaCorrection)
// with respect to albedo texels, per-material roughness and emission.
//
// Estimator.  Every indirect sample of the renderer contributes only at the
// vertex where its path ends (an emitter or sky hit, or the light samples of
// sampleRay's NEE-and-stop branch, src/render.cpp:121-312), so a sample's
// image contribution is
//     c = M(p0) * (w * E * prod_j beta_j)
// with p0 the first vertex's bsdfPdf, beta_j the sampled bsdfPdf of vertex j,
// E the terminal light value and M the split of accumulateInwardRadiance
// (src/image.cpp:630-659) recombined by shade (M(v) = v unless the first
// vertex's albedo was nudged, src/render.cpp:492-495,549).  Hence
//     d c / d beta_j = c / beta_j   (per channel),
// and the derivative of every vertex factor needs only C = sum of c over the
// sample's light samples.  One lane replays a sample twice from its Philox
// stream: pass A recomputes C (and the image as a check), pass B replays the
// same decisions and adds, at every vertex,
//   albedo:    (d beta / d albedo)^T (delta * C / beta)  -> texels (bilinear x trilinear)
//   roughness: (delta * C / beta) . d beta / d a  +  score * (delta . C)
// where the score is d log p / d a of every roughness-dependent draw of that
// vertex: each GGX half vector (density D_a(H) cos) and the cosine/GGX pick
// (probability p1/(p1+p2)).  This makes the roughness gradient an unbiased
// estimate of the derivative of the expected image of this (biased) estimator.
// Light-object and Russian-roulette probabilities are held fixed: both choose
// between unbiased alternatives, so the expectation does not depend on them.
// The primary-vertex light samples (k_direct) are differentiated directly.
#include <hip/hip_runtime.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "integrator.h"
#include "rm_frame.h"

namespace rm {

constexpr int ADJ_B = 256;

// ------------------------------------------------------------ mip levels
// ImageData::generateMipmaps (src/material.cpp:113-148) on the device: level
// lv texel = (((0 + a) + b) + c) + d) / 4 over its 2x2 parents, like the host.
__global__ __launch_bounds__(ADJ_B) void k_mip_level(float *texels, int64_t off_prev, int pw, int ph, int64_t off_cur,
                                                      int cw, int ch) {
    int64_t i = (int64_t)blockIdx.x * ADJ_B + threadIdx.x;
    if (i >= (int64_t)cw * ch) return;
    int x = (int)(i % cw), y = (int)(i / cw);
    const float *prev = texels + off_prev;
    float *cur = texels + off_cur;
    for (int c = 0; c < 3; ++c) {
        float s = 0.0f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                int px = (x * 2 + dx) % pw, py = (y * 2 + dy) % ph;
                s += prev[((int64_t)py * pw + px) * 3 + c];
            }
        cur[((int64_t)y * cw + x) * 3 + c] = s / 4.0f;
    }
}

int32_t launch_mip_build(const rm_scene *s, int t, hipStream_t st) {
    const TexRec &T = s->host.texs[t];
    float *tx = s->texels.as<float>();
    for (int lv = 1; lv < T.levels; ++lv) {
        int pw = T.width >> (lv - 1), ph = T.height >> (lv - 1), cw = T.width >> lv, ch = T.height >> lv;
        int64_t n = (int64_t)cw * ch;
        hipLaunchKernelGGL(k_mip_level, dim3((unsigned)((n + ADJ_B - 1) / ADJ_B)), dim3(ADJ_B), 0, st, tx,
                           T.offset[lv - 1], pw, ph, T.offset[lv], cw, ch);
    }
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// Adjoint of one mip step: parent texel (px, py) of level lv-1 gets a quarter
// of its child's gradient.  2x + 1 <= 2 (w >> 1) - 1 < w, so the build's wrap
// never fires and every parent has at most one child.
__global__ __launch_bounds__(ADJ_B) void k_mip_adjoint(float *g, int64_t off_prev, int pw, int ph, int64_t off_cur,
                                                        int cw, int ch) {
    int64_t i = (int64_t)blockIdx.x * ADJ_B + threadIdx.x;
    if (i >= (int64_t)pw * ph) return;
    int px = (int)(i % pw), py = (int)(i / pw);
    int x = px >> 1, y = py >> 1;
    if (x >= cw || y >= ch) return;
    const float *cur = g + off_cur + ((int64_t)y * cw + x) * 3;
    float *prev = g + off_prev + i * 3;
    for (int c = 0; c < 3; ++c) prev[c] += cur[c] / 4.0f;
}

__global__ void k_fill(float *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_split_mat(const float *gm, int n, float *rough, float *emis) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    if (rough) rough[m] = gm[m * 4];
    if (emis)
        for (int c = 0; c < 3; ++c) emis[m * 3 + c] = gm[m * 4 + 1 + c];
}

// ------------------------------------------------------------ pieces
struct AdjArgs {
    SceneView S;
    RArgs A;
    const float *Fbase;    // the frame's base colour as the render left it
    const int32_t *drop;   // the render's dropped outlier per pixel
    const float *dimg;     // d loss / d image [np][3]
    float *gtex;           // texel gradients, laid out like S.texels (every level), or null
    float *gmat;           // [n_mats][4]: d roughness, d emission rgb
    float *replay;         // [np][3] or null
    const float4 *rec;     // the render's per-sample record (sum dd | sum ds | sum w L), or null
    int32_t rec_stride;
    int64_t np;
    int split;
};

// clamp (src/sampling.cpp:238-243) maps p to p * 64 / c when c = lum(p) > 64:
// J u = (64/c)(u - p (W.u)/c),  J^T y = (64/c)(y - W (p.y)/c).
__device__ __forceinline__ V3 clampJ(V3 p, V3 u) {
    float c = dot(p, W_RGB());
    if (!(c > 64.0f)) return u;
    return (64.0f / c) * (u - p * (dot(W_RGB(), u) / c));
}
__device__ __forceinline__ V3 clampJT(V3 p, V3 y) {
    float c = dot(p, W_RGB());
    if (!(c > 64.0f)) return y;
    return (64.0f / c) * (y - W_RGB() * (dot(p, y) / c));
}

// M(v): how a first-vertex bsdfPdf v reaches the shaded image.
// accumulateInwardRadiance splits v against the nudged albedo bn into a diffuse
// part (divided by bn) and the rest; shade multiplies the diffuse part by the
// frame's base colour Fb.  With Fb == bn this recombines to v exactly.  When
// the nudge is undone (Fb = bn - 0.04 e_x):
//   generic split:  M = v + (Fb - bn) f(v),  f = (v.b0 - xy v.w) / ((1 - xy^2)|bn|)
//   near white:     M = Fb * v / bn
struct ImgW {
    int mode;  // 0: M = v, 1: generic, 2: near white
    V3 bn, Fb, D, b0, wv;
    float xy, n, den;
};
__device__ __forceinline__ ImgW img_weight(V3 bn, V3 Fb) {
    ImgW w;
    w.mode = 0;
    w.bn = bn;
    w.Fb = Fb;
    w.D = Fb - bn;
    if (w.D.x == 0.0f && w.D.y == 0.0f && w.D.z == 0.0f) return w;
    if (length(bn) < T_EPS) return w;
    w.b0 = normalize(bn);
    w.wv = normalize(v3(1.0f));
    w.xy = dot(w.b0, w.wv);
    if (w.xy > 0.99f) {
        w.mode = 2;
        return w;
    }
    w.mode = 1;
    w.n = length(bn);
    w.den = (1.0f - w.xy * w.xy) * w.n;
    return w;
}
__device__ __forceinline__ V3 img_apply(const ImgW &w, V3 v) {
    if (w.mode == 0) return v;
    if (w.mode == 2) return w.Fb * v / w.bn;
    float f = (dot(v, w.b0) - w.xy * dot(v, w.wv)) / w.den;
    return v + w.D * f;
}
// Given yM = d(delta.c)/dM: yv through v, and gb through the albedo itself
// (bn and Fb both move with it one for one).
__device__ __forceinline__ void img_adjoint(const ImgW &w, V3 v, V3 yM, V3 &yv, V3 &gb) {
    if (w.mode == 0) {
        yv = yM;
        gb = v3(0.0f);
        return;
    }
    if (w.mode == 2) {
        yv = yM * w.Fb / w.bn;
        gb = -(yM * v * w.D) / (w.bn * w.bn);
        return;
    }
    float s = dot(yM, w.D);
    yv = yM + s * ((w.b0 - w.xy * w.wv) / w.den);
    float A = dot(v, w.b0), Bw = dot(v, w.wv);
    float f = (A - w.xy * Bw) / w.den;
    V3 dA = (v - A * w.b0) / w.n;
    V3 dxy = (w.wv - w.xy * w.b0) / w.n;
    V3 dnum = dA - Bw * dxy;
    V3 dden = (-2.0f * w.xy * w.n) * dxy + (1.0f - w.xy * w.xy) * w.b0;
    gb = s * ((dnum - f * dden) / w.den);
}

// Recorder of one sample_reflection call.
struct RefTap {
    float a, score, ct2, lh2, p1, p2;
    V3 raw1;
    bool first;
    __device__ explicit RefTap(float r)
        : a(r), score(0.0f), ct2(0.0f), lh2(1.0f), p1(0.0f), p2(0.0f), raw1(v3(0.0f)), first(false) {}
    __device__ __forceinline__ void ggx_try(float ct) { score += gtr2_dlog(ct, a); }
    __device__ __forceinline__ void ggx_ok(float ct, float lh) { ct2 = ct; lh2 = lh; }
    __device__ __forceinline__ void cos_raw(V3 r) { raw1 = r; }
    __device__ __forceinline__ void pick(float q1, float q2, bool f) { p1 = q1; p2 = q2; first = f; }
};

// Derivatives of the value bp sample_reflection returned: d bp / d a (dbp),
// d bp / d albedo = kdp * J (J = the clamp Jacobian at raw for the cosine
// sample, identity for the GGX sample), and the score of the vertex's draws.
struct VtxD {
    V3 dbp, raw;
    float kdp, score;
    bool clamp;
};
__device__ __forceinline__ VtxD reflect_deriv(const Surf &s, V3 V, V3 nd, V3 bp, const RefTap &tp) {
    VtxD o;
    float kd;
    V3 dr;
    brdf_grad(s, V, nd, kd, dr);
    float dp2 = tp.p2 > 0.0f ? gtr2_da(tp.ct2, tp.a) / (4.0f * tp.lh2) : 0.0f;
    float q = tp.p1 + tp.p2;
    if (tp.first) {
        o.raw = tp.raw1;
        o.kdp = kd / tp.p1;
        o.dbp = clampJ(tp.raw1, dr / tp.p1);
        o.clamp = true;
        o.score = tp.score - dp2 / q;
    } else {
        o.raw = bp;
        o.kdp = kd / tp.p2;
        o.dbp = dr / tp.p2 - bp * (dp2 / tp.p2);
        o.clamp = false;
        o.score = tp.score + dp2 / tp.p2 - dp2 / q;
    }
    return o;
}
__device__ __forceinline__ V3 albedo_adj(const VtxD &o, V3 y) { return o.kdp * (o.clamp ? clampJT(o.raw, y) : y); }

// Scatter of an albedo gradient through tex_get (bilinear on two mip levels).
__device__ __forceinline__ void bilinear_scatter(float *g, int w, int h, float u, float v, V3 gv) {
    float x = u * float(w), y = v * float(h);
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float dx = x - float(x0), dy = y - float(y0);
    if (x0 < 0 || x0 >= w) { x0 %= w; if (x0 < 0) x0 += w; }
    if (y0 < 0 || y0 >= h) { y0 %= h; if (y0 < 0) y0 += h; }
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    const int64_t i00 = ((int64_t)y0 * w + x0) * 3, i01 = ((int64_t)y0 * w + x1) * 3;
    const int64_t i10 = ((int64_t)y1 * w + x0) * 3, i11 = ((int64_t)y1 * w + x1) * 3;
    const float w00 = (1.0f - dx) * (1.0f - dy), w01 = dx * (1.0f - dy), w10 = (1.0f - dx) * dy, w11 = dx * dy;
    const float gc[3] = {gv.x, gv.y, gv.z};
    for (int c = 0; c < 3; ++c) {
        atomicAdd(g + i00 + c, w00 * gc[c]);
        atomicAdd(g + i01 + c, w01 * gc[c]);
        atomicAdd(g + i10 + c, w10 * gc[c]);
        atomicAdd(g + i11 + c, w11 * gc[c]);
    }
}
__device__ __forceinline__ void tex_scatter(float *g, const SceneView &S, const TexFoot &f, V3 gv) {
    const TexRec &T = S.texs[f.tex];
    float v = 1.0f - f.v;
    float hi = float(T.levels - 1);
    float depth = gmax(gmin(f.depth, hi), 0.0f);
    int level = (int)depth;
    int next = level + 1 < T.levels - 1 ? level + 1 : T.levels - 1;
    float blend = depth - float(level);
    bilinear_scatter(g + T.offset[level], T.width >> level, T.height >> level, f.u, v, gv * (1.0f - blend));
    if (blend != 0.0f) bilinear_scatter(g + T.offset[next], T.width >> next, T.height >> next, f.u, v, gv * blend);
}

// d light / d emission of a light-object sample (per channel): the light is
// 2 pi power colour / (d^2 + 1e-3) / P with power colour = area * emission.
__device__ __forceinline__ float light_kappa(const SceneView &S, const DirectInfo &di, int &mat) {
    const LightRec &L = S.lights[di.li];
    mat = S.light_faces[L.first_face].material;
    V3 e = ld3(S.mats[mat].emission);
    float lum = dot(e, W_RGB());
    return (2.0f * PI_F * (L.power / lum)) / (di.dist * di.dist + 1e-3f) / di.P;
}

__device__ __forceinline__ V3 safe_div(V3 a, V3 b) {
    return V3{b.x != 0.0f ? a.x / b.x : 0.0f, b.y != 0.0f ? a.y / b.y : 0.0f, b.z != 0.0f ? a.z / b.z : 0.0f};
}
__device__ __forceinline__ bool nonzero(V3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }

// One indirect sample of an opaque first vertex (k_spawn + k_bounce of the
// renderer, same draws in the same order).  AD = false: Lam = sum w L and
// Cv = the image contribution (before exposure).  AD = true: the derivatives.
template <bool AD>
__device__ __forceinline__ void replay_path(const AdjArgs &G, const Surf &s0, V3 in, float hit_t, const Diff &cam,
                                            const ImgW &iw, int64_t p, uint32_t si, float inv, int32_t drop, V3 &Lam,
                                            V3 &Cv, V3 dl, V3 &gb, float &gr0, float *sh) {
    const SceneView &S = G.S;
    const RArgs &A = G.A;
    const V3 V0 = -in;
    if (!AD) {
        Lam = v3(0.0f);
        Cv = v3(0.0f);
    }
    Stream rng{A.seed, TAG_INDIRECT, (uint32_t)p, si, 0u};
    (void)rng.next();  // `rng.next() < P_reflect`, P_reflect = 1 for an opaque vertex
    Diff next;
    calc_dPdxy(in, hit_t, s0.shapeN, cam, next.dPdx, next.dPdy);
    calc_dDdxy(in, s0.surfN, cam, next.dDdx, next.dDdy);
    V3 nd, bp;
    int fails = 0;
    RefTap tap0(s0.roughness);
    if (AD)
        sample_reflection(s0, V0, rng, nd, bp, fails, tap0);
    else
        sample_reflection(s0, V0, rng, nd, bp, fails);
    bp /= 1.0f;
    if (!any_finite(nd)) return;
    const V3 p0 = bp;
    const V3 M0 = img_apply(iw, p0);
    float W = fails > 0 ? 1.0f / (float)(1 + fails) : 1.0f;
    float rf = s0.roughness * 1.0f;
    int src = s0.id;
    V3 T = v3(1.0f);
    bool exclude = true;
    int depth = 1;
    V3 o = s0.pos, d = nd;
    Diff diff = next;
    const float dC = AD ? dot(dl, Cv) : 0.0f;
    if (AD) {
        V3 yv, gbM;
        img_adjoint(iw, p0, dl * Lam, yv, gbM);
        VtxD vd = reflect_deriv(s0, V0, nd, p0, tap0);
        gb += albedo_adj(vd, yv) + gbM;
        gr0 += dot(yv, vd.dbp) + vd.score * dC;
    }
    const int maxD = A.max_depth;
    for (;;) {
        RayPre r = ray_pre(o, d);
        float t = INFINITY;
        int32_t slot = model_closest<false, 0>(S, o, d, r, t);
        if (slot < 0) {
            if (!exclude && S.sky) {
                V3 L = sky_get(S, d) * v3(1.0f);
                L = L * T;
                float w = 1.0f * W * inv;
                if ((int32_t)(si * 8u) != drop && !(length(L) < T_EPS) && !AD) {
                    Acc4 dd{0, 0, 0, 0}, ds{0, 0, 0, 0};
                    accumulate(iw.bn, LSample{p0, L, w}, dd, ds);
                    Cv += iw.Fb * V3{dd.r, dd.g, dd.b} + V3{ds.r, ds.g, ds.b};
                    Lam += w * L;
                }
            }
            return;
        }
        Surf s = surf_default();
        s.pos = o + d * t;
        V3 dPdx, dPdy;
        TexFoot tf;
        tf.tex = -1;
        hit_info<false>(S, slot, d, t, diff, dPdx, dPdy, s, AD ? &tf : nullptr);  // rm_adjoint refuses 8-bit maps
        V3 V = -d;
        if (rf < 1.0f * s.roughness) src = s.id;  // gmax keeps the first argument on ties
        rf = gmax(rf, 1.0f * s.roughness);
        s.roughness = gmax(s.roughness, rf);
        const float P_RR = 1.0f + (0.5f - 1.0f) * sqrtf(s.roughness);
        bool doNEE = P_RR < 0.9f && depth >= A.rr_min;
        if (length(s.emission) > T_EPS) {
            if (!exclude) {
                V3 L = s.emission * v3(1.0f);
                L = L * T;
                float w = 1.0f * W * inv;
                if ((int32_t)(si * 8u) != drop && !(length(L) < T_EPS)) {
                    if (!AD) {
                        Acc4 dd{0, 0, 0, 0}, ds{0, 0, 0, 0};
                        accumulate(iw.bn, LSample{p0, L, w}, dd, ds);
                        Cv += iw.Fb * V3{dd.r, dd.g, dd.b} + V3{ds.r, ds.g, ds.b};
                        Lam += w * L;
                    } else {
                        V3 ge = dl * M0 * (w * T);
                        atomicAdd(sh + s.id * 4 + 1, ge.x);
                        atomicAdd(sh + s.id * 4 + 2, ge.y);
                        atomicAdd(sh + s.id * 4 + 3, ge.z);
                    }
                }
            }
            return;
        }
        (void)rng.next();  // `rng.next() <= P_reflect`, P_reflect = 1
        doNEE = doNEE && s.entering;
        if (doNEE && (depth >= maxD || rng.next() > P_RR)) {
            float factor = 1.0f / 1.0f;
            if (depth < maxD) factor /= (1.0f - P_RR);
            const int cnt = sample_count(depth);
            float total = direct_total(S, s, V);
            if (total != 0.0f) {
                uint32_t k = 0;
                V3 ga = v3(0.0f);
                float gr = 0.0f;
                for (int Ti = 0; Ti < cnt; ++Ti) {
                    LSample x;
                    DirectInfo di;
                    TraceCounts tc{};
                    if (!direct_sample<false, false>(S, s, V, total, cnt, rng, x, tc, &di)) continue;
                    x.bsdf *= factor;
                    const uint32_t id = si * 8u + k;
                    ++k;
                    V3 L = x.light * x.bsdf;
                    L = L * T;
                    float w = x.weight * W * inv;
                    if ((int32_t)id == drop || length(L) < T_EPS) continue;
                    if (!AD) {
                        Acc4 dd{0, 0, 0, 0}, ds{0, 0, 0, 0};
                        accumulate(iw.bn, LSample{p0, L, w}, dd, ds);
                        Cv += iw.Fb * V3{dd.r, dd.g, dd.b} + V3{ds.r, ds.g, ds.b};
                        Lam += w * L;
                    } else {
                        V3 yb = dl * M0 * ((w * factor) * x.light * T);
                        V3 raw = bsdf_eval(s, V, di.dir);
                        float kd;
                        V3 dr;
                        brdf_grad(s, V, di.dir, kd, dr);
                        V3 yraw = clampJT(raw, yb);
                        ga += kd * yraw;
                        gr += dot(yraw, dr);
                        if (di.li >= 0) {
                            int lm;
                            float kap = light_kappa(S, di, lm);
                            V3 ge = dl * M0 * ((w * kap) * x.bsdf * T);
                            atomicAdd(sh + lm * 4 + 1, ge.x);
                            atomicAdd(sh + lm * 4 + 2, ge.y);
                            atomicAdd(sh + lm * 4 + 3, ge.z);
                        }
                    }
                }
                if (AD) {
                    atomicAdd(sh + src * 4, gr);
                    if (G.gtex && tf.tex >= 0 && nonzero(ga)) tex_scatter(G.gtex, S, tf, ga);
                }
            }
            return;
        }
        Diff nx;
        nx.dPdx = dPdx;
        nx.dPdy = dPdy;
        calc_dDdxy(d, s.surfN, diff, nx.dDdx, nx.dDdy);
        int fl = 0;
        RefTap tp(s.roughness);
        if (AD)
            sample_reflection(s, V, rng, nd, bp, fl, tp);
        else
            sample_reflection(s, V, rng, nd, bp, fl);
        const V3 bps = bp;
        bp /= 1.0f;
        if (doNEE) bp /= P_RR;
        if (!(any_finite(nd) && depth != maxD)) return;
        if (AD) {
            VtxD vd = reflect_deriv(s, V, nd, bps, tp);
            V3 y = safe_div(dl * Cv, bps);
            atomicAdd(sh + src * 4, dot(y, vd.dbp) + vd.score * dC);
            if (G.gtex && tf.tex >= 0) tex_scatter(G.gtex, S, tf, albedo_adj(vd, y));
        }
        bp *= v3(1.0f);  // absorb of air
        T = T * bp;
        if (fl > 0) W *= 1.0f / (float)(1 + fl);
        exclude = doNEE;
        depth = depth + 1;
        o = s.pos;
        d = nd;
        diff = nx;
    }
}

__device__ __forceinline__ void adjoint_pixel(const AdjArgs &G, int64_t p, int chunk, float *sh) {
    const SceneView &S = G.S;
    const RArgs &A = G.A;
    const int x = (int)(p % A.w), y = (int)(p / A.w);
    V3 Dir;
    Diff cam;
    camera_ray(A, x, y, Dir, cam);
    RayPre r = ray_pre(A.pos, Dir);
    float t = INFINITY;
    int32_t slot = model_closest<false, 0>(S, A.pos, Dir, r, t);
    const V3 dimg = ld3(G.dimg + p * 3);
    const float ex = A.exposure;
    V3 rep = v3(0.0f);
    if (slot < 0) {
        if (G.replay && chunk == 0) {
            V3 e = sky_get(S, Dir) * ex;
            atomicAdd(G.replay + p * 3, e.x);
            atomicAdd(G.replay + p * 3 + 1, e.y);
            atomicAdd(G.replay + p * 3 + 2, e.z);
        }
        return;
    }
    Surf s = surf_default();
    s.pos = A.pos + t * Dir;
    V3 dPdx, dPdy;
    TexFoot tf0;
    hit_info<false>(S, slot, Dir, t, cam, dPdx, dPdy, s, &tf0);
    const int mat0 = s.id;
    if (chunk == 0) {
        rep += s.emission * ex;
        if (length(s.emission) > 0.0f) {
            V3 ge = dimg * ex;
            atomicAdd(sh + mat0 * 4 + 1, ge.x);
            atomicAdd(sh + mat0 * 4 + 2, ge.y);
            atomicAdd(sh + mat0 * 4 + 3, ge.z);
        }
    }
    const bool want_grad = nonzero(dimg);
    if (length(s.emission) > 0.0f || s.opacity <= 1.0f - T_EPS || (!want_grad && !G.replay)) {
        if (G.replay) {
            atomicAdd(G.replay + p * 3, rep.x);
            atomicAdd(G.replay + p * 3 + 1, rep.y);
            atomicAdd(G.replay + p * 3 + 2, rep.z);
        }
        return;
    }
    nudge_base(s.base);
    const V3 bn = s.base;
    const V3 Fb = ld3(G.Fbase + p * 3);
    const ImgW iw = img_weight(bn, Fb);
    const V3 dl = dimg * ex;
    const int cnt = (int)(float(A.spp) * A.p_direct);
    const int n_ind = A.spp - cnt;
    V3 gb = v3(0.0f);
    float gr0 = 0.0f;
    const V3 in = normalize(s.pos - A.pos);
    const V3 V = -in;
    // the primary vertex's own light samples (k_direct)
    {
        float total = direct_total(S, s, V);
        if (total != 0.0f)
            for (int T = chunk; T < cnt; T += G.split) {
                Stream g{A.seed, TAG_DIRECT, (uint32_t)p, (uint32_t)T, 0u};
                LSample ls;
                DirectInfo di;
                TraceCounts tc{};
                if (!direct_sample<false, false>(S, s, V, total, cnt, g, ls, tc, &di)) continue;
                if (length(ls.light) < T_EPS) continue;
                if (G.replay) {
                    Acc4 dd{0, 0, 0, 0}, ds{0, 0, 0, 0};
                    accumulate(bn, ls, dd, ds);
                    rep += (Fb * V3{dd.r, dd.g, dd.b} + V3{ds.r, ds.g, ds.b}) * ex;
                }
                if (!want_grad) continue;
                V3 yv, gbM;
                img_adjoint(iw, ls.bsdf, dl * (ls.weight * ls.light), yv, gbM);
                V3 raw = bsdf_eval(s, V, di.dir);
                float kd;
                V3 dr;
                brdf_grad(s, V, di.dir, kd, dr);
                V3 yraw = clampJT(raw, yv);
                gb += kd * yraw + gbM;
                gr0 += dot(yraw, dr);
                if (di.li >= 0) {
                    int lm;
                    float kap = light_kappa(S, di, lm);
                    V3 ge = dl * img_apply(iw, ls.bsdf) * (ls.weight * kap);
                    atomicAdd(sh + lm * 4 + 1, ge.x);
                    atomicAdd(sh + lm * 4 + 2, ge.y);
                    atomicAdd(sh + lm * 4 + 3, ge.z);
                }
            }
    }
    // indirect samples (k_spawn's vertex uses the relative eta 1 / eta)
    Surf s0 = s;
    s0.eta = 1.0f / s0.eta;
    const float inv = n_ind > 0 ? 1.0f / (float)n_ind : 0.0f;
    const int32_t drop = G.drop[p];
    const float hit_t = length(s.pos - A.pos);
    for (int si = chunk; si < n_ind; si += G.split) {
        V3 Lam, Cv;
        if (G.rec && !(drop >= 0 && (drop >> 3) == si)) {
            // the render recorded this sample's sums (its dropped outlier, if
            // any, is not excluded there, so such a sample is replayed instead)
            const float4 *r = G.rec + ((int64_t)p * G.rec_stride + si) * 3;
            float4 r0 = r[0], r1 = r[1], r2 = r[2];
            Cv = Fb * V3{r0.x, r0.y, r0.z} + V3{r1.x, r1.y, r1.z};
            Lam = V3{r2.x, r2.y, r2.z};
        } else {
            replay_path<false>(G, s0, in, hit_t, cam, iw, p, (uint32_t)si, inv, drop, Lam, Cv, dl, gb, gr0, sh);
        }
        rep += Cv * ex;
        if (want_grad && nonzero(Cv))
            replay_path<true>(G, s0, in, hit_t, cam, iw, p, (uint32_t)si, inv, drop, Lam, Cv, dl, gb, gr0, sh);
    }
    if (G.replay) {
        atomicAdd(G.replay + p * 3, rep.x);
        atomicAdd(G.replay + p * 3 + 1, rep.y);
        atomicAdd(G.replay + p * 3 + 2, rep.z);
    }
    if (want_grad) {
        atomicAdd(sh + mat0 * 4, gr0);
        if (G.gtex && tf0.tex >= 0 && nonzero(gb)) tex_scatter(G.gtex, S, tf0, gb);
    }
}

// At most 128 VGPRs (four waves per SIMD): uncapped, the kernel takes 256
// VGPRs and 2 AGPRs, one wave per SIMD, and its loads and atomics wait
// exposed.  The spill the cap costs is cheaper than that latency (512x512,
// 128 spp: 50 ms with one wave, 28 with two, 24 with four).
__global__ __launch_bounds__(ADJ_B, 4) void k_adjoint(AdjArgs G) {
    extern __shared__ float sh[];
    const int nm4 = G.S.n_mats * 4;
    for (int i = threadIdx.x; i < nm4; i += ADJ_B) sh[i] = 0.0f;
    __syncthreads();
    const int64_t gid = (int64_t)blockIdx.x * ADJ_B + threadIdx.x;
    if (gid < G.np * G.split) adjoint_pixel(G, gid % G.np, (int)(gid / G.np), sh);
    __syncthreads();
    for (int i = threadIdx.x; i < nm4; i += ADJ_B) {
        float v = sh[i];
        if (v != 0.0f) atomicAdd(G.gmat + i, v);
    }
}

static int32_t fill(float *p, int64_t n, hipStream_t st) {
    if (n <= 0) return RM_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

}  // namespace rm

using namespace rm;

extern "C" int32_t rm_adjoint(const rm_scene *scene, const rm_render_args *a, rm_frame *f, const float *d_image,
                              const rm_grads *g, void *stream) {
    if (!scene || !a || !f || !d_image || !g) return fail(RM_ERR_INVALID, "rm_adjoint: NULL argument");
    if (!f->stamped || f->stamp_scene != scene || f->stamp_version != scene->version ||
        memcmp(&f->stamp_args, a, sizeof(rm_render_args)) != 0)
        return fail(RM_ERR_INVALID, "rm_adjoint: the frame does not hold rm_render(scene, args)");
    const HostScene &h = scene->host;
    for (size_t m = 0; m < h.mats.size(); ++m) {
        const MatRec &M = h.mats[m];
        if (!(M.opacity > 1.0f - T_EPS))
            return fail(RM_ERR_UNSUPPORTED, "rm_adjoint: material %zu transmits (opaque scenes only)", m);
        if (M.ntex >= 0 || M.etex >= 0 || M.stex >= 0 || (M.tex >= 0 && h.texs[M.tex].channels))
            return fail(RM_ERR_UNSUPPORTED, "rm_adjoint: material %zu has 8-bit texture maps (fp32 albedo only)", m);
    }
    // Every indirect sample's sums come from the render's record.  A frame
    // rendered without recording would need the two-replay form (each sample
    // replayed once for its sums, once for its derivatives); that form is
    // refused, because one GPU run of it ended in a memory fault whose cause
    // has not been found.
    const int n_ind = a->spp - (int)(float(a->spp) * a->p_direct);
    if (n_ind > 0 && !f->rec_valid)
        return fail(RM_ERR_UNSUPPORTED, "rm_adjoint: the frame was rendered without recording "
                                        "(rm_frame_set_recording(frame, 1) before rm_render)");
    const int n_mats = (int)h.mats.size();
    if (n_mats > 4096) return fail(RM_ERR_UNSUPPORTED, "rm_adjoint: %d materials (at most 4096)", n_mats);
    hipStream_t st = (hipStream_t)stream;
    const int64_t np = (int64_t)f->w * f->h;
    const int64_t pool = (int64_t)(scene->texels.bytes / sizeof(float));
    const size_t pool_bytes = ((size_t)pool * 4 + 255) & ~(size_t)255;
    const size_t need = pool_bytes + (size_t)n_mats * 16;
    if (f->adj.bytes < need) {
        int32_t rc = f->adj.alloc(need);
        if (rc) return rc;
    }
    float *gtex = f->adj.as<float>();
    float *gmat = reinterpret_cast<float *>(f->adj.as<char>() + pool_bytes);
    int32_t rc;
    if (g->textures && (rc = fill(gtex, pool, st))) return rc;
    if ((rc = fill(gmat, (int64_t)n_mats * 4, st))) return rc;
    if (g->replay && (rc = fill(g->replay, np * 3, st))) return rc;
    AdjArgs G;
    G.S = scene->view;
    // k_adjoint's dynamic LDS holds the material-gradient sums: its walks keep
    // their stacks in scratch (trace_walk<..., false>), and the view says so
    G.S.walk_lds = 0;
    G.A = rargs_from(a);
    G.Fbase = f->buf[RM_BUF_BASE_COLOR].as<float>();
    G.drop = f->drop.as<int32_t>();
    G.dimg = d_image;
    G.gtex = g->textures ? gtex : nullptr;
    G.gmat = gmat;
    G.replay = g->replay;
    G.rec = f->rec_valid ? f->rec.as<float4>() : nullptr;
    G.rec_stride = f->rec_stride;
    G.np = np;
    G.split = tuning().adjoint_split;
    const int64_t lanes = np * G.split;
    hipLaunchKernelGGL(k_adjoint, dim3((unsigned)((lanes + ADJ_B - 1) / ADJ_B)), dim3(ADJ_B),
                       (size_t)n_mats * 4 * sizeof(float), st, G);
    RM_HIP(hipGetLastError());
    if (g->textures) {
        int64_t out = 0;
        for (const TexRec &T : h.texs) {
            for (int lv = T.levels - 1; lv >= 1; --lv) {
                int pw = T.width >> (lv - 1), ph = T.height >> (lv - 1), cw = T.width >> lv, ch = T.height >> lv;
                int64_t n = (int64_t)pw * ph;
                hipLaunchKernelGGL(k_mip_adjoint, dim3((unsigned)((n + ADJ_B - 1) / ADJ_B)), dim3(ADJ_B), 0, st, gtex,
                                   T.offset[lv - 1], pw, ph, T.offset[lv], cw, ch);
            }
            const int64_t n0 = (int64_t)T.width * T.height * 3;
            RM_HIP(hipMemcpyAsync(g->textures + out, gtex + T.offset[0], n0 * sizeof(float), hipMemcpyDeviceToDevice,
                                  st));
            out += n0;
        }
        RM_HIP(hipGetLastError());
    }
    if ((g->roughness || g->emission) && n_mats > 0) {
        hipLaunchKernelGGL(k_split_mat, dim3((unsigned)((n_mats + 63) / 64)), dim3(64), 0, st, (const float *)gmat,
                           n_mats, g->roughness, g->emission);
        RM_HIP(hipGetLastError());
    }
    RM_HIP(hipStreamSynchronize(st));
    return RM_OK;
}

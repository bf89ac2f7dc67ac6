// gfx950 shading: hit information, texture fetch, the reference BSDF and its
// samplers, and next-event estimation.  Each function follows the reference
// function it names, in the same fp32 operation order (built with
// -ffp-contract=off); only transcendental functions (sin, cos, log, pow, acos,
// atan2, tanh) come from the ROCm device library instead of glibc.
#pragma once
#include <hip/hip_runtime.h>
#include "rm_layout.h"
#include "rm_vec.h"
#include "tex8.h"
#include "traverse.h"

namespace rm {

constexpr float PI_F = 3.14159265358979323846f;
__device__ __forceinline__ V3 W_RGB() { return V3{0.3f, 0.6f, 0.1f}; }  // RGB_Weight
__device__ __forceinline__ V3 ld3(const float *p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 vnan() { return V3{NAN, NAN, NAN}; }

struct V2 { float x, y; };
__device__ __forceinline__ V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ V2 operator*(float s, V2 a) { return V2{s * a.x, s * a.y}; }
__device__ __forceinline__ float len2d(V2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
__device__ __forceinline__ bool any_finite2(V2 v) { return isfinite(v.x) || isfinite(v.y); }

// Shading state at a vertex (HitInfo, include/model.h:12-18, defaults from
// src/model.cpp:4-7).
struct Surf {
    V3 shapeN, surfN, emission, base, pos;
    float specular, roughness, metallic, opacity, eta;
    int id;
    bool entering;
};

__device__ __forceinline__ Surf surf_default() {
    Surf s;
    s.pos = vnan();
    s.shapeN = vnan();
    s.surfN = vnan();
    s.opacity = 1.0f;
    s.specular = 0.04f;
    s.roughness = 0.8f;
    s.metallic = 0.0f;
    s.eta = 1.0f;
    s.emission = v3(0.0f);
    s.base = v3(0.0f);
    s.entering = true;
    s.id = 0;
    return s;
}

struct Diff { V3 dPdx, dPdy, dDdx, dDdy; };  // RayDifferential

// barycentric (src/geometry.cpp:89-103)
__device__ __forceinline__ V3 barycentric(V3 v0, V3 v1, V3 v2, V3 P) {
    V3 a = v1 - v0, b = v2 - v0;
    V3 n = cross(a, b);
    float denom = dot(n, n);
    V3 p = P - v0;
    float alpha = dot(cross(p, b), n) / denom;
    float beta = dot(cross(a, p), n) / denom;
    float gamma = 1.0f - alpha - beta;
    return V3{gamma, alpha, beta};
}

// getTangentSpace / getTangentSpaceWithInDir (src/geometry.cpp:105-123)
__device__ __forceinline__ void tangent_space(V3 n, V3 &t, V3 &b) {
    // per-component selects: a select between two vector temporaries would
    // be made in memory (per-lane scratch)
    const bool ax = fabsf(n.x) < 0.8f;
    const V3 v0{ax ? 1.0f : 0.0f, ax ? 0.0f : 1.0f, 0.0f};
    t = normalize(cross(v0, n));
    b = cross(n, t);
}
// Both frames are computed and the result selected per component: with a
// branch, the compiler merges the two branches' final stores (into t in one,
// b in the other) into one store through a selected address, which keeps
// the frames in per-lane scratch.  The selected values are the branch's.
__device__ __forceinline__ void tangent_space_in(V3 n, V3 in, V3 &t, V3 &b) {
    const bool opposite = dot(n, in) < -1.0f + 1e-3f;
    V3 t1, b1;
    tangent_space(n, t1, b1);
    const V3 b2 = normalize(cross(n, in));
    const V3 t2 = cross(b2, n);
    t = V3{opposite ? t1.x : t2.x, opposite ? t1.y : t2.y, opposite ? t1.z : t2.z};
    b = V3{opposite ? b1.x : b2.x, opposite ? b1.y : b2.y, opposite ? b1.z : b2.z};
}

// reverseFix (src/model.cpp:232-238)
__device__ __forceinline__ bool reverse_fix(V3 &v, V3 dir) {
    if (dot(v, dir) < 0.0f) {
        v = v * -1.0f;
        return false;
    }
    return true;
}

// getHitNormals (src/model.cpp:240-270)
__device__ __forceinline__ void hit_normals(const TriShade &F, V3 in_dir, V3 bc, V3 &shapeN, V3 &surfN, bool &entering) {
    V3 v0 = ld3(F.v[0]), v1 = ld3(F.v[1]), v2 = ld3(F.v[2]);
    V3 c = cross(v1 - v0, v2 - v0);
    shapeN = normalize(c);
    entering = reverse_fix(shapeN, -in_dir);
    surfN = shapeN;
    float area = length(c) / 2.0f;
    if (area > 1e-2f) return;
    V3 n0 = ld3(F.n[0]), n1 = ld3(F.n[1]), n2 = ld3(F.n[2]);
    reverse_fix(n0, shapeN);
    reverse_fix(n1, shapeN);
    reverse_fix(n2, shapeN);
    V3 a = bc.x * (dot(n0, shapeN) > 0.85f ? n0 : shapeN);
    V3 b = bc.y * (dot(n1, shapeN) > 0.85f ? n1 : shapeN);
    V3 d = bc.z * (dot(n2, shapeN) > 0.85f ? n2 : shapeN);
    surfN = normalize(a + b + d);
    if (!any_finite(surfN)) surfN = shapeN;
}

// ImageData::get_bilinear + get (src/material.cpp:48-94) on fp32 texels
__device__ __forceinline__ V3 tex_bilinear(const float *data, int w, int h, float u, float v) {
    float x = u * float(w), y = v * float(h);
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float dx = x - float(x0), dy = y - float(y0);
    if (x0 < 0 || x0 >= w) { x0 %= w; if (x0 < 0) x0 += w; }
    if (y0 < 0 || y0 >= h) { y0 %= h; if (y0 < 0) y0 += h; }
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    V3 c00 = ld3(data + ((int64_t)y0 * w + x0) * 3), c01 = ld3(data + ((int64_t)y0 * w + x1) * 3);
    V3 c10 = ld3(data + ((int64_t)y1 * w + x0) * 3), c11 = ld3(data + ((int64_t)y1 * w + x1) * 3);
    V3 c0 = c00 * (1.0f - dx) + c01 * dx;
    V3 c1 = c10 * (1.0f - dx) + c11 * dx;
    return c0 * (1.0f - dy) + c1 * dy;
}

__device__ __forceinline__ V3 tex_get(const SceneView &S, const TexRec &T, float u, float v, float depth) {
    v = 1.0f - v;
    float hi = float(T.levels - 1);
    depth = gmax(gmin(depth, hi), 0.0f);  // std::max(std::min(depth, hi), 0)
    int level = (int)depth;
    int next = level + 1 < T.levels - 1 ? level + 1 : T.levels - 1;
    float blend = depth - float(level);
    V3 a = tex_bilinear(S.texels + T.offset[level], T.width >> level, T.height >> level, u, v);
    V3 b = tex_bilinear(S.texels + T.offset[next], T.width >> next, T.height >> next, u, v);
    return a * (1.0f - blend) + b * blend;
}

// Component 3 of ImageData::get<vec4> at level 0 (get_bilinear,
// src/material.cpp:48-94) on an alpha plane of a / 255 values.
__device__ __forceinline__ float alpha_bilinear(const float *data, int w, int h, float u, float v) {
    float x = u * float(w), y = v * float(h);
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float dx = x - float(x0), dy = y - float(y0);
    if (x0 < 0 || x0 >= w) { x0 %= w; if (x0 < 0) x0 += w; }
    if (y0 < 0 || y0 >= h) { y0 %= h; if (y0 < 0) y0 += h; }
    int x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
    float c00 = data[(int64_t)y0 * w + x0], c01 = data[(int64_t)y0 * w + x1];
    float c10 = data[(int64_t)y1 * w + x0], c11 = data[(int64_t)y1 * w + x1];
    float c0 = c00 * (1.0f - dx) + c01 * dx;
    float c1 = c10 * (1.0f - dx) + c11 * dx;
    return c0 * (1.0f - dy) + c1 * dy;
}


// TransparentTest (src/model.cpp:217-230): the hit (t, slot) of the ray
// (o, d) lies on a fully transparent texel of an alpha cut-out.  The level-1
// term of get() is multiplied by a blend of 0, so the level-0 value is the
// result exactly.
__device__ __forceinline__ bool transparent_at(const SceneView &S, V3 o, V3 d, float t, int32_t slot) {
    const TriShade &F = S.shade[slot];
    const MatRec &M = S.mats[F.material];
    if (!M.cutout) return false;
    V3 P = o + d * t;
    V3 bc = barycentric(ld3(F.v[0]), ld3(F.v[1]), ld3(F.v[2]), P);
    V2 u0{F.uv[0][0], F.uv[0][1]}, u1{F.uv[1][0], F.uv[1][1]}, u2{F.uv[2][0], F.uv[2][1]};
    V2 uv = (bc.x * u0 + bc.y * u1) + bc.z * u2;
    const TexRec &T = S.texs[M.tex];
    if (T.channels) {  // getDiffuseColor(u, v, NaN)[3]: the level-0 sample's fourth component
        float c[4];
        tex8_get(S.bytes, T, uv.x, uv.y, 0.0f, 4, c);
        return c[3] < T_EPS;
    }
    return alpha_bilinear(S.texels + T.alpha, T.width, T.height, uv.x, 1.0f - uv.y) < T_EPS;
}

// getDuv (src/model.cpp:272-277)
__device__ __forceinline__ V2 get_duv(const TriShade &F, V3 dP) {
    V3 v0 = ld3(F.v[0]);
    V3 b = barycentric(v0, ld3(F.v[1]), ld3(F.v[2]), v0 + dP);
    V2 u0{F.uv[0][0], F.uv[0][1]}, u1{F.uv[1][0], F.uv[1][1]}, u2{F.uv[2][0], F.uv[2][1]};
    return ((b.x - 1.0f) * u0 + b.y * u1) + b.z * u2;
}

// calcSurfaceNormal (src/model.cpp:279-298): the normal map's x and y tilt
// the normal in the face's uv tangent frame (its z is not used)
__device__ __forceinline__ void calc_surface_normal(const TriShade &F, V3 nm, V3 shapeN, V3 &surfN) {
    V3 v0 = ld3(F.v[0]);
    V3 e1 = ld3(F.v[1]) - v0, e2 = ld3(F.v[2]) - v0;
    V2 d1{F.uv[1][0] - F.uv[0][0], F.uv[1][1] - F.uv[0][1]};
    V2 d2{F.uv[2][0] - F.uv[0][0], F.uv[2][1] - F.uv[0][1]};
    float f = 1.0f / (d1.x * d2.y - d2.x * d1.y);
    V3 tu = f * (d2.y * e1 - d1.y * e2), tv = f * (-d2.x * e1 + d1.x * e2);
    V3 t = normalize(tu - shapeN * dot(shapeN, tu));
    V3 b = normalize(tv - shapeN * dot(shapeN, tv) - t * dot(t, tv));
    V3 sav = surfN;
    surfN = normalize(t * nm.x + b * nm.y + surfN);
    if (!any_finite(surfN)) surfN = sav;
}


// The mip depth Material's getters use: log2(duv * width), 0 for duv = NaN.
__device__ __forceinline__ float tex_depth(float duv, int width) {
    return isnan(duv) ? 0.0f : log2f(duv * float(width));
}

// Where a vertex's albedo came from: texture `tex` at (u, v) and mip depth
// `depth` as tex_get read it, or tex = -1 for a constant colour.
struct TexFoot {
    int tex;
    float u, v, depth;
};

// getHitMaterial (src/model.cpp:300-328).  `tf`, when given, receives the
// albedo footprint (the adjoint scatters texel gradients through it).
// MAPS: the scene may have 8-bit texture maps (the lookups are compiled in
// only then: they would cost every other scene registers).
template <bool MAPS = true>
__device__ __forceinline__ void hit_material(const SceneView &S, const TriShade &F, V3 bc, V3 dPdx, V3 dPdy, Surf &h,
                                             TexFoot *tf = nullptr) {
    V2 u0{F.uv[0][0], F.uv[0][1]}, u1{F.uv[1][0], F.uv[1][1]}, u2{F.uv[2][0], F.uv[2][1]};
    V2 uv = (bc.x * u0 + bc.y * u1) + bc.z * u2;
    const MatRec &M = S.mats[F.material];
    h.id = F.material;
    h.roughness = M.roughness;  // getSurfaceData without a specular map
    h.metallic = M.metallic;
    h.opacity = M.opacity;
    h.eta = M.ior;
    if (h.opacity > 1.0f - T_EPS) h.entering = true;
    V2 ddx = get_duv(F, dPdx), ddy = get_duv(F, dPdy);
    float duv = any_finite2(ddx) ? (len2d(ddx) + len2d(ddy)) / 2.0f : NAN;
    if (tf) tf->tex = -1;
    int dtex = -1;  // an 8-bit diffuse texture
    if (h.opacity < T_EPS) {
        h.base = ld3(M.trans);
    } else if (M.tex >= 0) {
        const TexRec &T = S.texs[M.tex];
        if (T.channels) {
            dtex = M.tex;
        } else {
            float depth = tex_depth(duv, T.width);
            h.base = tex_get(S, T, uv.x, uv.y, depth);
            if (tf) *tf = TexFoot{M.tex, uv.x, uv.y, depth};
        }
    } else {
        h.base = ld3(M.base);
    }
    h.emission = ld3(M.emission);
    V3 nm = v3(0.0f);
    // The 8-bit maps, in one loop so that the kernel holds one copy of the
    // lookup: 0 getSurfaceData (the level-0 sample; metallic = min(b, 0.99),
    // roughness = max(g, 1e-3)), 1 getDiffuseColor and 2 getEmissiveColor
    // (gammaPow after filtering), 3 getNormal (a vec3 lookup, * 2 - 1)
    // (src/material.cpp:337-383).
    if (MAPS && (M.stex >= 0 || dtex >= 0 || M.etex >= 0 || M.ntex >= 0)) {
#pragma nounroll
        for (int k = 0; k < 4; ++k) {
            const int ti = k == 0 ? M.stex : k == 1 ? dtex : k == 2 ? M.etex : M.ntex;
            if (ti < 0) continue;
            const TexRec &T = S.texs[ti];
            float c[4];
            tex8_get(S.bytes, T, uv.x, uv.y, k == 0 ? 0.0f : tex_depth(duv, T.width), k == 3 ? 3 : 4, c);
            V3 x{c[0], c[1], c[2]};
            if (k == 1 || k == 2) {
                // powf(x, 2.2f), evaluated in fp64 and rounded once
                const double g = (double)2.2f;
#pragma nounroll
                for (int j = 0; j < 3; ++j) {
                    const float y = (float)pow((double)(j == 0 ? x.x : j == 1 ? x.y : x.z), g);
                    x.x = j == 0 ? y : x.x;
                    x.y = j == 1 ? y : x.y;
                    x.z = j == 2 ? y : x.z;
                }
            }
            if (k == 0) {
                h.metallic = gmin(c[2], 0.99f);   // std::min(b, 0.99)
                h.roughness = gmax(c[1], 1e-3f);  // std::max(g, 1e-3)
            } else if (k == 1) {
                h.base = x;
            } else if (k == 2) {
                h.emission = x;
            } else {
                nm = V3{x.x * 2.0f - 1.0f, x.y * 2.0f - 1.0f, x.z * 2.0f - 1.0f};
            }
        }
    }
    calc_surface_normal(F, nm, h.shapeN, h.surfN);
}

// calc_dPdxy (src/render.cpp:44-50)
__device__ __forceinline__ void calc_dPdxy(V3 dir, float t, V3 n, const Diff &b, V3 &dPdx, V3 &dPdy) {
    float dtdx = -dot(b.dPdx + t * b.dDdx, n) / dot(dir, n);
    float dtdy = -dot(b.dPdy + t * b.dDdy, n) / dot(dir, n);
    dPdx = (b.dPdx + dtdx * dir) + t * b.dDdx;
    dPdy = (b.dPdy + dtdy * dir) + t * b.dDdy;
}

// calc_dDdxy (src/render.cpp:52-60); the normal derivatives are zero
__device__ __forceinline__ void calc_dDdxy(V3 dir, V3 n, const Diff &b, V3 &dDdx, V3 &dDdy) {
    V3 z = v3(0.0f);
    float ax = dot(z, dir) + dot(b.dDdx, n);
    float ay = dot(z, dir) + dot(b.dDdy, n);
    dDdx = b.dDdx - 2.0f * (dot(dir, n) * z + ax * n);
    dDdy = b.dDdy - 2.0f * (dot(dir, n) * z + ay * n);
}

// getHitInfo (src/render.cpp:62-79); h.pos must be set by the caller
template <bool MAPS = true>
__device__ __forceinline__ void hit_info(const SceneView &S, int32_t slot, V3 dir, float t, const Diff &base, V3 &dPdx,
                                         V3 &dPdy, Surf &h, TexFoot *tf = nullptr) {
    const TriShade &F = S.shade[slot];
    V3 bc = barycentric(ld3(F.v[0]), ld3(F.v[1]), ld3(F.v[2]), h.pos);
    hit_normals(F, dir, bc, h.shapeN, h.surfN, h.entering);
    V3 raw = h.surfN;
    calc_dPdxy(dir, t, h.shapeN, base, dPdx, dPdy);
    hit_material<MAPS>(S, F, bc, dPdx, dPdy, h, tf);
    if (dot(h.surfN, dir) >= 0.0f) h.surfN = raw;
    if (dot(h.surfN, dir) >= 0.0f) h.surfN = h.shapeN;
}

// SkyBox::get (src/component.cpp:120-140)
__device__ __forceinline__ V3 sky_get(const SceneView &S, V3 dir) {
    if (!S.sky) return v3(0.0f);
    float theta = atan2f(-dir.x, dir.z);
    float phi = acosf(dir.y);
    if (theta < 0.0f) theta += 2.0f * PI_F;
    int u = (int)(theta / (2.0f * PI_F) * float(S.sky_w));
    int v = (int)(phi / PI_F * float(S.sky_h));
    if (u < 0) u = 0;
    if (u >= S.sky_w) u = S.sky_w - 1;
    if (v < 0) v = 0;
    if (v >= S.sky_h) v = S.sky_h - 1;
    phi = PI_F * (float(v) + 0.5f) / float(S.sky_h);
    float area = sinf(phi) * 2.0f * PI_F / float(S.sky_w * S.sky_h);
    return ld3(S.sky + ((int64_t)v * S.sky_w + u) * 3) / area;
}

// ------------------------------------------------------------------ BSDF
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1 - t) + b * t; }
__device__ __forceinline__ V3 mix3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float schlick(float u) {
    float m = 1 - u;
    m = m < 0 ? 0 : (m > 1 ? 1 : m);
    float m2 = m * m;
    return m2 * m2 * m;
}
__device__ __forceinline__ float gtr1(float NdotH, float a) {
    if (a >= 1) return 1 / PI_F;
    float a2 = a * a;
    float t = 1 + (a2 - 1) * NdotH * NdotH;
    return (a2 - 1) / (PI_F * logf(a2) * t);
}
__device__ __forceinline__ float gtr2(float NdotH, float a) {
    float a2 = a * a;
    float t = 1 + (a2 - 1) * NdotH * NdotH;
    return a2 / (PI_F * t * t);
}
__device__ __forceinline__ float smith_ggx(float NdotV, float alphaG) {
    float a = alphaG * alphaG;
    float b = NdotV * NdotV;
    return 1 / (NdotV + sqrtf(a + b - a * b));
}
__device__ __forceinline__ float sqrt_s(float x) {
    if (x < -T_EPS) return 0.0f;
    return x <= 0.0f ? 0.0f : sqrtf(x);
}

// BSDF::getBRDF (src/sampling.cpp:49-157); returns f * N.L
__device__ __forceinline__ V3 brdf(const Surf &s, V3 V, V3 L) {
    V3 N = s.surfN;
    float NdotL = dot(N, L), NdotV = dot(N, V);
    if (NdotL <= 0.0f || NdotV <= 0.0f) return v3(0.0f);
    V3 H = normalize(L + V);
    float NdotH = dot(N, H), LdotH = dot(L, H);
    V3 Cd = s.base;
    float Cdlum = dot(Cd, W_RGB());
    V3 Ctint = Cdlum > 0.0f ? Cd / Cdlum : v3(1.0f);
    V3 Cspec0 = mix3(s.specular * mix3(v3(1.0f), Ctint, 0.0f), Cd, s.metallic);
    V3 Csheen = mix3(v3(1.0f), Ctint, 0.0f);
    float FL = schlick(NdotL), FV = schlick(NdotV);
    float Fd90 = 0.5f + 2.0f * LdotH * LdotH * s.roughness;
    float Fd = mixf(1.0f, Fd90, FL) * mixf(1.0f, Fd90, FV);
    float Fss90 = LdotH * LdotH * s.roughness;
    float Fss = mixf(1.0f, Fss90, FL) * mixf(1.0f, Fss90, FV);
    float ss = 1.25f * (Fss * (1.0f / (NdotL + NdotV) - .5f) + .5f);
    float Ds = gtr2(NdotH, s.roughness);
    float FH;
    if (s.eta <= 1.0f) {
        FH = schlick(LdotH);
    } else {
        float ci = fabsf(LdotH);
        float si = sqrt_s(1.0f - ci * ci);
        float st = s.eta * si;
        if (st >= 1.0f) {
            FH = 1.0f;
        } else {
            float ct = sqrt_s(1.0f - st * st);
            float R0 = (s.eta - 1.0f) / (s.eta + 1.0f);
            R0 = R0 * R0;
            FH = mixf(R0, 1.0f, schlick(ct));
        }
    }
    V3 Fs = mix3(Cspec0, v3(1.0f), FH);
    float Gs = smith_ggx(NdotL, s.roughness);
    Gs *= smith_ggx(NdotV, s.roughness);
    V3 Fsheen = FH * 0.0f * Csheen;
    float Dr = gtr1(NdotH, mixf(.1f, .001f, 0.2f));
    float Fr = mixf(.04f, 1.0f, FH);
    float Gr = smith_ggx(NdotL, .25f) * smith_ggx(NdotV, .25f);
    V3 ret = (1.0f / PI_F) * mixf(Fd, ss, 0.0f) * Cd + Fsheen;
    ret *= (1.0f - s.metallic);
    ret += (0.25f * 1.5f * Gr * Fr * Dr) * mix3(v3(1.0f), Ctint, 0.0f);
    ret *= s.opacity;
    ret += Fs * Ds * Gs;
    return ret * NdotL;
}

// d gtr2 / d a and d smith_ggx / d alphaG (closed forms of the two functions above)
__device__ __forceinline__ float gtr2_da(float NdotH, float a) {
    float a2 = a * a, c2 = NdotH * NdotH;
    float t = 1 + (a2 - 1) * c2;
    return 2.0f * a / (PI_F * t * t) * (1.0f - 2.0f * a2 * c2 / t);
}
__device__ __forceinline__ float smith_ggx_da(float NdotV, float alphaG) {
    float a = alphaG * alphaG, b = NdotV * NdotV;
    float r = sqrtf(a + b - a * b);
    float den = NdotV + r;
    return -(alphaG * (1.0f - b) / r) / (den * den);
}
// d log D(H) / d a for the GGX distribution gtr2 at cos(theta_h) = c: the
// score of one half-vector draw of sample_ggx.
__device__ __forceinline__ float gtr2_dlog(float c, float a) {
    float c2 = c * c;
    float t = 1 + (a * a - 1) * c2;
    return 2.0f / a - 4.0f * a * c2 / t;
}

// Partial derivatives of brdf() at fixed directions.  The albedo enters
// linearly and identically in each channel (diffuse lobe and metallic
// specular tint), so d brdf / d base = kd * I; d brdf / d roughness = dr.
__device__ __forceinline__ void brdf_grad(const Surf &s, V3 V, V3 L, float &kd, V3 &dr) {
    V3 N = s.surfN;
    float NdotL = dot(N, L), NdotV = dot(N, V);
    kd = 0.0f;
    dr = v3(0.0f);
    if (NdotL <= 0.0f || NdotV <= 0.0f) return;
    V3 H = normalize(L + V);
    float NdotH = dot(N, H), LdotH = dot(L, H);
    const float r = s.roughness, m = s.metallic;
    float FL = schlick(NdotL), FV = schlick(NdotV);
    float Fd90 = 0.5f + 2.0f * LdotH * LdotH * r;
    float mL = mixf(1.0f, Fd90, FL), mV = mixf(1.0f, Fd90, FV);
    float Fd = mL * mV;
    float dFd = 2.0f * LdotH * LdotH * (FL * mV + mL * FV);
    float Ds = gtr2(NdotH, r), dDs = gtr2_da(NdotH, r);
    float FH;
    if (s.eta <= 1.0f) {
        FH = schlick(LdotH);
    } else {
        float ci = fabsf(LdotH);
        float si = sqrt_s(1.0f - ci * ci);
        float st = s.eta * si;
        if (st >= 1.0f) {
            FH = 1.0f;
        } else {
            float ct = sqrt_s(1.0f - st * st);
            float R0 = (s.eta - 1.0f) / (s.eta + 1.0f);
            R0 = R0 * R0;
            FH = mixf(R0, 1.0f, schlick(ct));
        }
    }
    V3 Cspec0 = mix3(v3(s.specular), s.base, m);
    V3 Fs = mix3(Cspec0, v3(1.0f), FH);
    float gl = smith_ggx(NdotL, r), gv = smith_ggx(NdotV, r);
    float Gs = gl * gv, dGs = smith_ggx_da(NdotL, r) * gv + gl * smith_ggx_da(NdotV, r);
    kd = ((1.0f / PI_F) * Fd * (1.0f - m) * s.opacity + m * (1.0f - FH) * Ds * Gs) * NdotL;
    dr = (((1.0f / PI_F) * dFd * (1.0f - m) * s.opacity) * s.base + Fs * (dDs * Gs + Ds * dGs)) * NdotL;
}

// BSDF::getBTDF (src/sampling.cpp:159-192)
__device__ __forceinline__ V3 btdf(const Surf &s, V3 V, V3 L) {
    V3 N = s.surfN;
    float NdotL = dot(N, L);
    if (NdotL >= 0.0f) return v3(0.0f);
    V3 H = normalize(L + s.eta * V);
    float D = gtr2(dot(N, H), s.roughness);
    float B = D * (-NdotL);
    float LdotH = dot(L, H), NdotV = dot(N, V), HdotV = dot(H, V);
    B *= fabsf(LdotH * HdotV) / (fabsf(NdotL * NdotV) + T_EPS);
    float k = s.eta * HdotV + LdotH;
    B /= k * k;
    return v3(B);
}

// BSDF::getBSDF (src/sampling.cpp:194-199)
__device__ __forceinline__ V3 bsdf_eval(const Surf &s, V3 V, V3 L) {
    if (s.opacity > 1.0f - T_EPS || s.entering) return brdf(s, V, L);
    return btdf(s, V, L);
}

// clamp (src/sampling.cpp:238-243)
__device__ __forceinline__ void clamp64(V3 &p) {
    float c = dot(p, W_RGB());
    if (c > 64.0f) p /= c / 64.0f;
}

// Observer of the reflection samplers.  The renderer passes NoTap (every call
// compiles away); the adjoint passes a recorder of the values its derivative
// needs: every GGX half-vector cosine drawn, the accepted GGX sample, the
// cosine sample's value before the clamp, and the pick between the two.
struct NoTap {
    __device__ __forceinline__ void ggx_try(float) {}
    __device__ __forceinline__ void ggx_ok(float, float) {}
    __device__ __forceinline__ void cos_raw(V3) {}
    __device__ __forceinline__ void pick(float, float, bool) {}
};

// BSDF::sampleBRDF (src/sampling.cpp:203-236)
template <class Tap>
__device__ __forceinline__ void sample_ggx(const Surf &s, V3 V, Stream &g, V3 &out, V3 &bp, float &pdf, int &fails,
                                           Tap &tap) {
    V3 t, b;
    tangent_space_in(s.surfN, V, t, b);
    float r2 = s.roughness * s.roughness;
    for (int T = 1; T <= 16; ++T) {
        float u = g.next(), phi = g.next() * 2.0f * PI_F;
        float ct = sqrtf((1.0f - u) / (1.0f + (r2 - 1.0f) * u)), st = sqrtf(1 - ct * ct);
        tap.ggx_try(ct);
        V3 H = (st * cosf(phi) * t + st * sinf(phi) * b) + ct * s.surfN;
        out = 2.0f * dot(V, H) * H - V;
        float LdotH = dot(out, H), LdotN = dot(out, s.surfN);
        if (LdotH <= 0.0f || LdotN <= 0.0f)
            pdf = 0.0f;
        else
            pdf = gtr2(ct, s.roughness) / (4.0f * LdotH);
        if (dot(out, s.shapeN) > 0.0f && pdf > 0.0f) {
            tap.ggx_ok(ct, LdotH);
            bp = brdf(s, V, out) / pdf;
            return;
        }
        fails++;
    }
    pdf = 0.0f;
    bp = v3(0.0f);
    out = vnan();
}

// BSDF::sampleCos (src/sampling.cpp:245-269)
template <class Tap>
__device__ __forceinline__ void sample_cos(const Surf &s, V3 V, Stream &g, V3 &out, V3 &bp, float &pdf, int &fails,
                                           Tap &tap) {
    V3 t, b;
    tangent_space_in(s.surfN, V, t, b);
    for (int T = 1; T <= 16; ++T) {
        float u = g.next(), phi = g.next() * 2.0f * PI_F;
        float d = sqrtf(u), z = sqrtf(1 - d * d);
        float x = d * cosf(phi), y = d * sinf(phi);
        out = (x * t + y * b) + z * s.surfN;
        pdf = dot(out, s.surfN) / PI_F;
        if (dot(out, s.shapeN) > 0.0f && pdf > 0.0f) {
            bp = brdf(s, V, out) / pdf;
            tap.cos_raw(bp);
            clamp64(bp);
            return;
        }
        fails++;
    }
    pdf = 0.0f;
    bp = v3(0.0f);
    out = vnan();
}

// BSDF::preciseRefraction (src/sampling.cpp:271-308)
__device__ __forceinline__ void precise_refraction(const Surf &s, V3 V, V3 &out, float &F) {
    V3 N = s.surfN;
    float eta = s.eta;
    float NdotV = dot(N, V);
    if (NdotV < 0.0f) { F = 1.0f; out = vnan(); return; }
    float delta = 1.0f - eta * eta * (1.0f - NdotV * NdotV);
    if (delta < 0.0f) { F = 1.0f; out = vnan(); return; }
    float k = eta * NdotV - sqrtf(delta);
    out = k * N - eta * V;
    float LdotN = -dot(out, N);
    float R0 = (1.0f - eta) / (1.0f + eta);
    R0 = R0 * R0;
    F = R0 + (1.0f - R0) * schlick(LdotN);
}

// BSDF::sampleReflection (src/sampling.cpp:310-340)
template <class Tap>
__device__ __forceinline__ void sample_reflection(const Surf &s, V3 V, Stream &g, V3 &dir, V3 &bp, int &fails,
                                                  Tap &tap) {
    V3 d1, d2, b1, b2;
    int f1 = 0, f2 = 0;
    float p1, p2;
    sample_cos(s, V, g, d1, b1, p1, f1, tap);
    sample_ggx(s, V, g, d2, b2, p2, f2, tap);
    float pc = p1 / (p1 + p2);
    bool first = g.next() < pc;
    tap.pick(p1, p2, first);
    if (first) {
        dir = d1; bp = b1; fails += f1;
    } else {
        dir = d2; bp = b2; fails += f2;
    }
}
__device__ __forceinline__ void sample_reflection(const Surf &s, V3 V, Stream &g, V3 &dir, V3 &bp, int &fails) {
    NoTap tap;
    sample_reflection(s, V, g, dir, bp, fails, tap);
}

// refract (src/sampling.cpp:342-349)
__device__ __forceinline__ V3 refract_dir(V3 V, V3 N, float eta) {
    float NdotV = dot(N, V);
    float delta = 1.0f - eta * eta * (1.0f - NdotV * NdotV);
    if (delta < 0.0f) return vnan();
    float k = eta * NdotV - sqrtf(delta);
    return k * N - eta * V;
}

// BSDF::sampleBTDF (src/sampling.cpp:351-391)
__device__ __forceinline__ void sample_btdf(const Surf &s, V3 V, Stream &g, V3 &out, V3 &bp, int &fails) {
    if (fabsf(s.eta - 1.0f) < T_EPS) {
        out = -V;
        bp = v3(1.0f);
        return;
    }
    V3 t, b;
    tangent_space_in(s.surfN, V, t, b);
    float r2 = s.roughness * s.roughness;
    for (int T = 1; T <= 16; ++T) {
        float u = g.next(), phi = g.next() * 2.0f * PI_F;
        float ct = sqrtf((1.0f - u) / (1.0f + (r2 - 1.0f) * u)), st = sqrtf(1 - ct * ct);
        V3 H = (st * cosf(phi) * t + st * sinf(phi) * b) + ct * s.surfN;
        out = refract_dir(V, H, s.eta);
        float w;
        if (!any_finite(out)) {
            w = 0.0f;
        } else {
            float VdotH = dot(V, H), VdotN = dot(V, s.surfN), HdotN = dot(H, s.surfN);
            w = fabsf(VdotH) / fabsf(VdotN * HdotN);
        }
        if (w > 0.0f && dot(out, s.shapeN) < 0.0f) {
            bp = v3(w);
            return;
        }
        fails++;
    }
    bp = v3(0.0f);
    out = vnan();
}

// ------------------------------------------------------------ light samples
struct LSample {
    V3 bsdf, light;
    float weight;
};

// RandomDistribution::operator() (src/component.cpp:20-24): lower_bound of
// total * u in the prefix sums [0, n)
__device__ __forceinline__ int cdf_pick(const float *cdf, int n, float r) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] < r) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The same lower_bound on the sky CDF, restricted to the bracket the guide
// table gives for r (the bucket of r computed in fp32 is off by at most one,
// so buckets b-1 .. b+2 are searched).
__device__ __forceinline__ int sky_pick(const SceneView &S, int n, float tot, float r) {
    int b = (int)(r * ((float)SKY_GUIDE / tot));
    b = b < 0 ? 0 : (b > SKY_GUIDE - 1 ? SKY_GUIDE - 1 : b);
    int lo = S.sky_guide[b > 0 ? b - 1 : 0];
    int hi = S.sky_guide[b + 2 < SKY_GUIDE ? b + 2 : SKY_GUIDE];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (S.sky_cdf[mid] < r) lo = mid + 1; else hi = mid;
    }
    (void)n;
    return lo;
}

// getLightObjectWeight (src/sampling.cpp:406-417) for one light
__device__ __forceinline__ float light_weight(const LightRec &L, const Surf &s, V3 V) {
    V3 c = ld3(L.center);
    V3 dir = normalize(c - s.pos);
    float dist = length(c - s.pos);
    V3 b = bsdf_eval(s, V, dir);
    float cl = dot(b, W_RGB());
    return cl * L.power / (dist * dist + 1e-3f);
}

// What an accepted light sample was made of: direction, distance to the light
// point, the light object's pick probability and index (-1 for the sky).
struct DirectInfo {
    V3 dir;
    float dist, P;
    int li;
};

// The shadow ray of a light sample: from the surface point along dir; the
// sample counts when the closest hit on (T_EPS, aim + T_EPS) is not below aim
// (aim = INFINITY toward the sky).
struct ShadowRay {
    V3 dir;
    float aim;
};

// One light sample of sampleDirectLight (src/sampling.cpp:467-527) up to its
// shadow ray: the ray, the sample as it counts when the ray is clear, and
// what it was made of.  Returns false when the reference `continue`s before
// tracing.  The shadow ray draws no random numbers, so a caller may trace it
// later.
__device__ __forceinline__ bool direct_ray(const SceneView &S, const Surf &s, V3 V, float total, int cnt, Stream &g,
                                           ShadowRay &sr, LSample &out, DirectInfo *info = nullptr) {
    if (S.sky) {
        // sampleSkyBox (src/sampling.cpp:450-465)
        const int n = S.sky_w * S.sky_h;
        const float tot = S.sky_cdf[n - 1];
        V3 dir = vnan(), light = v3(0.0f);
        bool ok = false;
        for (int T = 1; T <= 16; ++T) {
            int pi = sky_pick(S, n, tot, tot * g.next());
            int u = pi % S.sky_w, v = pi / S.sky_w;
            float phi = PI_F * (float(v) + 0.5f) / float(S.sky_h);
            float th = 2.0f * PI_F * (float(u) + 0.5f) / float(S.sky_w);
            float sp = sinf(phi);
            dir = V3{-sp * sinf(th), cosf(phi), sp * cosf(th)};
            if (dot(dir, s.surfN) > 0.0f) {
                float now = S.sky_cdf[pi];
                if (pi > 0) now -= S.sky_cdf[pi - 1];
                light = ld3(S.sky + (int64_t)pi * 3) / (now / tot);
                ok = true;
                break;
            }
        }
        if (!ok) return false;
        sr = ShadowRay{dir, INFINITY};
        V3 b = bsdf_eval(s, V, dir);
        clamp64(b);
        out = LSample{b, light, 1.0f / float(cnt)};
        if (info) *info = DirectInfo{dir, INFINITY, 1.0f, -1};
        return true;
    }
    // emissive light objects: sample() walks the weights (src/sampling.cpp:396-404)
    float r = total * g.next();
    int li = -1;
    float cum = 0.0f, wsel = 0.0f;
    for (int i = 0; i < S.n_lights; ++i) {
        float w = light_weight(S.lights[i], s, V);
        cum += w;
        if (r <= cum) { li = i; wsel = w; break; }
    }
    if (li < 0) return false;
    float P = wsel / total;
    const LightRec &L = S.lights[li];
    V3 lp = vnan();
    int fails = 0;
    // sampleLightFace (src/sampling.cpp:436-448) + generateRandomPointInLightFace (:419-434)
    bool found = false;
    for (int T = 1; T <= 16; ++T) {
        const float *cdf = S.light_cdf + L.first_face;
        int fi = cdf_pick(cdf, L.n_faces, cdf[L.n_faces - 1] * g.next());
        const TriShade &F = S.light_faces[L.first_face + fi];
        float a = g.next(), b = g.next();
        if (a + b > 1.0f) { a = 1.0f - a; b = 1.0f - b; }
        float c = 1.0f - a - b;
        V3 p = (a * ld3(F.v[0]) + b * ld3(F.v[1])) + c * ld3(F.v[2]);
        V3 ldir = normalize(p - s.pos);
        V3 sn, un;
        bool ent;
        hit_normals(F, ldir, V3{a, b, c}, sn, un, ent);
        float cphi = dot(un, -ldir);
        lp = p;
        if (cphi > 0.0f && g.next() < cphi) { found = true; break; }
        fails++;
    }
    if (!found) return false;
    V3 dir = normalize(lp - s.pos);
    float dist = length(lp - s.pos);
    sr = ShadowRay{dir, dist - T_EPS};
    V3 b = bsdf_eval(s, V, dir);
    clamp64(b);
    V3 light = ((2.0f * PI_F * L.power) * ld3(L.color)) / (dist * dist + 1e-3f) / P;
    out = LSample{b, light, 1.0f / float(cnt * (fails + 1))};
    if (info) *info = DirectInfo{dir, dist, P, li};
    return true;
}

// The rest of Model::rayHit (src/model.cpp:332-341) for a scene with alpha
// cut-outs, given the first walk's result (t, slot) over [tmin, tmax0]: while
// the hit is transparent, walk again over [t + T_EPS, tmax0]; 8 walks in all,
// after which the ray is a miss.  Returns the slot (-1: miss, t = tmax0).
// The reference's tmax0 is inf.
template <bool LDSW = true, int TPS = 0>
__device__ __noinline__ int32_t model_hit_rest(const SceneView &S, V3 o, V3 d, float &t, int32_t slot,
                                               float tmax0 = INFINITY) {
    const RayPre r = ray_pre(o, d);
    for (int T = 0; T < 8; ++T) {
        if (slot < 0 || !transparent_at(S, o, d, t, slot)) return slot;
        if (T == 7) break;
        const float tmin = t + T_EPS;
        t = tmax0;
        slot = trace_walk<false, LDSW, TPS>(S, r, tmin, t, -INFINITY);
    }
    t = tmax0;
    return -1;
}

// Model::rayHit (src/model.cpp:332-341) from the first walk on, around the
// caller's one inlined walk: without alpha cut-outs the loop ends after that
// walk, so kernels keep one copy of the traversal and no call.
template <bool LDSW = true, int TPS = 0>
__device__ __forceinline__ int32_t model_closest(const SceneView &S, V3 o, V3 d, const RayPre &r, float &t) {
    float tmin = T_EPS;
    for (int T = 0;; ++T) {
        t = INFINITY;
        const int32_t slot = trace_walk<false, LDSW, TPS>(S, r, tmin, t, -INFINITY);
        if (!S.cutout || slot < 0 || !transparent_at(S, o, d, t, slot)) return slot;
        if (T == 7) break;
        tmin = t + T_EPS;
    }
    t = INFINITY;
    return -1;
}

// Model::rayHit_test (src/model.cpp:343-354) as "the shadow ray is clear",
// around the caller's one inlined any-hit walk.  Without alpha cut-outs that
// walk stops at the first t below aim and decides.  With them it runs with
// aim -inf, which never stops it early, so it is the closest walk the loop
// needs (the walk's ANY flag only adds those stops).
template <bool LDSW = true, int TPS = 0>
__device__ __forceinline__ bool model_clear(const SceneView &S, V3 o, V3 d, const RayPre &r, float aim) {
    float tmin = T_EPS;
    const float stop = S.cutout ? -INFINITY : aim;
    for (int T = 0; T < 8; ++T) {
        float tmax = aim + T_EPS;
        const int32_t slot = trace_walk<true, LDSW, TPS>(S, r, tmin, tmax, stop);
        if (!(tmax < aim)) return true;
        if (!S.cutout || !transparent_at(S, o, d, tmax, slot)) return false;
        tmin = tmax + T_EPS;
    }
    return false;
}

// Model::rayHit_test (src/model.cpp:343-354) in full, for a shadow ray that
// the early-exit any-hit walk found blocked in a scene with alpha cut-outs:
// closest walks over [tmin, aim + T_EPS], each from the last hit + T_EPS,
// until a hit at or beyond aim (clear), an opaque hit (blocked) or 8 walks
// (blocked).
template <bool LDSW = true, int TPS = 0>
__device__ __noinline__ bool model_blocked(const SceneView &S, V3 o, V3 d, float tmin, float aim) {
    const RayPre r = ray_pre(o, d);
    for (int T = 0; T < 8; ++T) {
        float t = aim + T_EPS;
        const int32_t slot = trace_walk<false, LDSW, TPS>(S, r, tmin, t, -INFINITY);
        if (t >= aim) return false;
        if (!transparent_at(S, o, d, t, slot)) return true;
        tmin = t + T_EPS;
    }
    return true;
}

// One light sample with its shadow ray traced here.  Returns false when the
// reference `continue`s (no sample, or the shadow ray is blocked).  `nrays`,
// when given, counts the shadow rays traced.
template <bool COUNT, bool LDSW = true, int TPS = 0>
__device__ __forceinline__ bool direct_sample(const SceneView &S, const Surf &s, V3 V, float total, int cnt, Stream &g,
                                              LSample &out, TraceCounts &tc, DirectInfo *info = nullptr,
                                              uint32_t *nrays = nullptr) {
    ShadowRay sr;
    if (!direct_ray(S, s, V, total, cnt, g, sr, out, info)) return false;
    if (nrays) ++*nrays;
    RayPre r = ray_pre(s.pos, sr.dir);
    return model_clear<LDSW, TPS>(S, s.pos, sr.dir, r, sr.aim);
}

// Total light-object weight of sampleDirectLight (0 -> no samples); 1 with a sky.
__device__ __forceinline__ float direct_total(const SceneView &S, const Surf &s, V3 V) {
    if (S.sky) return 1.0f;
    float tot = 0.0f;
    for (int i = 0; i < S.n_lights; ++i) tot += light_weight(S.lights[i], s, V);
    return tot;
}

// accumulateInwardRadiance (src/image.cpp:615-659) as increments of the
// diffuse (d) and specular (s) accumulators (rgb + Var).
struct Acc4 { float r, g, b, v; };
__device__ __forceinline__ void acc_basic(Acc4 &a, V3 inr, float w) {
    if (!any_finite(inr) || !isfinite(w)) return;
    V3 x = inr * w;
    a.r += x.x; a.g += x.y; a.b += x.z;
    a.v += dot(inr, inr) * w;
}
// The increments are found first and added by two calls with fixed targets:
// one call whose target depends on the branch would address d or s through
// a selected pointer, which keeps them in per-lane scratch.
__device__ __forceinline__ void accumulate(V3 base, const LSample &smp, Acc4 &d, Acc4 &s) {
    if (length(smp.light) < T_EPS) return;
    V3 b0 = normalize(base);
    bool to_d = true, to_s = true;
    V3 xd = v3(0.0f), xs = v3(0.0f);
    if (length(base) < T_EPS) {
        to_d = false;
        xs = smp.light * smp.bsdf;
    } else {
        const V3 white = normalize(v3(1.0f));
        float xy = dot(b0, white);
        if (xy > 0.99f) {
            to_s = false;
            xd = smp.light * smp.bsdf / base;
        } else {
            V3 perp = normalize(cross(b0, white));
            V3 bp = smp.bsdf - perp * dot(perp, smp.bsdf);
            float d1 = dot(bp, white), d2 = dot(bp, b0);
            float apb = (d1 + d2) / (1 + xy), amb = (d1 - d2) / (1 - xy);
            float B = (apb - amb) / 2.0f;
            V3 bb = B * b0;
            xd = smp.light * B / length(base);
            xs = smp.light * (smp.bsdf - bb);
        }
    }
    if (to_d) acc_basic(d, xd, smp.weight);
    if (to_s) acc_basic(s, xs, smp.weight);
}

}  // namespace rm

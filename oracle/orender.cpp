// ORACLE -- test infrastructure only (see oracle.h).
// Restates the integrator of src/render.cpp:9-626 and the post path of
// src/image.cpp:215-246,358-479.
//
// The only structural change is the random source: each primary-vertex light
// sample T of pixel p draws from stream (seed, TAG_DIRECT, p, T) and each
// indirect sample T from stream (seed, TAG_INDIRECT, p, T), in the reference's
// draw order; the reference draws everything from one mt19937 per thread.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <thread>
#include "oracle.h"

namespace ref {

namespace {

struct MediumData {                                      // include/render.h:17-21
    float ior;
    vec3 absorb;
    MediumData(float i, const vec3 &a) : ior(i), absorb(a) {}
};

class Medium {                                           // src/render.cpp:13-42
    std::multimap<int, MediumData> mediums;
public:
    void Init() { mediums.clear(); mediums.emplace(-1, MediumData(1.0f, vec3(1.0f))); }
    void insert(int id, float ior, const vec3 &absorb) { mediums.emplace(id, MediumData(ior, absorb)); }
    void erase(int id) { mediums.erase(id); }
    float ior() const {
        float r = 1.0f;
        for (const auto &m : mediums) r = std::max(r, m.second.ior);
        return r;
    }
    vec3 absorb() const {
        vec3 a = vec3(1.0f);
        for (const auto &m : mediums) a *= m.second.absorb;
        return a;
    }
    int size() const { return int(mediums.size()); }
};

struct PathCtx {
    const Model &model;
    const RenderArgs &args;
    Generator gen;
    Medium mediums;
};

}  // namespace

// src/render.cpp:44-50
void calc_dPdxy(const Ray &ray, float hit_t, const vec3 &normal, const RayDifferential &base_diff, vec3 &hit_dPdx,
                vec3 &hit_dPdy) {
    float dtdx = -dot(base_diff.dPdx + hit_t * base_diff.dDdx, normal) / dot(ray.direction, normal);
    float dtdy = -dot(base_diff.dPdy + hit_t * base_diff.dDdy, normal) / dot(ray.direction, normal);
    hit_dPdx = base_diff.dPdx + dtdx * ray.direction + hit_t * base_diff.dDdx;
    hit_dPdy = base_diff.dPdy + dtdy * ray.direction + hit_t * base_diff.dDdy;
}

// src/render.cpp:52-60 (the normal derivatives are zero)
void calc_dDdxy(const Ray &ray, const vec3 &normal, const RayDifferential &base_diff, vec3 &hit_dDdx, vec3 &hit_dDdy) {
    const vec3 hit_dNdx(0.0f), hit_dNdy(0.0f);
    float dDNdx = dot(hit_dNdx, ray.direction) + dot(base_diff.dDdx, normal);
    float dDNdy = dot(hit_dNdy, ray.direction) + dot(base_diff.dDdy, normal);
    hit_dDdx = base_diff.dDdx - 2.0f * (dot(ray.direction, normal) * hit_dNdx + dDNdx * normal);
    hit_dDdy = base_diff.dDdy - 2.0f * (dot(ray.direction, normal) * hit_dNdy + dDNdy * normal);
}

// src/render.cpp:62-79
void getHitInfo(const HitRecord &hit, const Ray &ray, const RayDifferential &base_diff, vec3 &hit_dPdx,
                vec3 &hit_dPdy, HitInfo &hitInfo) {
    const Face &face = *hit.face;
    const vec3 baryCoords = barycentric(face.v[0], face.v[1], face.v[2], hitInfo.position);
    getHitNormals(face, ray.direction, baryCoords, hitInfo.shapeNormal, hitInfo.surfaceNormal, hitInfo.entering);
    vec3 surfaceNormal_raw = hitInfo.surfaceNormal;
    calc_dPdxy(ray, hit.t_max, hitInfo.shapeNormal, base_diff, hit_dPdx, hit_dPdy);
    getHitMaterial(face, baryCoords, hit_dPdx, hit_dPdy, hitInfo);
    if (dot(hitInfo.surfaceNormal, ray.direction) >= 0.0f) hitInfo.surfaceNormal = surfaceNormal_raw;
    if (dot(hitInfo.surfaceNormal, ray.direction) >= 0.0f) hitInfo.surfaceNormal = hitInfo.shapeNormal;
}

namespace {

// src/render.cpp:83-87
vec3 getAbsorb(const vec3 &absorb, float dis) {
    const float omega_absorb = 32.0f;
    float Clum = dot(absorb, RGB_Weight);
    return (Clum < 1.0f - eps_zero) ? pow_s(absorb, omega_absorb * dis) : vec3(1.0f);
}

// src/render.cpp:89-99
void calcEta(Medium &mediums, HitInfo &hitInfo) {
    float eta1 = mediums.ior(), eta2;
    if (hitInfo.entering)
        eta2 = std::max(eta1, hitInfo.eta);
    else {
        mediums.erase(hitInfo.id);
        eta2 = mediums.ior();
        mediums.insert(hitInfo.id, hitInfo.eta, hitInfo.baseColor);
    }
    hitInfo.eta = eta1 / eta2;
}

void scaling(std::vector<LightSample> &s, float f) { for (auto &x : s) x.bsdfPdf *= f; }        // :101-104
void passBsdf(std::vector<LightSample> &s, const vec3 &b) {                                         // :106-111
    for (auto &x : s) { x.light *= x.bsdfPdf; x.bsdfPdf = b; }
}
void mulWeight(std::vector<LightSample> &s, float w) { for (auto &x : s) x.weight *= w; }          // :113-116

const float P0_reflect = 0.24f;                 // src/render.cpp:118
const float regularizationFactor = 1.0f;        // src/render.cpp:119
const int sampleCount[17] = {0, 1, 2, 2, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 5, 6};  // src/render.cpp:126

// src/render.cpp:121-312
std::vector<LightSample> sampleRay(Ray ray, const RayDifferential &base_diff, PathCtx &ctx, float roughnessFactor,
                                   bool excludeDirectLight, int depth) {
    const Model &model = ctx.model;
    const int maxRayDepth = ctx.args.max_depth;
    OneGen og{&ctx.gen};

    HitRecord hit = model.rayHit(ray);
    if (hit.t_max == INFINITY) {
        if (excludeDirectLight || model.skyMap.empty()) return {};
        return {LightSample{vec3(1.0f), model.skyMap.get(ray.direction), 1.0f}};
    }
    BSDF bsdf(-ray.direction, ray.origin + ray.direction * hit.t_max);
    vec3 hit_dPdx, hit_dPdy;
    getHitInfo(hit, ray, base_diff, hit_dPdx, hit_dPdy, bsdf.surface);
    float ior = bsdf.surface.eta;
    roughnessFactor = std::max(roughnessFactor, regularizationFactor * bsdf.surface.roughness);
    bsdf.surface.roughness = std::max(bsdf.surface.roughness, roughnessFactor);
    const float P_RR_Rough = 0.5f, P_RR_Smooth = 1.0f;
    const float P_RR = P_RR_Smooth + (P_RR_Rough - P_RR_Smooth) * std::sqrt(bsdf.surface.roughness);
    const float P_RR_Therehold = 0.9f;
    bool doDirectLightSample = P_RR < P_RR_Therehold && depth >= ctx.args.rr_min_depth;  // EXT rr_min_depth
    vec3 absorb = getAbsorb(ctx.mediums.absorb(), hit.t_max);
    if (length(bsdf.surface.emission) > eps_zero) {
        if (excludeDirectLight) return {};
        return {LightSample{absorb, bsdf.surface.emission, 1.0f}};
    }
    float P_reflect = 1.0f, F = 0.0f;
    vec3 refractDir(NAN);
    if (bsdf.surface.opacity < eps_zero) {
        calcEta(ctx.mediums, bsdf.surface);
        bsdf.preciseRefraction(refractDir, F);
        if (ctx.mediums.size() == 1)
            P_reflect = P0_reflect + (1.0f - P0_reflect) * F;
        else
            P_reflect = F;
        P_reflect = std::max(P_reflect - 1e-3f, 0.0f);
    }
    vec3 newDir, bsdfPdf(NAN);
    RayDifferential next_diff;
    int fails = 0;
    if (ctx.gen() <= P_reflect) {
        doDirectLightSample &= bsdf.surface.entering && ctx.mediums.size() == 1;
        if (doDirectLightSample) {
            if (depth >= maxRayDepth || ctx.gen() > P_RR) {
                std::vector<LightSample> samples = sampleDirectLight(bsdf, model, og, sampleCount[depth]);
                float factor = 1.0f / P_reflect;
                if (depth < maxRayDepth) factor /= (1.0f - P_RR);
                scaling(samples, factor);
                return samples;
            }
        }
        bsdf.sampleReflection(ctx.gen, newDir, bsdfPdf, fails);
        bsdfPdf /= P_reflect;
        if (doDirectLightSample) bsdfPdf /= P_RR;
        vec3 hit_dDdx, hit_dDdy;
        calc_dDdxy(ray, bsdf.surface.surfaceNormal, base_diff, hit_dDdx, hit_dDdy);
        next_diff = {hit_dPdx, hit_dPdy, hit_dDdx, hit_dDdy};
    } else {
        doDirectLightSample &= !bsdf.surface.entering && ctx.mediums.size() == 2;
        if (doDirectLightSample) {
            if (depth >= maxRayDepth || ctx.gen() > P_RR) {
                std::vector<LightSample> samples = sampleDirectLight(bsdf, model, og, sampleCount[depth]);
                float factor = 1.0f / (1.0f - P_reflect);
                if (depth < maxRayDepth) factor /= (1.0f - P_RR);
                scaling(samples, factor);
                passBsdf(samples, absorb);
                return samples;
            }
        }
        if (std::abs(bsdf.surface.eta - 1.0f) < eps_zero) {
            newDir = ray.direction;
            bsdfPdf = vec3(1.0f);
            next_diff = base_diff;
        } else {
            bsdf.sampleBTDF(ctx.gen, newDir, bsdfPdf, fails);
            bsdfPdf *= (1.0f - F);
            next_diff = base_diff;
        }
        bsdfPdf /= (1.0f - P_reflect);
        if (doDirectLightSample) bsdfPdf /= P_RR;
        if (bsdf.surface.entering)
            ctx.mediums.insert(bsdf.surface.id, ior, bsdf.surface.baseColor);
        else
            ctx.mediums.erase(bsdf.surface.id);
    }
    if (!isfinite(newDir) || depth == maxRayDepth) return {};
    Ray newRay = {bsdf.surface.position, newDir};
    bsdfPdf *= absorb;
    std::vector<LightSample> samples = sampleRay(newRay, next_diff, ctx, roughnessFactor, doDirectLightSample, depth + 1);
    passBsdf(samples, bsdfPdf);
    if (fails > 0) mulWeight(samples, 1.0f / static_cast<float>(1 + fails));
    return samples;
}

// src/render.cpp:314-423
void sampleIndirectLightFromFirstIntersection(const HitInfo &hitInfo, const vec3 &origin,
                                              const RayDifferential &base_diff, PathCtx &ctx,
                                              std::vector<LightSample> &samples) {
    if (length(hitInfo.emission) > 0) return;
    ctx.mediums.Init();
    vec3 inDir = normalize(hitInfo.position - origin);
    Ray ray = {origin, inDir};
    float hit_t = length(hitInfo.position - origin);
    BSDF bsdf(-inDir, hitInfo);
    float roughnessFactor = hitInfo.roughness * regularizationFactor;
    float ior = bsdf.surface.eta;
    bsdf.surface.eta = 1.0f / bsdf.surface.eta;
    float P_reflect = 1.0f, F = 0.0f;
    vec3 refractDir;
    if (bsdf.surface.opacity < eps_zero) {
        bsdf.preciseRefraction(refractDir, F);
        P_reflect = P0_reflect + (1.0f - P0_reflect) * F;
    }
    vec3 newDir, bsdfPdf = vec3(NAN);
    RayDifferential next_diff;
    int fails = 0;
    if (ctx.gen() < P_reflect) {
        vec3 hit_dPdx, hit_dPdy;
        calc_dPdxy(ray, hit_t, bsdf.surface.shapeNormal, base_diff, hit_dPdx, hit_dPdy);
        vec3 hit_dDdx, hit_dDdy;
        calc_dDdxy(ray, bsdf.surface.surfaceNormal, base_diff, hit_dDdx, hit_dDdy);
        next_diff = {hit_dPdx, hit_dPdy, hit_dDdx, hit_dDdy};
        bsdf.sampleReflection(ctx.gen, newDir, bsdfPdf, fails);
        bsdfPdf /= P_reflect;
    } else {
        bsdf.sampleBTDF(ctx.gen, newDir, bsdfPdf, fails);
        bsdfPdf *= (1.0f - F);
        bsdfPdf /= (1.0f - P_reflect);
        next_diff = base_diff;
        if (bsdf.surface.entering)
            ctx.mediums.insert(bsdf.surface.id, ior, bsdf.surface.baseColor);
        else
            ctx.mediums.erase(bsdf.surface.id);
    }
    if (!isfinite(newDir)) return;
    Ray newRay = {bsdf.surface.position, newDir};
    std::vector<LightSample> samplesNew = sampleRay(newRay, next_diff, ctx, roughnessFactor, true, 1);
    passBsdf(samplesNew, bsdfPdf);
    if (fails > 0) mulWeight(samplesNew, 1.0f / static_cast<float>(1 + fails));
    for (const auto &s : samplesNew) samples.push_back(s);
}

}  // namespace

// src/render.cpp:439-446
void initRayDiff(const vec3 &d, const RenderArgs &args, RayDifferential &base_diff) {
    base_diff.dPdx = vec3(0.0f), base_diff.dPdy = vec3(0.0f);
    const vec3 dddx = args.accuracy * args.right, dddy = args.accuracy * args.up;
    float d_dot_d = dot(d, d), d_dot_dddx = dot(d, dddx), d_dot_dddy = dot(d, dddy);
    base_diff.dDdx = (d_dot_d * dddx - d * d_dot_dddx) / (std::sqrt(d_dot_d) * d_dot_d);
    base_diff.dDdy = (d_dot_d * dddy - d * d_dot_dddy) / (std::sqrt(d_dot_d) * d_dot_d);
}

// src/render.cpp:448-551
void renderPixel(const Model &model, const RenderArgs &args, Photo &image, int x, int y) {
    const int width = args.width, height = args.height;
    vec3 view = args.direction, right = args.right, up = args.up, position = args.position;
    const float accuracy = args.accuracy;
    int id = y * width + x;
    float rayX = float(x) - float(width) / 2.0f, rayY = float(y) - float(height) / 2.0f;
    vec3 d = view + accuracy * (rayX * right + rayY * up);
    HitInfo &Gbuffer = image.Gbuffer[id];
    RayDifferential base_diff;
    initRayDiff(d, args, base_diff);
    vec3 Dir = normalize(d);
    Ray ray = {position, Dir};
    HitRecord hit = model.rayHit(ray);
    if (hit.t_max == INFINITY) {
        Gbuffer.position = vec3(NAN);
        if (!model.skyMap.empty()) Gbuffer.emission = model.skyMap.get(Dir);
        return;
    }
    Gbuffer.position = ray.origin + hit.t_max * ray.direction;
    vec3 hit_dPdx, hit_dPdy;
    getHitInfo(hit, ray, base_diff, hit_dPdx, hit_dPdy, Gbuffer);
    bool opacity = (Gbuffer.opacity > 1.0f - eps_zero);
    vec3 sav_baseColor = Gbuffer.baseColor;
    float Clum0 = dot(Gbuffer.baseColor, RGB_Weight);
    if (Clum0 < 2e-2 || (Clum0 < 0.8f && length(Gbuffer.baseColor / Clum0 - vec3(1.0f)) < 2e-2))
        Gbuffer.baseColor[0] += 4e-2;
    const int MultiSampleForRefraction = 16;
    const int spp_direct = int(float(args.spp) * args.P_Direct),
              spp_indirect = (args.spp - spp_direct) * (opacity ? 1 : MultiSampleForRefraction);
    const float exposure = args.exposure;
    RadianceData &radiance_Dd = image.radiance_Dd[id], &radiance_Ds = image.radiance_Ds[id],
                 &radiance_Id = image.radiance_Id[id], &radiance_Is = image.radiance_Is[id];
    auto calcVar = [](RadianceData &radiance, float exp) {
        radiance.radiance *= exp;
        radiance.Var *= exp * exp;
        radiance.Var -= dot(radiance.radiance, radiance.radiance);
        if (radiance.Var < 0.0f) radiance.Var = 0.0f;
    };
    // sampleDirectLightFromFirstIntersection (src/render.cpp:425-437)
    if (isfinite(Gbuffer.position) && !(length(Gbuffer.emission) > 0)) {
        vec3 inDir = normalize(Gbuffer.position - args.position);
        BSDF bsdf(-inDir, Gbuffer);
        std::vector<Generator> gens;
        gens.reserve(std::max(1, spp_direct));
        PerSampleGen pg{args.seed, (uint32_t)id, &gens};
        std::vector<LightSample> samples = sampleDirectLight(bsdf, model, pg, spp_direct);
        for (const auto &sample : samples) accumulateInwardRadiance(Gbuffer.baseColor, sample, radiance_Dd, radiance_Ds);
    }
    calcVar(radiance_Dd, exposure);
    calcVar(radiance_Ds, exposure);
    std::vector<LightSample> samples;
    for (int T = 0; T < spp_indirect; T++) {
        PathCtx ctx{model, args, Generator(args.seed, TAG_INDIRECT, (uint32_t)id, (uint32_t)T), Medium()};
        sampleIndirectLightFromFirstIntersection(Gbuffer, args.position, base_diff, ctx, samples);
    }
    if (samples.empty()) {
        // the reference returns here without calcVar on the indirect buffers and
        // without restoring the base colour (src/render.cpp:529-530)
        return;
    }
    mulWeight(samples, 1.0f / static_cast<float>(spp_indirect));
    float meanClum = 0.0f;
    for (const auto &sample : samples) meanClum += dot(sample.bsdfPdf * sample.light, RGB_Weight) * sample.weight;
    const float clampThreshold = 16.0f;
    for (auto &sample : samples) {
        float Clum = dot(sample.bsdfPdf * sample.light, RGB_Weight) * sample.weight;
        if (Clum / (meanClum - Clum + eps_zero) > clampThreshold) continue;
        accumulateInwardRadiance(Gbuffer.opacity < eps_zero ? vec3(0.0f) : Gbuffer.baseColor, sample, radiance_Id,
                                 radiance_Is);
    }
    calcVar(radiance_Id, exposure);
    calcVar(radiance_Is, exposure);
    Gbuffer.baseColor = sav_baseColor;
}

// src/render.cpp:593-626: a pool of `threads` workers pulling whole rows.
// `rows` restricts the render to a subset of rows (CPU-baseline sampling).
void render_multiThread(const Model &model, const RenderArgs &args, Photo &photo, int threads,
                        const std::vector<int> *rows) {
    std::vector<int> all;
    if (!rows) {
        for (int y = 0; y < args.height; ++y) all.push_back(y);
        rows = &all;
    }
    std::atomic<size_t> next{0};
    auto worker = [&] {
        while (true) {
            size_t k = next.fetch_add(1);
            if (k >= rows->size()) break;
            int y = (*rows)[k];
            for (int x = 0; x < args.width; ++x) renderPixel(model, args, photo, x, y);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
}

// The same worker pool over an explicit pixel list (CPU-baseline sampling).
void render_pixels(const Model &model, const RenderArgs &args, Photo &photo, int threads, const std::vector<int> &pix) {
    std::atomic<size_t> next{0};
    auto worker = [&] {
        while (true) {
            size_t k = next.fetch_add(1);
            if (k >= pix.size()) break;
            renderPixel(model, args, photo, pix[k] % args.width, pix[k] / args.width);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
}

// ------------------------------------------------------------------ image
Photo::Photo(int w, int h)
    : width(w), height(h), Gbuffer((size_t)w * h), radiance_Dd((size_t)w * h), radiance_Ds((size_t)w * h),
      radiance_Id((size_t)w * h), radiance_Is((size_t)w * h), pixelarray((size_t)w * h) {}

// src/image.cpp:215-246
void Photo::shade(int options) {
    for (int i = 0; i < width * height; ++i) {
        const HitInfo &G = Gbuffer[i];
        if (options & shapeNormal) {
            pixelarray[i] = (G.shapeNormal + vec3(1.0f)) / 2.0f;
            continue;
        }
        if (options & surfaceNormal) {
            pixelarray[i] = (G.surfaceNormal + vec3(1.0f)) / 2.0f;
            continue;
        }
        vec3 radiance_d = vec3(0.0f), radiance_s = vec3(0.0f);
        if (options & DirectLight) {
            if (options & Diffuse) radiance_d += radiance_Dd[i].radiance;
            if (options & Specular) radiance_s += radiance_Ds[i].radiance;
        }
        if (options & IndirectLight) {
            if (options & Diffuse) radiance_d += radiance_Id[i].radiance;
            if (options & Specular) radiance_s += radiance_Is[i].radiance;
        }
        if (!(options & (DirectLight | IndirectLight))) radiance_d = vec3(1.0f);
        vec3 diffuseColor = (options & BaseColor) ? G.baseColor : vec3(1.0f);
        pixelarray[i] = diffuseColor * radiance_d + radiance_s;
        if (options & Emission) pixelarray[i] += G.emission * exposure;
    }
}

// src/image.cpp:358-452
void Photo::FXAA() {
    const float EDGE_THRESHOLD_MIN = 0.0312f, EDGE_THRESHOLD_MAX = 0.125f, SUBPIXEL_QUALITY = 0.75f;
    const int QUALITY = 12;
    const std::vector<vec3> data = pixelarray;
    std::vector<vec3> output((size_t)width * height);
    auto L = [](const vec3 &rgb) { return dot(rgb, RGB_Weight); };
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            float lumaM = L(data[y * width + x]);
            float lumaN = y > 0 ? L(data[(y - 1) * width + x]) : lumaM;
            float lumaS = y < height - 1 ? L(data[(y + 1) * width + x]) : lumaM;
            float lumaE = x < width - 1 ? L(data[y * width + x + 1]) : lumaM;
            float lumaW = x > 0 ? L(data[y * width + x - 1]) : lumaM;
            float rangeMin = std::min({lumaN, lumaS, lumaE, lumaW});
            float rangeMax = std::max({lumaN, lumaS, lumaE, lumaW});
            float range = rangeMax - rangeMin;
            if (range < std::max(EDGE_THRESHOLD_MIN, rangeMax * EDGE_THRESHOLD_MAX)) {
                output[y * width + x] = data[y * width + x];
                continue;
            }
            float lumaNW = (y > 0 && x > 0) ? L(data[(y - 1) * width + x - 1]) : lumaM;
            float lumaNE = (y > 0 && x < width - 1) ? L(data[(y - 1) * width + x + 1]) : lumaM;
            float lumaSW = (y < height - 1 && x > 0) ? L(data[(y + 1) * width + x - 1]) : lumaM;
            float lumaSE = (y < height - 1 && x < width - 1) ? L(data[(y + 1) * width + x + 1]) : lumaM;
            float edgeHorz = std::abs((lumaNW + lumaW + lumaSW) - (lumaNE + lumaE + lumaSE)) * (1.0f / 3.0f);
            float edgeVert = std::abs((lumaNW + lumaN + lumaNE) - (lumaSW + lumaS + lumaSE)) * (1.0f / 3.0f);
            bool isHorizontal = edgeHorz >= edgeVert;
            float stepLength = isHorizontal ? 1.0f / float(width) : 1.0f / float(height);
            float gradientStep = std::clamp((isHorizontal ? edgeHorz : edgeVert) / range, -2.0f, 2.0f);
            vec2 uv(float(x) / float(width), float(y) / float(height));
            vec3 finalColor = data[y * width + x];
            float bestDelta = 0.0f;
            for (int i = 0; i < QUALITY; i++) {
                vec2 offset(isHorizontal ? 0.0f : gradientStep * stepLength * float(i + 1),
                            isHorizontal ? gradientStep * stepLength * float(i + 1) : 0.0f);
                vec2 sampleUv = uv + offset;
                if (sampleUv.x < 0.0f || sampleUv.x > 1.0f || sampleUv.y < 0.0f || sampleUv.y > 1.0f) continue;
                int sx = int(sampleUv.x * float(width));
                int sy = int(sampleUv.y * float(height));
                sx = std::clamp(sx, 0, width - 1);
                sy = std::clamp(sy, 0, height - 1);
                vec3 sampleColor = data[sy * width + sx];
                float sampleLuma = L(sampleColor);
                float delta = std::abs(sampleLuma - lumaM);
                if (delta > bestDelta) {
                    bestDelta = delta;
                    finalColor = sampleColor;
                }
            }
            float subPixelOffset =
                std::min((std::abs(lumaN + lumaS - 2.0f * lumaM) * 2.0f + std::abs(lumaE + lumaW - 2.0f * lumaM)) * 0.25f,
                         1.0f);
            output[y * width + x] = mix(data[y * width + x], finalColor, subPixelOffset * SUBPIXEL_QUALITY);
        }
    }
    pixelarray = output;
}

// src/image.cpp:454-468
void Photo::gammaCorrection() {
    const float GammaFactor = 2.2f;
    auto lumBound = [](float Clum) -> float { return std::tanh(3.0f * (Clum - 0.75f)) / 3.0f + 0.75f; };
    for (int i = 0; i < width * height; ++i) {
        pixelarray[i] = max(pixelarray[i], 0.0f);
        float Clum = dot(pixelarray[i], RGB_Weight);
        if (Clum > 0.75f) pixelarray[i] = pixelarray[i] / Clum * lumBound(Clum);
        pixelarray[i] = min(pixelarray[i], 1.0f);
        pixelarray[i] = pow(pixelarray[i], vec3(1.0f / GammaFactor));
    }
}

// src/image.cpp:470-479 (bloom and depthFeildBlur in oimage.cpp).
// EXT: the bit 65536 (the product's RM_SHADE_LINEAR) skips gammaCorrection.
void Photo::postProcessing(int shadeOptions) {
    shade(shadeOptions);
    if (shadeOptions & DoDepthFieldBlur) depthFeildBlur();
    if (shadeOptions & DoBloom) bloom();
    if (!(shadeOptions & 65536)) gammaCorrection();
    if (shadeOptions & DoFXAA) FXAA();
}

}  // namespace ref

// ORACLE -- test infrastructure only (see oracle.h).
// Restates src/sampling.cpp (BSDF, samplers, next-event estimation) and
// accumulateInwardRadiance (src/image.cpp:615-659).
#include <cmath>
#include <vector>
#include "oracle.h"

namespace ref {

BSDF::BSDF(const vec3 &inDir_, const vec3 &hit_position) : inDir(inDir_) { surface.position = hit_position; }
BSDF::BSDF(const vec3 &inDir_, const HitInfo &hitInfo) : inDir(inDir_), surface(hitInfo) {}

static float sqr(float x) { return x * x; }
static float clampf(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
static float mixf(float a, float b, float t) { return a * (1 - t) + b * t; }                 // :16-18
static float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }
static float SchlickFresnel(float u) { float m = clampf(1 - u, 0, 1); return pow5(m); }     // :25-28
static float GTR1(float NdotH, float a) {                                                    // :30-35
    if (a >= 1) return 1 / PI;
    float a2 = a * a;
    float t = 1 + (a2 - 1) * NdotH * NdotH;
    return (a2 - 1) / (PI * std::log(a2) * t);
}
static float GTR2(float NdotH, float a) {                                                    // :37-41
    float a2 = a * a;
    float t = 1 + (a2 - 1) * NdotH * NdotH;
    return a2 / (PI * t * t);
}
static float smithG_GGX(float NdotV, float alphaG) {                                        // :43-47
    float a = alphaG * alphaG;
    float b = NdotV * NdotV;
    return 1 / (NdotV + std::sqrt(a + b - a * b));
}

// src/sampling.cpp:49-157
vec3 BSDF::getBRDF(vec3 L) const {
    const vec3 &V = inDir;
    const vec3 &N = surface.surfaceNormal;
    float NdotL = dot(N, L);
    float NdotV = dot(N, V);
    if (NdotL <= 0.0f || NdotV <= 0.0f) return vec3(0);
    vec3 H = normalize(L + V);
    float NdotH = dot(N, H);
    float LdotH = dot(L, H);
    vec3 Cdlin = surface.baseColor;
    float Cdlum = dot(Cdlin, RGB_Weight);
    const float subsurface = 0.0f, specularTint = 0.0f, sheen = 0.0f, sheenTint = 0.0f, clearcoat = 1.5f,
                clearcoatGloss = 0.2f, clearcoatTint = 0.0f;
    vec3 Ctint = Cdlum > 0.0f ? Cdlin / Cdlum : vec3(1.0f);
    vec3 Cspec0 = mix(surface.specular * mix(vec3(1.0f), Ctint, specularTint), Cdlin, surface.metallic);
    vec3 Csheen = mix(vec3(1.0f), Ctint, sheenTint);
    float FL = SchlickFresnel(NdotL), FV = SchlickFresnel(NdotV);
    float Fd90 = 0.5f + 2.0f * LdotH * LdotH * surface.roughness;
    float Fd = mixf(1.0f, Fd90, FL) * mixf(1.0f, Fd90, FV);
    float Fss90 = LdotH * LdotH * surface.roughness;
    float Fss = mixf(1.0f, Fss90, FL) * mixf(1.0f, Fss90, FV);
    float ss = 1.25f * (Fss * (1.0f / (NdotL + NdotV) - .5f) + .5f);
    float Ds = GTR2(NdotH, surface.roughness);
    float FH;
    if (surface.eta <= 1.0f) {
        FH = SchlickFresnel(LdotH);
    } else {
        float cosThetaI = std::fabs(LdotH);
        float sinThetaI = sqrt_s(1.0f - cosThetaI * cosThetaI);
        float sinThetaT = surface.eta * sinThetaI;
        if (sinThetaT >= 1.0f) {
            FH = 1.0f;
        } else {
            float cosThetaT = sqrt_s(1.0f - sinThetaT * sinThetaT);
            float R0 = (surface.eta - 1.0f) / (surface.eta + 1.0f);
            R0 = R0 * R0;
            FH = mixf(R0, 1.0f, SchlickFresnel(cosThetaT));
        }
    }
    vec3 Fs = mix(Cspec0, vec3(1.0f), FH);
    float Gs;
    Gs = smithG_GGX(NdotL, surface.roughness);
    Gs *= smithG_GGX(NdotV, surface.roughness);
    vec3 Fsheen = FH * sheen * Csheen;
    float Dr = GTR1(NdotH, mixf(.1f, .001f, clearcoatGloss));
    float Fr = mixf(.04f, 1.0f, FH);
    float Gr = smithG_GGX(NdotL, .25f) * smithG_GGX(NdotV, .25f);
    vec3 ret = ((1.0f / PI) * mixf(Fd, ss, subsurface) * Cdlin + Fsheen);
    ret *= (1.0f - surface.metallic);
    ret += (0.25f * clearcoat * Gr * Fr * Dr) * mix(vec3(1.0f), Ctint, clearcoatTint);
    ret *= surface.opacity;
    ret += Fs * Ds * Gs;
    return ret * NdotL;
}

// src/sampling.cpp:159-192
vec3 BSDF::getBTDF(vec3 L) const {
    const vec3 &N = surface.surfaceNormal;
    float NdotL = dot(surface.surfaceNormal, L);
    if (NdotL >= 0.0f) return vec3(0.0f);
    const vec3 &V = inDir;
    vec3 H = normalize(L + surface.eta * V);
    float D = GTR2(dot(N, H), surface.roughness);
    float BTDF = D * (-NdotL);
    float LdotH = dot(L, H);
    float NdotV = dot(N, V);
    float HdotV = dot(H, V);
    BTDF *= std::abs(LdotH * HdotV) / (std::abs(NdotL * NdotV) + eps_zero);
    float k = surface.eta * HdotV + LdotH;
    BTDF /= k * k;
    return vec3(BTDF);
}

// src/sampling.cpp:194-199
vec3 BSDF::getBSDF(vec3 outDir) const {
    if (surface.opacity > 1.0f - eps_zero || surface.entering) return getBRDF(outDir);
    return getBTDF(outDir);
}

static const int MaxTrys = 16;   // src/sampling.cpp:201

// src/sampling.cpp:203-236
void BSDF::sampleBRDF(Generator &gen, vec3 &outDir, vec3 &brdfPdf, float &pdf, int &fails) const {
    vec3 tangent, bitangent;
    getTangentSpaceWithInDir(surface.surfaceNormal, inDir, tangent, bitangent);
    float lastCt = 0.0f, lastLH = 1.0f;
    auto sampleGTR2 = [&](vec3 &od, float &p) {
        float u = gen(), phi = gen() * 2.0f * PI;
        float cosTheta = std::sqrt((1.0f - u) / (1.0f + (sqr(surface.roughness) - 1.0f) * u)),
              sinTheta = std::sqrt(1 - cosTheta * cosTheta);
        if (adj_tap) adj_tap->ggx_ct.push_back(cosTheta);
        vec3 H = sinTheta * std::cos(phi) * tangent + sinTheta * std::sin(phi) * bitangent + cosTheta * surface.surfaceNormal;
        od = 2.0f * dot(inDir, H) * H - inDir;
        float LDotH = dot(od, H);
        float LdotN = dot(od, surface.surfaceNormal);
        lastCt = cosTheta;
        lastLH = LDotH;
        if (LDotH <= 0.0f || LdotN <= 0.0f) {
            p = 0.0f;
            return;
        }
        p = GTR2(cosTheta, surface.roughness) / (4.0f * LDotH);
    };
    for (int T = 1; T <= MaxTrys; T++) {
        sampleGTR2(outDir, pdf);
        if (dot(outDir, surface.shapeNormal) > 0.0f && pdf > 0.0f) {
            if (adj_tap) {
                adj_tap->ct2 = lastCt;
                adj_tap->lh2 = lastLH;
            }
            brdfPdf = getBRDF(outDir) / pdf;
            return;
        }
        fails++;
    }
    pdf = 0.0f;
    brdfPdf = vec3(0.0f);
    outDir = vec3(NAN);
}

// src/sampling.cpp:238-243
void clampPdf(vec3 &pdf) {
    const float clampThreshold = 64.0f;
    float Clum = dot(pdf, RGB_Weight);
    if (Clum > clampThreshold) pdf /= Clum / clampThreshold;
}

// src/sampling.cpp:245-269
void BSDF::sampleCos(Generator &gen, vec3 &outDir, vec3 &brdfPdf, float &pdf, int &fails) const {
    vec3 tangent, bitangent;
    getTangentSpaceWithInDir(surface.surfaceNormal, inDir, tangent, bitangent);
    auto sampleC = [&]() -> vec3 {
        float u = gen(), phi = gen() * 2.0f * PI;
        float d = std::sqrt(u);
        float z = std::sqrt(1 - d * d);
        float x = d * std::cos(phi), y = d * std::sin(phi);
        return x * tangent + y * bitangent + z * surface.surfaceNormal;
    };
    for (int T = 1; T <= MaxTrys; T++) {
        outDir = sampleC();
        pdf = dot(outDir, surface.surfaceNormal) / PI;
        if (dot(outDir, surface.shapeNormal) > 0.0f && pdf > 0.0f) {
            brdfPdf = getBRDF(outDir) / pdf;
            if (adj_tap) adj_tap->raw1 = brdfPdf;
            clampPdf(brdfPdf);
            return;
        }
        fails++;
    }
    pdf = 0.0f;
    brdfPdf = vec3(0.0f);
    outDir = vec3(NAN);
}

// src/sampling.cpp:271-308
void BSDF::preciseRefraction(vec3 &outDir, float &F) const {
    const vec3 &V = inDir;
    const vec3 &N = surface.surfaceNormal;
    float eta = surface.eta;
    float NdotV = dot(N, V);
    if (NdotV < 0.0f) {
        F = 1.0f;
        outDir = vec3(NAN);
        return;
    }
    float delta = 1.0f - eta * eta * (1.0f - NdotV * NdotV);
    if (delta < 0.0f) {
        F = 1.0f;
        outDir = vec3(NAN);
        return;
    }
    float k = eta * NdotV - std::sqrt(delta);
    outDir = k * N - eta * V;
    float LdotN = -dot(outDir, N);
    float R0 = (1.0f - eta) / (1.0f + eta);
    R0 = R0 * R0;
    F = R0 + (1.0f - R0) * SchlickFresnel(LdotN);
}

// src/sampling.cpp:310-340
void BSDF::sampleReflection(Generator &gen, vec3 &Dir, vec3 &brdfPdf, int &fails) const {
    vec3 Dir1, Dir2, brdfPdf1, brdfPdf2;
    int fail1 = 0, fail2 = 0;
    float pdf1, pdf2;
    sampleCos(gen, Dir1, brdfPdf1, pdf1, fail1);
    sampleBRDF(gen, Dir2, brdfPdf2, pdf2, fail2);
    float p1 = pdf1 / (pdf1 + pdf2);
    bool first = gen() < p1;
    if (adj_tap) {
        adj_tap->pdf1 = pdf1;
        adj_tap->pdf2 = pdf2;
        adj_tap->first = first;
    }
    if (first) {
        Dir = Dir1;
        brdfPdf = brdfPdf1;
        fails += fail1;
    } else {
        Dir = Dir2;
        brdfPdf = brdfPdf2;
        fails += fail2;
    }
}

// src/sampling.cpp:342-349
static vec3 refract(const vec3 &V, const vec3 &N, float eta) {
    float NdotV = dot(N, V);
    float delta = 1.0f - eta * eta * (1.0f - NdotV * NdotV);
    if (delta < 0.0f) return vec3(NAN);
    float k = eta * NdotV - std::sqrt(delta);
    return k * N - eta * V;
}

// src/sampling.cpp:351-391
void BSDF::sampleBTDF(Generator &gen, vec3 &outDir, vec3 &btdfPdf, int &fails) const {
    if (std::abs(surface.eta - 1.0f) < eps_zero) {
        outDir = -inDir;
        btdfPdf = vec3(1.0f);
        return;
    }
    vec3 tangent, bitangent;
    getTangentSpaceWithInDir(surface.surfaceNormal, inDir, tangent, bitangent);
    const vec3 &V = inDir;
    auto sampleGTR2 = [&](vec3 &od, float &weight) {
        float u = gen(), phi = gen() * 2.0f * PI;
        float cosTheta = std::sqrt((1.0f - u) / (1.0f + (sqr(surface.roughness) - 1.0f) * u)),
              sinTheta = std::sqrt(1 - cosTheta * cosTheta);
        vec3 H = sinTheta * std::cos(phi) * tangent + sinTheta * std::sin(phi) * bitangent + cosTheta * surface.surfaceNormal;
        od = refract(V, H, surface.eta);
        if (!isfinite(od)) {
            weight = 0.0f;
            return;
        }
        float VdotH = dot(V, H);
        float VdotN = dot(V, surface.surfaceNormal);
        float HdotN = dot(H, surface.surfaceNormal);
        weight = std::abs(VdotH) / std::abs(VdotN * HdotN);
    };
    for (int T = 1; T <= MaxTrys; T++) {
        float weight;
        sampleGTR2(outDir, weight);
        if (weight > 0.0f && dot(outDir, surface.shapeNormal) < 0.0f) {
            btdfPdf = vec3(weight);
            return;
        }
        fails++;
    }
    btdfPdf = vec3(0.0f);
    outDir = vec3(NAN);
}

// src/sampling.cpp:396-404
static int sampleIndex(const std::vector<float> &weights, float randomValue) {
    float cumulativeWeight = 0.0f;
    for (size_t i = 0; i < weights.size(); ++i) {
        cumulativeWeight += weights[i];
        if (randomValue <= cumulativeWeight) return (int)i;
    }
    return -1;
}

// src/sampling.cpp:406-417
static void getLightObjectWeight(const vec3 &pos, const BSDF &bsdf, const Model &model, std::vector<float> &weights) {
    weights.reserve(model.lightObjects.size());
    for (const auto &lightObject : model.lightObjects) {
        vec3 lightDir = normalize(lightObject.center - pos);
        float distance = length(lightObject.center - pos);
        vec3 bsdfPdf = bsdf.getBSDF(lightDir);
        float Clum = dot(bsdfPdf, RGB_Weight);
        float weight = Clum * lightObject.power / (distance * distance + eps_lightRadius);
        weights.emplace_back(weight);
    }
}

// src/sampling.cpp:419-434
static void generateRandomPointInLightFace(const Face &face, const vec3 &pos, Generator &gen, vec3 &lightPos,
                                           float &cosPhi) {
    float a = gen(), b = gen();
    if (a + b > 1.0f) {
        a = 1.0f - a;
        b = 1.0f - b;
    }
    float c = 1.0f - a - b;
    lightPos = a * face.v[0] + b * face.v[1] + c * face.v[2];
    vec3 lightDir = normalize(lightPos - pos);
    vec3 shapeNormal, surfaceNormal;
    bool entering;
    getHitNormals(face, lightDir, vec3(a, b, c), shapeNormal, surfaceNormal, entering);
    cosPhi = dot(surfaceNormal, -lightDir);
}

// src/sampling.cpp:436-448
static void sampleLightFace(const vec3 &pos, const LightObject &lightObject, Generator &gen, vec3 &lightPos, int &fails) {
    for (int T = 1; T <= MaxTrys; T++) {
        int faceIndex = lightObject.faceDist(gen);
        const Face &lightFace = lightObject.faces[faceIndex];
        float cosPhi;
        generateRandomPointInLightFace(lightFace, pos, gen, lightPos, cosPhi);
        if (cosPhi > 0.0f && gen() < cosPhi) return;
        fails++;
    }
    lightPos = vec3(NAN);
}

// src/sampling.cpp:450-465
static void sampleSkyBox(const SkyBox &skyBox, const vec3 &shapeNormal, Generator &gen, vec3 &Dir, vec3 &light) {
    for (int T = 1; T <= MaxTrys; T++) {
        int pixelIndex = skyBox.dist(gen);
        int u = pixelIndex % skyBox.width, v = pixelIndex / skyBox.width;
        float phi = PI * (float(v) + 0.5f) / float(skyBox.height);
        float theta = 2.0f * PI * (float(u) + 0.5f) / float(skyBox.width);
        Dir = vec3(-std::sin(phi) * std::sin(theta), std::cos(phi), std::sin(phi) * std::cos(theta));
        float cosPhi = dot(Dir, shapeNormal);
        if (cosPhi > 0.0f) {
            light = skyBox.data[pixelIndex] / skyBox.dist.pdf(pixelIndex);
            return;
        }
    }
    Dir = vec3(NAN);
}

// src/sampling.cpp:467-527
template <class GenFor>
std::vector<LightSample> sampleDirectLight(const BSDF &bsdf, const Model &model, GenFor genFor, int sampleCnt) {
    const vec3 &pos = bsdf.surface.position;
    std::vector<LightSample> samples;
    if (model.skyMap.empty()) {
        std::vector<float> weights;
        getLightObjectWeight(pos, bsdf, model, weights);
        float totalWeight = 0.0f;
        for (float weight : weights) totalWeight += weight;
        if (totalWeight == 0.0f) return {};
        for (int T = 0; T < sampleCnt; T++) {
            Generator &gen = genFor(T);
            int lightIndex = sampleIndex(weights, totalWeight * gen());
            if (lightIndex == -1) continue;
            float P_lightObject = weights[lightIndex] / totalWeight;
            const auto &lightObject = model.lightObjects[lightIndex];
            vec3 lightPos;
            int fails = 0;
            sampleLightFace(pos, lightObject, gen, lightPos, fails);
            if (!isfinite(lightPos)) continue;
            vec3 lightDir = normalize(lightPos - pos);
            float distance = length(lightPos - pos);
            if (model.rayHit_test({pos, lightDir}, distance - eps_zero)) continue;
            vec3 bsdfPdf = bsdf.getBSDF(lightDir);
            clampPdf(bsdfPdf);
            vec3 light = 2.0f * PI * lightObject.power * lightObject.color / (distance * distance + eps_lightRadius) /
                         P_lightObject;
            samples.emplace_back(bsdfPdf, light, 1.0f / static_cast<float>(sampleCnt * (fails + 1)));
            if (adj_tap) adj_tap->lights.push_back({lightDir, distance, P_lightObject, lightIndex});
        }
    } else {
        for (int T = 0; T < sampleCnt; T++) {
            Generator &gen = genFor(T);
            vec3 Dir, light;
            sampleSkyBox(model.skyMap, bsdf.surface.surfaceNormal, gen, Dir, light);
            if (!isfinite(Dir)) continue;
            if (model.rayHit_test({pos, Dir}, INFINITY)) continue;
            vec3 bsdfPdf = bsdf.getBSDF(Dir);
            clampPdf(bsdfPdf);
            samples.emplace_back(bsdfPdf, light, 1.0f / static_cast<float>(sampleCnt));
            if (adj_tap) adj_tap->lights.push_back({Dir, INFINITY, 1.0f, -1});
        }
    }
    return samples;
}

template std::vector<LightSample> sampleDirectLight<OneGen>(const BSDF &, const Model &, OneGen, int);

// Probes for the tests: the file-local helpers above, one call each.
float brdf_term(int which, float x, float a) {
    return which == 0 ? GTR1(x, a) : (which == 1 ? GTR2(x, a) : smithG_GGX(x, a));
}
void probe_sky_sample(const SkyBox &sky, const vec3 &normal, Generator &gen, vec3 &Dir, vec3 &light) {
    sampleSkyBox(sky, normal, gen, Dir, light);
}
template std::vector<LightSample> sampleDirectLight<PerSampleGen>(const BSDF &, const Model &, PerSampleGen, int);

// src/image.cpp:615-628
static void accumulateInwardRadiance_basic(RadianceData &radiance, vec3 inradiance, float weight) {
    if (!isfinite(inradiance)) return;
    if (!std::isfinite(weight)) return;
    radiance.radiance += inradiance * weight;
    radiance.Var += dot(inradiance, inradiance) * weight;
}

// src/image.cpp:630-659
void accumulateInwardRadiance(const vec3 &baseColor, const LightSample &sample, RadianceData &radiance_d,
                              RadianceData &radiance_s) {
    if (length(sample.light) < eps_zero) return;
    vec3 baseColor0 = normalize(baseColor);
    if (length(baseColor) < eps_zero) {
        accumulateInwardRadiance_basic(radiance_s, sample.light * sample.bsdfPdf, sample.weight);
        return;
    }
    static const vec3 White = normalize(vec3(1.0f));
    float XdotY = dot(baseColor0, White);
    if (XdotY > 0.99f) {
        accumulateInwardRadiance_basic(radiance_d, sample.light * sample.bsdfPdf / baseColor, sample.weight);
        return;
    }
    vec3 perp = normalize(cross(baseColor0, White));
    vec3 bsdfPdf_p = sample.bsdfPdf - perp * dot(perp, sample.bsdfPdf);
    float d1 = dot(bsdfPdf_p, White);
    float d2 = dot(bsdfPdf_p, baseColor0);
    float AplusB = (d1 + d2) / (1 + XdotY);
    float AminusB = (d1 - d2) / (1 - XdotY);
    float B = (AplusB - AminusB) / 2.0f;
    vec3 brdfPdf_base = B * baseColor0;
    accumulateInwardRadiance_basic(radiance_d, sample.light * B / length(baseColor), sample.weight);
    accumulateInwardRadiance_basic(radiance_s, sample.light * (sample.bsdfPdf - brdfPdf_base), sample.weight);
}

}  // namespace ref

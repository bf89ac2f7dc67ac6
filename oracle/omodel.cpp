// ORACLE -- test infrastructure only (see oracle.h).
// Restates the material sampling of src/material.cpp:26-148,337-383 and the
// shading-point code of src/model.cpp:23-82,232-328.
#include <algorithm>
#include <cmath>
#include "oracle.h"

namespace ref {

// src/material.cpp:113-148 on fp32 texels (EXT: no integer truncation).
void ImageData::generateMipmaps() {
    map_depth = 8;
    for (int level = 1; level < 8; ++level) {
        int prevWidth = width >> (level - 1), prevHeight = height >> (level - 1);
        int currWidth = width >> level, currHeight = height >> level;
        if (currWidth == 0 || currHeight == 0) {
            map_depth = level;
            break;
        }
        data[level].resize((size_t)currWidth * currHeight * 3);
        for (int y = 0; y < currHeight; ++y)
            for (int x = 0; x < currWidth; ++x)
                for (int c = 0; c < 3; ++c) {
                    float sum = 0.0f;
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx) {
                            int px = (x * 2 + dx) % prevWidth;
                            int py = (y * 2 + dy) % prevHeight;
                            sum += data[level - 1][((size_t)py * prevWidth + px) * 3 + c];
                        }
                    data[level][((size_t)y * currWidth + x) * 3 + c] = sum / 4.0f;
                }
    }
}

// src/material.cpp:48-79 (wrap addressing, bilinear)
static vec3 get_bilinear(const std::vector<float> &data, int width, int height, float u, float v) {
    auto mod = [](int &x, int m) {
        if (x < 0 || x >= m) {
            x %= m;
            if (x < 0) x += m;
        }
    };
    float x = u * float(width), y = v * float(height);
    int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    float dx = x - float(x0), dy = y - float(y0);
    mod(x0, width);
    mod(y0, height);
    int x1 = (x0 + 1) % width, y1 = (y0 + 1) % height;
    auto at = [&](int xx, int yy) {
        size_t i = ((size_t)yy * width + xx) * 3;
        return vec3(data[i], data[i + 1], data[i + 2]);
    };
    vec3 c00 = at(x0, y0), c01 = at(x1, y0), c10 = at(x0, y1), c11 = at(x1, y1);
    vec3 c0 = c00 * (1.0f - dx) + c01 * dx;
    vec3 c1 = c10 * (1.0f - dx) + c11 * dx;
    return c0 * (1.0f - dy) + c1 * dy;
}

// src/material.cpp:81-94
vec3 ImageData::get(float u, float v, float depth) const {
    v = 1.0f - v;
    depth = std::max(std::min(depth, static_cast<float>(map_depth - 1)), 0.0f);
    int level = static_cast<int>(depth);
    int nextLevel = std::min(level + 1, map_depth - 1);
    float levelBlend = depth - float(level);
    vec3 ret1 = get_bilinear(data[level], width >> level, height >> level, u, v);
    vec3 ret2 = get_bilinear(data[nextLevel], width >> nextLevel, height >> nextLevel, u, v);
    return ret1 * (1.0f - levelBlend) + ret2 * levelBlend;
}

// src/material.cpp:102-107: some alpha byte below 255.
bool ImageData::hasTransparentPart() const {
    for (float a : alpha)
        if (a < 1.0f) return true;
    return false;
}

// The alpha component of get<vec4>(u, v, 0) (src/material.cpp:48-94): v is
// flipped, level 0 is sampled bilinearly with wrap-around, and the level-1
// sample enters with blend weight 0, which leaves the level-0 value as it is.
float ImageData::getAlpha(float u, float v) const {
    v = 1.0f - v;
    float x = u * float(width), y = v * float(height);
    int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    float dx = x - float(x0), dy = y - float(y0);
    auto wrap = [](int &q, int m) {
        if (q < 0 || q >= m) {
            q %= m;
            if (q < 0) q += m;
        }
    };
    wrap(x0, width);
    wrap(y0, height);
    int x1 = (x0 + 1) % width, y1 = (y0 + 1) % height;
    float a00 = alpha[(size_t)y0 * width + x0], a01 = alpha[(size_t)y0 * width + x1];
    float a10 = alpha[(size_t)y1 * width + x0], a11 = alpha[(size_t)y1 * width + x1];
    float a0 = a00 * (1.0f - dx) + a01 * dx;
    float a1 = a10 * (1.0f - dx) + a11 * dx;
    return a0 * (1.0f - dy) + a1 * dy;
}

// ---------------------------------------------------------------- 8-bit
// src/material.cpp:113-148: 2 x 2 integer means (sum / 4) in the texture's
// channel layout, wrapping at odd edges, until a side would reach 0.
void ImageData8::generateMipmaps() {
    map_depth = 8;
    for (int level = 1; level < 8; ++level) {
        int prevWidth = width >> (level - 1), prevHeight = height >> (level - 1);
        int currWidth = width >> level, currHeight = height >> level;
        if (currWidth == 0 || currHeight == 0) {
            map_depth = level;
            break;
        }
        data[level].resize((size_t)currWidth * currHeight * channels);
        for (int y = 0; y < currHeight; ++y)
            for (int x = 0; x < currWidth; ++x)
                for (int c = 0; c < channels; ++c) {
                    int sum = 0;
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx) {
                            int px = (x * 2 + dx) % prevWidth;
                            int py = (y * 2 + dy) % prevHeight;
                            sum += data[level - 1][((size_t)py * prevWidth + px) * channels + c];
                        }
                    data[level][((size_t)y * currWidth + x) * channels + c] = (uint8_t)(sum / 4);
                }
    }
}

// src/material.cpp:102-107: every 4th byte from the 4th, below 255.
bool ImageData8::hasTransparentPart() const {
    for (size_t i = 3; i < data[0].size(); i += 4)
        if (data[0][i] < 255) return true;
    return false;
}

namespace {
// get_basic (src/material.cpp:26-46): BitWidth bytes at `index`, / 255.
template <int BitWidth>
void get_basic8(const std::vector<uint8_t> &d, size_t index, float out[4]) {
    for (int c = 0; c < 4; ++c) out[c] = 0.0f;
    for (int c = 0; c < BitWidth; ++c) {
        size_t i = index + (size_t)c;
        out[c] = static_cast<float>(i < d.size() ? d[i] : 0) / 255.0f;
    }
}

// get_bilinear (src/material.cpp:48-79) with the sampler's stride BitWidth.
template <int BitWidth>
void get_bilinear8(const std::vector<uint8_t> &data, int width, int height, float u, float v, float out[4]) {
    auto mod = [](int &x, int m) {
        if (x < 0 || x >= m) {
            x %= m;
            if (x < 0) x += m;
        }
    };
    float x = u * float(width), y = v * float(height);
    int x0 = int(std::floor(x)), y0 = int(std::floor(y));
    float dx = x - float(x0), dy = y - float(y0);
    mod(x0, width);
    mod(y0, height);
    int x1 = (x0 + 1) % width, y1 = (y0 + 1) % height;
    float c00[4], c01[4], c10[4], c11[4];
    get_basic8<BitWidth>(data, ((size_t)y0 * width + x0) * BitWidth, c00);
    get_basic8<BitWidth>(data, ((size_t)y0 * width + x1) * BitWidth, c01);
    get_basic8<BitWidth>(data, ((size_t)y1 * width + x0) * BitWidth, c10);
    get_basic8<BitWidth>(data, ((size_t)y1 * width + x1) * BitWidth, c11);
    for (int c = 0; c < BitWidth; ++c) {
        float c0 = c00[c] * (1.0f - dx) + c01[c] * dx;
        float c1 = c10[c] * (1.0f - dx) + c11[c] * dx;
        out[c] = c0 * (1.0f - dy) + c1 * dy;
    }
    for (int c = BitWidth; c < 4; ++c) out[c] = 0.0f;
}

// gammaPow (src/material.cpp:337-347) on the first three components
void gammaPow(float v[4]) {
    const float gamma = 2.2f;
    for (int c = 0; c < 3; ++c) v[c] = std::pow(v[c], gamma);
}
}  // namespace

// src/material.cpp:81-94
template <int BitWidth>
void ImageData8::get(float u, float v, float depth, float out[4]) const {
    v = 1.0f - v;
    depth = std::max(std::min(depth, static_cast<float>(map_depth - 1)), 0.0f);
    int level = static_cast<int>(depth);
    int nextLevel = std::min(level + 1, map_depth - 1);
    float levelBlend = depth - float(level);
    float r1[4], r2[4];
    get_bilinear8<BitWidth>(data[level], width >> level, height >> level, u, v, r1);
    get_bilinear8<BitWidth>(data[nextLevel], width >> nextLevel, height >> nextLevel, u, v, r2);
    for (int c = 0; c < 4; ++c) out[c] = r1[c] * (1.0f - levelBlend) + r2[c] * levelBlend;
}
template void ImageData8::get<3>(float, float, float, float *) const;
template void ImageData8::get<4>(float, float, float, float *) const;

// src/material.cpp:349-356.  EXT: constant albedo without a texture, fp32
// textures linear (no gamma); an 8-bit texture as the reference samples it.
vec4 Material::getDiffuseColor(float u, float v, float duv) const {
    if (!diffuse8.empty()) {
        float map_depth = std::isnan(duv) ? 0.0f : std::log2(duv * float(diffuse8.width));
        float c[4];
        diffuse8.get<4>(u, v, map_depth, c);
        gammaPow(c);
        return vec4(c[0], c[1], c[2], c[3]);
    }
    if (diffuse.empty()) return vec4(baseColor.x, baseColor.y, baseColor.z, 1.0f);
    float map_depth = std::isnan(duv) ? 0.0f : std::log2(duv * float(diffuse.width));
    vec3 c = diffuse.get(u, v, map_depth);
    return vec4(c.x, c.y, c.z, 1.0f);
}

// src/material.cpp:358-363
vec3 Material::getNormal(float u, float v, float duv) const {
    if (normals8.empty()) return vec3(0.0f);
    float map_depth = std::isnan(duv) ? 0.0f : std::log2(duv * float(normals8.width));
    float c[4];
    normals8.get<3>(u, v, map_depth, c);
    return vec3(c[0], c[1], c[2]) * 2.0f - 1.0f;
}

// src/material.cpp:365-372.  EXT: constant emission without an emissive texture.
vec3 Material::getEmissiveColor(float u, float v, float duv) const {
    if (emissive8.empty()) return emission;
    float map_depth = std::isnan(duv) ? 0.0f : std::log2(duv * float(emissive8.width));
    float c[4];
    emissive8.get<4>(u, v, map_depth, c);
    gammaPow(c);
    return vec3(c[0], c[1], c[2]);
}

// src/material.cpp:374-383.  EXT: constant metallic without a specular map.
void Material::getSurfaceData(float u, float v, float &r, float &m) const {
    if (specular8.empty()) {
        m = metallic;
        r = roughness;
        return;
    }
    float c[4];
    specular8.get<4>(u, v, 0, c);
    m = std::min(c[2], 0.99f);
    r = std::max(c[1], 1e-3f);
}

// src/model.cpp:232-238
static bool reverseFix(vec3 &v, const vec3 &Dir) {
    if (dot(v, Dir) < 0.0f) {
        v *= -1.0f;
        return false;
    }
    return true;
}

// src/model.cpp:240-270
void getHitNormals(const Face &face, const vec3 &inDir, const vec3 &baryCoords, vec3 &shapeNormal,
                   vec3 &surfaceNormal_raw, bool &entering) {
    vec3 crossV0 = cross(face.v[1] - face.v[0], face.v[2] - face.v[0]);
    shapeNormal = normalize(crossV0);
    entering = reverseFix(shapeNormal, -inDir);
    surfaceNormal_raw = shapeNormal;
    float area = length(crossV0) / 2.0f;
    const float areaThreshold = 1e-2f;
    if (area > areaThreshold) return;
    vec3 normalV0 = face.data[0]->normal, normalV1 = face.data[1]->normal, normalV2 = face.data[2]->normal;
    reverseFix(normalV0, shapeNormal);
    reverseFix(normalV1, shapeNormal);
    reverseFix(normalV2, shapeNormal);
    const float dotThreshold = 0.85f;
    surfaceNormal_raw = normalize(baryCoords[0] * (dot(normalV0, shapeNormal) > dotThreshold ? normalV0 : shapeNormal) +
                                  baryCoords[1] * (dot(normalV1, shapeNormal) > dotThreshold ? normalV1 : shapeNormal) +
                                  baryCoords[2] * (dot(normalV2, shapeNormal) > dotThreshold ? normalV2 : shapeNormal));
    if (!isfinite(surfaceNormal_raw)) surfaceNormal_raw = shapeNormal;
}

// src/model.cpp:272-277
static vec2 getDuv(const Face &face, const vec3 &hit_dPdx) {
    vec3 b = barycentric(face.v[0], face.v[1], face.v[2], face.v[0] + hit_dPdx);
    return (b[0] - 1.0f) * face.data[0]->uv + b[1] * face.data[1]->uv + b[2] * face.data[2]->uv;
}

// src/model.cpp:279-298
void calcSurfaceNormal(const Face &face, const vec3 &normalMap, const vec3 &shapeNormal, vec3 &surfaceNormal) {
    vec3 edge1 = face.v[1] - face.v[0], edge2 = face.v[2] - face.v[0];
    vec2 deltaUV1 = face.data[1]->uv - face.data[0]->uv, deltaUV2 = face.data[2]->uv - face.data[0]->uv;
    float f = 1.0f / (deltaUV1.x * deltaUV2.y - deltaUV2.x * deltaUV1.y);
    vec3 tangentBasisU = f * (deltaUV2.y * edge1 - deltaUV1.y * edge2),
         tangentBasisV = f * (-deltaUV2.x * edge1 + deltaUV1.x * edge2);
    vec3 tangent = normalize(tangentBasisU - shapeNormal * dot(shapeNormal, tangentBasisU)),
         bitangent = normalize(tangentBasisV - shapeNormal * dot(shapeNormal, tangentBasisV) -
                               tangent * dot(tangent, tangentBasisV));
    vec3 sav = surfaceNormal;
    surfaceNormal = normalize(tangent * normalMap.x + bitangent * normalMap.y + surfaceNormal);
    if (!isfinite(surfaceNormal)) surfaceNormal = sav;
}

// src/model.cpp:300-328
void getHitMaterial(const Face &face, const vec3 &baryCoords, const vec3 &hit_dPdx, const vec3 &hit_dPdy,
                    HitInfo &hitInfo) {
    vec2 texUV = baryCoords[0] * face.data[0]->uv + baryCoords[1] * face.data[1]->uv + baryCoords[2] * face.data[2]->uv;
    const Material &material = *face.material;
    hitInfo.id = material.id;
    material.getSurfaceData(texUV[0], texUV[1], hitInfo.roughness, hitInfo.metallic);
    hitInfo.opacity = material.opacity;
    hitInfo.eta = material.ior;
    if (hitInfo.opacity > 1.0f - eps_zero) hitInfo.entering = true;
    vec2 dUVdx = getDuv(face, hit_dPdx), dUVdy = getDuv(face, hit_dPdy);
    float duv = (isfinite(dUVdx) && isfinite(dUVdx)) ? (length(dUVdx) + length(dUVdy)) / 2.0f : NAN;
    if (adj_tap) {
        adj_tap->u = texUV[0];
        adj_tap->v = texUV[1];
        adj_tap->duv = duv;
    }
    if (hitInfo.opacity < eps_zero)
        hitInfo.baseColor = material.transmittingColor;
    else {
        vec4 c = material.getDiffuseColor(texUV[0], texUV[1], duv);
        hitInfo.baseColor = vec3(c.x, c.y, c.z);
    }
    hitInfo.emission = material.getEmissiveColor(texUV[0], texUV[1], duv);
    vec3 normalMap = material.getNormal(texUV[0], texUV[1], duv);
    calcSurfaceNormal(face, normalMap, hitInfo.shapeNormal, hitInfo.surfaceNormal);
}

}  // namespace ref

// ORACLE -- test infrastructure only (see oracle.h).
// Restates src/geometry.cpp, src/bvh.cpp, src/component.cpp:5-67,120-140 and
// the hit API of src/model.cpp:217-354 (without alpha textures).
#include <algorithm>
#include <cmath>
#include <iostream>
#include "oracle.h"

namespace ref {

// src/geometry.cpp:4-10 (OR semantics)
bool isfinite(vec2 v) { return std::isfinite(v.x) || std::isfinite(v.y); }
bool isfinite(vec3 v) { return std::isfinite(v.x) || std::isfinite(v.y) || std::isfinite(v.z); }

// src/geometry.cpp:12-19
float sqrt_s(float x) {
    if (x < -eps_zero) return 0.0f;
    return x <= 0.0f ? 0.0f : std::sqrt(x);
}

// src/geometry.cpp:21-27
vec3 pow_s(vec3 v, float k) {
    if (v.x < 0.0f) v.x = 0.0f;
    if (v.y < 0.0f) v.y = 0.0f;
    if (v.z < 0.0f) v.z = 0.0f;
    return vec3(std::pow(v.x, k), std::pow(v.y, k), std::pow(v.z, k));
}

// src/geometry.cpp:33-38
Box operator+(const Box &A, const Box &B) {
    return Box(vec3(std::fmin(A.v0[0], B.v0[0]), std::fmin(A.v0[1], B.v0[1]), std::fmin(A.v0[2], B.v0[2])),
               vec3(std::fmax(A.v1[0], B.v1[0]), std::fmax(A.v1[1], B.v1[1]), std::fmax(A.v1[2], B.v1[2])));
}

// src/geometry.cpp:40-61
void rayInBox(const Ray &ray, const Box &box, float &tL, float &tR) {
    for (int i = 0; i < 3; ++i) {
        if (std::abs(ray.direction[i]) < eps_zero) {
            if (ray.origin[i] < box.v0[i] || ray.origin[i] > box.v1[i]) {
                tR = -1.0;
                return;
            }
        } else {
            float invD = 1.0f / ray.direction[i];
            if (invD >= 0) {
                tL = std::fmax(tL, (box.v0[i] - ray.origin[i]) * invD);
                tR = std::fmin(tR, (box.v1[i] - ray.origin[i]) * invD);
            } else {
                tL = std::fmax(tL, (box.v1[i] - ray.origin[i]) * invD);
                tR = std::fmin(tR, (box.v0[i] - ray.origin[i]) * invD);
            }
            tR += eps_zero;
            if (tL > tR) return;
        }
    }
}

// src/geometry.cpp:63-87
float RayTriangleIntersection(const Ray &ray, const vec3 &v0, const vec3 &v1, const vec3 &v2) {
    vec3 edge1 = v1 - v0;
    vec3 edge2 = v2 - v0;
    vec3 h = cross(ray.direction, edge2);
    float a = dot(edge1, h);
    if (std::abs(a) / length(edge1) < eps_zero) return INFINITY;
    float f = 1.0f / a;
    vec3 s = ray.origin - v0;
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return INFINITY;
    vec3 q = cross(s, edge1);
    float v = f * dot(ray.direction, q);
    if (v < 0.0f || u + v > 1.0f) return INFINITY;
    return f * dot(edge2, q);
}

// src/geometry.cpp:89-103
vec3 barycentric(const vec3 &v0, const vec3 &v1, const vec3 &v2, const vec3 &P) {
    vec3 v0v1 = v1 - v0;
    vec3 v0v2 = v2 - v0;
    vec3 n = cross(v0v1, v0v2);
    float denom = dot(n, n);
    vec3 v0P = P - v0;
    float alpha = dot(cross(v0P, v0v2), n) / denom;
    float beta = dot(cross(v0v1, v0P), n) / denom;
    float gamma = 1.0f - alpha - beta;
    return vec3(gamma, alpha, beta);
}

// src/geometry.cpp:105-111
void getTangentSpace(const vec3 &normal, vec3 &tangent, vec3 &bitangent) {
    vec3 v0 = std::abs(normal.x) < 0.8f ? vec3(1.0f, 0.0f, 0.0f) : vec3(0.0f, 1.0f, 0.0f);
    tangent = normalize(cross(v0, normal));
    bitangent = cross(normal, tangent);
}

// src/geometry.cpp:113-123
void getTangentSpaceWithInDir(const vec3 &normal, const vec3 &inDir, vec3 &tangent, vec3 &bitangent) {
    const float eps_direction = 1e-3f;
    float cosTheta = dot(normal, inDir);
    if (cosTheta < -1.0f + eps_direction) {
        getTangentSpace(normal, tangent, bitangent);
        return;
    }
    bitangent = normalize(cross(normal, inDir));
    tangent = cross(bitangent, normal);
}

// ------------------------------------------------------------- component
vec3 Face::center() const { return (v[0] + v[1] + v[2]) / 3.0f; }
Box Face::aabb() const { return Box(min(v[0], min(v[1], v[2])), max(v[0], max(v[1], v[2]))); }

// src/component.cpp:12-18
void RandomDistribution::Init(const std::vector<float> &distribution) {
    prefixSums.resize(distribution.size());
    prefixSums[0] = distribution[0];
    for (size_t i = 1; i < distribution.size(); ++i) prefixSums[i] = prefixSums[i - 1] + distribution[i];
}
// src/component.cpp:20-24
int RandomDistribution::operator()(Generator &gen) const {
    float randomValue = prefixSums.back() * gen();
    return static_cast<int>(std::lower_bound(prefixSums.begin(), prefixSums.end(), randomValue) - prefixSums.begin());
}
// src/component.cpp:26-31
float RandomDistribution::pdf(int index) const {
    float now = prefixSums[index];
    if (index > 0) now -= prefixSums[index - 1];
    return now / prefixSums.back();
}

// src/component.cpp:54-67
void SkyBox::Init() {
    std::vector<float> weights;
    weights.reserve((size_t)width * height);
    for (int v = 0; v < height; v++)
        for (int u = 0; u < width; u++) {
            float phi = PI * (float(v) + 0.5f) / float(height);
            float area = std::sin(phi) * 2.0f * PI / float(width * height);
            int id = v * width + u;
            data[id] *= area;
            float Clum = dot(data[id], RGB_Weight);
            weights.emplace_back(Clum);
        }
    dist.Init(weights);
}

// src/component.cpp:120-140
vec3 SkyBox::get(vec3 dir) const {
    if (data.empty()) return vec3(0.0f);
    float theta = std::atan2(-dir.x, dir.z);
    float phi = std::acos(dir.y);
    if (theta < 0.0f) theta += 2.0f * PI;
    int u = static_cast<int>(theta / (2.0f * PI) * float(width));
    int v = static_cast<int>(phi / PI * float(height));
    if (u < 0) u = 0;
    if (u >= width) u = width - 1;
    if (v < 0) v = 0;
    if (v >= height) v = height - 1;
    phi = PI * (float(v) + 0.5f) / float(height);
    float area = std::sin(phi) * 2.0f * PI / float(width * height);
    return data[v * width + u] / area;
}

// ------------------------------------------------------------- BVH
static const unsigned int LeafBagSize = 10;   // src/bvh.cpp:16

template <int axis>
static bool cmp(const Face &A, const Face &B) { return A.center()[axis] < B.center()[axis]; }
static bool (*cmpFunc[3])(const Face &, const Face &) = {cmp<0>, cmp<1>, cmp<2>};

// src/bvh.cpp:18-42
void BVH::dfs_build(int u, int faceL, int faceR) {
    if ((unsigned)(faceR - faceL) <= LeafBagSize) {
        node[u].faceL = faceL;
        node[u].faceR = faceR;
        node[u].box = Box(vec3(INFINITY), vec3(-INFINITY));
        for (int i = faceL; i < faceR; i++) node[u].box = node[u].box + faces[i].aabb();
        return;
    }
    vec3 Em = vec3(0), Em2 = vec3(0);
    for (int i = faceL; i < faceR; i++) {
        vec3 m = faces[i].center();
        Em += m;
        Em2 += m * m;
    }
    vec3 D = Em2 - Em * Em / (float)(faceR - faceL);
    int axis = 0;
    if (D[1] > D[0]) axis = 1;
    if (D[2] > D[0] && D[2] > D[1]) axis = 2;
    int faceM = (faceR + faceL) / 2;
    std::nth_element(faces + faceL, faces + faceM, faces + faceR, cmpFunc[axis]);
    dfs_build(u << 1, faceL, faceM);
    dfs_build(u << 1 | 1, faceM, faceR);
    node[u].box = node[u << 1].box + node[u << 1 | 1].box;
}

static int nodeCount(int u, int n) { return ((unsigned)n <= LeafBagSize) ? u : nodeCount(u << 1 | 1, (n + 1) >> 1); }

// src/bvh.cpp:48-54 (zero faces is undefined there; callers reject it)
void BVH::build(std::vector<Face> &faces0) {
    size = nodeCount(1, int(faces0.size())) + 1;
    node.assign(size, BVH_Node{});
    faces = &faces0.front();
    dfs_build(1, 0, int(faces0.size()));
}

// src/bvh.cpp:56-88, with counters: one interior visit = two box tests.
void BVH::dfs_rayHit(int u, const Ray &ray, HitRecord &closest_hit, Counters *c) const {
    if (node[u].faceR) {
        if (c) { c->leaves++; c->tris += (uint64_t)(node[u].faceR - node[u].faceL); }
        for (int i = node[u].faceL; i < node[u].faceR; i++) {
            const Face &face = faces[i];
            float t = RayTriangleIntersection(ray, face.v[0], face.v[1], face.v[2]);
            if (closest_hit.t_min < t && t < closest_hit.t_max) {
                closest_hit.t_max = t;
                closest_hit.face = &face;
            }
        }
        return;
    }
    if (c) { c->interior++; c->boxes += 2; }
    float tL0 = closest_hit.t_min, tR0 = closest_hit.t_max, tL1 = closest_hit.t_min, tR1 = closest_hit.t_max;
    rayInBox(ray, node[u << 1].box, tL0, tR0);
    rayInBox(ray, node[u << 1 | 1].box, tL1, tR1);
    if (tL0 < tL1) {
        if (tL0 < tR0) dfs_rayHit(u << 1, ray, closest_hit, c);
        if (tL1 < tR1 && tL1 < closest_hit.t_max) dfs_rayHit(u << 1 | 1, ray, closest_hit, c);
    } else {
        if (tL1 < tR1) dfs_rayHit(u << 1 | 1, ray, closest_hit, c);
        if (tL0 < tR0 && tL0 < closest_hit.t_max) dfs_rayHit(u << 1, ray, closest_hit, c);
    }
}

void BVH::rayHit(const Ray &ray, HitRecord &closest_hit, Counters *c) const { dfs_rayHit(1, ray, closest_hit, c); }

// src/model.cpp:217-230
static bool TransparentTest(const Ray &ray, const HitRecord &hit) {
    const Face &face = *hit.face;
    const Material &material = *face.material;
    if (!material.hasFullyTransparentPart) return false;
    vec3 intersection = ray.origin + ray.direction * hit.t_max;
    vec3 b = barycentric(face.v[0], face.v[1], face.v[2], intersection);
    vec2 texUV = b.x * face.data[0]->uv + b.y * face.data[1]->uv + b.z * face.data[2]->uv;
    if (!material.diffuse8.empty()) return material.getDiffuseColor(texUV.x, texUV.y, NAN)[3] < eps_zero;
    return material.diffuse.getAlpha(texUV.x, texUV.y) < eps_zero;
}

// src/model.cpp:330-341
static const int maxRayDepth_hit = 8;

HitRecord Model::rayHitFrom(Ray ray, float tmin, float tmax, Counters *c) const {
    HitRecord hit(tmin, tmax);
    for (int T = 0; T < maxRayDepth_hit; T++) {
        bvh.rayHit(ray, hit, c);
        if (hit.t_max == tmax || !TransparentTest(ray, hit)) return hit;
        hit = HitRecord(hit.t_max + eps_zero, tmax);
    }
    return hit;
}

HitRecord Model::rayHit(Ray ray, Counters *c) const { return rayHitFrom(ray, eps_zero, INFINITY, c); }

// src/model.cpp:343-354
bool Model::rayHitTestFrom(Ray ray, float tmin, float aimDepth, Counters *c) const {
    HitRecord hit(tmin, aimDepth + eps_zero);
    for (int T = 0; T < maxRayDepth_hit; T++) {
        bvh.rayHit(ray, hit, c);
        if (hit.t_max >= aimDepth) return false;
        if (!TransparentTest(ray, hit)) return true;
        hit = HitRecord(hit.t_max + eps_zero, aimDepth + eps_zero);
    }
    return true;
}

bool Model::rayHit_test(Ray ray, float aimDepth, Counters *c) const { return rayHitTestFrom(ray, eps_zero, aimDepth, c); }

}  // namespace ref

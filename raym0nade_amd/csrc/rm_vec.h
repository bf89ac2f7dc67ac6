// Three-component fp32 vector for host and gfx950 code.
//
// Operation order follows the reference's vector library (glm) so that fp32
// results are reproducible bit for bit when compiled with -ffp-contract=off:
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   length(v)    = sqrt(dot(v,v))
//   normalize(v) = v * (1 / sqrt(dot(v,v)))
//   v / s        = per-component division
//   mix(a,b,t)   = a*(1-t) + b*t
#pragma once
#include <math.h>
#include "philox.h"

namespace rm {

struct V3 {
    float x, y, z;
};

RM_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RM_HD V3 v3(float s) { return V3{s, s, s}; }
RM_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RM_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RM_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
RM_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RM_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
RM_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
RM_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
RM_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RM_HD V3 &operator+=(V3 &a, V3 b) { a = a + b; return a; }
RM_HD V3 &operator*=(V3 &a, V3 b) { a = a * b; return a; }
RM_HD V3 &operator*=(V3 &a, float s) { a = a * s; return a; }
RM_HD V3 &operator/=(V3 &a, float s) { a = a / s; return a; }

RM_HD float dot(V3 a, V3 b) {
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return (x + y) + z;
}
RM_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
RM_HD float length(V3 v) { return sqrtf(dot(v, v)); }
RM_HD V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
// glm::min(x,y) = (y < x) ? y : x and glm::max(x,y) = (x < y) ? y : x per component.
RM_HD float gmin(float x, float y) { return (y < x) ? y : x; }
RM_HD float gmax(float x, float y) { return (x < y) ? y : x; }
RM_HD V3 vmin(V3 a, V3 b) { return V3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
RM_HD V3 vmax(V3 a, V3 b) { return V3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
// std::fmin / std::fmax as glibc defines them: the non-NaN operand, else
// fmin(x,y) = x < y ? x : y and fmax(x,y) = x > y ? x : y.
RM_HD float cfmin(float x, float y) { return (x < y || y != y) ? x : y; }
RM_HD float cfmax(float x, float y) { return (x > y || y != y) ? x : y; }
RM_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// The reference's isfinite(vec3) is true if ANY component is finite
// (src/geometry.cpp:4-10); only an all-non-finite vector counts as invalid.
RM_HD bool any_finite(V3 v) { return isfinite(v.x) || isfinite(v.y) || isfinite(v.z); }

}  // namespace rm

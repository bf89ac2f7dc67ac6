// ORACLE -- test infrastructure only.  Nothing in the product may include,
// link or call anything under oracle/; only tests/, the smoke check and the
// benchmark's CPU-baseline leg use it, as the checker.
//
// Minimal stand-in for the glm vector types the reference uses
// (include/geometry.h:9-11), with glm's evaluation order, written for this
// restatement:
//   dot  : (x*x' + y*y') + z*z'          cross: (y*z' - y'*z, z*x' - z'*x, x*y' - x'*y)
//   length = sqrt(dot(v,v))              normalize = v * (1/sqrt(dot(v,v)))
//   min(x,y) = y < x ? y : x             max(x,y) = x < y ? y : x
//   mix(a,b,t) = a*(1-t) + b*t
// Build with -ffp-contract=off.
#pragma once
#include <cmath>

namespace ref {

struct vec2 {
    float x = 0, y = 0;
    vec2() = default;
    vec2(float s) : x(s), y(s) {}
    vec2(float a, float b) : x(a), y(b) {}
    float &operator[](int i) { return i == 0 ? x : y; }
    float operator[](int i) const { return i == 0 ? x : y; }
};
struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    vec4() = default;
    explicit vec4(float s) : x(s), y(s), z(s), w(s) {}
    vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
};

inline vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
inline vec2 operator*(float s, vec2 a) { return {s * a.x, s * a.y}; }
inline vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }

inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3 operator/(vec3 a, vec3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline vec3 operator+(vec3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
inline vec3 operator-(vec3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
inline vec3 &operator+=(vec3 &a, vec3 b) { return a = a + b; }
inline vec3 &operator-=(vec3 &a, vec3 b) { return a = a - b; }
inline vec3 &operator*=(vec3 &a, vec3 b) { return a = a * b; }
inline vec3 &operator*=(vec3 &a, float s) { return a = a * s; }
inline vec3 &operator/=(vec3 &a, float s) { return a = a / s; }

inline vec4 operator*(vec4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline vec4 operator+(vec4 a, vec4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }

inline float dot(vec3 a, vec3 b) {
    vec3 t = a * b;
    return t.x + t.y + t.z;
}
inline float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
inline vec3 cross(vec3 x, vec3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline float length(vec3 v) { return std::sqrt(dot(v, v)); }
inline float length(vec2 v) { return std::sqrt(dot(v, v)); }
inline vec3 normalize(vec3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline vec3 min(vec3 a, vec3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline vec3 max(vec3 a, vec3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
inline vec3 min(vec3 a, float s) { return min(a, vec3(s)); }
inline vec3 max(vec3 a, float s) { return max(a, vec3(s)); }
inline vec3 mix(vec3 a, vec3 b, float t) { return a * (1.0f - t) + b * t; }
inline vec3 mix(vec3 a, vec3 b, vec3 t) { return a * (vec3(1.0f) - t) + b * t; }
inline vec3 pow(vec3 v, vec3 e) { return {std::pow(v.x, e.x), std::pow(v.y, e.y), std::pow(v.z, e.z)}; }

}  // namespace ref

// ORACLE -- test infrastructure only (see oracle.h).
//
// The reference's denoise path over Photo's four demodulated radiance
// buffers: Photo::spatialClamp (src/image.cpp:30-82) and Photo::filter
// (src/image.cpp:84-213): filterVar, then five edge-aware a-trous passes at
// strides 1, 2, 4, 8, 16.  Every arithmetic step keeps the reference's
// operation order (glm's dot = (x x' + y y') + z z', length = sqrt(dot),
// normalize = v * (1 / sqrt(dot))) so the restatement rounds as it does.
#include <algorithm>
#include <cmath>
#include <thread>
#include "oracle.h"

namespace ref {

namespace {

// src/image.cpp:30-75: a 7-tap binomial blur of the luminance (horizontal,
// then vertical, taps outside the image skipped), the centre tap's own share
// taken out, and a pixel brighter than 36x that neighbourhood estimate scaled
// down to it.  Only the radiance is scaled; Var is left as it is.
void clamp_buffer(std::vector<RadianceData> &rad, int width, int height) {
    const int R = 3;
    const float h[7] = {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f};
    const size_t n = (size_t)width * height;
    std::vector<float> lum(n), b1(n), b2(n);
    for (size_t i = 0; i < n; ++i) lum[i] = dot(rad[i].radiance, RGB_Weight);       // :38-40
    for (int y = 0; y < height; ++y)                                                   // :42-52
        for (int x = 0; x < width; ++x) {
            float sum = 0.0f;
            for (int dx = -R; dx <= R; ++dx) {
                int nx = x + dx;
                if (nx >= 0 && nx < width) sum += lum[(size_t)y * width + nx] * h[dx + R];
            }
            b1[(size_t)y * width + x] = sum;
        }
    for (int y = 0; y < height; ++y)                                                   // :54-64
        for (int x = 0; x < width; ++x) {
            float sum = 0.0f;
            for (int dy = -R; dy <= R; ++dy) {
                int ny = y + dy;
                if (ny >= 0 && ny < height) sum += b1[(size_t)ny * width + x] * h[dy + R];
            }
            b2[(size_t)y * width + x] = sum;
        }
    const float threshold = 36.0f;                                                     // :66-74
    for (size_t i = 0; i < n; ++i) {
        float l = lum[i];
        float blurred = b2[i] - h[R] * h[R] * l;
        if (l > threshold * blurred) rad[i].radiance *= blurred / (l + eps_zero) / (1.0f - h[R] * h[R]);
    }
}

}  // namespace

// src/image.cpp:84-107: Var <- the 3x3 binomial-weighted mean of Var plus the
// spatial variance of the radiance over the same window (weights are not
// renormalised at the border).
void filter_var(std::vector<RadianceData> &rad, int width, int height) {
    const int R = 1;
    const float h[3] = {0.25f, 0.5f, 0.25f};
    std::vector<float> out((size_t)width * height);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            vec3 E(0.0f);
            float E2 = 0.0f, V = 0.0f;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    int nx = x + dx, ny = y + dy;
                    if (nx < 0 || nx >= width || ny < 0 || ny >= height) continue;
                    const RadianceData &q = rad[(size_t)ny * width + nx];
                    float w = h[dx + R] * h[dy + R];
                    E += q.radiance * w;
                    E2 += dot(q.radiance, q.radiance) * w;
                    V += q.Var * w;
                }
            out[(size_t)y * width + x] = V + E2 - dot(E, E);
        }
    for (size_t i = 0; i < out.size(); ++i) rad[i].Var = out[i];
}

// src/image.cpp:109-156: the edge-stopping weight of neighbour q for centre p.
float denoise_weight(const HitInfo &Gp, const HitInfo &Gq, const RadianceData &Lp, const RadianceData &Lq,
                     bool specular) {
    const float sigma_z = 1.0f, sigma_n = 1024.0f, sigma_l = 1.0f, sigma_m = 1.0f, eps = 1e-2f, eps_r = 4e-2f;
    if (!isfinite(Gq.position)) return 0.0f;
    float weight = std::pow(std::max(0.0f, dot(Gp.surfaceNormal, Gq.surfaceNormal)), sigma_n);
    if (weight < 1e-6f) return 0.0f;
    float k = 0.0f;
    vec3 dir = normalize(Gq.position - Gp.position);
    float sinTheta = std::fabs(dot(Gp.shapeNormal, dir));
    float tanTheta = sinTheta / (std::sqrt(1.0f - sinTheta * sinTheta) + eps_zero);
    k += -tanTheta / sigma_z;
    float radianceDiff = length(Lp.radiance - Lq.radiance) / (sigma_l * std::sqrt(Lp.Var) + eps);
    k += -radianceDiff;
    float materialDiff = length(vec3(Gp.metallic - Gq.metallic, Gp.specular - Gq.specular, Gp.opacity - Gq.opacity));
    k += -materialDiff / sigma_m;
    if (k < -7.5f) return 0.0f;
    weight *= std::exp(k);
    if (specular) weight *= std::max(Gp.roughness, eps_r);
    if (!isfinite(weight)) return 0.0f;
    return weight;
}

// src/image.cpp:158-189: one a-trous pass at `step` (5x5 B3-spline taps at
// stride `step`), centre tap unweighted; Var propagates with squared weights.
// A pixel whose primary ray missed gets radiance 0 and Var 0.
void filter_pass(std::vector<RadianceData> &rad, const std::vector<HitInfo> &G, int width, int height, bool specular,
                 int step) {
    const int R = 2;
    const float h[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    const size_t n = (size_t)width * height;
    std::vector<vec3> rb(n, vec3(0.0f));
    std::vector<float> vb(n, 0.0f);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t p = (size_t)y * width + x;
            float wsum = 0.0f;
            const RadianceData &Lp = rad[p];
            const HitInfo &Gp = G[p];
            if (!isfinite(Gp.position)) continue;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    int nx = x + dx * step, ny = y + dy * step;
                    if (nx < 0 || nx >= width || ny < 0 || ny >= height) continue;
                    const size_t q = (size_t)ny * width + nx;
                    float w = h[dx + R] * h[dy + R];
                    if (dx != 0 || dy != 0) w *= denoise_weight(Gp, G[q], Lp, rad[q], specular);
                    wsum += w;
                    rb[p] += rad[q].radiance * w;
                    vb[p] += rad[q].Var * w * w;
                }
            rb[p] /= wsum;
            vb[p] /= wsum * wsum;
        }
    for (size_t i = 0; i < n; ++i) {
        rad[i].radiance = rb[i];
        rad[i].Var = vb[i];
    }
}

// src/image.cpp:191-198
void filter_radiance(std::vector<RadianceData> &rad, const std::vector<HitInfo> &G, int width, int height,
                     bool specular) {
    filter_var(rad, width, height);
    for (int step = 1; step <= 16; step *= 2) filter_pass(rad, G, width, height, specular, step);
}

// src/image.cpp:248-283: the bright part of the image (luminance >= 1,
// compressed by Clum^0.65, minus 1 per channel) blurred by five a-trous
// B3-spline passes (zero padding, weights not renormalised); after every
// pass the image gains a sixth of that level.
void Photo::bloom() {
    const int R = 2;
    const float h[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    const size_t n = (size_t)width * height;
    std::vector<vec3> level(n, vec3(0.0f));  // value-initialised like std::vector<vec3>(n)
    const float omega_bloom = 0.65f;
    for (size_t id = 0; id < n; ++id) {
        float Clum = dot(pixelarray[id], RGB_Weight);
        if (Clum < 1.0f) continue;
        level[id] = pixelarray[id] / std::pow(Clum, omega_bloom);
        level[id].x = gmax(level[id].x - 1.0f, 0.0f);
        level[id].y = gmax(level[id].y - 1.0f, 0.0f);
        level[id].z = gmax(level[id].z - 1.0f, 0.0f);
    }
    for (int step = 1; step <= 16; step *= 2) {
        const std::vector<vec3> prev = level;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                vec3 col(0.0f);
                for (int ky = -R; ky <= R; ++ky)
                    for (int kx = -R; kx <= R; ++kx) {
                        int nx = x + kx * step, ny = y + ky * step;
                        if (nx < 0 || nx >= width || ny < 0 || ny >= height) continue;
                        col += prev[(size_t)ny * width + nx] * h[ky + R] * h[kx + R];
                    }
                level[(size_t)y * width + x] = col;
            }
        for (size_t id = 0; id < n; ++id) pixelarray[id] += level[id] / 6.0f;
    }
}

// src/image.cpp:285-356: every pixel, nearest first (stable order by distance
// to the camera), spreads its colour over a disc of radius
// CoC0 = min(CoC |1 - focus / depth| + 1e-4, 96) with weights
// min(CoC0 - r, 1) normalised over the whole disc; a receiving pixel takes at
// most 0.99 of weight in total, first come first served, and ends as its
// weighted mean.
// EXT: a pixel whose primary ray missed (NaN position) is at depth +inf
// (CoC0 = min(CoC + 1e-4, 96)).  In the reference its NaN depth breaks the
// sort's ordering and int(NaN) is undefined.
void Photo::depthFeildBlur() {
    struct PixelPos {
        int x, y;
        float depth;
    };
    const size_t n = (size_t)width * height;
    std::vector<PixelPos> pixels;
    pixels.reserve(n);
    std::vector<vec3> blurred(n, vec3(0.0f));
    std::vector<float> gained(n, 0.0f);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            float depth = length(Gbuffer[(size_t)y * width + x].position - cameraPosition);
            if (depth != depth) depth = INFINITY;
            pixels.push_back(PixelPos{x, y, depth});
        }
    std::stable_sort(pixels.begin(), pixels.end(), [](const PixelPos &a, const PixelPos &b) { return a.depth < b.depth; });
    const float gainThreshold = 0.99f, MaxCoC = 96.0f;
    for (size_t i = 0; i < n; ++i) {
        const PixelPos &p = pixels[i];
        float CoC0 = std::min(CoC * std::fabs(1.0f - focus / p.depth) + eps_zero, MaxCoC);
        int radius = int(CoC0);
        vec3 color = pixelarray[(size_t)p.y * width + p.x];
        float totalWeight = 0.0f;
        for (int dy = -radius; dy <= radius; ++dy) {
            int xLen = int(std::sqrt(CoC0 * CoC0 - dy * dy));
            int xl, xr;
            for (xl = -xLen; xl <= 0; xl++) {
                float w = std::min(CoC0 - std::sqrt(float(xl * xl + dy * dy)), 1.0f);
                if (w == 1.0f) break;
                totalWeight += w;
            }
            for (xr = xLen; xr > 0; xr--) {
                float w = std::min(CoC0 - std::sqrt(float(xr * xr + dy * dy)), 1.0f);
                if (w == 1.0f) break;
                totalWeight += w;
            }
            totalWeight += xr - xl + 1;
        }
        for (int dy = -radius; dy <= radius; ++dy) {
            int xLen = int(std::sqrt(CoC0 * CoC0 - dy * dy));
            for (int dx = -xLen; dx <= xLen; ++dx) {
                int nx = p.x + dx, ny = p.y + dy;
                if (nx < 0 || nx >= width || ny < 0 || ny >= height) continue;
                float w = std::min(CoC0 - std::sqrt(float(dx * dx + dy * dy)), 1.0f) / totalWeight;
                if (w < eps_zero) continue;
                size_t id = (size_t)ny * width + nx;
                if (gained[id] + w > gainThreshold) w = gainThreshold - gained[id];
                if (w < eps_zero) continue;
                blurred[id] += color * w;
                gained[id] += w;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) pixelarray[i] = blurred[i] / gained[i];
}

// src/image.cpp:77-82
void Photo::spatialClamp() {
    clamp_buffer(radiance_Dd, width, height);
    clamp_buffer(radiance_Ds, width, height);
    clamp_buffer(radiance_Id, width, height);
    clamp_buffer(radiance_Is, width, height);
}

// src/image.cpp:200-213: the four buffers on four threads (diffuse buffers
// unweighted by roughness, specular ones weighted).
void Photo::filter() {
    std::thread t1([this] { filter_radiance(radiance_Dd, Gbuffer, width, height, false); });
    std::thread t2([this] { filter_radiance(radiance_Ds, Gbuffer, width, height, true); });
    std::thread t3([this] { filter_radiance(radiance_Id, Gbuffer, width, height, false); });
    std::thread t4([this] { filter_radiance(radiance_Is, Gbuffer, width, height, true); });
    t1.join();
    t2.join();
    t3.join();
    t4.join();
}

}  // namespace ref

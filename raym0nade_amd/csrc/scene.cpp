// Host side of rm_scene_create: the reference BVH build, its flattening into
// the HBM layout, light objects, sky CDF and texture mips, then the upload.
//
// The tree must be the reference's own tree (box extents feed the traversal's
// slab quirks, and leaf order decides exact-t ties), so the build follows
// BVH::dfs_build (src/bvh.cpp:18-42) decision for decision: fp32 centroid
// sums in the current face order, variance axis, std::nth_element at the
// midpoint, leaves of <= 10 faces.  It permutes an index array instead of
// 72-B Face records; nth_element's moves depend only on the comparison
// outcomes, so the permutation is the same.  Disjoint subtrees are built on
// separate threads, which cannot change any decision.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <thread>
#include "rm_internal.h"
#include "rm_vec.h"
#include "tex8.h"

namespace rm {
namespace {

constexpr float EPS = 1e-4f;  // eps_zero, include/geometry.h:13
constexpr float PI_F = 3.14159265358979323846f;

struct Builder {
    int64_t n;
    const float *verts;             // original order [n][9]
    // One record per slot: the centroid travels with the face id, so the
    // selection below compares contiguous records instead of looking the
    // centroid up through the id.  The comparator sees the same floats in
    // the same order either way, so std::nth_element makes the same moves
    // and the id permutation is the one the reference's index vector gets.
    struct Rec {
        float c[3];
        int32_t id;
    };
    HostArray<Rec> rec;             // slot -> (centroid, original id), permuted in place
    std::vector<float> lo, hi;      // per heap node [size][3]
    std::vector<int32_t> fl, fr;    // per heap node face range (fr==0: interior)
    std::vector<uint8_t> kind;      // 0 unused, 1 interior, 2 leaf
    int64_t size;

    V3 vert(int32_t id, int k) const {
        const float *p = verts + (int64_t)id * 9 + k * 3;
        return V3{p[0], p[1], p[2]};
    }

    void set_box(int64_t u, V3 a, V3 b) {
        lo[u * 3 + 0] = a.x; lo[u * 3 + 1] = a.y; lo[u * 3 + 2] = a.z;
        hi[u * 3 + 0] = b.x; hi[u * 3 + 1] = b.y; hi[u * 3 + 2] = b.z;
    }
    V3 get_lo(int64_t u) const { return V3{lo[u * 3], lo[u * 3 + 1], lo[u * 3 + 2]}; }
    V3 get_hi(int64_t u) const { return V3{hi[u * 3], hi[u * 3 + 1], hi[u * 3 + 2]}; }

    // Box operator+ (src/geometry.cpp:29-34) uses std::fmin / std::fmax.
    static V3 fmin3(V3 a, V3 b) { return V3{cfmin(a.x, b.x), cfmin(a.y, b.y), cfmin(a.z, b.z)}; }
    static V3 fmax3(V3 a, V3 b) { return V3{cfmax(a.x, b.x), cfmax(a.y, b.y), cfmax(a.z, b.z)}; }

    void dfs(int64_t u, int64_t L, int64_t R, int par_depth) {
        if (R - L <= LEAF_BAG) {
            kind[u] = 2;
            fl[u] = (int32_t)L;
            fr[u] = (int32_t)R;
            V3 a = v3(INFINITY), b = v3(-INFINITY);
            for (int64_t i = L; i < R; ++i) {
                int32_t f = rec[i].id;
                V3 v0 = vert(f, 0), v1 = vert(f, 1), v2 = vert(f, 2);
                // Face::aabb (src/component.cpp:41-46): glm::min/max
                V3 flo = vmin(v0, vmin(v1, v2)), fhi = vmax(v0, vmax(v1, v2));
                a = fmin3(a, flo);
                b = fmax3(b, fhi);
            }
            set_box(u, a, b);
            return;
        }
        kind[u] = 1;
        V3 Em = v3(0.0f), Em2 = v3(0.0f);
        for (int64_t i = L; i < R; ++i) {
            const float *c = rec[i].c;
            V3 m{c[0], c[1], c[2]};
            Em += m;
            Em2 += m * m;
        }
        V3 D = Em2 - Em * Em / (float)(R - L);
        int axis = 0;
        if (D.y > D.x) axis = 1;
        if (D.z > D.x && D.z > D.y) axis = 2;
        int64_t M = (R + L) / 2;
        std::nth_element(rec.begin() + L, rec.begin() + M, rec.begin() + R,
                         [axis](const Rec &A, const Rec &B) { return A.c[axis] < B.c[axis]; });
        if (par_depth > 0 && R - L > 65536) {
            std::thread t([&] { dfs(u << 1, L, M, par_depth - 1); });
            dfs(u << 1 | 1, M, R, par_depth - 1);
            t.join();
        } else {
            dfs(u << 1, L, M, 0);
            dfs(u << 1 | 1, M, R, 0);
        }
        set_box(u, fmin3(get_lo(u << 1), get_lo(u << 1 | 1)), fmax3(get_hi(u << 1), get_hi(u << 1 | 1)));
    }
};

// fn(i0, i1) over [0, n) cut into one contiguous range per hardware thread
// (one range below 65536 items); for loops whose iterations are independent.
template <class F>
void parallel_ranges(int64_t n, F fn) {
    int64_t parts = n >= 65536 ? (int64_t)std::max(1u, std::thread::hardware_concurrency()) : 1;
    int64_t step = (n + parts - 1) / parts;
    std::vector<std::thread> pool;
    for (int64_t p = 1; p < parts && p * step < n; ++p)
        pool.emplace_back(fn, p * step, std::min(n, (p + 1) * step));
    fn((int64_t)0, std::min(n, step));
    for (std::thread &t : pool) t.join();
}

// nodeCount (src/bvh.cpp:46-48).
int64_t node_count(int64_t u, int64_t n) { return n <= LEAF_BAG ? u : node_count(u << 1 | 1, (n + 1) >> 1); }

// Smallest x >= 0 with fl(x / L) >= 1e-4f.  Correctly rounded division is
// monotone in x, so  fl(|a| / L) < 1e-4f  <=>  |a| < thr  for every |a|,
// including NaN (both false).  L == 0 (degenerate edge1) gives a NaN or an
// out-of-range u in the reference, never a hit; thr = inf rejects it outright.
// The search starts at fl(EPS * L), within an ulp or two of the answer, and
// walks one representable value at a time in the direction the test gives;
// monotonicity makes the first change of the test the boundary.
float reject_threshold(float L) {
    if (L != L) return L;
    if (!(L > 0.0f) || isinf(L)) return INFINITY;
    auto passes = [L](uint32_t bits) {
        float x;
        memcpy(&x, &bits, 4);
        volatile float q = x / L;
        return q >= EPS;
    };
    float x0 = EPS * L;
    uint32_t b;
    memcpy(&b, &x0, 4);
    if (passes(b)) {
        while (b > 0 && passes(b - 1)) --b;
    } else {
        while (!passes(b)) ++b;  // ends at +inf at the latest (inf / L = inf)
    }
    float r;
    memcpy(&r, &b, 4);
    return r;
}

uint32_t leaf_code(int32_t first, int32_t count) { return ((uint32_t)first << LEAF_SHIFT) | (uint32_t)count; }

MatRec mat_rec(const rm_material &M, bool has_sky) {
    MatRec R;
    memset(&R, 0, sizeof(R));
    for (int k = 0; k < 3; ++k) {
        R.base[k] = M.base_color[k];
        // with a sky the reference never loads emissive textures (src/model.cpp:140)
        R.emission[k] = has_sky ? 0.0f : M.emission[k];
        R.trans[k] = M.transmitting_color[k];
    }
    R.roughness = M.roughness;
    R.metallic = M.metallic;
    R.opacity = M.opacity;
    R.ior = M.ior;
    R.tex = M.albedo_tex;
    R.ntex = M.normal_tex;
    R.etex = has_sky ? -1 : M.emissive_tex;  // no emissive maps with a sky (src/model.cpp:140)
    R.stex = M.specular_tex;
    return R;
}

// getAverageEmissiveColor (src/model.cpp:23-42) with a constant emission:
// the mean of 64 identical lookups.
V3 emissive_avg(V3 em) {
    V3 avg = v3(0.0f);
    for (int q = 0; q < 64; ++q) avg += em;
    return avg / 64.0f;
}
float emissive_lum(V3 em) { return dot(emissive_avg(em), v3(0.3f, 0.6f, 0.1f)); }

// getAverageEmissiveColor (src/model.cpp:23-42) with an emissive map: the
// mean of getEmissiveColor(uv, NaN) (level 0, then the power 2.2 with glibc's
// powf, src/material.cpp:337-347,365-372) over an 8 x 8 barycentric grid of
// the face's uvs.
V3 emissive_avg_tex(const HostScene &hs, const TexRec &T, const TriShade &F) {
    V3 avg = v3(0.0f);
    const int g = 8;
    for (int i = 0; i < g; ++i)
        for (int j = 0; j < g; ++j) {
            float a = static_cast<float>(i) / (g - 1), b = static_cast<float>(j) / (g - 1);
            if (a + b > 1.0f) {
                a = 1.0f - a;
                b = 1.0f - b;
            }
            float c = 1.0f - a - b;
            float u = (a * F.uv[0][0] + b * F.uv[1][0]) + c * F.uv[2][0];
            float v = (a * F.uv[0][1] + b * F.uv[1][1]) + c * F.uv[2][1];
            float t[4];
            tex8_get(hs.bytes.data(), T, u, v, 0.0f, 4, t);
            avg += V3{powf(t[0], 2.2f), powf(t[1], 2.2f), powf(t[2], 2.2f)};
        }
    return avg / static_cast<float>(g * g);
}

// The per-face mean emission of light face k (set at build; faces of
// constant-emission materials follow rm_scene_update's new values).
V3 light_face_avg(const HostScene &hs, int k) {
    const TriShade &F = hs.light_faces[k];
    const MatRec &M = hs.mats[F.material];
    if (M.etex >= 0) return hs.light_face_avg[k];
    return emissive_avg(V3{M.emission[0], M.emission[1], M.emission[2]});
}

// The rest of checkLightObject (src/model.cpp:44-82) for every light object
// whose faces are chosen: colour = area-weighted mean emission normalised to
// unit luminance, power = sum of area * luminance, centre = power-weighted
// centroid, and the face CDF (RandomDistribution::Init, src/component.cpp:12-19).
void light_params(HostScene &hs) {
    for (LightRec &L : hs.lights) {
        V3 color = v3(0.0f);
        std::vector<float> w(L.n_faces);
        for (int k = 0; k < L.n_faces; ++k) {
            const TriShade &F = hs.light_faces[L.first_face + k];
            V3 avg = light_face_avg(hs, L.first_face + k);
            float Clum = dot(avg, v3(0.3f, 0.6f, 0.1f));
            V3 v0{F.v[0][0], F.v[0][1], F.v[0][2]}, v1{F.v[1][0], F.v[1][1], F.v[1][2]},
                v2{F.v[2][0], F.v[2][1], F.v[2][2]};
            float area = length(cross(v1 - v0, v2 - v0)) / 2.0f;
            color += avg * area;
            w[k] = area * Clum;
        }
        color = color / dot(color, v3(0.3f, 0.6f, 0.1f));
        float power = 0.0f;
        V3 center = v3(0.0f);
        for (int k = 0; k < L.n_faces; ++k) {
            power += w[k];
            const TriShade &F = hs.light_faces[L.first_face + k];
            V3 v0{F.v[0][0], F.v[0][1], F.v[0][2]}, v1{F.v[1][0], F.v[1][1], F.v[1][2]},
                v2{F.v[2][0], F.v[2][1], F.v[2][2]};
            V3 c = (v0 + v1 + v2) / 3.0f;
            center += c * w[k];
        }
        center /= power;
        L.center[0] = center.x; L.center[1] = center.y; L.center[2] = center.z;
        L.power = power;
        L.color[0] = color.x; L.color[1] = color.y; L.color[2] = color.z;
        float acc = 0.0f;
        for (int k = 0; k < L.n_faces; ++k) {
            acc = k == 0 ? w[0] : acc + w[k];
            hs.light_cdf[L.first_face + k] = acc;
        }
    }
}

}  // namespace

int32_t build_host_scene(const rm_scene_desc *d, HostScene &hs) {
    if (!d) return fail(RM_ERR_INVALID, "rm_scene_create: desc is NULL");
    if (d->n_tris <= 0) return fail(RM_ERR_EMPTY, "rm_scene_create: scene has no triangles");
    if (!d->verts || !d->material) return fail(RM_ERR_INVALID, "rm_scene_create: verts/material are required");
    if (d->n_tris >= (int64_t(1) << 27)) return fail(RM_ERR_UNSUPPORTED, "rm_scene_create: at most 2^27-1 triangles");
    if (d->n_materials <= 0 || !d->materials) return fail(RM_ERR_INVALID, "rm_scene_create: no materials");
    const int64_t n = d->n_tris;
    for (int64_t i = 0; i < n; ++i)
        if (d->material[i] < 0 || d->material[i] >= d->n_materials)
            return fail(RM_ERR_INVALID, "rm_scene_create: material id %d of triangle %lld out of range", d->material[i], (long long)i);
    for (int32_t t = 0; t < d->n_textures; ++t) {
        const rm_texture &X = d->textures[t];
        if (X.width <= 0 || X.height <= 0 || (!X.rgb && !X.bytes))
            return fail(RM_ERR_INVALID, "rm_scene_create: texture %d is empty", t);
        if (X.bytes && X.channels != 3 && X.channels != 4)
            return fail(RM_ERR_INVALID, "rm_scene_create: 8-bit texture %d has %d channels (3 or 4)", t, X.channels);
    }
    for (int32_t m = 0; m < d->n_materials; ++m) {
        const rm_material &M = d->materials[m];
        int32_t t = M.albedo_tex;
        if (t >= d->n_textures || t < -1) return fail(RM_ERR_INVALID, "rm_scene_create: material %d texture %d out of range", m, t);
        for (int32_t q : {M.normal_tex, M.emissive_tex, M.specular_tex})
            if (q != -1 && (q < 0 || q >= d->n_textures || !d->textures[q].bytes))
                return fail(RM_ERR_INVALID, "rm_scene_create: material %d map %d is not an 8-bit texture of the scene", m, q);
    }
    const bool has_sky = d->sky_rgb && d->sky_width > 0 && d->sky_height > 0;

    hs.n = n;
    // ---- BVH (src/bvh.cpp:18-54)
    Builder B;
    B.n = n;
    B.verts = d->verts;
    B.rec.resize(n);
    parallel_ranges(n, [&B](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            V3 c = (B.vert((int32_t)i, 0) + B.vert((int32_t)i, 1) + B.vert((int32_t)i, 2)) / 3.0f;  // Face::center
            B.rec[i] = Builder::Rec{{c.x, c.y, c.z}, (int32_t)i};
        }
    });
    B.size = node_count(1, n) + 1;
    B.lo.assign(B.size * 3, 0.0f);
    B.hi.assign(B.size * 3, 0.0f);
    B.fl.assign(B.size, 0);
    B.fr.assign(B.size, 0);
    B.kind.assign(B.size, 0);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int par_depth = 0;
    while ((1u << par_depth) < hw && par_depth < 6) ++par_depth;
    B.dfs(1, 0, n, par_depth);
    hs.n_nodes = B.size;
    hs.perm.resize(n);
    for (int64_t i = 0; i < n; ++i) hs.perm[i] = B.rec[i].id;
    HostArray<Builder::Rec>().swap(B.rec);

    // ---- flatten: pair records per interior node
    int32_t depth = 0;
    for (int64_t u = 1; u < B.size; ++u)
        if (B.kind[u]) {
            int32_t lv = 0;
            for (int64_t x = u; x > 1; x >>= 1) ++lv;
            depth = std::max(depth, lv + 1);
        }
    hs.depth = depth;
    hs.pairs.assign(B.size / 2 + 1, PairNode{});
    for (int64_t u = 1; u < B.size; ++u) {
        if (B.kind[u] != 1) continue;
        PairNode &P = hs.pairs[u];
        V3 l0 = B.get_lo(2 * u), h0 = B.get_hi(2 * u), l1 = B.get_lo(2 * u + 1), h1 = B.get_hi(2 * u + 1);
        float a[4] = {l0.x, l0.y, l0.z, h0.x}, b[4] = {h0.y, h0.z, l1.x, l1.y}, c[4] = {l1.z, h1.x, h1.y, h1.z};
        memcpy(P.a, a, 16); memcpy(P.b, b, 16); memcpy(P.c, c, 16);
        for (int k = 0; k < 2; ++k) {
            int64_t ch = 2 * u + k;
            P.leaf[k] = B.kind[ch] == 2 ? leaf_code(B.fl[ch], B.fr[ch] - B.fl[ch]) : 0u;
        }
    }
    hs.root_leaf = B.kind[1] == 2 ? leaf_code(B.fl[1], B.fr[1] - B.fl[1]) : 0u;
    for (int k = 0; k < 3; ++k) {
        hs.root_lo[k] = B.lo[3 + k];
        hs.root_hi[k] = B.hi[3 + k];
    }

    // ---- triangle records in slot order (every slot independent)
    hs.tris.resize(n);
    hs.shade.resize(n);
    auto tri_range = [&](int64_t s0, int64_t s1) {
        for (int64_t s = s0; s < s1; ++s) {
            int32_t f = hs.perm[s];
            V3 v0 = B.vert(f, 0), v1 = B.vert(f, 1), v2 = B.vert(f, 2);
            V3 e1 = v1 - v0, e2 = v2 - v0;
            TriRec &T = hs.tris[s];
            T.p[0] = v0.x; T.p[1] = v0.y; T.p[2] = v0.z;
            T.p[3] = reject_threshold(sqrtf(dot(e1, e1)));
            T.e1[0] = e1.x; T.e1[1] = e1.y; T.e1[2] = e1.z;
            memcpy(&T.e1[3], &f, 4);
            T.e2[0] = e2.x; T.e2[1] = e2.y; T.e2[2] = e2.z; T.e2[3] = 0.0f;
            TriShade &S = hs.shade[s];
            memset(&S, 0, sizeof(S));
            for (int k = 0; k < 3; ++k) {
                V3 v = B.vert(f, k);
                S.v[k][0] = v.x; S.v[k][1] = v.y; S.v[k][2] = v.z;
                if (d->normals) {
                    const float *q = d->normals + (int64_t)f * 9 + k * 3;
                    S.n[k][0] = q[0]; S.n[k][1] = q[1]; S.n[k][2] = q[2];
                } else {
                    V3 g = normalize(cross(e1, e2));
                    S.n[k][0] = g.x; S.n[k][1] = g.y; S.n[k][2] = g.z;
                }
                if (d->uvs) {
                    S.uv[k][0] = d->uvs[(int64_t)f * 6 + k * 2];
                    S.uv[k][1] = d->uvs[(int64_t)f * 6 + k * 2 + 1];
                }
            }
            S.material = d->material[f];
            S.mesh = d->mesh ? d->mesh[f] : 0;
        }
    };
    parallel_ranges(n, tri_range);

    // ---- materials and textures
    hs.mats.resize(d->n_materials);
    for (int32_t m = 0; m < d->n_materials; ++m) hs.mats[m] = mat_rec(d->materials[m], has_sky);
    for (int32_t t = 0; t < d->n_textures; ++t) {
        const rm_texture &X = d->textures[t];
        TexRec R{};
        R.width = X.width;
        R.height = X.height;
        R.levels = 8;
        R.alpha = -1;
        if (X.bytes) {
            // ImageData::generateMipmaps (src/material.cpp:113-148) on the
            // reference's bytes: integer 2x2 sums / 4 in the channel layout
            const int ch = X.channels;
            R.channels = ch;
            hs.maps8 = 1;
            R.offset[0] = (int64_t)hs.bytes.size();
            R.level_bytes[0] = (int64_t)X.width * X.height * ch;
            hs.bytes.insert(hs.bytes.end(), X.bytes, X.bytes + R.level_bytes[0]);
            for (int lv = 1; lv < 8; ++lv) {
                int pw = X.width >> (lv - 1), ph = X.height >> (lv - 1);
                int cw = X.width >> lv, chh = X.height >> lv;
                if (cw == 0 || chh == 0) { R.levels = lv; break; }
                R.offset[lv] = (int64_t)hs.bytes.size();
                R.level_bytes[lv] = (int64_t)cw * chh * ch;
                hs.bytes.resize(hs.bytes.size() + (size_t)R.level_bytes[lv]);
                const uint8_t *prev = hs.bytes.data() + R.offset[lv - 1];
                uint8_t *cur = hs.bytes.data() + R.offset[lv];
                for (int y = 0; y < chh; ++y)
                    for (int x = 0; x < cw; ++x)
                        for (int c = 0; c < ch; ++c) {
                            int sum = 0;
                            for (int dy = 0; dy < 2; ++dy)
                                for (int dx = 0; dx < 2; ++dx) {
                                    int px = (x * 2 + dx) % pw, py = (y * 2 + dy) % ph;
                                    sum += prev[((int64_t)py * pw + px) * ch + c];
                                }
                            cur[((int64_t)y * cw + x) * ch + c] = (uint8_t)(sum / 4);
                        }
            }
            hs.texs.push_back(R);
            continue;
        }
        // ImageData::generateMipmaps (src/material.cpp:113-148) on fp32 texels:
        // up to 8 levels, 2x2 box average with wrap, stop when a side reaches 0.
        R.offset[0] = (int64_t)hs.texels.size();
        hs.texels.insert(hs.texels.end(), X.rgb, X.rgb + (int64_t)X.width * X.height * 3);
        for (int lv = 1; lv < 8; ++lv) {
            int pw = X.width >> (lv - 1), ph = X.height >> (lv - 1);
            int cw = X.width >> lv, ch = X.height >> lv;
            if (cw == 0 || ch == 0) { R.levels = lv; break; }
            R.offset[lv] = (int64_t)hs.texels.size();
            hs.texels.resize(hs.texels.size() + (size_t)cw * ch * 3);
            const float *prev = hs.texels.data() + R.offset[lv - 1];
            float *cur = hs.texels.data() + R.offset[lv];
            for (int y = 0; y < ch; ++y)
                for (int x = 0; x < cw; ++x)
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
        // TransparentTest reads level 0 only (duv = NaN), so the alpha plane
        // has no mip chain.
        R.alpha = -1;
        if (X.alpha) {
            R.alpha = (int64_t)hs.texels.size();
            hs.texels.insert(hs.texels.end(), X.alpha, X.alpha + (int64_t)X.width * X.height);
        }
        hs.texs.push_back(R);
    }
    // Material::hasFullyTransparentPart (src/material.cpp:330-333): an alpha
    // value below 1 (an 8-bit alpha below 255) anywhere in the albedo texture;
    // for the reference's bytes, every 4th byte from the 4th
    // (ImageData::hasTransparentPart, src/material.cpp:102-107).
    for (MatRec &M : hs.mats) {
        if (M.tex >= 0 && d->textures[M.tex].bytes) {
            const TexRec &T = hs.texs[M.tex];
            for (int64_t i = 3; i < T.level_bytes[0] && !M.cutout; i += 4)
                M.cutout = hs.bytes[T.offset[0] + i] < 255 ? 1 : 0;
            hs.cutout |= M.cutout;
            continue;
        }
        if (M.tex < 0 || !d->textures[M.tex].alpha) continue;
        const rm_texture &X = d->textures[M.tex];
        const int64_t np = (int64_t)X.width * X.height;
        for (int64_t i = 0; i < np && !M.cutout; ++i) M.cutout = X.alpha[i] < 1.0f ? 1 : 0;
        hs.cutout |= M.cutout;
    }

    // ---- light objects (Model::checkLightObject, src/model.cpp:23-82), per
    // mesh in mesh order, only without a sky (src/model.cpp:121).
    if (!has_sky) {
        int64_t i = 0;
        while (i < n) {
            int64_t j = i;
            int32_t mesh = d->mesh ? d->mesh[i] : 0;
            while (j < n && (d->mesh ? d->mesh[j] : 0) == mesh) ++j;
            // a mesh has one material in the reference (aiMesh::mMaterialIndex);
            // it is a light object with an emissive map (src/model.cpp:121) or,
            // as an extension, a constant emission
            const MatRec &M = hs.mats[d->material[i]];
            bool emissive = M.etex >= 0 || M.emission[0] != 0.0f || M.emission[1] != 0.0f || M.emission[2] != 0.0f;
            if (emissive) {
                LightRec L{};
                std::vector<int64_t> faces;
                std::vector<V3> avgs;
                V3 em{M.emission[0], M.emission[1], M.emission[2]};
                for (int64_t f = i; f < j; ++f) {
                    V3 avg = emissive_avg(em);
                    if (M.etex >= 0) {
                        TriShade F{};
                        for (int k = 0; k < 3 && d->uvs; ++k) {
                            F.uv[k][0] = d->uvs[f * 6 + k * 2];
                            F.uv[k][1] = d->uvs[f * 6 + k * 2 + 1];
                        }
                        avg = emissive_avg_tex(hs, hs.texs[M.etex], F);
                    }
                    float Clum = dot(avg, v3(0.3f, 0.6f, 0.1f));
                    V3 v0 = B.vert((int32_t)f, 0), v1 = B.vert((int32_t)f, 1), v2 = B.vert((int32_t)f, 2);
                    float area = length(cross(v1 - v0, v2 - v0)) / 2.0f;
                    if (Clum < EPS || area < EPS) continue;
                    faces.push_back(f);
                    avgs.push_back(avg);
                }
                if (!faces.empty()) {
                    // the faces only; power, colour, centre and face CDF come from light_params
                    L.first_face = (int32_t)hs.light_faces.size();
                    L.n_faces = (int32_t)faces.size();
                    for (size_t k = 0; k < faces.size(); ++k) {
                        hs.light_cdf.push_back(0.0f);
                        hs.light_faces.push_back(TriShade{});
                        hs.light_face_avg.push_back(avgs[k]);
                    }
                    // fill copies from original-order data
                    for (size_t k = 0; k < faces.size(); ++k) {
                        int32_t f = (int32_t)faces[k];
                        TriShade &S = hs.light_faces[L.first_face + k];
                        memset(&S, 0, sizeof(S));
                        V3 e1 = B.vert(f, 1) - B.vert(f, 0), e2 = B.vert(f, 2) - B.vert(f, 0);
                        for (int kk = 0; kk < 3; ++kk) {
                            V3 v = B.vert(f, kk);
                            S.v[kk][0] = v.x; S.v[kk][1] = v.y; S.v[kk][2] = v.z;
                            if (d->normals) {
                                const float *q = d->normals + (int64_t)f * 9 + kk * 3;
                                S.n[kk][0] = q[0]; S.n[kk][1] = q[1]; S.n[kk][2] = q[2];
                            } else {
                                V3 g = normalize(cross(e1, e2));
                                S.n[kk][0] = g.x; S.n[kk][1] = g.y; S.n[kk][2] = g.z;
                            }
                            if (d->uvs) {
                                S.uv[kk][0] = d->uvs[(int64_t)f * 6 + kk * 2];
                                S.uv[kk][1] = d->uvs[(int64_t)f * 6 + kk * 2 + 1];
                            }
                        }
                        S.material = d->material[f];
                        S.mesh = mesh;
                    }
                    hs.lights.push_back(L);
                }
            }
            i = j;
        }
        light_params(hs);
    }

    // ---- sky (SkyBox::Init, src/component.cpp:54-67)
    if (has_sky) {
        int w = d->sky_width, h = d->sky_height;
        hs.sky_w = w;
        hs.sky_h = h;
        hs.sky.assign(d->sky_rgb, d->sky_rgb + (int64_t)w * h * 3);
        hs.sky_cdf.resize((int64_t)w * h);
        float acc = 0.0f;
        for (int v = 0; v < h; ++v)
            for (int u = 0; u < w; ++u) {
                float phi = PI_F * (float(v) + 0.5f) / float(h);
                float area = sinf(phi) * 2.0f * PI_F / float(w * h);
                int64_t id = (int64_t)v * w + u;
                V3 c{hs.sky[id * 3], hs.sky[id * 3 + 1], hs.sky[id * 3 + 2]};
                c = c * area;
                hs.sky[id * 3] = c.x; hs.sky[id * 3 + 1] = c.y; hs.sky[id * 3 + 2] = c.z;
                float Clum = dot(c, v3(0.3f, 0.6f, 0.1f));
                acc = id == 0 ? Clum : acc + Clum;
                hs.sky_cdf[id] = acc;
            }
        // guide[b] = lower_bound(cdf, total * b / G): a query r with
        // total*(b-1)/G <= r <= total*(b+2)/G has its lower_bound inside
        // [guide[b-1], guide[b+2]] (lower_bound is monotone in r)
        const int64_t n = (int64_t)w * h;
        const double tot = hs.sky_cdf[n - 1];
        hs.sky_guide.resize(SKY_GUIDE + 1);
        int64_t i = 0;
        for (int b = 0; b <= SKY_GUIDE; ++b) {
            double t = tot * (double)b / (double)SKY_GUIDE;
            while (i < n && (double)hs.sky_cdf[i] < t) ++i;
            hs.sky_guide[b] = (int32_t)i;
        }
        hs.sky_guide[SKY_GUIDE] = (int32_t)n;
    }
    return RM_OK;
}

}  // namespace rm

using namespace rm;

extern "C" int32_t rm_scene_create(const rm_scene_desc *desc, rm_scene **out) {
    if (!out) return fail(RM_ERR_INVALID, "rm_scene_create: out is NULL");
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    rm_scene *s = new (std::nothrow) rm_scene();
    if (!s) return fail(RM_ERR_OOM, "rm_scene_create: host allocation");
    int32_t st = build_host_scene(desc, s->host);
    if (st) { delete s; return st; }
    auto t1 = std::chrono::steady_clock::now();
    st = ensure_device();
    if (st) { delete s; return st; }
    HostScene &h = s->host;
#define UP(buf, vec) \
    if ((st = s->buf.upload(h.vec.data(), h.vec.size() * sizeof(h.vec[0])))) { delete s; return st; }
    UP(nodes, pairs);
    UP(tris, tris);
    UP(shade, shade);
    UP(mats, mats);
    UP(texs, texs);
    UP(texels, texels);
    UP(bytes, bytes);
    UP(lights, lights);
    UP(light_faces, light_faces);
    UP(light_cdf, light_cdf);
    UP(sky, sky);
    UP(sky_cdf, sky_cdf);
    UP(sky_guide, sky_guide);
#undef UP
    auto t2 = std::chrono::steady_clock::now();
    SceneView &v = s->view;
    v.nodes = s->nodes.as<PairNode>();
    v.tris = s->tris.as<TriRec>();
    v.shade = s->shade.as<TriShade>();
    v.mats = s->mats.as<MatRec>();
    v.texs = s->texs.as<TexRec>();
    v.texels = s->texels.as<float>();
    v.bytes = s->bytes.as<uint8_t>();
    v.cutout = h.cutout;
    v.lights = s->lights.as<LightRec>();
    v.light_faces = s->light_faces.as<TriShade>();
    v.light_cdf = s->light_cdf.as<float>();
    v.sky = h.sky.empty() ? nullptr : s->sky.as<float>();
    v.sky_cdf = s->sky_cdf.as<float>();
    v.sky_guide = s->sky_guide.as<int32_t>();
    v.n_tris = h.n;
    for (int k = 0; k < 3; ++k) {
        v.lo[k] = h.root_lo[k];
        v.hi[k] = h.root_hi[k];
    }
    v.root_leaf = (int32_t)h.root_leaf;
    v.depth = h.depth;
    v.n_lights = (int32_t)h.lights.size();
    v.sky_w = h.sky_w;
    v.sky_h = h.sky_h;
    v.n_mats = (int32_t)h.mats.size();
    rm_scene_info &I = s->info;
    I.n_tris = h.n;
    I.n_nodes = h.n_nodes;
    I.depth = h.depth;
    I.n_light_objects = (int32_t)h.lights.size();
    I.n_light_faces = (int32_t)h.light_faces.size();
    I.has_sky = h.sky.empty() ? 0 : 1;
    I.device_bytes = (int64_t)(s->nodes.bytes + s->tris.bytes + s->shade.bytes + s->mats.bytes + s->texs.bytes +
                               s->texels.bytes + s->bytes.bytes + s->lights.bytes + s->light_faces.bytes +
                               s->light_cdf.bytes +
                               s->sky.bytes + s->sky_cdf.bytes + s->sky_guide.bytes);
    I.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    I.upload_seconds = std::chrono::duration<double>(t2 - t1).count();
    // keep only what host calls need
    std::vector<PairNode>().swap(h.pairs);
    HostArray<TriRec>().swap(h.tris);
    HostArray<TriShade>().swap(h.shade);
    std::vector<float>().swap(h.texels);
    std::vector<uint8_t>().swap(h.bytes);
    *out = s;
    return RM_OK;
}

extern "C" int32_t rm_bvh_build(const float *verts, int64_t n, int32_t *perm, int64_t *n_nodes, int32_t *depth) {
    if (!verts || !perm) return fail(RM_ERR_INVALID, "rm_bvh_build: NULL argument");
    if (n <= 0) return fail(RM_ERR_EMPTY, "rm_bvh_build: scene has no triangles");
    std::vector<int32_t> mat((size_t)n, 0);
    rm_material m{};
    m.roughness = 0.8f;
    m.opacity = 1.0f;
    m.ior = 1.0f;
    m.albedo_tex = m.normal_tex = m.emissive_tex = m.specular_tex = -1;
    rm_scene_desc d{};
    d.n_tris = n;
    d.verts = verts;
    d.material = mat.data();
    d.n_materials = 1;
    d.materials = &m;
    HostScene hs;
    int32_t st = build_host_scene(&d, hs);
    if (st) return st;
    memcpy(perm, hs.perm.data(), (size_t)n * sizeof(int32_t));
    if (n_nodes) *n_nodes = hs.n_nodes;
    if (depth) *depth = hs.depth;
    return RM_OK;
}

extern "C" int32_t rm_scene_destroy(rm_scene *scene) {
    delete scene;
    return RM_OK;
}

extern "C" int32_t rm_scene_get_info(const rm_scene *scene, rm_scene_info *info) {
    if (!scene || !info) return fail(RM_ERR_INVALID, "rm_scene_get_info: NULL argument");
    *info = scene->info;
    return RM_OK;
}

extern "C" int32_t rm_scene_get_perm(const rm_scene *scene, int32_t *perm) {
    if (!scene || !perm) return fail(RM_ERR_INVALID, "rm_scene_get_perm: NULL argument");
    memcpy(perm, scene->host.perm.data(), scene->host.perm.size() * sizeof(int32_t));
    return RM_OK;
}

extern "C" int32_t rm_scene_update(rm_scene *scene, const rm_material *materials, const float *texels, void *stream) {
    if (!scene) return fail(RM_ERR_INVALID, "rm_scene_update: scene is NULL");
    HostScene &h = scene->host;
    hipStream_t st = (hipStream_t)stream;
    const bool has_sky = !h.sky.empty();
    if (materials) {
        std::vector<MatRec> next(h.mats.size());
        for (size_t m = 0; m < h.mats.size(); ++m) {
            next[m] = mat_rec(materials[m], has_sky);
            const MatRec &o = h.mats[m], &n = next[m];
            if (n.tex != o.tex || n.ntex != o.ntex || n.etex != o.etex || n.stex != o.stex)
                return fail(RM_ERR_INVALID, "rm_scene_update: material %zu changes a texture binding", m);
            next[m].cutout = o.cutout;  // the alpha planes do not change
            bool e0 = o.emission[0] != 0.0f || o.emission[1] != 0.0f || o.emission[2] != 0.0f;
            bool e1 = n.emission[0] != 0.0f || n.emission[1] != 0.0f || n.emission[2] != 0.0f;
            if (e0 != e1) return fail(RM_ERR_INVALID, "rm_scene_update: material %zu changes whether it emits", m);
            if (e1 && emissive_lum(V3{n.emission[0], n.emission[1], n.emission[2]}) < EPS)
                return fail(RM_ERR_INVALID, "rm_scene_update: emission of material %zu is below the light threshold", m);
        }
        h.mats.swap(next);
        light_params(h);
        RM_HIP(hipMemcpy(scene->mats.p, h.mats.data(), h.mats.size() * sizeof(MatRec), hipMemcpyHostToDevice));
        if (!h.lights.empty()) {
            RM_HIP(hipMemcpy(scene->lights.p, h.lights.data(), h.lights.size() * sizeof(LightRec),
                             hipMemcpyHostToDevice));
            RM_HIP(hipMemcpy(scene->light_cdf.p, h.light_cdf.data(), h.light_cdf.size() * sizeof(float),
                             hipMemcpyHostToDevice));
        }
    }
    if (texels) {
        for (const TexRec &T : h.texs)
            if (T.channels)
                return fail(RM_ERR_UNSUPPORTED, "rm_scene_update: texels of a scene with 8-bit textures");
        int64_t off = 0;
        for (size_t t = 0; t < h.texs.size(); ++t) {
            const TexRec &T = h.texs[t];
            const int64_t n = (int64_t)T.width * T.height * 3;
            RM_HIP(hipMemcpyAsync(scene->texels.as<float>() + T.offset[0], texels + off, n * sizeof(float),
                                  hipMemcpyDeviceToDevice, st));
            int32_t rc = launch_mip_build(scene, (int)t, st);
            if (rc) return rc;
            off += n;
        }
        RM_HIP(hipStreamSynchronize(st));
    }
    ++scene->version;
    return RM_OK;
}

// ORACLE -- test infrastructure only (see oracle.h).
//
// C entry points that let the tests and the CPU-baseline leg drive the
// restatement: scene assembly (what Model::Model does after assimp,
// src/model.cpp:88-123,172-215, incl. light objects src/model.cpp:23-82 and
// the sky, src/component.cpp:54-67), threaded ray batches through
// BVH::rayHit / Model::rayHit_test, single-primitive probes, and rendering.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include "../include/raym0nade_hip.h"
#include "oracle.h"

using namespace ref;

struct oracle_scene {
    Model model;
    std::vector<int32_t> perm;
    std::vector<int32_t> mat_tex;                 // texture index per material (-1: none)
    std::vector<std::pair<int, int>> tex_dims;    // (width, height) per texture of the description
};

namespace {

// Model::checkLightObject (src/model.cpp:44-82) for one mesh
void checkLightObject(Model &M, const std::vector<Face> &meshFaces, const Material &material) {
    LightObject lo;
    vec3 color(0.0f);
    std::vector<float> faceWeights;
    for (const Face &face : meshFaces) {
        // getAverageEmissiveColor (src/model.cpp:23-42)
        vec3 averageColor(0.0f);
        const int gridResolution = 8;
        for (int i = 0; i < gridResolution; ++i)
            for (int j = 0; j < gridResolution; ++j) {
                float a = static_cast<float>(i) / (gridResolution - 1);
                float b = static_cast<float>(j) / (gridResolution - 1);
                if (a + b > 1.0f) {
                    a = 1.0f - a;
                    b = 1.0f - b;
                }
                float c = 1.0f - a - b;
                vec2 uv = a * face.data[0]->uv + b * face.data[1]->uv + c * face.data[2]->uv;
                averageColor += material.getEmissiveColor(uv[0], uv[1], NAN);
            }
        vec3 textureColor = averageColor / static_cast<float>(gridResolution * gridResolution);
        float Clum = dot(textureColor, RGB_Weight),
              area = length(cross(face.v[1] - face.v[0], face.v[2] - face.v[0])) / 2.0f;
        if (Clum < eps_zero || area < eps_zero) continue;
        color += textureColor * area;
        float power = area * Clum;
        lo.faces.emplace_back(face);
        faceWeights.emplace_back(power);
    }
    if (lo.faces.empty()) return;
    lo.color = color / dot(color, RGB_Weight);
    for (size_t j = 0; j < lo.faces.size(); j++) {
        lo.power += faceWeights[j];
        lo.center += lo.faces[j].center() * faceWeights[j];
    }
    lo.faceDist.Init(faceWeights);
    lo.center /= lo.power;
    M.lightObjects.push_back(std::move(lo));
}

template <class F>
void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 1024) {
        for (int64_t i = 0; i < n; ++i) f(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    const int64_t chunk = 1024;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            while (true) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; ++i) f(i, t);
            }
        });
    for (auto &th : pool) th.join();
}

Ray load_ray(const float *rays, int64_t i) {
    const float *p = rays + i * 6;
    return Ray{vec3(p[0], p[1], p[2]), vec3(p[3], p[4], p[5])};
}

}  // namespace

extern "C" {

// Returns NULL on an invalid or empty scene.
oracle_scene *oracle_scene_create(const rm_scene_desc *d) {
    if (!d || d->n_tris <= 0 || !d->verts || !d->material || d->n_materials <= 0) return nullptr;
    oracle_scene *S = new oracle_scene();
    Model &M = S->model;
    const int64_t n = d->n_tris;
    const bool has_sky = d->sky_rgb && d->sky_width > 0 && d->sky_height > 0;
    M.materials.resize(d->n_materials);
    for (int t = 0; t < d->n_textures; ++t) S->tex_dims.emplace_back(d->textures[t].width, d->textures[t].height);
    S->mat_tex.assign(d->n_materials, -1);
    for (int m = 0; m < d->n_materials; ++m) {
        const rm_material &R = d->materials[m];
        Material &X = M.materials[m];
        if (R.albedo_tex >= 0 && R.albedo_tex < d->n_textures) S->mat_tex[m] = R.albedo_tex;
        X.id = m;
        X.baseColor = vec3(R.base_color[0], R.base_color[1], R.base_color[2]);
        X.emission = has_sky ? vec3(0.0f) : vec3(R.emission[0], R.emission[1], R.emission[2]);
        X.transmittingColor = vec3(R.transmitting_color[0], R.transmitting_color[1], R.transmitting_color[2]);
        X.roughness = R.roughness;
        X.metallic = R.metallic;
        X.opacity = R.opacity;
        X.ior = R.ior;
        auto load8 = [&](int32_t t, ImageData8 &img) {
            if (t < 0 || t >= d->n_textures || !d->textures[t].bytes) return;
            const rm_texture &T = d->textures[t];
            img.width = T.width;
            img.height = T.height;
            img.channels = T.channels;
            img.data[0].assign(T.bytes, T.bytes + (size_t)T.width * T.height * T.channels);
            img.generateMipmaps();
        };
        if (R.albedo_tex >= 0 && R.albedo_tex < d->n_textures) {
            const rm_texture &T = d->textures[R.albedo_tex];
            if (T.bytes) {
                load8(R.albedo_tex, X.diffuse8);
                X.hasFullyTransparentPart = X.diffuse8.hasTransparentPart();
            } else {
                X.diffuse.width = T.width;
                X.diffuse.height = T.height;
                X.diffuse.data[0].assign(T.rgb, T.rgb + (size_t)T.width * T.height * 3);
                X.diffuse.generateMipmaps();
                if (T.alpha) X.diffuse.alpha.assign(T.alpha, T.alpha + (size_t)T.width * T.height);
                X.hasFullyTransparentPart = X.diffuse.hasTransparentPart();
            }
        }
        load8(R.normal_tex, X.normals8);
        if (!has_sky) load8(R.emissive_tex, X.emissive8);  // src/model.cpp:140: no emissive maps with a sky
        load8(R.specular_tex, X.specular8);
    }
    M.vertexDatas.reserve(n * 3);
    M.faces.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        Face f;
        const float *v = d->verts + i * 9;
        for (int k = 0; k < 3; ++k) f.v[k] = vec3(v[k * 3], v[k * 3 + 1], v[k * 3 + 2]);
        vec3 g = normalize(cross(f.v[1] - f.v[0], f.v[2] - f.v[0]));
        for (int k = 0; k < 3; ++k) {
            VertexData vd;
            vd.uv = d->uvs ? vec2(d->uvs[i * 6 + k * 2], d->uvs[i * 6 + k * 2 + 1]) : vec2(0.0f);
            vd.normal = d->normals ? vec3(d->normals[i * 9 + k * 3], d->normals[i * 9 + k * 3 + 1],
                                          d->normals[i * 9 + k * 3 + 2])
                                   : g;
            M.vertexDatas.push_back(vd);
            f.data[k] = &M.vertexDatas.back();
        }
        f.material = &M.materials[d->material[i]];
        f.id = (int32_t)i;
        M.faces.push_back(f);
    }
    // light objects per mesh, before the BVH permutes `faces` (src/model.cpp:121,214)
    if (!has_sky) {
        int64_t i = 0;
        while (i < n) {
            int64_t j = i;
            int32_t mesh = d->mesh ? d->mesh[i] : 0;
            while (j < n && (d->mesh ? d->mesh[j] : 0) == mesh) ++j;
            const Material &mat = M.materials[d->material[i]];
            if (mat.hasEmission()) {
                std::vector<Face> mf(M.faces.begin() + i, M.faces.begin() + j);
                checkLightObject(M, mf, mat);
            }
            i = j;
        }
    }
    if (has_sky) {
        SkyBox &K = M.skyMap;
        K.width = d->sky_width;
        K.height = d->sky_height;
        K.data.resize((size_t)K.width * K.height);
        for (size_t p = 0; p < K.data.size(); ++p)
            K.data[p] = vec3(d->sky_rgb[p * 3], d->sky_rgb[p * 3 + 1], d->sky_rgb[p * 3 + 2]);
        K.Init();
    }
    M.bvh.build(M.faces);
    S->perm.resize(n);
    for (int64_t s = 0; s < n; ++s) S->perm[s] = M.faces[s].id;
    return S;
}

void oracle_scene_destroy(oracle_scene *S) { delete S; }

int64_t oracle_scene_node_count(const oracle_scene *S) { return S->model.bvh.size; }
int32_t oracle_scene_light_count(const oracle_scene *S) { return (int32_t)S->model.lightObjects.size(); }
void oracle_scene_get_perm(const oracle_scene *S, int32_t *perm) {
    memcpy(perm, S->perm.data(), S->perm.size() * sizeof(int32_t));
}
// Node u: box (6 floats) and face range; used by the BVH invariant tests.
void oracle_scene_get_nodes(const oracle_scene *S, float *boxes, int32_t *ranges) {
    const BVH &B = S->model.bvh;
    for (int u = 0; u < B.size; ++u) {
        const BVH_Node &N = B.node[u];
        float b[6] = {N.box.v0.x, N.box.v0.y, N.box.v0.z, N.box.v1.x, N.box.v1.y, N.box.v1.z};
        memcpy(boxes + (size_t)u * 6, b, sizeof(b));
        ranges[u * 2] = N.faceL;
        ranges[u * 2 + 1] = N.faceR;
    }
}
// Light object k: center, color, power, face count, first face CDF entries.
void oracle_scene_get_light(const oracle_scene *S, int32_t k, float *out7, float *cdf) {
    const LightObject &L = S->model.lightObjects[k];
    float v[7] = {L.center.x, L.center.y, L.center.z, L.color.x, L.color.y, L.color.z, L.power};
    memcpy(out7, v, sizeof(v));
    if (cdf) memcpy(cdf, L.faceDist.prefixSums.data(), L.faceDist.prefixSums.size() * sizeof(float));
}

// Closest hit for a batch (Model::rayHit's loop from HitRecord(tmin, tmax)).  counts:
// uint64[4] = interior visits, leaf visits, box tests, triangle tests.
void oracle_trace_closest(const oracle_scene *S, const float *rays, int64_t n, float tmin, float tmax, float *t_out,
                          int32_t *id_out, int32_t threads, uint64_t *counts) {
    std::vector<Counters> per(std::max(1, threads));
    parallel_for(n, threads, [&](int64_t i, int t) {
        HitRecord hit = S->model.rayHitFrom(load_ray(rays, i), tmin, tmax, counts ? &per[t] : nullptr);
        t_out[i] = hit.t_max;
        id_out[i] = hit.t_max != tmax && hit.face ? hit.face->id : -1;
    });
    if (counts) {
        memset(counts, 0, 4 * sizeof(uint64_t));
        for (auto &c : per) { counts[0] += c.interior; counts[1] += c.leaves; counts[2] += c.boxes; counts[3] += c.tris; }
    }
}

// Model::rayHit_test for a batch with t_min = tmin (the reference uses 1e-4).
void oracle_trace_any(const oracle_scene *S, const float *rays, int64_t n, float tmin, float aim_depth,
                      const float *aim, uint8_t *occ, int32_t threads, uint64_t *counts) {
    std::vector<Counters> per(std::max(1, threads));
    parallel_for(n, threads, [&](int64_t i, int t) {
        float a = aim ? aim[i] : aim_depth;
        occ[i] = S->model.rayHitTestFrom(load_ray(rays, i), tmin, a, counts ? &per[t] : nullptr) ? 1 : 0;
    });
    if (counts) {
        memset(counts, 0, 4 * sizeof(uint64_t));
        for (auto &c : per) { counts[0] += c.interior; counts[1] += c.leaves; counts[2] += c.boxes; counts[3] += c.tris; }
    }
}

// Brute force over all faces in original order with the same triangle test.
void oracle_brute_closest(const oracle_scene *S, const float *rays, int64_t n, float tmin, float tmax, float *t_out,
                          int32_t *id_out, int32_t threads) {
    const Model &M = S->model;
    std::vector<const Face *> byId(M.faces.size());
    for (const Face &f : M.faces) byId[f.id] = &f;
    parallel_for(n, threads, [&](int64_t i, int) {
        Ray r = load_ray(rays, i);
        float best = tmax;
        int32_t id = -1;
        for (const Face *f : byId) {
            float t = RayTriangleIntersection(r, f->v[0], f->v[1], f->v[2]);
            if (tmin < t && t < best) { best = t; id = f->id; }
        }
        t_out[i] = best;
        id_out[i] = id;
    });
}

// Single-primitive probes.
void oracle_ray_in_box(const float *ray6, const float *box6, float *tL, float *tR) {
    Ray r{vec3(ray6[0], ray6[1], ray6[2]), vec3(ray6[3], ray6[4], ray6[5])};
    Box b(vec3(box6[0], box6[1], box6[2]), vec3(box6[3], box6[4], box6[5]));
    rayInBox(r, b, *tL, *tR);
}
float oracle_ray_triangle(const float *ray6, const float *tri9) {
    Ray r{vec3(ray6[0], ray6[1], ray6[2]), vec3(ray6[3], ray6[4], ray6[5])};
    return RayTriangleIntersection(r, vec3(tri9[0], tri9[1], tri9[2]), vec3(tri9[3], tri9[4], tri9[5]),
                                   vec3(tri9[6], tri9[7], tri9[8]));
}
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox4x32_10(ctr, key, out); }
float oracle_gen(uint32_t seed, uint32_t tag, uint32_t pix, uint32_t smp, int32_t skip) {
    Generator g(seed, tag, pix, smp);
    for (int i = 0; i < skip; ++i) g();
    return g();
}

int32_t oracle_hw_threads(void) { return (int32_t)std::max(1u, std::thread::hardware_concurrency()); }

}  // extern "C"

// ----------------------------------------------------------------- render
namespace {
RenderArgs to_args(const rm_render_args *a) {
    RenderArgs r;
    r.position = vec3(a->position[0], a->position[1], a->position[2]);
    r.direction = vec3(a->direction[0], a->direction[1], a->direction[2]);
    r.up = vec3(a->up[0], a->up[1], a->up[2]);
    r.right = vec3(a->right[0], a->right[1], a->right[2]);
    r.accuracy = a->accuracy;
    r.focus = a->focus;
    r.CoC = a->coc;
    r.exposure = a->exposure;
    r.P_Direct = a->p_direct;
    r.width = a->width;
    r.height = a->height;
    r.spp = a->spp;
    r.threads = a->threads;
    r.seed = a->seed;
    r.max_depth = a->max_depth > 0 ? a->max_depth : 16;
    r.rr_min_depth = a->rr_min_depth > 0 ? a->rr_min_depth : 1;
    return r;
}
void put3(float *p, const vec3 &v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
vec3 get3(const float *p) { return vec3(p[0], p[1], p[2]); }

// One G-buffer row: position, shapeNormal, surfaceNormal, emission,
// baseColor, then (metallic, roughness, opacity).
constexpr int GB = 18;
void put_gbuf(float *g, const HitInfo &G) {
    put3(g + 0, G.position);
    put3(g + 3, G.shapeNormal);
    put3(g + 6, G.surfaceNormal);
    put3(g + 9, G.emission);
    put3(g + 12, G.baseColor);
    g[15] = G.metallic;
    g[16] = G.roughness;
    g[17] = G.opacity;
}
void get_gbuf(const float *g, HitInfo &G) {
    G.position = get3(g);
    G.shapeNormal = get3(g + 3);
    G.surfaceNormal = get3(g + 6);
    G.emission = get3(g + 9);
    G.baseColor = get3(g + 12);
    G.metallic = g[15];
    G.roughness = g[16];
    G.opacity = g[17];
}
std::vector<RadianceData> *photo_bufs(Photo &p, int b) {
    std::vector<RadianceData> *bufs[4] = {&p.radiance_Dd, &p.radiance_Ds, &p.radiance_Id, &p.radiance_Is};
    return bufs[b];
}
}  // namespace

extern "C" {

// Renders the rows listed in `rows` (all rows if NULL) with `threads` workers.
// gbuf: [h][w][18] = position, shapeNormal, surfaceNormal, emission,
// baseColor, (metallic, roughness, opacity).
// rad:  [4][h][w][4] = Dd, Ds, Id, Is as (rgb, Var).  Rows not rendered keep
// the defaults of a fresh Photo.
int32_t oracle_render(const oracle_scene *S, const rm_render_args *a, int32_t threads, const int32_t *rows,
                      int32_t n_rows, float *gbuf, float *rad) {
    if (!S || !a || a->width <= 0 || a->height <= 0 || a->max_depth > 16) return 1;
    RenderArgs args = to_args(a);
    Photo photo(args.width, args.height);
    photo.exposure = args.exposure;
    std::vector<int> rs;
    if (rows)
        for (int i = 0; i < n_rows; ++i) rs.push_back(rows[i]);
    render_multiThread(S->model, args, photo, threads, rows ? &rs : nullptr);
    const size_t np = (size_t)args.width * args.height;
    if (gbuf)
        for (size_t i = 0; i < np; ++i) put_gbuf(gbuf + i * GB, photo.Gbuffer[i]);
    if (rad)
        for (int b = 0; b < 4; ++b)
            for (size_t i = 0; i < np; ++i) {
                const RadianceData &R = (*photo_bufs(photo, b))[i];
                float *p = rad + ((size_t)b * np + i) * 4;
                put3(p, R.radiance);
                p[3] = R.Var;
            }
    return 0;
}

// Renders only the listed pixels (index y*width+x); outputs are compact:
// gbuf [n][18], rad [4][n][4] in list order.
int32_t oracle_render_pixels(const oracle_scene *S, const rm_render_args *a, int32_t threads, const int32_t *pix,
                             int32_t n, float *gbuf, float *rad) {
    if (!S || !a || a->width <= 0 || a->height <= 0 || a->max_depth > 16 || !pix) return 1;
    RenderArgs args = to_args(a);
    Photo photo(args.width, args.height);
    photo.exposure = args.exposure;
    std::vector<int> list(pix, pix + n);
    render_pixels(S->model, args, photo, threads, list);
    for (int k = 0; k < n; ++k) {
        if (gbuf) put_gbuf(gbuf + (size_t)k * GB, photo.Gbuffer[list[k]]);
        if (rad)
            for (int b = 0; b < 4; ++b) {
                const RadianceData &R = (*photo_bufs(photo, b))[list[k]];
                float *p = rad + ((size_t)b * n + k) * 4;
                put3(p, R.radiance);
                p[3] = R.Var;
            }
    }
    return 0;
}

namespace {
void load_photo(Photo &photo, const float *gbuf, const float *rad) {
    const size_t np = (size_t)photo.width * photo.height;
    for (size_t i = 0; i < np; ++i) get_gbuf(gbuf + i * GB, photo.Gbuffer[i]);
    for (int b = 0; b < 4; ++b)
        for (size_t i = 0; i < np; ++i) {
            const float *p = rad + ((size_t)b * np + i) * 4;
            (*photo_bufs(photo, b))[i].radiance = get3(p);
            (*photo_bufs(photo, b))[i].Var = p[3];
        }
}
}  // namespace

// Photo::postProcessing on given buffers (same layouts as oracle_render);
// focus, coc and cam (3 floats, may be NULL) are Photo's depth-of-field
// fields (src/render.cpp:666-668).
void oracle_postprocess(int32_t w, int32_t h, const float *gbuf, const float *rad, int32_t options, float exposure,
                        float focus, float coc, const float *cam, float *pixels) {
    Photo photo(w, h);
    photo.exposure = exposure;
    photo.focus = focus;
    photo.CoC = coc;
    if (cam) photo.cameraPosition = get3(cam);
    load_photo(photo, gbuf, rad);
    photo.postProcessing(options);
    const size_t np = (size_t)w * h;
    for (size_t i = 0; i < np; ++i) put3(pixels + i * 3, photo.pixelarray[i]);
}

// The denoise path on given buffers (layouts as oracle_render), in place on
// rad: steps & 1 runs Photo::spatialClamp, steps & 2 then Photo::filter
// (four threads, one per buffer, as the reference runs it); steps & 4 runs
// filterVar alone on each buffer.
void oracle_denoise(int32_t w, int32_t h, const float *gbuf, float *rad, int32_t steps) {
    Photo photo(w, h);
    load_photo(photo, gbuf, rad);
    if (steps & 1) photo.spatialClamp();
    if (steps & 2) photo.filter();
    if (steps & 4)
        for (int b = 0; b < 4; ++b) filter_var(*photo_bufs(photo, b), w, h);
    const size_t np = (size_t)w * h;
    for (int b = 0; b < 4; ++b)
        for (size_t i = 0; i < np; ++i) {
            float *p = rad + ((size_t)b * np + i) * 4;
            put3(p, (*photo_bufs(photo, b))[i].radiance);
            p[3] = (*photo_bufs(photo, b))[i].Var;
        }
}

// One a-trous pass (src/image.cpp:158-189) of one buffer, in place on rad
// [h][w][4], for the pass-level tests.
void oracle_filter_pass(int32_t w, int32_t h, const float *gbuf, float *rad, int32_t specular, int32_t step) {
    const size_t np = (size_t)w * h;
    std::vector<HitInfo> G(np);
    std::vector<RadianceData> R(np);
    for (size_t i = 0; i < np; ++i) {
        get_gbuf(gbuf + i * GB, G[i]);
        R[i].radiance = get3(rad + i * 4);
        R[i].Var = rad[i * 4 + 3];
    }
    filter_pass(R, G, w, h, specular != 0, step);
    for (size_t i = 0; i < np; ++i) {
        put3(rad + i * 4, R[i].radiance);
        rad[i * 4 + 3] = R[i].Var;
    }
}

// getWeight (src/image.cpp:109-156) of one neighbour: g* are G-buffer rows,
// l* are (rgb, Var).
float oracle_denoise_weight(const float *gp, const float *gq, const float *lp, const float *lq, int32_t specular) {
    HitInfo Gp, Gq;
    get_gbuf(gp, Gp);
    get_gbuf(gq, Gq);
    RadianceData Lp, Lq;
    Lp.radiance = get3(lp);
    Lp.Var = lp[3];
    Lq.radiance = get3(lq);
    Lq.Var = lq[3];
    return denoise_weight(Gp, Gq, Lp, Lq, specular != 0);
}

// accumulateInwardRadiance on one sample, for its invariant tests.
void oracle_accumulate(const float *base3, const float *bsdf3, const float *light3, float weight, float *d4, float *s4) {
    RadianceData d, s;
    d.radiance = get3(d4); d.Var = d4[3];
    s.radiance = get3(s4); s.Var = s4[3];
    accumulateInwardRadiance(get3(base3), LightSample(get3(bsdf3), get3(light3), weight), d, s);
    put3(d4, d.radiance); d4[3] = d.Var;
    put3(s4, s.radiance); s4[3] = s.Var;
}

// BSDF::getBRDF / getBTDF at a hand-chosen configuration: params =
// {N(3), shapeN(3), V(3), base(3), roughness, metallic, opacity, eta, entering}.
void oracle_bsdf_eval(const float *params, const float *L3, int32_t which, float *out3) {
    HitInfo h;
    h.surfaceNormal = get3(params);
    h.shapeNormal = get3(params + 3);
    h.baseColor = get3(params + 9);
    h.roughness = params[12];
    h.metallic = params[13];
    h.opacity = params[14];
    h.eta = params[15];
    h.entering = params[16] != 0.0f;
    BSDF b(get3(params + 6), h);
    vec3 r = which == 0 ? b.getBRDF(get3(L3)) : (which == 1 ? b.getBTDF(get3(L3)) : b.getBSDF(get3(L3)));
    put3(out3, r);
}

// One BSDF sample from stream (seed, tag, pix, smp): which 0 = sampleCos,
// 1 = sampleBRDF (GGX), 2 = sampleReflection, 3 = sampleBTDF.
// out: dir(3), bsdfPdf(3), pdf, fails.
void oracle_bsdf_sample(const float *params, int32_t which, uint32_t seed, uint32_t pix, uint32_t smp, float *out8) {
    HitInfo h;
    h.surfaceNormal = get3(params);
    h.shapeNormal = get3(params + 3);
    h.baseColor = get3(params + 9);
    h.roughness = params[12];
    h.metallic = params[13];
    h.opacity = params[14];
    h.eta = params[15];
    h.entering = params[16] != 0.0f;
    BSDF b(get3(params + 6), h);
    Generator g(seed, TAG_INDIRECT, pix, smp);
    vec3 dir, pdf3;
    float pdf = 0.0f;
    int fails = 0;
    if (which == 0) b.sampleCos(g, dir, pdf3, pdf, fails);
    else if (which == 1) b.sampleBRDF(g, dir, pdf3, pdf, fails);
    else if (which == 2) b.sampleReflection(g, dir, pdf3, fails);
    else b.sampleBTDF(g, dir, pdf3, fails);
    put3(out8, dir);
    put3(out8 + 3, pdf3);
    out8[6] = pdf;
    out8[7] = (float)fails;
}

// Forward-mode FP64 adjoint (oadjoint.cpp) along one parameter direction:
// tan_tex = level-0 texel tangents of every texture of the description,
// concatenated ([h][w][3] each; NULL = 0), tan_rough[n_materials],
// tan_emis[n_materials][3].  Writes the linear Full image and its directional
// derivative, [h][w][3] doubles each.  Returns 2 for a scene with transmission.
int32_t oracle_adjoint_dir(const oracle_scene *S, const rm_render_args *a, const float *tan_tex,
                           const double *tan_rough, const double *tan_emis, int32_t threads, double *image,
                           double *dimage) {
    if (!S || !a || a->width <= 0 || a->height <= 0 || a->max_depth > 16 || !tan_rough || !tan_emis) return 1;
    const Model &M = S->model;
    const size_t nm = M.materials.size();
    AdjDirection dir;
    dir.rough.assign(tan_rough, tan_rough + nm);
    dir.emis.assign(tan_emis, tan_emis + nm * 3);
    dir.tex.resize(nm);
    std::vector<size_t> off(S->tex_dims.size() + 1, 0);
    for (size_t t = 0; t < S->tex_dims.size(); ++t)
        off[t + 1] = off[t] + (size_t)S->tex_dims[t].first * S->tex_dims[t].second * 3;
    for (size_t m = 0; m < nm && tan_tex; ++m) {
        int t = S->mat_tex[m];
        const ImageData &img = M.materials[m].diffuse;
        if (t < 0 || img.empty()) continue;
        // the tangent's mip pyramid, built like ImageData::generateMipmaps
        auto &lv = dir.tex[m];
        lv.resize(img.map_depth);
        lv[0].assign(tan_tex + off[t], tan_tex + off[t + 1]);
        for (int level = 1; level < img.map_depth; ++level) {
            int pw = img.width >> (level - 1), ph = img.height >> (level - 1);
            int cw = img.width >> level, ch = img.height >> level;
            lv[level].assign((size_t)cw * ch * 3, 0.0);
            for (int y = 0; y < ch; ++y)
                for (int x = 0; x < cw; ++x)
                    for (int c = 0; c < 3; ++c) {
                        double s = 0.0;
                        for (int dy = 0; dy < 2; ++dy)
                            for (int dx = 0; dx < 2; ++dx)
                                s += lv[level - 1][((size_t)((y * 2 + dy) % ph) * pw + (x * 2 + dx) % pw) * 3 + c];
                        lv[level][((size_t)y * cw + x) * 3 + c] = s / 4.0;
                    }
        }
    }
    return adjoint_render(M, to_args(a), dir, threads, image, dimage);
}

}  // extern "C"

// ----------------------------------------------------------------- probes
// Single functions of the restatement, for the tests that pin them against
// closed forms, quadrature and independent recomputations.
namespace {
HitInfo surface_from(const float *p) {
    // position(3) shapeNormal(3) surfaceNormal(3) baseColor(3) roughness metallic opacity eta entering
    HitInfo h;
    h.position = get3(p);
    h.shapeNormal = get3(p + 3);
    h.surfaceNormal = get3(p + 6);
    h.baseColor = get3(p + 9);
    h.roughness = p[12];
    h.metallic = p[13];
    h.opacity = p[14];
    h.eta = p[15];
    h.entering = p[16] != 0.0f;
    return h;
}
}  // namespace

extern "C" {

// RandomDistribution (src/component.cpp:12-31) over weights[n]: `count` picks
// from stream (seed, TAG_INDIRECT, 0, 0) into out, the uniform draw of each
// pick into u_out, and pdf(i) of every index into pdf_out (each may be NULL).
void oracle_dist_probe(const float *weights, int32_t n, uint32_t seed, int32_t count, int32_t *out, float *u_out,
                       float *pdf_out) {
    RandomDistribution d;
    d.Init(std::vector<float>(weights, weights + n));
    Generator g(seed, TAG_INDIRECT, 0, 0), gu(seed, TAG_INDIRECT, 0, 0);
    for (int32_t i = 0; i < count; ++i) {
        if (out) out[i] = d(g);
        if (u_out) u_out[i] = gu();
    }
    if (pdf_out)
        for (int32_t i = 0; i < n; ++i) pdf_out[i] = d.pdf(i);
}

// The scene's sky (SkyBox, src/component.cpp:54-67,120-140).  what 0:
// get(dir) for n directions (in: [n][3]) into out [n][3]; 1: dist.pdf(i) of
// every pixel into out [h*w]; 2: the premultiplied data into out [h*w][3].
// Returns the pixel count (0 without a sky).
int64_t oracle_sky_probe(const oracle_scene *S, int32_t what, const float *in, int64_t n, float *out) {
    const SkyBox &sky = S->model.skyMap;
    const int64_t np = (int64_t)sky.width * sky.height;
    if (sky.empty()) return 0;
    if (what == 0)
        for (int64_t i = 0; i < n; ++i) put3(out + i * 3, sky.get(get3(in + i * 3)));
    else if (what == 1)
        for (int64_t i = 0; i < np; ++i) out[i] = sky.dist.pdf((int)i);
    else
        for (int64_t i = 0; i < np; ++i) put3(out + i * 3, sky.data[i]);
    return np;
}

// sampleSkyBox (src/sampling.cpp:450-465) about `normal`, one call per
// stream (seed, TAG_INDIRECT, s, 0): out [count][6] = Dir, light (NaN Dir
// when all 16 tries fell below the normal's plane).
void oracle_sky_sample(const oracle_scene *S, const float *normal, uint32_t seed, int32_t count, float *out) {
    for (int32_t s = 0; s < count; ++s) {
        Generator g(seed, TAG_INDIRECT, (uint32_t)s, 0);
        vec3 dir, light(0.0f);
        probe_sky_sample(S->model.skyMap, get3(normal), g, dir, light);
        put3(out + s * 6, dir);
        put3(out + s * 6 + 3, light);
    }
}

// sampleDirectLight (src/sampling.cpp:467-527) at a surface (surface_from
// layout, 17 floats) seen along -in_dir, `cnt` samples from one stream
// (seed, TAG_INDIRECT, s, 0) per s < streams: out [streams][3] = the sum over
// the samples of weight * light * bsdfPdf.
void oracle_direct_light(const oracle_scene *S, const float *surface, const float *in_dir, int32_t cnt, uint32_t seed,
                         int32_t streams, int32_t threads, float *out) {
    const HitInfo h = surface_from(surface);
    const BSDF bsdf(get3(in_dir), h);
    parallel_for(streams, threads, [&](int64_t s, int) {
        Generator g(seed, TAG_INDIRECT, (uint32_t)s, 0);
        std::vector<LightSample> v = sampleDirectLight(bsdf, S->model, OneGen{&g}, cnt);
        vec3 sum(0.0f);
        for (const LightSample &x : v) sum += x.weight * x.light * x.bsdfPdf;
        put3(out + s * 3, sum);
    });
}

// getBRDF / getBTDF / getBSDF (which 0 / 1 / 2) at a surface (surface_from
// layout) seen along -in_dir, for n directions L [n][3].
void oracle_bsdf_eval_n(const float *surface, const float *in_dir, const float *L, int64_t n, int32_t which,
                        float *out) {
    const BSDF b(get3(in_dir), surface_from(surface));
    for (int64_t i = 0; i < n; ++i) {
        const vec3 l = get3(L + i * 3);
        put3(out + i * 3, which == 0 ? b.getBRDF(l) : (which == 1 ? b.getBTDF(l) : b.getBSDF(l)));
    }
}

// GTR1 / GTR2 / smithG_GGX (which 0 / 1 / 2, src/sampling.cpp:30-47) at n points.
void oracle_brdf_term(int32_t which, const float *x, int64_t n, float a, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = brdf_term(which, x[i], a);
}

// renderPixel's camera ray and initRayDiff (src/render.cpp:439-477) for
// pixel (x, y): out = Dir(3), dDdx(3), dDdy(3).
void oracle_camera(const rm_render_args *a, int32_t x, int32_t y, float *out9) {
    const RenderArgs args = to_args(a);
    float rayX = float(x) - float(args.width) / 2.0f, rayY = float(y) - float(args.height) / 2.0f;
    vec3 d = args.direction + args.accuracy * (rayX * args.right + rayY * args.up);
    RayDifferential diff;
    initRayDiff(d, args, diff);
    put3(out9, normalize(d));
    put3(out9 + 3, diff.dDdx);
    put3(out9 + 6, diff.dDdy);
}

// barycentric (src/geometry.cpp:89-103) of P in the triangle tri9, and
// getHitNormals (src/model.cpp:240-270) with vertex normals n9 for a ray
// along dir3 at barycentric coordinates bc3: out = bary(3), shapeNormal(3),
// surfaceNormal(3), entering.
void oracle_hit_normals(const float *tri9, const float *n9, const float *P3, const float *dir3, const float *bc3,
                        float *out10) {
    VertexData vd[3];
    Face f;
    for (int k = 0; k < 3; ++k) {
        f.v[k] = get3(tri9 + 3 * k);
        vd[k].uv = vec2(0.0f);
        vd[k].normal = get3(n9 + 3 * k);
        f.data[k] = &vd[k];
    }
    f.material = nullptr;
    f.id = 0;
    put3(out10, barycentric(f.v[0], f.v[1], f.v[2], get3(P3)));
    vec3 sn, un;
    bool entering = false;
    getHitNormals(f, get3(dir3), get3(bc3), sn, un, entering);
    put3(out10 + 3, sn);
    put3(out10 + 6, un);
    out10[9] = entering ? 1.0f : 0.0f;
}

// ImageData::get (src/material.cpp:81-94, fp32 texels) of a w x h texture
// with its mip chain (generateMipmaps, :113-148) at n points uvd [n][3] =
// (u, v, depth): out [n][3].
void oracle_tex_get(int32_t w, int32_t h, const float *rgb, const float *uvd, int64_t n, float *out) {
    ImageData img;
    img.width = w;
    img.height = h;
    img.data[0].assign(rgb, rgb + (size_t)w * h * 3);
    img.generateMipmaps();
    for (int64_t i = 0; i < n; ++i) put3(out + i * 3, img.get(uvd[i * 3], uvd[i * 3 + 1], uvd[i * 3 + 2]));
}

// 8-bit ImageData (src/material.cpp:26-148) of a w x h texture with
// `channels` bytes per texel: get<vec3> (bw 3) or get<vec4> (bw 4) at n points
// uvd [n][3] = (u, v, depth): out [n][4].
void oracle_tex8_get(int32_t w, int32_t h, int32_t channels, const uint8_t *bytes, int32_t bw, const float *uvd,
                     int64_t n, float *out) {
    ImageData8 img;
    img.width = w;
    img.height = h;
    img.channels = channels;
    img.data[0].assign(bytes, bytes + (size_t)w * h * channels);
    img.generateMipmaps();
    for (int64_t i = 0; i < n; ++i) {
        if (bw == 3)
            img.get<3>(uvd[i * 3], uvd[i * 3 + 1], uvd[i * 3 + 2], out + i * 4);
        else
            img.get<4>(uvd[i * 3], uvd[i * 3 + 1], uvd[i * 3 + 2], out + i * 4);
    }
}

// generateMipmaps of the same texture: the level count; when `out` is given,
// every level's bytes, level after level.  transparent: hasTransparentPart.
int32_t oracle_tex8_mips(int32_t w, int32_t h, int32_t channels, const uint8_t *bytes, uint8_t *out,
                         int32_t *transparent) {
    ImageData8 img;
    img.width = w;
    img.height = h;
    img.channels = channels;
    img.data[0].assign(bytes, bytes + (size_t)w * h * channels);
    img.generateMipmaps();
    if (out)
        for (int l = 0; l < img.map_depth; ++l) {
            memcpy(out, img.data[l].data(), img.data[l].size());
            out += img.data[l].size();
        }
    if (transparent) *transparent = img.hasTransparentPart() ? 1 : 0;
    return img.map_depth;
}

// A material whose one 8-bit map is the given texture: which 0
// getDiffuseColor (vec4), 1 getNormal, 2 getEmissiveColor, 3 getSurfaceData
// (roughness, metallic) at n points uvd [n][3] = (u, v, duv; NaN duv: level
// 0): out [n][4].
void oracle_material8(int32_t w, int32_t h, int32_t channels, const uint8_t *bytes, int32_t which, const float *uvd,
                      int64_t n, float *out) {
    Material m;
    ImageData8 &img = which == 0 ? m.diffuse8 : which == 1 ? m.normals8 : which == 2 ? m.emissive8 : m.specular8;
    img.width = w;
    img.height = h;
    img.channels = channels;
    img.data[0].assign(bytes, bytes + (size_t)w * h * channels);
    img.generateMipmaps();
    for (int64_t i = 0; i < n; ++i) {
        const float u = uvd[i * 3], v = uvd[i * 3 + 1], duv = uvd[i * 3 + 2];
        float *o = out + i * 4;
        o[3] = 0.0f;
        if (which == 0) {
            vec4 c = m.getDiffuseColor(u, v, duv);
            for (int k = 0; k < 4; ++k) o[k] = c[k];
        } else if (which == 1 || which == 2) {
            put3(o, which == 1 ? m.getNormal(u, v, duv) : m.getEmissiveColor(u, v, duv));
        } else {
            m.getSurfaceData(u, v, o[0], o[1]);
            o[2] = 0.0f;
        }
    }
}

// calcSurfaceNormal (src/model.cpp:279-298): the surface normal surfN tilted
// by the normal-map value nm in the uv tangent frame of the triangle tri9
// with uvs uv6.
void oracle_surface_normal(const float *tri9, const float *uv6, const float *nm3, const float *shapeN3,
                           const float *surfN3, float *out3) {
    VertexData vd[3];
    Face f;
    for (int k = 0; k < 3; ++k) {
        f.v[k] = get3(tri9 + 3 * k);
        vd[k].uv = vec2(uv6[2 * k], uv6[2 * k + 1]);
        vd[k].normal = vec3(0.0f);
        f.data[k] = &vd[k];
    }
    f.material = nullptr;
    f.id = 0;
    vec3 sn = get3(surfN3);
    calcSurfaceNormal(f, get3(nm3), get3(shapeN3), sn);
    put3(out3, sn);
}

}  // extern "C"

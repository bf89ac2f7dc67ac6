// ORACLE -- test infrastructure only.  Nothing in the product may include,
// link or call anything under oracle/; only tests/, the smoke check and the
// benchmark's CPU-baseline leg use it, as the checker.
//
// A line-faithful CPU restatement of the reference renderer's hot path
// (lemonchu/Raym0nade).  Each function cites the reference file:line it
// restates.  Deviations, all forced by the workloads or by the counter RNG,
// are marked "EXT:".
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>
#include "orng.h"
#include "ovec.h"

namespace ref {

constexpr float eps_zero = 1e-4f;                       // include/geometry.h:13
const float PI = 3.14159265358979323846f;               // include/geometry.h:14
const vec3 RGB_Weight(0.3f, 0.6f, 0.1f);               // include/geometry.h:16

// ---------------------------------------------------------------- geometry
struct Ray { vec3 origin, direction; };                 // include/geometry.h:24-26
struct RayDifferential {                                // include/geometry.h:28-31
    vec3 dPdx{0.0f}, dPdy{0.0f}, dDdx{0.0f}, dDdy{0.0f};
};
struct Box { vec3 v0, v1; Box() = default; Box(vec3 a, vec3 b) : v0(a), v1(b) {} };

bool isfinite(vec2 v);
bool isfinite(vec3 v);
float sqrt_s(float x);
vec3 pow_s(vec3 v, float k);
Box operator+(const Box &A, const Box &B);
void rayInBox(const Ray &ray, const Box &box, float &tL, float &tR);
float RayTriangleIntersection(const Ray &ray, const vec3 &v0, const vec3 &v1, const vec3 &v2);
vec3 barycentric(const vec3 &v0, const vec3 &v1, const vec3 &v2, const vec3 &P);
void getTangentSpace(const vec3 &normal, vec3 &tangent, vec3 &bitangent);
void getTangentSpaceWithInDir(const vec3 &normal, const vec3 &inDir, vec3 &tangent, vec3 &bitangent);

// ---------------------------------------------------------------- material
// EXT: fp32 linear RGB textures (reference: uint8 with gamma 2.2), constant
// albedo / emission / metallic when no texture is bound.
struct ImageData {                                      // include/material.h:12-24
    int width = 0, height = 0, map_depth = 0;
    std::vector<float> data[8];
    void generateMipmaps();                             // src/material.cpp:113-148
    vec3 get(float u, float v, float depth) const;      // src/material.cpp:81-94
    bool empty() const { return data[0].empty(); }
    // EXT: the alpha channel as its own level-0 plane of a / 255 values (the
    // reference keeps RGBA bytes); only TransparentTest reads it, at level 0.
    std::vector<float> alpha;
    bool hasTransparentPart() const;                    // src/material.cpp:102-107
    float getAlpha(float u, float v) const;             // component 3 of get<vec4>(u, v, 0), :81-94
};

// The reference's 8-bit texture (include/material.h:12-24): bytes as its
// loaders store them ([h][w][channels]), integer mips, samples of bytes / 255
// read with the sampler's own stride (4 for vec4 lookups, 3 for vec3).  The
// reference's reads past a level's last byte are undefined; here they read 0.
struct ImageData8 {
    int width = 0, height = 0, channels = 0, map_depth = 0;
    std::vector<uint8_t> data[8];
    void generateMipmaps();                             // src/material.cpp:113-148
    template <int BitWidth>
    void get(float u, float v, float depth, float out[4]) const;   // src/material.cpp:48-94
    bool empty() const { return data[0].empty(); }
    bool hasTransparentPart() const;                    // src/material.cpp:102-107
};

struct Material {                                       // include/material.h:29-52
    int id = 0;
    float opacity = 1.0f, ior = 1.0f, roughness = 0.8f; // src/material.cpp:109-111
    vec3 transmittingColor{0.0f};
    vec3 baseColor{1.0f};     // EXT constant albedo (reference: white)
    vec3 emission{0.0f};      // EXT constant emission
    float metallic = 0.0f;    // EXT constant metallic
    ImageData diffuse;        // EXT fp32 albedo texture
    // the reference's textures: aiTextureType_DIFFUSE, NORMALS, EMISSIVE, SPECULAR
    ImageData8 diffuse8, normals8, emissive8, specular8;
    bool hasFullyTransparentPart = false;               // src/material.cpp:330-333
    // EXT: constant emission, or (the reference's light objects) an emissive texture
    bool hasEmission() const {
        return !emissive8.empty() || emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f;
    }
    vec4 getDiffuseColor(float u, float v, float duv) const;             // src/material.cpp:349-356
    vec3 getNormal(float u, float v, float duv) const;                   // src/material.cpp:358-363
    vec3 getEmissiveColor(float u, float v, float duv) const;            // src/material.cpp:365-372
    void getSurfaceData(float u, float v, float &roughness, float &metallic) const;  // :374-383
};

// ---------------------------------------------------------------- scene
struct VertexData { vec2 uv; vec3 normal; };            // include/component.h:8-13
struct Face {                                           // include/component.h:15-22
    vec3 v[3];
    const VertexData *data[3];
    const Material *material;
    int32_t id;       // original index (the reference identifies faces by address)
    vec3 center() const;                                // src/component.cpp:37-39
    Box aabb() const;                                   // src/component.cpp:41-46
};
struct HitRecord {                                      // include/component.h:24-29
    float t_min = eps_zero, t_max = INFINITY;
    const Face *face = nullptr;
    HitRecord() = default;
    HitRecord(float a, float b) : t_min(a), t_max(b) {}
};

struct RandomDistribution {                             // src/component.cpp:12-31
    std::vector<float> prefixSums;
    void Init(const std::vector<float> &d);
    int operator()(Generator &gen) const;
    float pdf(int index) const;
};
struct LightObject {                                    // include/component.h:47-54
    vec3 center{0.0f}, color{0.0f};
    float power = 0.0f;
    std::vector<Face> faces;
    RandomDistribution faceDist;
};
struct SkyBox {                                         // include/component.h:56-67
    int width = 0, height = 0;
    std::vector<vec3> data;
    RandomDistribution dist;
    void Init();                                        // src/component.cpp:54-67
    bool empty() const { return data.empty(); }
    vec3 get(vec3 dir) const;                           // src/component.cpp:120-140
};

struct Counters { uint64_t interior = 0, leaves = 0, boxes = 0, tris = 0; };

struct BVH_Node { Box box; int faceL = 0, faceR = 0; };
struct BVH {                                            // include/bvh.h:13-29
    std::vector<BVH_Node> node;
    Face *faces = nullptr;
    int size = 0;
    void dfs_build(int u, int faceL, int faceR);        // src/bvh.cpp:18-42
    void build(std::vector<Face> &faces);               // src/bvh.cpp:48-54
    void dfs_rayHit(int u, const Ray &ray, HitRecord &hit, Counters *c) const;  // :56-88
    void rayHit(const Ray &ray, HitRecord &hit, Counters *c = nullptr) const;     // :90-92
};

struct HitInfo {                                        // include/model.h:12-18, src/model.cpp:4-7
    vec3 shapeNormal{NAN}, surfaceNormal{NAN}, emission{0.0f}, baseColor{0.0f}, position{NAN};
    float specular = 0.04f, roughness = 0.8f, metallic = 0.0f, opacity = 1.0f, eta = 1.0f;
    int id = 0;
    bool entering = true;
};

struct Model {                                          // include/model.h:25-40
    std::vector<Material> materials;
    std::vector<Face> faces;
    std::vector<VertexData> vertexDatas;
    std::vector<LightObject> lightObjects;
    BVH bvh;
    SkyBox skyMap;
    HitRecord rayHit(Ray ray, Counters *c = nullptr) const;                       // src/model.cpp:332-341
    bool rayHit_test(Ray ray, float aimDepth, Counters *c = nullptr) const;       // src/model.cpp:343-354
    // The same loops with the first t_min (and rayHit's t_max) as arguments;
    // the reference's are eps_zero and INFINITY.
    HitRecord rayHitFrom(Ray ray, float tmin, float tmax, Counters *c = nullptr) const;
    bool rayHitTestFrom(Ray ray, float tmin, float aimDepth, Counters *c = nullptr) const;
};

void getHitNormals(const Face &face, const vec3 &inDir, const vec3 &baryCoords, vec3 &shapeNormal,
                   vec3 &surfaceNormal_raw, bool &entering);                      // src/model.cpp:240-270
void getHitMaterial(const Face &face, const vec3 &baryCoords, const vec3 &hit_dPdx, const vec3 &hit_dPdy,
                    HitInfo &hitInfo);                                            // src/model.cpp:300-328
void calcSurfaceNormal(const Face &face, const vec3 &normalMap, const vec3 &shapeNormal,
                       vec3 &surfaceNormal);                                      // src/model.cpp:279-298

// ---------------------------------------------------------------- sampling
class BSDF {                                            // include/sampling.h:8-26
public:
    vec3 inDir;
    HitInfo surface;
    BSDF(const vec3 &inDir, const vec3 &hit_position);
    BSDF(const vec3 &inDir, const HitInfo &hitInfo);
    vec3 getBRDF(vec3 L) const;
    vec3 getBTDF(vec3 L) const;
    vec3 getBSDF(vec3 outDir) const;
    void sampleBRDF(Generator &gen, vec3 &outDir, vec3 &brdfPdf, float &pdf, int &fails) const;
    void sampleCos(Generator &gen, vec3 &outDir, vec3 &brdfPdf, float &pdf, int &fails) const;
    void preciseRefraction(vec3 &outDir, float &F) const;
    void sampleReflection(Generator &gen, vec3 &Dir, vec3 &bsdfPdf, int &fails) const;
    void sampleBTDF(Generator &gen, vec3 &Dir, vec3 &btdfPdf, int &fails) const;
};
struct LightSample {                                    // include/sampling.h:28-32
    vec3 bsdfPdf, light;
    float weight;
    LightSample(const vec3 &b, const vec3 &l, float w) : bsdfPdf(b), light(l), weight(w) {}
};
const float eps_lightRadius = 1e-3f;                    // include/sampling.h:34
// Recorder for the forward-mode adjoint (oadjoint.cpp).  While a thread's
// adj_tap is set, getHitMaterial, the reflection samplers and
// sampleDirectLight also report the values the derivative needs; they
// compute nothing differently.
struct AdjTap {
    float u = 0, v = 0, duv = 0;           // getHitMaterial: texture coordinate and footprint
    std::vector<float> ggx_ct;             // sampleBRDF: cos(theta_h) of every draw
    float ct2 = 0, lh2 = 1;                // sampleBRDF: the accepted draw
    vec3 raw1{0.0f};                       // sampleCos: brdfPdf before clampPdf
    float pdf1 = 0, pdf2 = 0;              // sampleReflection: the two pdfs and the pick
    bool first = false;
    struct Light { vec3 dir; float dist, P; int index; };
    std::vector<Light> lights;             // sampleDirectLight: accepted samples in order
};
extern thread_local AdjTap *adj_tap;
void clampPdf(vec3 &pdf);                               // src/sampling.cpp:238-243 (`clamp`)
// EXT: `genFor(T)` returns the generator used by light sample T.  The
// reference passes one generator; the primary-vertex call gives every sample
// its own counter stream so samples can be computed independently.
template <class GenFor>
std::vector<LightSample> sampleDirectLight(const BSDF &bsdf, const Model &model, GenFor genFor, int sampleCnt);
// Probes of file-local pieces for the tests: brdf_term(0|1|2, x, a) is
// GTR1(x, a) | GTR2(x, a) | smithG_GGX(x, a) (src/sampling.cpp:30-47);
// probe_sky_sample is sampleSkyBox (src/sampling.cpp:450-465).
float brdf_term(int which, float x, float a);
void probe_sky_sample(const SkyBox &sky, const vec3 &normal, Generator &gen, vec3 &Dir, vec3 &light);
struct OneGen {          // every light sample continues the caller's stream
    Generator *g;
    Generator &operator()(int) const { return *g; }
};
struct PerSampleGen {    // light sample T of pixel `pix` uses stream (seed, TAG_DIRECT, pix, T)
    uint32_t seed, pix;
    std::vector<Generator> *store;
    Generator &operator()(int T) const {
        store->emplace_back(seed, TAG_DIRECT, pix, (uint32_t)T);
        return store->back();
    }
};

// ---------------------------------------------------------------- image
struct RadianceData { vec3 radiance{0.0f}; float Var = 0.0f; };   // include/image.h:9-13
void accumulateInwardRadiance(const vec3 &baseColor, const LightSample &sample, RadianceData &d,
                              RadianceData &s);                    // src/image.cpp:630-659

struct RenderArgs {                                     // include/render.h:8-15 (+EXT)
    vec3 position, direction, up, right;
    float accuracy = 0, focus = 0, CoC = 0, exposure = 1, P_Direct = 0.7f;
    int width = 0, height = 0, spp = 0, threads = 1;
    uint32_t seed = 1;      // EXT counter-RNG seed
    int max_depth = 16;     // EXT maxRayDepth (src/render.cpp:125)
    int rr_min_depth = 1;   // EXT first depth allowed to terminate by NEE/RR
};

struct Photo {                                          // include/image.h:15-67
    enum ShadeOption {
        BaseColor = 1, Emission = 2, DirectLight = 4, IndirectLight = 8, Diffuse = 16, Specular = 32,
        shapeNormal = 64, surfaceNormal = 128,
        Full = DirectLight | IndirectLight | Diffuse | Specular | BaseColor | Emission,
        DoBloom = 256, DoFXAA = 512, DoDepthFieldBlur = 1024
    };
    int width, height;
    float exposure = 1.0f;
    float focus = 0.0f, CoC = 0.0f;                     // include/image.h:36-37
    vec3 cameraPosition{0.0f};
    std::vector<HitInfo> Gbuffer;
    std::vector<RadianceData> radiance_Dd, radiance_Ds, radiance_Id, radiance_Is;
    std::vector<vec3> pixelarray;
    Photo(int w, int h);
    void shade(int options);                            // src/image.cpp:215-246
    void FXAA();                                        // src/image.cpp:363-452
    void gammaCorrection();                             // src/image.cpp:456-468
    void postProcessing(int shadeOptions);              // src/image.cpp:470-479
    void spatialClamp();                                // src/image.cpp:77-82   (oimage.cpp)
    void bloom();                                       // src/image.cpp:248-283 (oimage.cpp)
    void depthFeildBlur();                              // src/image.cpp:285-356 (oimage.cpp)
    void filter();                                      // src/image.cpp:200-213 (oimage.cpp)
};

void calc_dPdxy(const Ray &ray, float hit_t, const vec3 &normal, const RayDifferential &base_diff, vec3 &hit_dPdx,
                vec3 &hit_dPdy);                                                                // src/render.cpp:44-50
void calc_dDdxy(const Ray &ray, const vec3 &normal, const RayDifferential &base_diff, vec3 &hit_dDdx,
                vec3 &hit_dDdy);                                                                // src/render.cpp:52-60
void getHitInfo(const HitRecord &hit, const Ray &ray, const RayDifferential &base_diff, vec3 &hit_dPdx,
                vec3 &hit_dPdy, HitInfo &hitInfo);                                              // src/render.cpp:62-79
void initRayDiff(const vec3 &d, const RenderArgs &args, RayDifferential &base_diff);           // src/render.cpp:439-446
void renderPixel(const Model &model, const RenderArgs &args, Photo &image, int x, int y);   // src/render.cpp:448-551
// Forward-mode FP64 adjoint (oadjoint.cpp): per-pixel linear Full image and
// its derivative along one parameter direction (texel tangents per material
// as mip pyramids, roughness and emission tangents per material).
struct AdjDirection {
    std::vector<std::vector<std::vector<double>>> tex;   // [material][level] texel tangents, empty = none
    std::vector<double> rough;                            // [material]
    std::vector<double> emis;                             // [material * 3]
};
int adjoint_render(const Model &model, const RenderArgs &args, const AdjDirection &dir, int threads, double *image,
                   double *dimage);
void render_multiThread(const Model &model, const RenderArgs &args, Photo &photo, int threads,
                        const std::vector<int> *rows = nullptr);                            // src/render.cpp:593-626
void render_pixels(const Model &model, const RenderArgs &args, Photo &photo, int threads, const std::vector<int> &pix);
// Denoise (oimage.cpp).  The reference's unqualified pow/exp/sqrt/abs on
// floats resolve to the float overloads: model.h pulls in Python.h, whose
// <math.h> is libstdc++'s wrapper that brings std's overloads into the
// global namespace.
void filter_var(std::vector<RadianceData> &rad, int width, int height);                     // src/image.cpp:84-107
float denoise_weight(const HitInfo &Gp, const HitInfo &Gq, const RadianceData &Lp, const RadianceData &Lq,
                     bool specular);                                                            // src/image.cpp:109-156
void filter_pass(std::vector<RadianceData> &rad, const std::vector<HitInfo> &G, int width, int height, bool specular,
                 int step);                                                                     // src/image.cpp:158-189
void filter_radiance(std::vector<RadianceData> &rad, const std::vector<HitInfo> &G, int width, int height,
                     bool specular);                                                            // src/image.cpp:191-198

}  // namespace ref

/*
 * raym0nade_hip.h -- C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * The reference renderer (lemonchu/Raym0nade) has no plugin or FFI layer; its
 * hot path sits behind in-process C++ calls.  Each entry point below names the
 * reference interface it replaces (file:line in the reference tree).
 *
 * Conventions
 *   - Every function returns an rm_status; on failure rm_last_error() returns
 *     a thread-local message.  Nothing aborts the process.
 *   - Arguments named *_dev are device pointers (HBM), caller-owned.  Host
 *     arrays are only read during the call.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Functions that take a stream are asynchronous unless stated otherwise.
 *   - A scene is immutable after creation; concurrent trace/render calls on
 *     distinct streams and distinct frames are safe.
 */
#ifndef RAYM0NADE_HIP_H
#define RAYM0NADE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rm_status {
    RM_OK = 0,
    RM_ERR_INVALID = 1,     /* bad argument */
    RM_ERR_OOM = 2,         /* device or host allocation failed */
    RM_ERR_HIP = 3,         /* HIP runtime error (message has the HIP string) */
    RM_ERR_EMPTY = 4,       /* zero-triangle scene (undefined behaviour in the reference, src/bvh.cpp:51) */
    RM_ERR_UNSUPPORTED = 5, /* feature outside this build */
    RM_ERR_NO_DEVICE = 6    /* no HIP device visible */
};

/* Thread-local description of the last failure in this thread. */
const char *rm_last_error(void);
/* ABI version, bumped on any layout change of the structs below. */
int32_t rm_abi_version(void);

/* ------------------------------------------------------------------ scene */

/* Material.  Replaces the reference's Material (include/material.h:29-52);
 * fields with a reference counterpart keep its defaults, the rest are
 * extensions used when no texture is bound:
 *   base_color   constant diffuse albedo (reference: white without a diffuse map,
 *                src/material.cpp:349-351)
 *   emission     constant emitted radiance (reference: emissive texture only,
 *                src/material.cpp:365-372)
 *   metallic     constant metallic (reference: 0 without a specular map,
 *                src/material.cpp:374-378)
 *   opacity/ior/roughness/transmitting_color as Material (src/material.cpp:109-111,300-335)
 *   albedo_tex   diffuse texture (rm_texture) or -1: fp32 linear RGB (extension),
 *                or 8-bit as the reference's Material::getDiffuseColor reads it
 *                (src/material.cpp:349-356: bilinear x trilinear filtering of
 *                bytes / 255, then the power 2.2)
 * and the reference's other maps, each an 8-bit texture or -1 (every -1 when
 * the material has none; the fields are read, so a caller sets them):
 *   normal_tex   aiTextureType_NORMALS: getNormal = sample * 2 - 1, tilting
 *                the shading normal in the face's uv tangent frame
 *                (src/material.cpp:358-363, src/model.cpp:279-298)
 *   emissive_tex aiTextureType_EMISSIVE: emission = sample ^ 2.2, and the
 *                mesh becomes a light object whose colour and power are the
 *                texture averaged over an 8 x 8 grid of each face
 *                (src/material.cpp:365-372, src/model.cpp:23-82)
 *   specular_tex aiTextureType_SPECULAR: metallic = min(b, 0.99), roughness =
 *                max(g, 1e-3) of the level-0 sample (src/material.cpp:374-383) */
typedef struct rm_material {
    float base_color[3];
    float emission[3];
    float transmitting_color[3];
    float roughness;
    float metallic;
    float opacity;
    float ior;
    int32_t albedo_tex;
    int32_t normal_tex;
    int32_t emissive_tex;
    int32_t specular_tex;
} rm_material;

/* fp32 linear RGB texture, row-major [height][width][3].  A mip chain of 2x2
 * box averages is built like ImageData::generateMipmaps (src/material.cpp:113-148).
 *
 * alpha (optional, NULL for none): [height][width] coverage, a / 255 for an
 * 8-bit alpha byte a.  A material whose albedo texture has an alpha value
 * below 1 is an alpha cut-out (Material::hasFullyTransparentPart,
 * src/material.cpp:330-333).  Every closest-hit and any-hit walk of the
 * library (rm_trace_closest, rm_trace_any, rm_render, rm_adjoint) then skips
 * its faces where the bilinear level-0 alpha at the hit is below 1e-4
 * (TransparentTest, src/model.cpp:217-230): it walks again from t + 1e-4, at
 * most 8 walks per ray.  A closest-hit ray whose 8 walks all end on such
 * points is a miss (Model::rayHit, src/model.cpp:332-341), and a shadow ray
 * that meets 8 of them is blocked (Model::rayHit_test, src/model.cpp:343-354).
 *
 * bytes (optional): an 8-bit texture in the reference's own storage
 * (ImageData, include/material.h:12-24): [height][width][channels] bytes as its
 * loaders produce them (channels 3 for PNG, 4 for DDS).  rgb and alpha are
 * then unused.  Mips are integer 2 x 2 means (sum / 4, src/material.cpp:113-148)
 * and a sample reads channels-strided texels with the sampler's own stride
 * (4 bytes for colour, emissive and specular lookups, 3 for normal maps,
 * src/material.cpp:48-79) as the reference does; a texel read that runs past
 * the level's last byte (a 4-byte read of 3-channel data, undefined in the
 * reference) reads zeros.  A diffuse texture with some alpha byte (every 4th
 * byte from the 4th) below 255 makes its material an alpha cut-out
 * (src/material.cpp:102-107,330-333), tested on the level-0 sample's 4th
 * component (src/model.cpp:217-230). */
typedef struct rm_texture {
    int32_t width, height;
    const float *rgb;
    const float *alpha;
    const uint8_t *bytes;
    int32_t channels;
} rm_texture;

/* Flat scene.  Replaces what Model's constructor assembles from assimp
 * (src/model.cpp:88-123,172-215): faces with per-vertex uv/normal and a
 * material, grouped into meshes (light objects are built per mesh, in mesh
 * order, like Model::checkLightObject, src/model.cpp:44-82), and an optional
 * HDR sky (SkyBox, src/component.cpp:54-67,120-140; with a sky, emissive
 * materials are ignored exactly as the reference skips emissive textures,
 * src/model.cpp:121,140). */
typedef struct rm_scene_desc {
    int64_t n_tris;
    const float *verts;      /* [n][3][3] */
    const float *uvs;        /* [n][3][2] or NULL (all zero) */
    const float *normals;    /* [n][3][3] or NULL (geometric normal) */
    const int32_t *material; /* [n] */
    const int32_t *mesh;     /* [n] mesh id, meshes contiguous; NULL = one mesh */
    int32_t n_materials;
    const rm_material *materials;
    int32_t n_textures;
    const rm_texture *textures;
    int32_t sky_width, sky_height;
    const float *sky_rgb;    /* [h][w][3] radiance or NULL */
} rm_scene_desc;

typedef struct rm_scene rm_scene;

typedef struct rm_scene_info {
    int64_t n_tris;
    int64_t n_nodes;        /* nodeCount(1,n)+1, src/bvh.cpp:46-50 */
    int32_t depth;          /* levels of the implicit heap */
    int32_t n_light_objects;
    int32_t n_light_faces;
    int32_t has_sky;
    int64_t device_bytes;   /* HBM held by the scene */
    double build_seconds;   /* host BVH build + flatten */
    double upload_seconds;
} rm_scene_info;

/* Builds the reference BVH (variance-axis median split, <=10 triangles per
 * leaf, implicit heap: BVH::build, src/bvh.cpp:18-54) on the host, flattens it
 * into the HBM layout and uploads it together with shading data, light
 * objects and the sky CDF.  Replaces Model::Model + BVH::build. */
int32_t rm_scene_create(const rm_scene_desc *desc, rm_scene **out);
int32_t rm_scene_destroy(rm_scene *scene);
int32_t rm_scene_get_info(const rm_scene *scene, rm_scene_info *info);
/* perm[slot] = original triangle id after the build's in-place permutation
 * (the reference permutes `faces`, src/bvh.cpp:38). Host array of n_tris. */
int32_t rm_scene_get_perm(const rm_scene *scene, int32_t *perm);

/* Model::Model(model_folder, model_name, skyMap_name) (src/model.cpp:172-215):
 * rm_model_read, then rm_scene_create. */
int32_t rm_scene_load(const char *model_folder, const char *model_name, const char *sky_name, rm_scene **out);

/* The host half of rm_scene_load (the reference's assimp import and
 * Model::processMaterial / processMesh, src/model.cpp:88-170): reads
 * model_folder + model_name, its texture maps and, unless sky_name is "null",
 * the Radiance HDR sky model_folder + sky_name into host memory, and fills
 * *desc with pointers into it, valid until rm_model_destroy.  Touches no
 * device.
 *   - Wavefront OBJ + MTL: faces fan-triangulated; meshes are the
 *     (group, material) pairs in order of appearance; MTL d / Tr below 0.99
 *     makes the reference's glass (src/material.cpp:300-333); Kd, Ke, Pr and
 *     Pm set the constant albedo, emission, roughness and metallic
 *     (extensions); map_Kd, norm / map_Kn, map_Ke / map_emissive and map_Ks
 *     are the diffuse, normal, emissive and specular maps.
 *   - FBX (binary or ASCII, version 7100 and later), as assimp imports it with the
 *     reference's flags (Triangulate | PreTransformVertices, hierarchy kept |
 *     SortByPType | FixInfacingNormals): one mesh per (geometry, material)
 *     in order of first use, vertices in world space, quads split at their
 *     concave corner, inward-facing meshes flipped; the maps connected as
 *     DiffuseColor, NormalMap, EmissiveColor / EmissiveFactor and
 *     SpecularColor / SpecularFactor; a material Opacity below 0.99 makes
 *     the glass.
 * Other model formats are RM_ERR_UNSUPPORTED. */
typedef struct rm_model rm_model;
int32_t rm_model_read(const char *model_folder, const char *model_name, const char *sky_name, rm_scene_desc *desc,
                      rm_model **out);
int32_t rm_model_destroy(rm_model *model);

/* An image as linear floats [h][w][3] (malloc'd; free with rm_image_free):
 * PNG as c / 255 after libpng's conversions in the reference's loader
 * (src/material.cpp:214-267: 16-bit to the high byte, grey and palette to
 * RGB, alpha dropped); DDS (uncompressed RGB(A), BC1/BC2/BC3 by FourCC or
 * DX10 header; the reference's loadImageFromDDS, src/material.cpp:190-212)
 * as c / 255 of the decoded 8-bit texels, alpha dropped; Radiance HDR (RGBE)
 * as stored (SkyBox::load, src/component.cpp:69-114). */
int32_t rm_image_load(const char *path, int32_t *width, int32_t *height, float **rgb);
void rm_image_free(float *rgb);

/* Host-only BVH build (BVH::build, src/bvh.cpp:48-54) of verts[n][3][3]:
 * writes the permutation perm[n] (slot -> original id), the heap size
 * nodeCount(1,n)+1 and the depth.  Needs no device. */
int32_t rm_bvh_build(const float *verts, int64_t n, int32_t *perm, int64_t *n_nodes, int32_t *depth);

/* ------------------------------------------------------------- traversal */

/* Closest hit.  Replaces BVH::rayHit (src/bvh.cpp:90-92) called as in
 * Model::rayHit without alpha textures (src/model.cpp:332-341): search
 * (tmin, tmax), strict on both ends.  rays_dev: [n][6] = origin xyz, direction
 * xyz.  t_out_dev: final t_max (tmax when missed).  id_out_dev: ORIGINAL
 * triangle id or -1.  Triangle ids are bit-exact with the reference
 * traversal; t is bit-exact as well. */
int32_t rm_trace_closest(const rm_scene *scene, const float *rays_dev, int64_t n,
                         float tmin, float tmax, float *t_out_dev, int32_t *id_out_dev,
                         void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  rm_trace_closest with the same results, whose
 * ray-reordering workspace is a static device array of the library (64 MiB per
 * GPU, held until the process ends) instead of allocated memory; it calls no
 * allocating routine and ignores RM_TUNE_KERNEL_TIMING.  At most
 * RM_NOALLOC_MAX_RAYS rays per call (RM_ERR_INVALID beyond).  Every call on a
 * GPU shares that GPU's array: calls that overlap in time on one GPU (on
 * different streams or threads) race, so the caller serialises them. */
#define RM_NOALLOC_MAX_RAYS 2097152
int32_t rm_trace_closest_noalloc(const rm_scene *scene, const float *rays_dev, int64_t n, float tmin, float tmax,
                                 float *t_out_dev, int32_t *id_out_dev, void *stream);

/* Occlusion.  Replaces Model::rayHit_test (src/model.cpp:343-354): true iff the
 * closest hit on (tmin, aim+1e-4) has t < aim.  aim_dev (per ray, may be NULL)
 * overrides aim_depth.  Exits as soon as the running closest t drops below
 * aim, which gives the same boolean.  occluded_dev: 0/1 bytes. */
int32_t rm_trace_any(const rm_scene *scene, const float *rays_dev, int64_t n,
                     float tmin, float aim_depth, const float *aim_dev,
                     uint8_t *occluded_dev, void *stream);

/* Instrumented closest/any traversal (same results) that also counts, in the
 * reference traversal order, interior visits, leaf visits, box tests and
 * triangle tests: counts_dev = uint64[4].  Used to price algorithmic bytes.
 * It counts the first walk of each ray: the further walks past alpha
 * cut-outs (rm_texture) are not counted. */
int32_t rm_trace_count(const rm_scene *scene, const float *rays_dev, int64_t n,
                       float tmin, float tmax, int32_t any_hit,
                       uint64_t *counts_dev, void *stream);

/* ------------------------------------------------------------------ render */

/* RenderArgs (include/render.h:8-15) plus the extensions the workloads need.
 * seed: counter-RNG seed.  max_depth: maxRayDepth (16 in the reference,
 * src/render.cpp:125; 1..16).  rr_min_depth: first depth at which NEE/RR may
 * terminate a path (1 reproduces the reference, src/render.cpp:147-150).
 * paths_per_round: indirect samples per pixel launched per round (0 = auto). */
typedef struct rm_render_args {
    float position[3], direction[3], up[3], right[3];
    float accuracy, focus, coc, exposure, p_direct;
    int32_t width, height, spp, threads;
    uint32_t seed;
    int32_t max_depth;
    int32_t rr_min_depth;
    int32_t paths_per_round;
} rm_render_args;

typedef struct rm_frame rm_frame;

/* Frame buffers (device), the GPU form of Photo (include/image.h:15-67). */
enum rm_buffer {
    RM_BUF_POSITION = 0,       /* G-buffer position, NaN on miss      [h][w][3] */
    RM_BUF_SHAPE_NORMAL = 1,   /*                                      [h][w][3] */
    RM_BUF_SURFACE_NORMAL = 2, /*                                      [h][w][3] */
    RM_BUF_EMISSION = 3,       /*                                      [h][w][3] */
    RM_BUF_BASE_COLOR = 4,     /*                                      [h][w][3] */
    RM_BUF_DD = 5,             /* radiance_Dd: rgb + Var               [h][w][4] */
    RM_BUF_DS = 6,
    RM_BUF_ID = 7,
    RM_BUF_IS = 8,
    RM_BUF_PIXELS = 9,         /* last post-processed image           [h][w][3] */
    RM_BUF_MATERIAL = 10       /* G-buffer (metallic, roughness, opacity) of
                                  the primary hit; HitInfo's defaults
                                  (0, 0.8, 1) on a miss               [h][w][3] */
};

int32_t rm_frame_create(int32_t width, int32_t height, rm_frame **out);
int32_t rm_frame_destroy(rm_frame *frame);
/* Device pointer of one buffer (stable for the frame's lifetime). */
int32_t rm_frame_buffer(rm_frame *frame, int32_t which, float **dev_ptr);
/* Synchronous copy of one buffer to host memory. */
int32_t rm_frame_download(const rm_frame *frame, int32_t which, float *host);

/* Renders one frame.  Replaces render_multiThread's pixel loop
 * (src/render.cpp:593-633 -> renderPixel, src/render.cpp:448-551). */
int32_t rm_render(const rm_scene *scene, const rm_render_args *args, rm_frame *frame,
                  void *stream);

/* Counters of the last render of a frame (no reference counterpart; the
 * reference prints only its wall time, src/render.cpp:632).  Every render
 * resets and fills them:
 *   PRIMARY_RAYS     camera rays (one closest-hit walk per pixel)
 *   EXTENSION_RAYS   closest-hit walks of indirect path vertices
 *   SHADOW_RAYS      any-hit walks of light samples (primary vertex and NEE)
 *   HITS             shading vertices: primary and extension hits
 *   PATHS            indirect samples started at the primary vertex
 *   MEDIUM_OVERFLOW  dielectric entries a path's medium stack could not hold:
 *                    it keeps 16 entered media, which holds the reference's
 *                    unbounded Medium (src/render.cpp:13-42) for every path
 *                    of max_depth <= 16, except an entry made at a path's
 *                    last vertex, which nothing reads afterwards
 *   MEDIUM_MAX       the deepest medium stack (entered media) of any path
 *   WALK_*           with RM_TUNE_RENDER_COUNT on: interior visits, leaf
 *                    visits, box tests and triangle tests of the extension and
 *                    shadow walks, in the reference traversal order (as
 *                    rm_trace_count counts them); 0 otherwise */
enum rm_render_stat {
    RM_STAT_PRIMARY_RAYS = 0,
    RM_STAT_EXTENSION_RAYS = 1,
    RM_STAT_SHADOW_RAYS = 2,
    RM_STAT_HITS = 3,
    RM_STAT_PATHS = 4,
    RM_STAT_MEDIUM_OVERFLOW = 5,
    RM_STAT_WALK_INTERIOR = 6,
    RM_STAT_WALK_LEAVES = 7,
    RM_STAT_WALK_BOXES = 8,
    RM_STAT_WALK_TRIS = 9,
    RM_STAT_MEDIUM_MAX = 10,
    RM_STAT_COUNT = 16
};
/* Synchronous copy of the frame's RM_STAT_COUNT counters (waits for the device). */
int32_t rm_frame_get_stats(const rm_frame *frame, uint64_t *stats);

/* Batched form (no reference counterpart; the reference renders one frame per
 * render_multiThread call).  All jobs run through one set of kernel launches
 * over their concatenated pixels -- each launch carries paths of every job --
 * and with use_graph != 0 the whole sequence after the primary pass is
 * captured into one hipGraph and launched once.  Each job needs its own frame;
 * post_options != 0 also runs rm_postprocess(frame, post_options) for it.
 * Results equal those of rm_render on each job.  Synchronous. */
typedef struct rm_batch_job {
    const rm_scene *scene;
    rm_render_args args;
    rm_frame *frame;
    int32_t post_options;
} rm_batch_job;
int32_t rm_render_batch(const rm_batch_job *jobs, int32_t n, int32_t use_graph, void *stream);

/* Photo::postProcessing (src/image.cpp:470-479): shade(options) ->
 * depthFeildBlur if options & 1024 -> bloom if options & 256 ->
 * gammaCorrection -> FXAA if options & 512.  Options are bit-compatible with
 * Photo::ShadeOption (include/image.h:17-34); the extra bit RM_SHADE_LINEAR
 * skips gammaCorrection (the linear image the adjoint differentiates).
 * Writes RM_BUF_PIXELS.  Stream-ordered, except that depth of field waits
 * once on the stream for the size of its work lists (so it is refused in an
 * rm_render_batch job's post_options).  A pixel whose primary ray missed is
 * at distance +inf for depth of field (undefined in the reference). */
#define RM_SHADE_FULL 63
#define RM_SHADE_LINEAR 65536
/* Extra option bits (no Photo::ShadeOption counterpart): run spatialClamp
 * and/or filter on the frame's radiance buffers, in place, before shading --
 * what the reference does between its exports (src/render.cpp:646,655).
 * Launch-only, so they may be given in an rm_render_batch job's
 * post_options, where they run inside the batch's graph. */
#define RM_POST_SPATIAL_CLAMP (1 << 17)
#define RM_POST_FILTER (1 << 18)
int32_t rm_postprocess(rm_frame *frame, int32_t shade_options, float exposure, void *stream);

/* Photo's depth-of-field fields (include/image.h:36-37, set from the render
 * arguments at src/render.cpp:666-668).  rm_render sets them from its
 * arguments (focus, coc, position); this call overrides them, e.g. for a
 * frame filled with rm_frame_upload.  Used by rm_postprocess with option
 * 1024 (DoDepthFieldBlur); CoC must be finite and >= 0. */
int32_t rm_frame_set_dof(rm_frame *frame, float focus, float coc, const float camera_position[3]);

/* Device-to-device copy of one frame buffer into `dst` (device memory of
 * width*height*channels floats), stream-ordered. */
int32_t rm_frame_copy(const rm_frame *frame, int32_t which, float *dst, void *stream);

/* Synchronous copy of host memory into one frame buffer (width*height*channels
 * floats): a caller that produced its own G-buffer and radiance (the
 * reference's Photo arrays) hands them to the denoiser this way. */
int32_t rm_frame_upload(rm_frame *frame, int32_t which, const float *host);

/* Photo::spatialClamp (src/image.cpp:30-82): in place on the four radiance
 * buffers, each pixel brighter than 36x its blurred neighbourhood luminance
 * scaled down to it.  Stream-ordered; launches only. */
int32_t rm_spatial_clamp(rm_frame *frame, void *stream);

/* Photo::filter (src/image.cpp:84-213): in place on the four radiance
 * buffers (rgb and Var), filterVar then the five edge-aware a-trous passes at
 * strides 1..16, guided by the G-buffer (position, normals, material); the
 * specular buffers (Ds, Is) are weighted by the centre's roughness.  Pixels
 * whose primary ray missed end at radiance 0, Var 0, as in the reference.
 * The first call on a frame allocates its ping-pong buffer (64 bytes per
 * pixel); later calls allocate nothing.  Stream-ordered. */
int32_t rm_filter(rm_frame *frame, void *stream);

/* ------------------------------------------------------- differentiation */

/* No reference counterpart (the reference renders only).  Replaces the
 * values the adjoint differentiates without rebuilding geometry or the BVH:
 * materials (host array of the scene's n_materials, NULL = keep) and the
 * level-0 texels of every texture (device memory, the textures of the scene
 * description concatenated in order, [h][w][3] each; NULL = keep).  Mip levels
 * and light objects are rebuilt exactly as rm_scene_create builds them.  A
 * material may change every value except its texture binding and whether it
 * emits.  Frames rendered before the update no longer match the scene. */
int32_t rm_scene_update(rm_scene *scene, const rm_material *materials, const float *texels, void *stream);

/* With recording on, rm_render also keeps per indirect sample of every pixel
 * the sums of its light samples (48 B per sample, scenes without transmission
 * only), so the adjoint replays each sample's path once.  rm_adjoint needs it
 * for every frame with indirect samples.  Off by default; off frees the record. */
int32_t rm_frame_set_recording(rm_frame *frame, int32_t on);

/* Gradient outputs of rm_adjoint, all device memory, overwritten; any may be NULL. */
typedef struct rm_grads {
    float *textures;   /* d loss / d level-0 texel, same layout as rm_scene_update's texels */
    float *roughness;  /* [n_materials] */
    float *emission;   /* [n_materials][3] */
    float *replay;     /* [h][w][3]: the linear image recomputed by the replay (a check of
                          the replay against the render it differentiates) */
} rm_grads;

/* Path-replay adjoint of the linear Full image (RM_SHADE_FULL | RM_SHADE_LINEAR)
 * of rm_render(scene, args, frame): given d_image = d loss / d image (device,
 * [h][w][3]), writes d loss / d parameters.  `frame` must hold that render,
 * made with recording on (rm_frame_set_recording; RM_ERR_UNSUPPORTED
 * otherwise).  Every sample path is recomputed from its counter-RNG stream,
 * not stored: the record keeps only each sample's sums (48 B), and the
 * adjoint replays the path once to form the derivatives.
 * Albedo textures and emission get the exact derivative of the seeded
 * estimator; roughness gets the pathwise derivative plus the score of every
 * roughness-dependent draw (GGX half vectors, the cosine/GGX pick), an unbiased
 * estimate of the derivative of the expected image.  Light-object and
 * Russian-roulette probabilities are held fixed (both are unbiased selections).
 * Scenes with transmissive materials are not supported.  Synchronous. */
int32_t rm_adjoint(const rm_scene *scene, const rm_render_args *args, rm_frame *frame, const float *d_image,
                   const rm_grads *grads, void *stream);

/* ------------------------------------------------------ synthetic inputs */

/* Seeded procedural workloads (no reference counterpart: the reference only
 * loads assimp models).  config 0 = Cornell box, 1 = triangle soup,
 * 2 = displaced geodesic spheres under an HDR sky, 4 = textured Cornell box. */
typedef struct rm_gen_options {
    int64_t n_tris;     /* config 1 triangle count (0 -> 1,000,000) */
    int32_t grid;       /* config 2 spheres per side (0 -> 8) */
    int32_t freq;       /* config 2 geodesic frequency (0 -> 90: 162000 triangles per sphere) */
    int32_t sky_width;  /* config 2 sky (0 -> 2048x1024, -1 -> no sky) */
    int32_t sky_height;
    int32_t tex_size;   /* config 4 texture side (0 -> 256) */
} rm_gen_options;

typedef struct rm_generated rm_generated;
int32_t rm_generate_scene(int32_t config, uint32_t seed, const rm_gen_options *opt, rm_generated **out);
/* Borrowed views of the generated arrays (valid until rm_generated_destroy)
 * and the workload's RenderArgs. */
int32_t rm_generated_desc(const rm_generated *gen, rm_scene_desc *desc, rm_render_args *args);
int32_t rm_generated_destroy(rm_generated *gen);
/* Ray batch into host memory rays[n][6]: kind 0 = pinhole primary rays of
 * `cam` (n == width*height, formed as renderPixel does, src/render.cpp:466-476);
 * kind 1 = origins U[0,1]^3, directions uniform on the sphere. */
int32_t rm_generate_rays(int32_t kind, uint32_t seed, int64_t n, const rm_render_args *cam, float *rays);

/* ------------------------------------------------------------- tuning */

/* Launch parameters (no reference counterpart: the reference's only knob is
 * its thread count).  All settings give identical results; they trade speed.
 * Process-wide; set them before launching work. */
enum rm_tune_key {
    RM_TUNE_TRACE_VARIANT = 0,          /* 0 stack walk (one ray per lane), 2 persistent step walk,
                                           3 persistent step walk with the first stack entries in LDS
                                           (default; in a render launch: every walk of its kernels);
                                           1 and 4 (a one-ray step walk and two walks per lane) are
                                           retired and rejected */
    RM_TUNE_PERSISTENT_BLOCKS_PER_CU = 1, /* resident 256-lane blocks per CU for variants 2, 3 (default 8) */
    RM_TUNE_FETCH_MIN = 2,              /* idle lanes before a wave claims new rays (default 48) */
    RM_TUNE_RAY_SORT = 3,               /* ray reordering for trace calls: 0 never, 1 when consecutive
                                           origins are far apart, 2 always, 3 as 1 with a radix sort
                                           of the bins (rocPRIM) instead of the counting sort, 4 as 3
                                           with the library's own radix sort, whose kernels do no work
                                           on a batch left unsorted, 5 as 4 with a closest-hit batch of
                                           shared origins (a camera's) sorted by direction into small
                                           tiles instead (default) */
    RM_TUNE_RENDER_BLOCKS_PER_CU = 4,   /* blocks per CU of the persistent bounce kernel (default 4) */
    RM_TUNE_ADJOINT_SPLIT = 5,          /* lanes per pixel of the adjoint, each replaying a share of
                                           the pixel's samples (default 4) */
    RM_TUNE_KERNEL_TIMING = 6,          /* 1: trace calls record HIP events around their reordering
                                           and traversal kernels (default 0) */
    RM_TUNE_RENDER_MIN_WAVES = 7,       /* 1 (default), 3 or 4: resident waves per SIMD the render's
                                           shading and NEE kernels are compiled for (3 caps them
                                           at 168 VGPRs, 4 at 128, spilling the rest) */
    RM_TUNE_LDS_STACK_ENTRIES = 8,      /* trace variant 3: walk-stack entries per lane kept in LDS
                                           (1..32, default 10: 160 KiB of LDS at 8 blocks/CU; deeper
                                           entries go to scratch; 32 holds the whole stack of any tree) */
    RM_TUNE_TRIS_PER_STEP = 9           /* trace variant 3 and render walks: triangles one walk step
                                           tests (0 = a whole leaf (default); 1, 2 or 4: a leaf over
                                           several steps; the render's shading kernels use 0 or 2) */,
    RM_TUNE_NEE_SPLIT = 10,             /* render: 1 traces the shadow rays of light samples (the
                                           primary vertex's and the NEE records') in their own
                                           persistent passes, 0 inside the sampling kernels, 2 (default)
                                           splits when the launch's deepest tree has >= 12 levels */
    RM_TUNE_RENDER_COUNT = 11           /* 1: renders count the extension and shadow walks' node
                                           visits, box and triangle tests into the frame's
                                           RM_STAT_WALK_* counters (every shadow ray then goes
                                           through the queued pass; results are unchanged, the
                                           render is slower); default 0 */,
    RM_TUNE_TRACE_UTIL = 12             /* 1: trace calls of variant 3 count the lanes that walk an
                                           interior node or test triangles in each step of each wave
                                           (rm_last_trace_util; results unchanged); default 0 */,
    RM_TUNE_WALK_PHASES = 13            /* p in 1..64: in each iteration of a persistent walk (trace
                                           variant 3) only one kind of step is taken -- triangle tests
                                           when at least p/64 of the walking lanes stand at a leaf,
                                           else interior nodes -- so the wave runs one kind at a time;
                                           each ray's steps and results are unchanged; 0: every
                                           walking lane steps (default 32) */
};
int32_t rm_set_tuning(int32_t key, int32_t value);
int32_t rm_get_tuning(int32_t key, int32_t *value);
/* Durations of the last rm_trace_closest / rm_trace_any call that ran with
 * kernel timing on: its ray-reordering kernels and its traversal kernel (HIP
 * events on the call's stream; waits for them). */
int32_t rm_last_trace_ms(float *reorder_ms, float *traverse_ms);
/* Lane utilisation of the last trace call that ran with RM_TUNE_TRACE_UTIL
 * on (waits for it): counts[0] wave steps in which some lane walked an
 * interior node and counts[1] the lanes that did; counts[2] wave steps in
 * which some lane tested triangles and counts[3] those lanes.  A wave step
 * offers 64 lanes, so counts[1] / (64 counts[0]) is the interior steps'
 * utilisation. */
int32_t rm_last_trace_util(uint64_t *counts);

/* ------------------------------------------------------------ utilities */

/* Philox-4x32-10 block (the counter RNG shared by host and device code). */
void rm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif /* RAYM0NADE_HIP_H */

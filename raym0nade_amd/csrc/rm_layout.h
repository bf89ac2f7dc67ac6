// HBM layout of a flattened scene, shared by the host builder and the kernels.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace rm {

// The reference stores BVH_Node{Box; faceL, faceR} (32 B) in an implicit heap
// (children of u at 2u, 2u+1; src/bvh.cpp:44-50, include/bvh.h:7-11) and tests
// BOTH child boxes when it visits an interior node (src/bvh.cpp:71-77).  The
// GPU keeps the heap numbering but stores, per interior node u, the two child
// boxes together plus a leaf code per child, so one visit is one aligned 64-B
// read (half a 128-B cache line).
//   a = child0.lo.xyz, child0.hi.x
//   b = child0.hi.yz,  child1.lo.xy
//   c = child1.lo.z,   child1.hi.xyz
//   leaf[k] = 0 if child k is interior (its pair record is at index 2u+k),
//             else (first_face << 4) | face_count   (count in 1..10)
struct alignas(16) PairNode {
    float a[4];
    float b[4];
    float c[4];
    uint32_t leaf[2];
    uint32_t pad[2];
};
static_assert(sizeof(PairNode) == 64, "PairNode is 64 B");

// Intersection record of one triangle in BVH (leaf) order, 48 B:
//   p = v0.xyz, thr      thr: |a| below which Moller-Trumbore rejects, exact
//                        stand-in for |a| / length(edge1) < 1e-4
//                        (src/geometry.cpp:70); see flatten code.
//   e1 = v1 - v0, orig   orig: original triangle id as int bits
//   e2 = v2 - v0, 0
// Edges are the same fp32 differences the reference computes per test
// (src/geometry.cpp:65-66), so every test result is bit-identical.
struct alignas(16) TriRec {
    float p[4];
    float e1[4];
    float e2[4];
};
static_assert(sizeof(TriRec) == 48, "TriRec is 48 B");

// Walk-stack capacity: pending entries <= tree depth - 1 (2^27 faces -> 25).
constexpr int MAX_STACK = 32;

// Buckets of the sky CDF guide table.
constexpr int SKY_GUIDE = 65536;

// Leaf-code helpers.
constexpr uint32_t LEAF_SHIFT = 4;
constexpr uint32_t LEAF_MASK = 15;
constexpr int LEAF_BAG = 10;  // LeafBagSize, src/bvh.cpp:16

// Shading data of one triangle in BVH order (what the reference reaches
// through Face::data[] and Face::material, include/component.h:15-22).
struct alignas(16) TriShade {
    float v[3][3];   // vertices
    float n[3][3];   // vertex normals
    float uv[3][2];  // texture coordinates
    int32_t material;
    int32_t mesh;
    int32_t pad[2];
};
static_assert(sizeof(TriShade) % 16 == 0, "TriShade aligned");

// Material as the kernels read it.
struct alignas(16) MatRec {
    float base[3];
    float roughness;
    float emission[3];
    float metallic;
    float trans[3];
    float opacity;
    float ior;
    int32_t tex;       // -1 or the diffuse texture (fp32 or 8-bit)
    int32_t cutout;    // 1: the diffuse texture has an alpha below 1 (an alpha byte below 255)
    int32_t ntex;      // -1 or the 8-bit normal map (aiTextureType_NORMALS)
    int32_t etex;      // -1 or the 8-bit emissive map (aiTextureType_EMISSIVE)
    int32_t stex;      // -1 or the 8-bit specular map (aiTextureType_SPECULAR)
    int32_t pad[2];
};

// Texture with a mip chain (level l at offset[l], size w>>l x h>>l).
// channels 0: fp32 RGB texels, offsets in floats into the texel pool;
// channels 3 or 4: the reference's 8-bit texels in that channel layout,
// offsets in bytes into the byte pool.
struct TexRec {
    int32_t width, height, levels, channels;
    int64_t offset[8];
    int64_t alpha;      // fp32 textures: level-0 alpha plane [h][w] in the texel pool, or -1
    int64_t level_bytes[8];  // 8-bit textures: bytes of each level (reads past them read 0)
};

// One light object (LightObject, include/component.h:47-54).
struct LightRec {
    float center[3];
    float power;
    float color[3];
    int32_t first_face;  // into light_faces / light_cdf
    int32_t n_faces;
    int32_t pad[3];
};

// Device view of a scene (all pointers are device pointers).
struct SceneView {
    const PairNode *nodes;
    const TriRec *tris;
    const TriShade *shade;
    const MatRec *mats;
    const TexRec *texs;
    const float *texels;
    const LightRec *lights;
    const TriShade *light_faces;  // copies of light faces (reference copies Face into LightObject)
    const float *light_cdf;       // per light object, prefix sums of face weights
    const float *sky;             // premultiplied by pixel solid angle, [h][w][3]
    const float *sky_cdf;         // prefix sums of luminance weights
    const int32_t *sky_guide;     // [SKY_GUIDE+1] lower_bound(cdf, total*b/SKY_GUIDE)
    int64_t n_tris;
    float lo[3], hi[3];           // root box (used to bin ray origins for reordering)
    int32_t root_leaf;            // 0, or leaf code when the whole scene is one leaf
    int32_t depth;
    int32_t n_lights;
    int32_t sky_w, sky_h;
    int32_t n_mats;
    // n > 0: trace_walk keeps its first n pending entries in the launching
    // kernel's dynamic LDS (n * blockDim.x * 8 bytes); set only on the job
    // table's copies of a view for launches that allocate that LDS, 0 in the
    // scene's own view.
    int32_t walk_lds;
    // 1 when a material is an alpha cut-out: walks then apply TransparentTest.
    int32_t cutout;
    const uint8_t *bytes;         // 8-bit texels of every 8-bit texture, all mip levels
};

}  // namespace rm

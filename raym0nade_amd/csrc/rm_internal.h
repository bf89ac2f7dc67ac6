// Internal declarations shared by the host translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <memory>
#include <string>
#include <utility>
#include <vector>
#include "../../include/raym0nade_hip.h"
#include "rm_layout.h"
#include "rm_vec.h"

namespace rm {

// Bijection blockIdx -> logical block over a grid of G blocks, giving each
// XCD (the hardware deals blocks round-robin over the 8 XCDs) one contiguous
// range, so that neighbouring work shares that XCD's L2.
__host__ __device__ inline uint32_t xcd_block(uint32_t b, uint32_t G) {
    uint32_t q = G >> 3, r = G & 7u, x = b & 7u;
    uint32_t start = x * q + (x < r ? x : r);
    return start + (b >> 3);
}

// Records a failure for rm_last_error() and returns `code`.
int32_t fail(int32_t code, const char *fmt, ...);
// Maps a HIP error to RM_ERR_HIP / RM_ERR_OOM with context.
int32_t hip_fail(hipError_t e, const char *what);

#define RM_HIP(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return ::rm::hip_fail(e_, #call); \
    } while (0)

// Ensures a device is present and current; caches the answer.
int32_t ensure_device();

// Owned device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    size_t window_max = 0;  // largest need of the current window of uses (see render_jobs)
    int window_calls = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int32_t alloc(size_t n) {
        release();
        window_max = 0;
        window_calls = 0;
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            // the runtime keeps a failed allocation as its last error; take it
            // so that the next launch check does not report it again
            (void)hipGetLastError();
            p = nullptr;
            return hip_fail(e, "hipMalloc");
        }
        bytes = n;
        return RM_OK;
    }
    int32_t upload(const void *src, size_t n) {
        int32_t s = alloc(n);
        if (s) return s;
        if (n) {
            hipError_t e = hipMemcpy(p, src, n, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
        }
        return RM_OK;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Host-built, flattened scene before upload (also used by host-only tests of
// the builder through rm_scene_create's info).
// Allocator whose resize(n) leaves trivially constructible elements
// uninitialised, for record arrays whose every byte the builder writes: the
// zero fill std::vector would do first is a serial pass over gigabytes.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U>
    void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A>
    void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
};
template <class T>
using HostArray = std::vector<T, NoInitAlloc<T>>;

struct HostScene {
    int64_t n = 0;
    int32_t depth = 0;
    int64_t n_nodes = 0;      // nodeCount(1,n)+1
    uint32_t root_leaf = 0;
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    std::vector<int32_t> perm;          // slot -> original id
    std::vector<PairNode> pairs;        // indexed by interior node u
    HostArray<TriRec> tris;             // slot order
    HostArray<TriShade> shade;          // slot order
    std::vector<MatRec> mats;
    std::vector<TexRec> texs;
    std::vector<float> texels;
    std::vector<uint8_t> bytes;         // 8-bit textures, every mip level
    std::vector<V3> light_face_avg;     // per light face: getAverageEmissiveColor (src/model.cpp:23-42)
    std::vector<LightRec> lights;
    std::vector<TriShade> light_faces;
    std::vector<float> light_cdf;
    std::vector<float> sky;             // premultiplied by solid angle
    std::vector<float> sky_cdf;
    std::vector<int32_t> sky_guide;
    int32_t sky_w = 0, sky_h = 0;
    int32_t cutout = 0;                 // some material is an alpha cut-out
    int32_t maps8 = 0;                  // the scene holds 8-bit textures
};

int32_t build_host_scene(const rm_scene_desc *d, HostScene &hs);

// Launch parameters chosen by the autotuner (rm_set_tuning).
struct Tuning {
    int trace_variant = 3;        // 0 stack walk, 2 persistent step walk, 3 same with LDS stacks
    int persistent_blocks_per_cu = 8;
    int fetch_min = 48;           // idle lanes before a wave claims new rays
    int ray_sort = 5;             // 0 never, 1 when consecutive origins are far apart, 2 always,
                                  // 3 as 1 with a radix sort (rocPRIM) instead of the counting sort,
                                  // 4 as 3 with the library's own radix sort (skipped on the device
                                  // for a batch left unsorted), 5 as 4 with closest-hit batches
                                  // of shared origins sorted by direction
    int render_blocks_per_cu = 4; // resident blocks per CU of the persistent bounce kernel
    int adjoint_split = 4;        // lanes per pixel of the adjoint kernel
    int kernel_timing = 0;        // record HIP events around the trace kernels (rm_last_trace_ms)
    int render_min_waves = 1;     // 1 or 4: register cap of the shading and NEE kernels
    int lds_stack_entries = 10;   // trace variant 3: walk-stack entries per lane kept in LDS (160 KiB at 8 blocks/CU)
    int tris_per_step = 0;        // trace variant 3: triangles tested per step (0: a whole leaf)
    int nee_split = 2;            // render: shadow rays of light samples in their own passes (2: deep trees)
    int render_count = 0;         // render: count the extension and shadow walks (RM_STAT_WALK_*)
    int trace_util = 0;           // trace variant 3: count lane utilisation (rm_last_trace_util)
    int walk_phases = 32;         // persistent walks: one kind of step per iteration (RM_TUNE_WALK_PHASES)
};
Tuning &tuning();
int device_cu_count();

}  // namespace rm

struct rm_scene {
    rm::HostScene host;  // perm + sizes kept; big arrays released after upload
    rm::SceneView view{};
    rm::DevBuf nodes, tris, shade, mats, texs, texels, bytes, lights, light_faces, light_cdf, sky, sky_cdf, sky_guide;
    rm_scene_info info{};
    uint64_t version = 0;  // bumped by rm_scene_update
};

namespace rm {
// Box-filtered mip levels 1.. of texture `t` from its level 0 in the device
// texel pool, the same arithmetic as the host build (stream-ordered).
int32_t launch_mip_build(const rm_scene *s, int t, hipStream_t st);
}  // namespace rm

// Ray-batch traversal kernels behind rm_trace_closest / rm_trace_any /
// rm_trace_count.
//
// One lane per ray; rays arrive as [n][6] fp32 (origin, direction) in HBM.
// Blocks are renumbered so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) works on one contiguous range of rays: neighbouring rays of a
// coherent batch then share that XCD's L2.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <type_traits>
#include "rm_internal.h"
#include "shade.h"
#include "traverse.h"

namespace rm {

constexpr int TRACE_BLOCK = 256;
static_assert(TRACE_BLOCK == WALK_BLOCK, "LDS walk stacks assume blocks of WALK_BLOCK lanes");

__device__ __forceinline__ void load_ray(const float *rays, int64_t i, V3 &o, V3 &d) {
    const float2 *p = reinterpret_cast<const float2 *>(rays + i * 6);
    float2 a = p[0], b = p[1], c = p[2];
    o = V3{a.x, a.y, b.x};
    d = V3{b.y, c.x, c.y};
}

// What a traversal kernel writes for a closest hit in slot `slot` (BVH
// order): the original triangle id, -1 for a miss; in scenes with alpha
// cut-outs the slot itself, which k_trace_cutout continues from and maps.
__device__ __forceinline__ int32_t out_id(const SceneView &S, int32_t slot) {
    if (slot < 0) return -1;
    return S.cutout ? slot : __float_as_int(S.tris[slot].e1[3]);
}

// KIND only distinguishes instantiations by name in profiles (0 = generic).
// Ray processed by slot j: order[j] when the batch was reordered, else j.
__device__ __forceinline__ int64_t ray_of(int64_t j, const uint32_t *order, const int *use_order) {
    return (order && *use_order) ? (int64_t)order[j] : j;
}

template <int KIND>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_closest(SceneView S, const float *__restrict__ rays, int64_t n,
                                                               float tmin, float tmax0, float *__restrict__ t_out,
                                                               int32_t *__restrict__ id_out,
                                                               const uint32_t *__restrict__ order,
                                                               const int *__restrict__ use_order) {
    int64_t j = (int64_t)xcd_block(blockIdx.x, gridDim.x) * TRACE_BLOCK + threadIdx.x;
    if (j >= n) return;
    int64_t i = ray_of(j, order, use_order);
    V3 o, d;
    load_ray(rays, i, o, d);
    RayPre r = ray_pre(o, d);
    float tmax = tmax0;
    TraceCounts c{};
    int32_t best = trace<false, false>(S, r, tmin, tmax, -INFINITY, c);
    t_out[i] = tmax;
    id_out[i] = out_id(S, best);
}

template <int KIND>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_any(SceneView S, const float *__restrict__ rays, int64_t n,
                                                           float tmin, float aim0, const float *__restrict__ aim_arr,
                                                           uint8_t *__restrict__ occ, const uint32_t *__restrict__ order,
                                                           const int *__restrict__ use_order) {
    int64_t j = (int64_t)xcd_block(blockIdx.x, gridDim.x) * TRACE_BLOCK + threadIdx.x;
    if (j >= n) return;
    int64_t i = ray_of(j, order, use_order);
    V3 o, d;
    load_ray(rays, i, o, d);
    RayPre r = ray_pre(o, d);
    float aim = aim_arr ? aim_arr[i] : aim0;
    // HitRecord hit(eps_zero, aimDepth + eps_zero) and `t_max >= aimDepth -> false`
    float tmax = aim + T_EPS;
    TraceCounts c{};
    trace<true, false>(S, r, tmin, tmax, aim, c);
    occ[i] = (tmax < aim) ? 1 : 0;
}

// Alpha cut-outs after the walk of a batch (TransparentTest,
// src/model.cpp:217-230): a ray whose closest hit is transparent continues
// through the rest of Model::rayHit (src/model.cpp:332-341); a blocked
// shadow ray runs Model::rayHit_test (src/model.cpp:343-354).  Launched only for scenes
// with a cut-out material, so the traversal kernels stay as they are.
template <bool ANY>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_cutout(SceneView S, const float *__restrict__ rays, int64_t n,
                                                              float tmin, float t0, const float *__restrict__ aim_arr,
                                                              float *t_out, int32_t *id_out, uint8_t *occ) {
    const int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x;
    if (i >= n) return;
    V3 o, d;
    load_ray(rays, i, o, d);
    if constexpr (ANY) {
        if (!occ[i]) return;
        const float aim = aim_arr ? aim_arr[i] : t0;
        occ[i] = model_blocked<false, 0>(S, o, d, tmin, aim) ? 1 : 0;
    } else {
        // the batch walk's closest hit (its slot): only a transparent one
        // walks again
        const int32_t first = id_out[i];
        if (first < 0) return;
        float t = t_out[i];
        const int32_t slot = model_hit_rest<false, 0>(S, o, d, t, first, t0);
        t_out[i] = t;
        id_out[i] = slot >= 0 ? __float_as_int(S.tris[slot].e1[3]) : -1;
    }
}

// Persistent lanes with dynamic ray fetch: a lane whose ray finished takes the
// next unclaimed ray; a wave claims rays for its idle lanes with one atomic
// once at least `fetch_min` lanes are idle (or all of its live lanes are).
// Every wave leaves the loop when the batch is exhausted and its lanes are
// done, so the grid always drains.  LDS: the walker's first lds_entries
// pending entries live in the block's dynamic LDS (lds_entries * TRACE_BLOCK *
// 8 bytes), deeper ones in per-lane scratch.  TPS: triangles per step (0: a
// whole leaf per step).
// UTIL: count lane utilisation into util[4]: loop iterations of a wave in
// which some lane walks an interior node, the lanes doing so, iterations in
// which some lane tests triangles, and those lanes.
template <bool ANY, bool LDS, int TPS, bool UTIL = false>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_persistent(SceneView S, const float *__restrict__ rays, int64_t n,
                                                                  float tmin, float tmax0,
                                                                  const float *__restrict__ aim_arr,
                                                                  float *__restrict__ t_out, int32_t *__restrict__ id_out,
                                                                  uint8_t *__restrict__ occ,
                                                                  unsigned long long *__restrict__ next, int fetch_min,
                                                                  const uint32_t *__restrict__ order,
                                                                  const int *__restrict__ use_order, int lds_entries,
                                                                  unsigned long long *__restrict__ util = nullptr,
                                                                  int phases = 0) {
    const uint32_t lane = threadIdx.x & 63u;
    ScratchEntries ent;
    Walker<ANY, false, typename std::conditional<LDS, LdsStack, PrivateStack>::type, TPS> w;
    if constexpr (LDS)
        w.stk.bind(&ent, lds_entries);
    else
        w.stk.bind(&ent);
    RayPre r;
    float aim = -INFINITY;
    int64_t ray = -1;
    bool exhausted = false;
    TraceCounts c{};
    unsigned long long u[4] = {0, 0, 0, 0};
    while (true) {
        bool idle = ray < 0;
        uint64_t idle_mask = __ballot(idle);
        uint64_t busy_mask = __ballot(!idle);
        if (exhausted && busy_mask == 0) break;
        if (!exhausted && idle_mask && (__popcll(idle_mask) >= fetch_min || busy_mask == 0)) {
            uint32_t cnt = (uint32_t)__popcll(idle_mask);
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle_mask) - 1u;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(next, (unsigned long long)cnt);
            base = __shfl(base, (int)leader, 64);
            if (base + cnt >= (unsigned long long)n) exhausted = true;
            if (idle) {
                int64_t slot = (int64_t)base + __popcll(idle_mask & ((1ull << lane) - 1ull));
                if (slot < n) {
                    int64_t k = ray_of(slot, order, use_order);
                    ray = k;
                    V3 o, d;
                    load_ray(rays, k, o, d);
                    r = ray_pre(o, d);
                    aim = ANY ? (aim_arr ? aim_arr[k] : tmax0) : -INFINITY;
                    w.init(S, r, ANY ? aim + T_EPS : tmax0);
                }
            }
        }
        const bool turn = walk_turn(w.cur, phases, ray >= 0);
        const bool go = ray >= 0 && turn;
        const bool leaf = ray >= 0 && (w.cur & 0x80000000u);
        if constexpr (UTIL) {
            const uint64_t mi = __ballot(go && !leaf), ml = __ballot(go && leaf);
            if (mi) { u[0] += 1; u[1] += (unsigned long long)__popcll(mi); }
            if (ml) { u[2] += 1; u[3] += (unsigned long long)__popcll(ml); }
        }
        if (go && w.step(S, r, tmin, aim, c)) {
            if (ANY) {
                occ[ray] = (w.tmax < aim) ? 1 : 0;
            } else {
                t_out[ray] = w.tmax;
                id_out[ray] = out_id(S, w.best);
            }
            ray = -1;
        }
    }
    if constexpr (UTIL) {
        if (lane == 0)
            for (int k = 0; k < 4; ++k)
                if (u[k]) atomicAdd(util + k, u[k]);
    }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// ---------------------------------------------------------------- reordering
// Each ray's walk depends only on the ray, so the batch may be processed in
// any order.  Rays whose origins lie close together walk the same subtrees;
// when consecutive rays of a batch are far apart, the batch is processed in
// the order of a counting sort by (Morton cell of the origin in the root box,
// 32^3 cells) x (direction octant).  Results are written back by original
// index.  The decision is made on the device (k_coherence), so no host
// round trip is needed.
constexpr int CELL_BITS = 5;
constexpr uint32_t NBINS = 1u << (3 * CELL_BITS + 3);

__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    uint32_t r = 0;
    for (int b = 0; b < CELL_BITS; ++b) r |= ((x >> b) & 1u) << (3 * b);
    return r;
}

__device__ __forceinline__ uint32_t ray_bin(const SceneView &S, V3 o, V3 d) {
    const float c[3] = {o.x, o.y, o.z};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float ext = S.hi[k] - S.lo[k];
        float f = ext > 0.0f ? (c[k] - S.lo[k]) / ext : 0.0f;
        int v = (int)(f * float(1 << CELL_BITS));
        v = v < 0 ? 0 : (v > (1 << CELL_BITS) - 1 ? (1 << CELL_BITS) - 1 : v);
        q[k] = (uint32_t)v;
    }
    uint32_t m = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    // the octants of a cell in Gray-code order (0, 1, 3, 2, 6, 7, 5, 4):
    // consecutive ones differ in one direction sign
    oct = oct ^ (oct >> 1) ^ (oct >> 2);
    return (m << 3) | oct;
}

// Octahedral map of a direction into [-1, 1]^2.
__device__ __forceinline__ void oct_map(V3 d, float &u, float &v) {
    const float a = fabsf(d.x) + fabsf(d.y) + fabsf(d.z);
    u = a > 0.0f ? d.x / a : 0.0f;
    v = a > 0.0f ? d.y / a : 0.0f;
    if (d.z < 0.0f) {
        const float uu = (1.0f - fabsf(v)) * (u < 0.0f ? -1.0f : 1.0f);
        v = (1.0f - fabsf(u)) * (v < 0.0f ? -1.0f : 1.0f);
        u = uu;
    }
}

// Direction bin of a ray whose batch shares its origins: the Hilbert index of
// its octahedral map over the batch's range `box` (u0, v0, 512 / width,
// 512 / height) on a 512 x 512 grid, so that consecutive bins are
// neighbouring directions (a camera batch falls into small square tiles).
__device__ __forceinline__ uint32_t dir_bin(V3 d, const float *box) {
    float u, v;
    oct_map(d, u, v);
    int qu = (int)((u - box[0]) * box[2]), qv = (int)((v - box[1]) * box[3]);
    qu = qu < 0 ? 0 : (qu > 511 ? 511 : qu);
    qv = qv < 0 ? 0 : (qv > 511 ? 511 : qv);
    // Hilbert index on the 512 x 512 grid: consecutive bins always share an
    // edge (a Morton order jumps at every power-of-two boundary), so the
    // 64 rays of a wave, two or three bins, form a compact tile
    uint32_t x = (uint32_t)qu, y = (uint32_t)qv, h = 0;
    for (uint32_t sq = 256; sq > 0; sq >>= 1) {
        const uint32_t rx = (x & sq) ? 1u : 0u, ry = (y & sq) ? 1u : 0u;
        h += sq * sq * ((3u * rx) ^ ry);
        if (!ry) {
            if (rx) {
                x = 511u - x;
                y = 511u - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return h;
}

// flag = 1 when the mean distance between consecutive ray origins (4096
// evenly spaced pairs) exceeds `frac` of the root box diagonal; otherwise 2
// when `box` is given (the batch is then sorted by direction: box receives
// the range of the sampled directions' octahedral map for dir_bin), else 0.
__global__ __launch_bounds__(256) void k_coherence(SceneView S, const float *__restrict__ rays, int64_t n, float frac,
                                                   int *__restrict__ flag, float *__restrict__ box = nullptr) {
    __shared__ float part[4], lo_u[4], hi_u[4], lo_v[4], hi_v[4];
    float acc = 0.0f, u0 = INFINITY, u1 = -INFINITY, v0 = INFINITY, v1 = -INFINITY;
    if (n >= 2) {
        for (int s = threadIdx.x; s < 4096; s += 256) {
            int64_t k = (int64_t)s * (n - 2) / 4095;
            V3 o0, d0, o1, d1;
            load_ray(rays, k, o0, d0);
            load_ray(rays, k + 1, o1, d1);
            acc += length(o1 - o0);
            if (box) {
                float u, v;
                oct_map(d0, u, v);
                u0 = fminf(u0, u), u1 = fmaxf(u1, u), v0 = fminf(v0, v), v1 = fmaxf(v1, v);
                oct_map(d1, u, v);
                u0 = fminf(u0, u), u1 = fmaxf(u1, u), v0 = fminf(v0, v), v1 = fmaxf(v1, v);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc += __shfl_down(acc, off, 64);
        u0 = fminf(u0, __shfl_down(u0, off, 64));
        u1 = fmaxf(u1, __shfl_down(u1, off, 64));
        v0 = fminf(v0, __shfl_down(v0, off, 64));
        v1 = fmaxf(v1, __shfl_down(v1, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        part[w] = acc, lo_u[w] = u0, hi_u[w] = u1, lo_v[w] = v0, hi_v[w] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float mean = (part[0] + part[1] + part[2] + part[3]) / 4096.0f;
        V3 ext{S.hi[0] - S.lo[0], S.hi[1] - S.lo[1], S.hi[2] - S.lo[2]};
        const bool spread = n >= 2 && mean > frac * length(ext);
        *flag = spread ? 1 : (box && n >= 2 ? 2 : 0);
        if (box) {
            for (int w = 1; w < 4; ++w) {
                lo_u[0] = fminf(lo_u[0], lo_u[w]), hi_u[0] = fmaxf(hi_u[0], hi_u[w]);
                lo_v[0] = fminf(lo_v[0], lo_v[w]), hi_v[0] = fmaxf(hi_v[0], hi_v[w]);
            }
            // 512 cells over the sampled range; a degenerate range gets one cell
            const float du = hi_u[0] - lo_u[0], dv = hi_v[0] - lo_v[0];
            box[0] = lo_u[0];
            box[1] = lo_v[0];
            box[2] = du > 0.0f ? 512.0f / du : 0.0f;
            box[3] = dv > 0.0f ? 512.0f / dv : 0.0f;
        }
    }
}

__global__ __launch_bounds__(256) void k_zero_bins(uint32_t *__restrict__ hist) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < NBINS) hist[i] = 0u;
}

__global__ __launch_bounds__(256) void k_bin_hist(SceneView S, const float *__restrict__ rays, int64_t n,
                                                  const int *__restrict__ flag, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ hist) {
    if (!*flag) return;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    V3 o, d;
    load_ray(rays, i, o, d);
    uint32_t b = ray_bin(S, o, d);
    keys[i] = b;
    atomicAdd(hist + b, 1u);
}

// Keys for the radix-sort reordering (ray_sort 3): the bin when the batch is
// incoherent, else 0 (a stable sort then keeps the batch's order).
// skip: leave the pairs unwritten on a batch left unsorted (sort mode 4,
// whose kernels then read nothing)
__global__ __launch_bounds__(256) void k_bin_keys(SceneView S, const float *__restrict__ rays, int64_t n,
                                                  const int *__restrict__ flag, uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ vals, int skip = 0,
                                                  const float *__restrict__ box = nullptr) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || (skip && !*flag)) return;
    uint32_t b = 0;
    if (*flag) {
        V3 o, d;
        load_ray(rays, i, o, d);
        b = *flag == 2 ? dir_bin(d, box) : ray_bin(S, o, d);
    }
    keys[i] = b;
    vals[i] = (uint32_t)i;
}

// exclusive scan of NBINS counters in place, one block of 1024 lanes
__global__ __launch_bounds__(1024) void k_bin_scan(const int *__restrict__ flag, uint32_t *__restrict__ hist) {
    if (!*flag) return;
    __shared__ uint32_t sh[1024];
    constexpr uint32_t PER = NBINS / 1024;
    uint32_t base = threadIdx.x * PER, sum = 0;
    for (uint32_t k = 0; k < PER; ++k) sum += hist[base + k];
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int s = 1; s < 1024; s <<= 1) {
        uint32_t add = threadIdx.x >= (unsigned)s ? sh[threadIdx.x - s] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - sum;
    for (uint32_t k = 0; k < PER; ++k) {
        uint32_t c = hist[base + k];
        hist[base + k] = run;
        run += c;
    }
}

__global__ __launch_bounds__(256) void k_bin_scatter(int64_t n, const int *__restrict__ flag,
                                                     const uint32_t *__restrict__ keys, uint32_t *__restrict__ hist,
                                                     uint32_t *__restrict__ order) {
    if (!*flag) return;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t pos = atomicAdd(hist + keys[i], 1u);
    order[pos] = (uint32_t)i;
}

// Sort mode 4: a stable least-significant-digit radix sort of the
// (bin, index) pairs in two passes of 9 bits, whose kernels all return at
// once on a batch that k_coherence left unsorted.  Tiles of RS_TILE pairs,
// one per block: per pass a digit histogram of each tile (digit-major,
// hist[d * nt + tile]), an exclusive scan of that array in chunks of
// RS_CHUNK entries plus a scan of the chunk totals, and a scatter.  The
// scatter ranks the tile stably (each wave a contiguous quarter, with
// wave-private digit counters and ballots), sorts it by digit in LDS and
// writes each digit's run contiguously; consecutive tiles go to blocks of
// one XCD so that the runs of neighbouring tiles meet in one L2.
constexpr int RS_BITS = 9;
constexpr uint32_t RS_DIGITS = 1u << RS_BITS;
constexpr int RS_TILE = 4096;      // 256 threads x 16 pairs
constexpr int RS_CHUNK = 4096;     // scan chunk (256 threads x 16 entries)
static_assert(2 * RS_BITS >= 3 * CELL_BITS + 3, "two radix passes cover the bin key");
static_assert(RS_DIGITS == 2 * 256 && RS_TILE == 16 * 256, "k_rs_scatter: two digits and 16 pairs per lane");

__global__ __launch_bounds__(256) void k_rs_hist(const int *__restrict__ flag, const uint32_t *__restrict__ keys,
                                                 int64_t n, int shift, uint32_t nt, uint32_t *__restrict__ hist) {
    if (!*flag) return;
    __shared__ uint32_t cnt[RS_DIGITS];
    for (uint32_t d = threadIdx.x; d < RS_DIGITS; d += 256) cnt[d] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_TILE / 256; ++k) {
        const int64_t i = base + k * 256 + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (RS_DIGITS - 1)], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < RS_DIGITS; d += 256) hist[(size_t)d * nt + blockIdx.x] = cnt[d];
}

// Exclusive scan of a[0, m) within chunks of RS_CHUNK entries (in place);
// part[c] = the total of chunk c.
__global__ __launch_bounds__(256) void k_rs_scan_chunks(const int *__restrict__ flag, uint32_t *__restrict__ a,
                                                        size_t m, uint32_t *__restrict__ part) {
    if (!*flag) return;
    __shared__ uint32_t sh[256];
    const size_t c0 = (size_t)blockIdx.x * RS_CHUNK, b = c0 + (size_t)threadIdx.x * 16;
    uint32_t v[16], sum = 0;
    for (int k = 0; k < 16; ++k) {
        v[k] = b + k < m ? a[b + k] : 0u;
        sum += v[k];
    }
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - sum;
    for (int k = 0; k < 16; ++k)
        if (b + k < m) {
            a[b + k] = run;
            run += v[k];
        }
    if (threadIdx.x == 255) part[blockIdx.x] = sh[255];
}

// Exclusive scan of part[0, nc) in place, one block.
__global__ __launch_bounds__(256) void k_rs_scan_parts(const int *__restrict__ flag, uint32_t *__restrict__ part,
                                                       uint32_t nc) {
    if (!*flag) return;
    __shared__ uint32_t sh[256];
    const uint32_t per = (nc + 255) / 256, b = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nc) sum += part[b + k];
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - sum;
    for (uint32_t k = 0; k < per; ++k)
        if (b + k < nc) {
            const uint32_t c = part[b + k];
            part[b + k] = run;
            run += c;
        }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Tile handled by block b of nt: blocks are dealt to the 8 XCDs round-robin,
// so block b runs on XCD b % 8; the first 8 * (nt / 8) blocks take tiles
// such that XCD x holds the consecutive tiles [x * q, (x + 1) * q).
__device__ __forceinline__ uint32_t rs_tile(uint32_t b, uint32_t nt) {
    const uint32_t q = nt >> 3;
    return b < 8u * q ? (b & 7u) * q + (b >> 3) : b;
}

__global__ __launch_bounds__(256) void k_rs_scatter(const int *__restrict__ flag, const uint32_t *__restrict__ kin,
                                                    const uint32_t *__restrict__ vin, int64_t n, int shift,
                                                    uint32_t nt, const uint32_t *__restrict__ hist,
                                                    const uint32_t *__restrict__ part, uint32_t *__restrict__ kout,
                                                    uint32_t *__restrict__ vout) {
    if (!*flag) return;
    constexpr int PER = RS_TILE / 256;  // pairs per lane
    __shared__ uint32_t cnt[4][RS_DIGITS];  // per wave: running count, then its first local slot
    __shared__ int64_t gbase[RS_DIGITS];    // global slot of local slot 0 for each digit
    __shared__ uint32_t sh[256];
    __shared__ uint32_t lk[RS_TILE], lv[RS_TILE];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t tile = rs_tile(blockIdx.x, nt);
    const int64_t base = (int64_t)tile * RS_TILE;
    const int len = (int)(n - base < RS_TILE ? n - base : RS_TILE);
    for (uint32_t d = threadIdx.x; d < RS_DIGITS; d += 256) cnt[0][d] = cnt[1][d] = cnt[2][d] = cnt[3][d] = 0u;
    __syncthreads();
    // rank within the wave's quarter: pair w * 1024 + k * 64 + lane
    uint32_t key[PER], val[PER], rank[PER];
    for (int k = 0; k < PER; ++k) {
        const int j = (int)w * (RS_TILE / 4) + k * 64 + (int)lane;
        const bool valid = j < len;
        key[k] = valid ? kin[base + j] : 0u;
        val[k] = valid ? vin[base + j] : 0u;
    }
    for (int k = 0; k < PER; ++k) {
        const int j = (int)w * (RS_TILE / 4) + k * 64 + (int)lane;
        const bool valid = j < len;
        const uint32_t d = (key[k] >> shift) & (RS_DIGITS - 1);
        uint64_t m = __ballot(valid);
        for (int b = 0; b < RS_BITS; ++b) {
            const uint64_t bb = __ballot(valid && ((d >> b) & 1u));
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t r = valid ? cnt[w][d] + (uint32_t)__popcll(m & below) : 0u;
        wave_sync();
        if (valid && (m >> lane) == 1ull) cnt[w][d] = r + 1u;  // the last lane of its digit
        wave_sync();
        rank[k] = r;
    }
    __syncthreads();
    // tile-local digit offsets (exclusive scan over digits, two per thread)
    const uint32_t d0 = 2u * threadIdx.x, d1 = d0 + 1u;
    const uint32_t t0 = cnt[0][d0] + cnt[1][d0] + cnt[2][d0] + cnt[3][d0];
    const uint32_t t1 = cnt[0][d1] + cnt[1][d1] + cnt[2][d1] + cnt[3][d1];
    sh[threadIdx.x] = t0 + t1;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    const uint32_t off0 = sh[threadIdx.x] - (t0 + t1), off1 = off0 + t0;
    {
        const size_t e0 = (size_t)d0 * nt + tile, e1 = (size_t)d1 * nt + tile;
        gbase[d0] = (int64_t)hist[e0] + part[e0 / RS_CHUNK] - (int64_t)off0;
        gbase[d1] = (int64_t)hist[e1] + part[e1 / RS_CHUNK] - (int64_t)off1;
        uint32_t a0 = off0, a1 = off1;
        for (int v = 0; v < 4; ++v) {
            const uint32_t c0 = cnt[v][d0], c1 = cnt[v][d1];
            cnt[v][d0] = a0;
            cnt[v][d1] = a1;
            a0 += c0;
            a1 += c1;
        }
    }
    __syncthreads();
    for (int k = 0; k < PER; ++k) {
        const int j = (int)w * (RS_TILE / 4) + k * 64 + (int)lane;
        if (j < len) {
            const uint32_t d = (key[k] >> shift) & (RS_DIGITS - 1);
            const uint32_t p = cnt[w][d] + rank[k];
            lk[p] = key[k];
            lv[p] = val[k];
        }
    }
    __syncthreads();
    for (int k = 0; k < PER; ++k) {
        const int j = k * 256 + (int)threadIdx.x;
        if (j < len) {
            const uint32_t kk = lk[j];
            const int64_t g = gbase[(kk >> shift) & (RS_DIGITS - 1)] + j;
            if (g >= 0 && g < n) {
                if (kout) kout[g] = kk;  // the last pass needs the values only
                vout[g] = lv[j];
            }
        }
    }
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_count(SceneView S, const float *__restrict__ rays, int64_t n,
                                                             float tmin, float tmax0, int any_hit,
                                                             unsigned long long *__restrict__ counts) {
    int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x;
    TraceCounts c{};
    if (i < n) {
        V3 o, d;
        load_ray(rays, i, o, d);
        RayPre r = ray_pre(o, d);
        if (any_hit) {
            float aim = tmax0;
            float tmax = aim + T_EPS;
            trace<true, true>(S, r, tmin, tmax, aim, c);
        } else {
            float tmax = tmax0;
            trace<false, true>(S, r, tmin, tmax, -INFINITY, c);
        }
    }
    // wave reduction then one atomic per wave
    unsigned long long v[4] = {c.interior, c.leaves, c.boxes, c.tris};
    for (int k = 0; k < 4; ++k) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(counts + k, x);
    }
}

}  // namespace rm

using namespace rm;

static int32_t check_trace_args(const rm_scene *s, const float *rays, int64_t n, const void *out1, const char *fn) {
    if (!s) return fail(RM_ERR_INVALID, "%s: scene is NULL", fn);
    if (n < 0) return fail(RM_ERR_INVALID, "%s: negative ray count", fn);
    if (n > 0 && (!rays || !out1)) return fail(RM_ERR_INVALID, "%s: NULL buffer", fn);
    if ((int64_t)((n + TRACE_BLOCK - 1) / TRACE_BLOCK) > 0x7fffffffLL) return fail(RM_ERR_INVALID, "%s: too many rays", fn);
    if (s->view.depth > MAX_STACK) return fail(RM_ERR_UNSUPPORTED, "%s: tree depth %d exceeds %d", fn, s->view.depth, MAX_STACK);
    return RM_OK;
}

// Claim counters of the persistent kernels: one slot per call in flight
// (rotating over 64), cleared by a kernel on the call's stream before use.
// Calls on one device that overlap in time use different slots unless more
// than 64 are in flight at once.
__device__ __align__(256) unsigned long long g_ray_next[64];

static int32_t ray_counter(unsigned long long **out) {
    static unsigned long long *cached[64] = {nullptr};
    static std::atomic<uint32_t> slot{0};
    int dev = 0;
    RM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(RM_ERR_UNSUPPORTED, "device index %d", dev);
    if (!cached[dev]) {
        void *p = nullptr;
        RM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_ray_next)));
        cached[dev] = static_cast<unsigned long long *>(p);
    }
    *out = cached[dev] + (slot.fetch_add(1) & 63u);
    return RM_OK;
}

// Lane-utilisation counters of the last trace call that ran with
// RM_TUNE_TRACE_UTIL on (variant 3), one set per device.
__device__ __align__(256) unsigned long long g_trace_util[4];
__global__ void k_zero_util(unsigned long long *u) {
    if (threadIdx.x < 4) u[threadIdx.x] = 0ull;
}
static bool g_util_valid = false;

// HIP events bracketing the reordering kernels and the traversal kernel of
// the last trace call (recorded when kernel timing is on; rm_last_trace_ms).
struct TraceEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    bool recorded = false;
};
static TraceEvents &trace_events() {
    static TraceEvents ev;
    return ev;
}
static int32_t mark(int k, hipStream_t st) {
    TraceEvents &ev = trace_events();
    if (!ev.e[k]) RM_HIP(hipEventCreate(&ev.e[k]));
    RM_HIP(hipEventRecord(ev.e[k], st));
    return RM_OK;
}

// Reordering workspace per (device, stream): calls queued on one stream run in
// order, so they can share it; the lock covers the enqueueing only.
struct TraceWork {
    std::mutex mu;
    DevBuf buf;
};
static TraceWork *trace_work(hipStream_t st) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    static std::mutex mu;
    static auto *table = new std::map<std::pair<int, uintptr_t>, TraceWork *>();
    std::lock_guard<std::mutex> lock(mu);
    TraceWork *&w = (*table)[{dev, (uintptr_t)st}];
    if (!w) w = new TraceWork();
    return w;
}

// A caller-provided reordering workspace (rm_trace_closest_noalloc): the
// call then allocates nothing, records no events, and fails when the
// workspace is too small.
struct FixedWorkspace {
    char *base;
    size_t bytes;
};

template <bool ANY>
static int32_t launch_trace(const rm_scene *s, const float *rays, int64_t n, float tmin, float t0, const float *aim_dev,
                            float *t_out, int32_t *id_out, uint8_t *occ, hipStream_t st,
                            const FixedWorkspace *fixed = nullptr) {
    const Tuning &T = tuning();
    unsigned grid = (unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK);
    const bool timing = T.kernel_timing != 0 && !fixed;
    int32_t mrc;
    if (timing && (mrc = mark(0, st))) return mrc;
    // reordering workspace: keys[n], order[n], bins, flag
    uint32_t *order = nullptr;
    int *flag = nullptr;
    char *ws = nullptr;
    std::unique_lock<std::mutex> ws_lock;
    if ((T.ray_sort == 4 || T.ray_sort == 5) && n >= 4096) {
        // the library's radix sort: keys, keys2, values, values2, flag, tile histograms, chunk totals
        const size_t plane = ((size_t)n * 4 + 255) & ~(size_t)255;
        const uint32_t nt = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
        const size_t m = (size_t)RS_DIGITS * nt;
        const uint32_t nc = (uint32_t)((m + RS_CHUNK - 1) / RS_CHUNK);
        const size_t hbytes = (m * 4 + 255) & ~(size_t)255;
        const size_t bytes = 4 * plane + 256 + hbytes + (size_t)nc * 4;
        char *base = nullptr;
        if (fixed) {
            if (fixed->bytes < bytes)
                return fail(RM_ERR_INVALID, "trace: %lld rays need %zu workspace bytes, %zu available", (long long)n,
                            bytes, fixed->bytes);
            base = fixed->base;
        } else {
            TraceWork *w = trace_work(st);
            ws_lock = std::unique_lock<std::mutex>(w->mu);
            if (w->buf.bytes < bytes) {
                if (w->buf.p) RM_HIP(hipStreamSynchronize(st));  // earlier calls on this stream may still read it
                int32_t rc = w->buf.alloc(bytes);
                if (rc) return rc;
            }
            base = (char *)w->buf.p;
        }
        uint32_t *keys = (uint32_t *)base, *keys2 = (uint32_t *)(base + plane);
        uint32_t *vals = (uint32_t *)(base + 2 * plane), *vals2 = (uint32_t *)(base + 3 * plane);
        flag = (int *)(base + 4 * plane);
        uint32_t *hist = (uint32_t *)(base + 4 * plane + 256);
        uint32_t *part = (uint32_t *)(base + 4 * plane + 256 + hbytes);
        const unsigned gn = (unsigned)((n + 255) / 256);
        // mode 5: flag, then the direction range of a coherent batch, in the flag's 256-byte slot;
        // closest-hit calls only (any-hit walks stop at the first hit and do not pay for the sort)
        float *box = (T.ray_sort == 5 && !ANY) ? reinterpret_cast<float *>(flag + 1) : nullptr;
        hipLaunchKernelGGL(k_coherence, dim3(1), dim3(256), 0, st, s->view, rays, n, 0.02f, flag, box);
        hipLaunchKernelGGL(k_bin_keys, dim3(gn), dim3(256), 0, st, s->view, rays, n, (const int *)flag, keys, vals, 1,
                           (const float *)box);
        for (int pass = 0; pass < 2; ++pass) {
            const int shift = pass * RS_BITS;
            const uint32_t *kin = pass ? keys2 : keys, *vin = pass ? vals2 : vals;
            uint32_t *kout = pass ? nullptr : keys2, *vout = pass ? vals : vals2;
            hipLaunchKernelGGL(k_rs_hist, dim3(nt), dim3(256), 0, st, (const int *)flag, kin, n, shift, nt, hist);
            hipLaunchKernelGGL(k_rs_scan_chunks, dim3(nc), dim3(256), 0, st, (const int *)flag, hist, m, part);
            hipLaunchKernelGGL(k_rs_scan_parts, dim3(1), dim3(256), 0, st, (const int *)flag, part, nc);
            hipLaunchKernelGGL(k_rs_scatter, dim3(nt), dim3(256), 0, st, (const int *)flag, kin, vin, n, shift, nt,
                               (const uint32_t *)hist, (const uint32_t *)part, kout, vout);
        }
        order = vals;  // after the second pass
    } else if (T.ray_sort == 3 && n >= 4096) {
        // radix sort of (bin, index) pairs: keys, sorted keys, indices, order, flag, sort storage
        size_t tmp_bytes = 0;
        RM_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                         (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0u,
                                         (unsigned)(3 * CELL_BITS + 3), st));
        const size_t plane = ((size_t)n * 4 + 255) & ~(size_t)255;
        const size_t bytes = 4 * plane + 256 + tmp_bytes;
        char *base = nullptr;
        if (fixed) {
            if (fixed->bytes < bytes)
                return fail(RM_ERR_INVALID, "trace: %lld rays need %zu workspace bytes, %zu available", (long long)n,
                            bytes, fixed->bytes);
            base = fixed->base;
        } else {
            TraceWork *w = trace_work(st);
            ws_lock = std::unique_lock<std::mutex>(w->mu);
            if (w->buf.bytes < bytes) {
                if (w->buf.p) RM_HIP(hipStreamSynchronize(st));  // earlier calls on this stream may still read it
                int32_t rc = w->buf.alloc(bytes);
                if (rc) return rc;
            }
            base = (char *)w->buf.p;
        }
        uint32_t *keys = (uint32_t *)base, *keys2 = (uint32_t *)(base + plane), *vals = (uint32_t *)(base + 2 * plane);
        order = (uint32_t *)(base + 3 * plane);
        flag = (int *)(base + 4 * plane);
        void *tmp = base + 4 * plane + 256;
        unsigned gn = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_coherence, dim3(1), dim3(256), 0, st, s->view, rays, n, 0.02f, flag);
        hipLaunchKernelGGL(k_bin_keys, dim3(gn), dim3(256), 0, st, s->view, rays, n, (const int *)flag, keys, vals);
        RM_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, (const uint32_t *)keys, keys2, (const uint32_t *)vals, order,
                                         (size_t)n, 0u, (unsigned)(3 * CELL_BITS + 3), st));
    } else if (T.ray_sort != 0 && n >= 4096) {
        size_t bytes = (size_t)n * 8 + (size_t)NBINS * 4 + 256;
        if (fixed) {
            if (fixed->bytes < bytes)
                return fail(RM_ERR_INVALID, "trace: %lld rays need %zu workspace bytes, %zu available", (long long)n,
                            bytes, fixed->bytes);
            ws = fixed->base;
        } else {
            TraceWork *w = trace_work(st);
            ws_lock = std::unique_lock<std::mutex>(w->mu);
            if (w->buf.bytes < bytes) {
                if (w->buf.p) RM_HIP(hipStreamSynchronize(st));  // earlier calls on this stream may still read it
                int32_t rc = w->buf.alloc(bytes);
                if (rc) return rc;
            }
            ws = (char *)w->buf.p;
        }
        uint32_t *keys = reinterpret_cast<uint32_t *>(ws);
        order = keys + n;
        uint32_t *hist = order + n;
        flag = reinterpret_cast<int *>(hist + NBINS);
        float frac = T.ray_sort == 2 ? -1.0f : 0.02f;
        unsigned gn = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_coherence, dim3(1), dim3(256), 0, st, s->view, rays, n, frac, flag);
        hipLaunchKernelGGL(k_zero_bins, dim3(NBINS / 256), dim3(256), 0, st, hist);
        hipLaunchKernelGGL(k_bin_hist, dim3(gn), dim3(256), 0, st, s->view, rays, n, (const int *)flag, keys, hist);
        hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, (const int *)flag, hist);
        hipLaunchKernelGGL(k_bin_scatter, dim3(gn), dim3(256), 0, st, n, (const int *)flag, (const uint32_t *)keys, hist,
                           order);
    }
    if (timing && (mrc = mark(1, st))) return mrc;
    if (T.trace_variant >= 2) {
        unsigned long long *next = nullptr;
        int32_t rc = ray_counter(&next);
        if (rc) return rc;
        unsigned pgrid = (unsigned)std::min<int64_t>(grid, (int64_t)device_cu_count() * T.persistent_blocks_per_cu);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, next);
        const int nl = std::min(std::max(1, (int)s->view.depth), T.lds_stack_entries);
        if (T.trace_variant == 3) {
            const bool util = T.trace_util != 0 && !fixed;
            unsigned long long *ud = nullptr;
            if (util) {
                void *p = nullptr;
                RM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_trace_util)));
                ud = static_cast<unsigned long long *>(p);
                hipLaunchKernelGGL(k_zero_util, dim3(1), dim3(64), 0, st, ud);
                g_util_valid = true;
            }
            auto kern = util ? (T.tris_per_step == 1   ? k_trace_persistent<ANY, true, 1, true>
                                : T.tris_per_step == 2 ? k_trace_persistent<ANY, true, 2, true>
                                : T.tris_per_step == 4 ? k_trace_persistent<ANY, true, 4, true>
                                                       : k_trace_persistent<ANY, true, 0, true>)
                             : (T.tris_per_step == 1   ? k_trace_persistent<ANY, true, 1>
                                : T.tris_per_step == 2 ? k_trace_persistent<ANY, true, 2>
                                : T.tris_per_step == 4 ? k_trace_persistent<ANY, true, 4>
                                                       : k_trace_persistent<ANY, true, 0>);
            hipLaunchKernelGGL(kern, dim3(pgrid), dim3(TRACE_BLOCK), (size_t)nl * TRACE_BLOCK * 8, st, s->view, rays, n,
                               tmin, t0, aim_dev, t_out, id_out, occ, next, T.fetch_min, (const uint32_t *)order,
                               (const int *)flag, nl, ud, T.walk_phases);
        } else
            hipLaunchKernelGGL((k_trace_persistent<ANY, false, 0>), dim3(pgrid), dim3(TRACE_BLOCK), 0, st, s->view, rays,
                               n, tmin, t0, aim_dev, t_out, id_out, occ, next, T.fetch_min, (const uint32_t *)order,
                               (const int *)flag, 0);
    } else if (ANY) {
        hipLaunchKernelGGL(k_trace_any<0>, dim3(grid), dim3(TRACE_BLOCK), 0, st, s->view, rays, n, tmin, t0, aim_dev, occ,
                           (const uint32_t *)order, (const int *)flag);
    } else {
        hipLaunchKernelGGL(k_trace_closest<0>, dim3(grid), dim3(TRACE_BLOCK), 0, st, s->view, rays, n, tmin, t0, t_out,
                           id_out, (const uint32_t *)order, (const int *)flag);
    }
    RM_HIP(hipGetLastError());
    if (s->host.cutout) {
        hipLaunchKernelGGL(k_trace_cutout<ANY>, dim3(grid), dim3(TRACE_BLOCK), 0, st, s->view, rays, n, tmin, t0, aim_dev,
                           t_out, id_out, occ);
        RM_HIP(hipGetLastError());
    }
    if (timing) {
        if ((mrc = mark(2, st))) return mrc;
        trace_events().recorded = true;
    }
    return RM_OK;
}

extern "C" int32_t rm_last_trace_ms(float *reorder_ms, float *traverse_ms) {
    TraceEvents &ev = trace_events();
    if (!ev.recorded) return fail(RM_ERR_INVALID, "rm_last_trace_ms: no trace call ran with kernel timing on");
    RM_HIP(hipEventSynchronize(ev.e[2]));
    float a = 0.0f, b = 0.0f;
    RM_HIP(hipEventElapsedTime(&a, ev.e[0], ev.e[1]));
    RM_HIP(hipEventElapsedTime(&b, ev.e[1], ev.e[2]));
    if (reorder_ms) *reorder_ms = a;
    if (traverse_ms) *traverse_ms = b;
    return RM_OK;
}

extern "C" int32_t rm_last_trace_util(uint64_t *counts) {
    if (!counts) return fail(RM_ERR_INVALID, "rm_last_trace_util: NULL");
    if (!g_util_valid) return fail(RM_ERR_INVALID, "rm_last_trace_util: no trace call ran with RM_TUNE_TRACE_UTIL on");
    RM_HIP(hipDeviceSynchronize());
    RM_HIP(hipMemcpyFromSymbol(counts, HIP_SYMBOL(g_trace_util), 4 * sizeof(uint64_t), 0, hipMemcpyDeviceToHost));
    return RM_OK;
}

extern "C" int32_t rm_trace_closest(const rm_scene *s, const float *rays, int64_t n, float tmin, float tmax,
                                    float *t_out, int32_t *id_out, void *stream) {
    int32_t st = check_trace_args(s, rays, n, t_out, "rm_trace_closest");
    if (st) return st;
    if (n > 0 && !id_out) return fail(RM_ERR_INVALID, "rm_trace_closest: id_out is NULL");
    if (n == 0) return RM_OK;
    return launch_trace<false>(s, rays, n, tmin, tmax, nullptr, t_out, id_out, nullptr, (hipStream_t)stream);
}

// rm_trace_closest_noalloc's workspace: one static array per GPU, of which
// host code only ever uses the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in, not a device
// address).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static int32_t noalloc_scratch(char **out) {
    static std::mutex mu;
    static char *ptr[64] = {nullptr};
    int dev = 0;
    RM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(RM_ERR_UNSUPPORTED, "rm_trace_closest_noalloc: device index %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        ptr[dev] = static_cast<char *>(p);
    }
    *out = ptr[dev];
    return RM_OK;
}

extern "C" int32_t rm_trace_closest_noalloc(const rm_scene *s, const float *rays, int64_t n, float tmin, float tmax,
                                            float *t_out, int32_t *id_out, void *stream) {
    int32_t st = check_trace_args(s, rays, n, t_out, "rm_trace_closest_noalloc");
    if (st) return st;
    if (n > 0 && !id_out) return fail(RM_ERR_INVALID, "rm_trace_closest_noalloc: id_out is NULL");
    if (n > RM_NOALLOC_MAX_RAYS)
        return fail(RM_ERR_INVALID, "rm_trace_closest_noalloc: %lld rays exceed the maximum of %d", (long long)n,
                    (int)RM_NOALLOC_MAX_RAYS);
    if (n == 0) return RM_OK;
    FixedWorkspace ws{nullptr, NOALLOC_SCRATCH_BYTES};
    if ((st = noalloc_scratch(&ws.base))) return st;
    return launch_trace<false>(s, rays, n, tmin, tmax, nullptr, t_out, id_out, nullptr, (hipStream_t)stream, &ws);
}

extern "C" int32_t rm_trace_any(const rm_scene *s, const float *rays, int64_t n, float tmin, float aim_depth,
                                const float *aim_dev, uint8_t *occ, void *stream) {
    int32_t st = check_trace_args(s, rays, n, occ, "rm_trace_any");
    if (st) return st;
    if (n == 0) return RM_OK;
    return launch_trace<true>(s, rays, n, tmin, aim_depth, aim_dev, nullptr, nullptr, occ, (hipStream_t)stream);
}

extern "C" int32_t rm_trace_count(const rm_scene *s, const float *rays, int64_t n, float tmin, float tmax,
                                  int32_t any_hit, uint64_t *counts, void *stream) {
    int32_t st = check_trace_args(s, rays, n, counts, "rm_trace_count");
    if (st) return st;
    if (!counts) return fail(RM_ERR_INVALID, "rm_trace_count: counts is NULL");
    RM_HIP(hipMemsetAsync(counts, 0, 4 * sizeof(uint64_t), (hipStream_t)stream));
    if (n == 0) return RM_OK;
    unsigned grid = (unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK);
    hipLaunchKernelGGL(k_trace_count, dim3(grid), dim3(TRACE_BLOCK), 0, (hipStream_t)stream, s->view, rays, n, tmin, tmax,
                       (int)any_hit, (unsigned long long *)counts);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

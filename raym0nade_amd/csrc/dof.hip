// Photo::depthFeildBlur (src/image.cpp:285-356) on the GPU, with the same
// result as the reference's sequential scatter.
//
// The reference visits pixels nearest first (a stable sort by distance to the
// camera) and spreads each over its disc.  A receiving pixel's weight is
// capped at 0.99 in total, first come first served.  What a receiving pixel
// ends with depends only on the sources that reach it, folded in that visit
// order.  Receivers are independent of each other.  So each lane folds one
// receiving pixel over its sources, taken in the order (distance, pixel
// index) -- the stable sort's order:
//   k_dof_src     per source pixel: distance, CoC0, radius, the disc's total
//                 weight (the reference's loops verbatim), a 64-bit sort key,
//                 and a count for every 16x16 tile its disc's box touches;
//   k_dof_scan    tile offsets (one block);
//   k_dof_fill    each source's key into the lists of those tiles;
//   segmented radix sort of every tile's list (rocPRIM);
//   k_dof_gather  one block per tile: the sorted list is staged through LDS
//                 256 entries at a time and every lane folds its pixel.  The
//                 block stops once every lane's pixel holds its full 0.99.
// The list sizes depend on the image, so the launch waits once for the total.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include "rm_frame.h"
#include "rm_vec.h"

namespace rm {

constexpr int DT = 16;   // tile side: 256 receiving pixels per block
constexpr int DB = 256;

struct DofSrc {
    float coc0, total;
    int32_t radius, pad;
};

// std::min(a, b) = (b < a) ? b : a
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }

__device__ __forceinline__ void tile_box(int x, int y, int r, int w, int h, int &tx0, int &tx1, int &ty0, int &ty1) {
    tx0 = (x - r > 0 ? x - r : 0) / DT;
    tx1 = (x + r < w - 1 ? x + r : w - 1) / DT;
    ty0 = (y - r > 0 ? y - r : 0) / DT;
    ty1 = (y + r < h - 1 ? y + r : h - 1) / DT;
}

// src/image.cpp:293-330 for one source pixel.  A missed pixel (NaN position)
// is at distance +inf: CoC0 = min(CoC + 1e-4, 96).
__global__ __launch_bounds__(DB) void k_dof_src(const float *__restrict__ pos, int w, int h, float focus, float coc,
                                               V3 cam, int ntx, DofSrc *__restrict__ src,
                                               uint64_t *__restrict__ key, uint32_t *__restrict__ tile_count) {
    uint32_t i = blockIdx.x * DB + threadIdx.x;
    if (i >= (uint32_t)w * (uint32_t)h) return;
    V3 p{pos[(size_t)i * 3], pos[(size_t)i * 3 + 1], pos[(size_t)i * 3 + 2]};
    float depth = length(p - cam);
    if (depth != depth) depth = INFINITY;
    const float coc0 = smin(coc * fabsf(1.0f - focus / depth) + 1e-4f, 96.0f);
    // (int)NaN is undefined in the reference (a pixel at the camera with
    // focus 0); here such a pixel keeps only itself
    const int radius = (coc0 == coc0) ? (int)coc0 : 0;
    float total = 0.0f;
    for (int dy = -radius; dy <= radius; ++dy) {
        int xLen = (int)sqrtf(coc0 * coc0 - (float)(dy * dy));
        int xl, xr;
        for (xl = -xLen; xl <= 0; xl++) {
            float wt = smin(coc0 - sqrtf((float)(xl * xl + dy * dy)), 1.0f);
            if (wt == 1.0f) break;
            total += wt;
        }
        for (xr = xLen; xr > 0; xr--) {
            float wt = smin(coc0 - sqrtf((float)(xr * xr + dy * dy)), 1.0f);
            if (wt == 1.0f) break;
            total += wt;
        }
        total += (float)(xr - xl + 1);
    }
    src[i] = DofSrc{coc0, total, radius, 0};
    // distance >= 0 (or +inf): its bits order like the value
    key[i] = ((uint64_t)__float_as_uint(depth) << 32) | i;
    int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
    int tx0, tx1, ty0, ty1;
    tile_box(x, y, radius, w, h, tx0, tx1, ty0, ty1);
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) atomicAdd(&tile_count[ty * ntx + tx], 1u);
}

__global__ __launch_bounds__(DB) void k_dof_zero(uint32_t *__restrict__ a, int n) {
    int i = blockIdx.x * DB + threadIdx.x;
    if (i < n) a[i] = 0u;
}

// Exclusive scan of the tile counts into off[0..nt] (off[nt] = total); the
// fill cursors start at zero.  One block of 1024 lanes.
__global__ __launch_bounds__(1024) void k_dof_scan(const uint32_t *__restrict__ count, int nt, uint32_t *__restrict__ off,
                                                  uint32_t *__restrict__ cursor) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (nt + 1023) / 1024;
    const int b = t * per, e = (b + per < nt) ? b + per : nt;
    uint32_t s = 0;
    for (int k = b; k < e; ++k) s += count[k];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        uint32_t v = (t >= d) ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int k = b; k < e; ++k) {
        off[k] = run;
        run += count[k];
        cursor[k] = 0u;
    }
    if (t == 1023) off[nt] = part[1023];
}

__global__ __launch_bounds__(DB) void k_dof_fill(const DofSrc *__restrict__ src, const uint64_t *__restrict__ key, int w,
                                                int h, int ntx, const uint32_t *__restrict__ off,
                                                uint32_t *__restrict__ cursor, uint64_t *__restrict__ list) {
    uint32_t i = blockIdx.x * DB + threadIdx.x;
    if (i >= (uint32_t)w * (uint32_t)h) return;
    int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
    int tx0, tx1, ty0, ty1;
    tile_box(x, y, src[i].radius, w, h, tx0, tx1, ty0, ty1);
    const uint64_t k = key[i];
    for (int ty = ty0; ty <= ty1; ++ty)
        for (int tx = tx0; tx <= tx1; ++tx) {
            const int t = ty * ntx + tx;
            list[off[t] + atomicAdd(&cursor[t], 1u)] = k;
        }
}

// src/image.cpp:332-355 for the 256 receiving pixels of one tile.
__global__ __launch_bounds__(DB) void k_dof_gather(const uint64_t *__restrict__ list, const uint32_t *__restrict__ off,
                                                  const DofSrc *__restrict__ src, const float *__restrict__ img, int w,
                                                  int h, int ntx, int nt, float *__restrict__ out) {
    __shared__ float s_coc0[DB], s_total[DB], s_col[3 * DB];
    __shared__ int s_x[DB], s_y[DB], s_r[DB];
    const int t = (int)xcd_block(blockIdx.x, gridDim.x);
    if (t >= nt) return;  // uniform over the block
    const int tid = threadIdx.x;
    const int x = (t % ntx) * DT + (tid % DT), y = (t / ntx) * DT + (tid / DT);
    const bool mine = x < w && y < h;
    float gained = 0.0f;
    V3 acc = v3(0.0f);
    bool done = !mine;
    const uint32_t beg = off[t], end = off[t + 1];
    for (uint32_t c = beg; c < end; c += DB) {
        const uint32_t j = c + tid;
        if (j < end) {
            const uint32_t s = (uint32_t)list[j];
            const DofSrc S = src[s];
            s_coc0[tid] = S.coc0;
            s_total[tid] = S.total;
            s_r[tid] = S.radius;
            s_x[tid] = (int)(s % (uint32_t)w);
            s_y[tid] = (int)(s / (uint32_t)w);
            s_col[3 * tid] = img[(size_t)s * 3];
            s_col[3 * tid + 1] = img[(size_t)s * 3 + 1];
            s_col[3 * tid + 2] = img[(size_t)s * 3 + 2];
        }
        __syncthreads();
        const int m = (end - c < (uint32_t)DB) ? (int)(end - c) : DB;
        if (!done)
            for (int e = 0; e < m; ++e) {
                const int r = s_r[e];
                const int dy = y - s_y[e];
                if (dy < -r || dy > r) continue;
                const int dx = x - s_x[e];
                if (dx < -r || dx > r) continue;
                const float c0 = s_coc0[e];
                const int xLen = (int)sqrtf(c0 * c0 - (float)(dy * dy));
                if (dx < -xLen || dx > xLen) continue;
                float wt = smin(c0 - sqrtf((float)(dx * dx + dy * dy)), 1.0f) / s_total[e];
                if (wt < 1e-4f) continue;
                if (gained + wt > 0.99f) wt = 0.99f - gained;
                if (wt < 1e-4f) continue;
                acc += V3{s_col[3 * e], s_col[3 * e + 1], s_col[3 * e + 2]} * wt;
                gained += wt;
            }
        // every later weight >= 1e-4 would exceed the cap and be cut below 1e-4
        done = done || ((gained + 1e-4f > 0.99f) && (0.99f - gained < 1e-4f));
        if (__syncthreads_and(done)) break;
    }
    if (mine) {
        V3 r = acc / gained;
        const size_t q = (size_t)y * w + x;
        out[q * 3] = r.x;
        out[q * 3 + 1] = r.y;
        out[q * 3 + 2] = r.z;
    }
}

}  // namespace rm

using namespace rm;

namespace {
size_t pad256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int32_t rm::dof_launch(rm_frame *f, const float *in, float *out, hipStream_t st) {
    const int w = f->w, h = f->h;
    const uint32_t n = (uint32_t)((size_t)w * h);
    const int ntx = (w + DT - 1) / DT, nty = (h + DT - 1) / DT, nt = ntx * nty;
    if (!(f->dof_coc >= 0.0f) || !isfinite(f->dof_coc) || !isfinite(f->dof_focus))
        return fail(RM_ERR_INVALID, "depth of field: CoC %g and focus %g must be finite, CoC >= 0", f->dof_coc, f->dof_focus);
    // a disc of radius <= 96 touches at most 14 x 14 tiles: the list offsets must fit 32 bits
    if ((uint64_t)n * 196u > 0xffffffffull)
        return fail(RM_ERR_UNSUPPORTED, "depth of field: frames above %u pixels are not supported", 0xffffffffu / 196u);
    // per pixel: source record and key; per tile: count, offsets, cursor; the total on the host side
    const size_t b_src = pad256((size_t)n * sizeof(DofSrc)), b_key = pad256((size_t)n * 8),
                 b_tile = pad256((size_t)nt * 4), b_off = pad256((size_t)(nt + 1) * 4);
    const size_t fixed = b_src + b_key + 2 * b_tile + b_off;
    int32_t rc;
    if (f->dof.bytes < fixed && (rc = f->dof.alloc(fixed))) return rc;
    char *base = (char *)f->dof.p;
    DofSrc *src = (DofSrc *)base;
    uint64_t *key = (uint64_t *)(base + b_src);
    uint32_t *count = (uint32_t *)(base + b_src + b_key);
    uint32_t *cursor = (uint32_t *)(base + b_src + b_key + b_tile);
    uint32_t *off = (uint32_t *)(base + b_src + b_key + 2 * b_tile);
    const unsigned g = (n + DB - 1) / DB;
    hipLaunchKernelGGL(k_dof_zero, dim3((nt + DB - 1) / DB), dim3(DB), 0, st, count, nt);
    V3 cam{f->dof_cam[0], f->dof_cam[1], f->dof_cam[2]};
    hipLaunchKernelGGL(k_dof_src, dim3(g), dim3(DB), 0, st, f->buf[RM_BUF_POSITION].as<float>(), w, h, f->dof_focus,
                       f->dof_coc, cam, ntx, src, key, count);
    hipLaunchKernelGGL(k_dof_scan, dim3(1), dim3(1024), 0, st, (const uint32_t *)count, nt, off, cursor);
    RM_HIP(hipGetLastError());
    uint32_t total = 0;
    RM_HIP(hipMemcpyAsync(&total, off + nt, 4, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    size_t tmp_bytes = 0;
    RM_HIP(rocprim::segmented_radix_sort_keys(nullptr, tmp_bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr, total,
                                              (unsigned)nt, (const uint32_t *)off, (const uint32_t *)off + 1, 0, 64, st));
    const size_t b_list = pad256((size_t)total * 8);
    const size_t need = 2 * b_list + pad256(tmp_bytes);
    if (f->dofl.bytes < need && (rc = f->dofl.alloc(need))) return rc;
    uint64_t *list = (uint64_t *)f->dofl.p;
    uint64_t *sorted = (uint64_t *)((char *)f->dofl.p + b_list);
    void *tmp = (char *)f->dofl.p + 2 * b_list;
    hipLaunchKernelGGL(k_dof_fill, dim3(g), dim3(DB), 0, st, (const DofSrc *)src, (const uint64_t *)key, w, h, ntx,
                       (const uint32_t *)off, cursor, list);
    RM_HIP(hipGetLastError());
    RM_HIP(rocprim::segmented_radix_sort_keys(tmp, tmp_bytes, (const uint64_t *)list, sorted, total, (unsigned)nt,
                                              (const uint32_t *)off, (const uint32_t *)off + 1, 0, 64, st));
    hipLaunchKernelGGL(k_dof_gather, dim3(nt), dim3(DB), 0, st, (const uint64_t *)sorted, (const uint32_t *)off,
                       (const DofSrc *)src, in, w, h, ntx, nt, out);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

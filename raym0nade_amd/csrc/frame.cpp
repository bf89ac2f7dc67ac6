// Frame buffers: the device form of Photo's per-pixel arrays
// (include/image.h:15-67, src/image.cpp:9-16).
#include <string.h>
#include "rm_frame.h"

using namespace rm;

extern "C" int32_t rm_frame_create(int32_t width, int32_t height, rm_frame **out) {
    if (!out) return fail(RM_ERR_INVALID, "rm_frame_create: out is NULL");
    *out = nullptr;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)1 << 28)
        return fail(RM_ERR_INVALID, "rm_frame_create: bad size %dx%d", width, height);
    int32_t st = ensure_device();
    if (st) return st;
    rm_frame *f = new (std::nothrow) rm_frame();
    if (!f) return fail(RM_ERR_OOM, "rm_frame_create: host allocation");
    f->w = width;
    f->h = height;
    const size_t np = (size_t)width * height;
    for (int b = 0; b < RM_NUM_BUFFERS; ++b) {
        size_t bytes = np * frame_channels(b) * sizeof(float);
        if ((st = f->buf[b].alloc(bytes))) { delete f; return st; }
        if (hipMemset(f->buf[b].p, 0, bytes) != hipSuccess) { delete f; return fail(RM_ERR_HIP, "hipMemset"); }
    }
    if ((st = f->stats.alloc(RM_STAT_COUNT * sizeof(uint64_t)))) { delete f; return st; }
    if (hipMemset(f->stats.p, 0, RM_STAT_COUNT * sizeof(uint64_t)) != hipSuccess) {
        delete f;
        return fail(RM_ERR_HIP, "hipMemset");
    }
    *out = f;
    return RM_OK;
}

extern "C" int32_t rm_frame_destroy(rm_frame *f) {
    delete f;
    return RM_OK;
}

extern "C" int32_t rm_frame_buffer(rm_frame *f, int32_t which, float **dev_ptr) {
    if (!f || !dev_ptr || which < 0 || which >= RM_NUM_BUFFERS) return fail(RM_ERR_INVALID, "rm_frame_buffer: bad argument");
    *dev_ptr = f->buf[which].as<float>();
    return RM_OK;
}

extern "C" int32_t rm_frame_set_recording(rm_frame *f, int32_t on) {
    if (!f) return fail(RM_ERR_INVALID, "rm_frame_set_recording: frame is NULL");
    f->record = on != 0;
    if (!f->record) {
        f->rec.release();
        f->rec_valid = false;
    }
    return RM_OK;
}

extern "C" int32_t rm_frame_copy(const rm_frame *f, int32_t which, float *dst, void *stream) {
    if (!f || !dst || which < 0 || which >= RM_NUM_BUFFERS) return fail(RM_ERR_INVALID, "rm_frame_copy: bad argument");
    RM_HIP(hipMemcpyAsync(dst, f->buf[which].p, (size_t)f->w * f->h * frame_channels(which) * sizeof(float),
                          hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return RM_OK;
}

extern "C" int32_t rm_frame_upload(rm_frame *f, int32_t which, const float *host) {
    if (!f || !host || which < 0 || which >= RM_NUM_BUFFERS) return fail(RM_ERR_INVALID, "rm_frame_upload: bad argument");
    RM_HIP(hipDeviceSynchronize());
    RM_HIP(hipMemcpy(f->buf[which].p, host, (size_t)f->w * f->h * frame_channels(which) * sizeof(float),
                     hipMemcpyHostToDevice));
    // the frame no longer holds the render the adjoint would differentiate
    f->stamped = false;
    f->rec_valid = false;
    return RM_OK;
}

extern "C" int32_t rm_frame_get_stats(const rm_frame *f, uint64_t *stats) {
    if (!f || !stats) return fail(RM_ERR_INVALID, "rm_frame_get_stats: NULL argument");
    RM_HIP(hipDeviceSynchronize());
    RM_HIP(hipMemcpy(stats, f->stats.p, RM_STAT_COUNT * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return RM_OK;
}

extern "C" int32_t rm_frame_download(const rm_frame *f, int32_t which, float *host) {
    if (!f || !host || which < 0 || which >= RM_NUM_BUFFERS) return fail(RM_ERR_INVALID, "rm_frame_download: bad argument");
    RM_HIP(hipDeviceSynchronize());
    RM_HIP(hipMemcpy(host, f->buf[which].p, (size_t)f->w * f->h * frame_channels(which) * sizeof(float),
                     hipMemcpyDeviceToHost));
    return RM_OK;
}

// Error reporting, device checks and small utilities of the C ABI.
#include <stdio.h>
#include <string.h>
#include "rm_internal.h"
#include "philox.h"

namespace rm {

static thread_local char g_err[1024] = "";

int32_t fail(int32_t code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int32_t hip_fail(hipError_t e, const char *what) {
    int32_t code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RM_ERR_OOM : RM_ERR_HIP;
    if (e == hipErrorNoDevice) code = RM_ERR_NO_DEVICE;
    return fail(code, "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int32_t ensure_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(RM_ERR_NO_DEVICE, "no HIP device visible (%s)", hipGetErrorString(e));
    return RM_OK;
}

Tuning &tuning() {
    static Tuning t;
    return t;
}

int device_cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

}  // namespace rm

extern "C" int32_t rm_set_tuning(int32_t key, int32_t value) {
    rm::Tuning &t = rm::tuning();
    switch (key) {
        case RM_TUNE_TRACE_VARIANT:
            if (value != 0 && value != 2 && value != 3)
                return rm::fail(RM_ERR_INVALID, "rm_set_tuning: trace variant %d (0, 2 or 3)", value);
            t.trace_variant = value;
            return RM_OK;
        case RM_TUNE_PERSISTENT_BLOCKS_PER_CU:
            if (value < 1 || value > 64) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: blocks per CU %d", value);
            t.persistent_blocks_per_cu = value;
            return RM_OK;
        case RM_TUNE_FETCH_MIN:
            if (value < 1 || value > 64) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: fetch_min %d", value);
            t.fetch_min = value;
            return RM_OK;
        case RM_TUNE_RAY_SORT:
            if (value < 0 || value > 5) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: ray sort %d", value);
            t.ray_sort = value;
            return RM_OK;
        case RM_TUNE_RENDER_BLOCKS_PER_CU:
            if (value < 1 || value > 64) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: render blocks per CU %d", value);
            t.render_blocks_per_cu = value;
            return RM_OK;
        case RM_TUNE_ADJOINT_SPLIT:
            if (value < 1 || value > 64) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: adjoint split %d", value);
            t.adjoint_split = value;
            return RM_OK;
        case RM_TUNE_KERNEL_TIMING:
            if (value < 0 || value > 1) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: kernel timing %d", value);
            t.kernel_timing = value;
            return RM_OK;
        case RM_TUNE_RENDER_MIN_WAVES:
            if (value != 1 && value != 3 && value != 4)
                return rm::fail(RM_ERR_INVALID, "rm_set_tuning: render min waves %d (1, 3 or 4)", value);
            t.render_min_waves = value;
            return RM_OK;
        case RM_TUNE_TRIS_PER_STEP:
            if (value != 0 && value != 1 && value != 2 && value != 4)
                return rm::fail(RM_ERR_INVALID, "rm_set_tuning: triangles per step %d (0, 1, 2 or 4)", value);
            t.tris_per_step = value;
            return RM_OK;
        case RM_TUNE_NEE_SPLIT:
            if (value < 0 || value > 2) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: shadow split %d", value);
            t.nee_split = value;
            return RM_OK;
        case RM_TUNE_RENDER_COUNT:
            if (value < 0 || value > 1) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: render count %d", value);
            t.render_count = value;
            return RM_OK;
        case RM_TUNE_TRACE_UTIL:
            if (value < 0 || value > 1) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: trace utilisation %d", value);
            t.trace_util = value;
            return RM_OK;
        case RM_TUNE_WALK_PHASES:
            if (value < 0 || value > 64) return rm::fail(RM_ERR_INVALID, "rm_set_tuning: walk phases %d", value);
            t.walk_phases = value;
            return RM_OK;
        case RM_TUNE_LDS_STACK_ENTRIES:
            if (value < 1 || value > rm::MAX_STACK)
                return rm::fail(RM_ERR_INVALID, "rm_set_tuning: LDS stack entries %d (1..%d)", value, rm::MAX_STACK);
            t.lds_stack_entries = value;
            return RM_OK;
        default:
            return rm::fail(RM_ERR_INVALID, "rm_set_tuning: unknown key %d", key);
    }
}

extern "C" int32_t rm_get_tuning(int32_t key, int32_t *value) {
    if (!value) return rm::fail(RM_ERR_INVALID, "rm_get_tuning: NULL");
    const rm::Tuning &t = rm::tuning();
    switch (key) {
        case RM_TUNE_TRACE_VARIANT: *value = t.trace_variant; return RM_OK;
        case RM_TUNE_PERSISTENT_BLOCKS_PER_CU: *value = t.persistent_blocks_per_cu; return RM_OK;
        case RM_TUNE_FETCH_MIN: *value = t.fetch_min; return RM_OK;
        case RM_TUNE_RAY_SORT: *value = t.ray_sort; return RM_OK;
        case RM_TUNE_RENDER_BLOCKS_PER_CU: *value = t.render_blocks_per_cu; return RM_OK;
        case RM_TUNE_ADJOINT_SPLIT: *value = t.adjoint_split; return RM_OK;
        case RM_TUNE_KERNEL_TIMING: *value = t.kernel_timing; return RM_OK;
        case RM_TUNE_RENDER_MIN_WAVES: *value = t.render_min_waves; return RM_OK;
        case RM_TUNE_LDS_STACK_ENTRIES: *value = t.lds_stack_entries; return RM_OK;
        case RM_TUNE_TRIS_PER_STEP: *value = t.tris_per_step; return RM_OK;
        case RM_TUNE_NEE_SPLIT: *value = t.nee_split; return RM_OK;
        case RM_TUNE_RENDER_COUNT: *value = t.render_count; return RM_OK;
        case RM_TUNE_TRACE_UTIL: *value = t.trace_util; return RM_OK;
        case RM_TUNE_WALK_PHASES: *value = t.walk_phases; return RM_OK;
        default: return rm::fail(RM_ERR_INVALID, "rm_get_tuning: unknown key %d", key);
    }
}

extern "C" const char *rm_last_error(void) { return rm::g_err; }

extern "C" int32_t rm_abi_version(void) { return 2; }

extern "C" void rm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    rm::U4 r = rm::philox4x32_10(rm::U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

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

}  // namespace rm

extern "C" const char *rm_last_error(void) { return rm::g_err; }

extern "C" int32_t rm_abi_version(void) { return 1; }

extern "C" void rm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    rm::U4 r = rm::philox4x32_10(rm::U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}

// Post-processing kernels (work in progress).
#include "rm_frame.h"
using namespace rm;
extern "C" int32_t rm_postprocess(rm_frame *, int32_t, float, void *) {
    return fail(RM_ERR_UNSUPPORTED, "rm_postprocess: not built yet");
}

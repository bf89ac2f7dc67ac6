// Wavefront integrator (work in progress).
#include "rm_frame.h"
using namespace rm;
extern "C" int32_t rm_render(const rm_scene *, const rm_render_args *, rm_frame *, void *) {
    return fail(RM_ERR_UNSUPPORTED, "rm_render: not built yet");
}

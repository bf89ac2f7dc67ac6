// Frame object shared by the render and post-processing translation units.
#pragma once
#include "rm_internal.h"

constexpr int RM_NUM_BUFFERS = 11;

inline int frame_channels(int which) { return (which >= RM_BUF_DD && which <= RM_BUF_IS) ? 4 : 3; }

struct rm_frame {
    int32_t w = 0, h = 0;
    rm::DevBuf buf[RM_NUM_BUFFERS];
    rm::DevBuf post;  // tone-mapped image read by FXAA
    rm::DevBuf filt;  // denoise: the four radiance buffers' other half of the ping-pong
    rm::DevBuf drop;  // per pixel: the dropped outlier light sample (render -> adjoint)
    rm::DevBuf adj;   // adjoint scratch: texel gradients of every mip level, material gradients
    // per indirect sample (rm_frame_set_recording): sums over its light samples
    // of the diffuse and specular increments and of w L, [pixel][sample] x 3 float4
    rm::DevBuf rec;
    bool record = false;
    bool rec_valid = false;  // the last render recorded (all pixels opaque)
    int32_t rec_stride = 0;  // indirect samples per pixel
    // the render the buffers hold (the adjoint differentiates exactly that one)
    const rm_scene *stamp_scene = nullptr;
    uint64_t stamp_version = 0;
    rm_render_args stamp_args{};
    bool stamped = false;
};

namespace rm {
// rm_postprocess in two halves so a captured graph contains only launches.
int32_t post_prepare(rm_frame *f, int opt);
int32_t post_launch(rm_frame *f, int opt, float exposure, hipStream_t st);
}  // namespace rm

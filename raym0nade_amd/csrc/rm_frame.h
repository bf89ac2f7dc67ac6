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
    rm::DevBuf lin;   // bloom / depth of field: the linear shaded image and its copy [2][h][w][3]
    rm::DevBuf lvl;   // bloom: the blur levels' ping-pong [2][h][w][3]
    rm::DevBuf dof;   // depth of field: per-pixel sources and keys, per-tile counts and offsets
    rm::DevBuf dofl;  // depth of field: the per-tile lists (unsorted, sorted) and the sort's storage
    // Photo's depth-of-field fields (include/image.h:36-37): set by rm_render
    // from its arguments (src/render.cpp:666-668) or by rm_frame_set_dof
    float dof_focus = 0.0f, dof_coc = 0.0f;
    float dof_cam[3] = {0.0f, 0.0f, 0.0f};
    rm::DevBuf drop;  // per pixel: the dropped outlier light sample (render -> adjoint)
    rm::DevBuf stats; // RM_STAT_COUNT uint64 counters of the last render
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
// Depth of field sizes its tile lists from the image, with a host wait, so
// it is refused where the launch half will be captured (in_batch).
int32_t post_prepare(rm_frame *f, int opt, bool in_batch = false);
int32_t post_launch(rm_frame *f, int opt, float exposure, hipStream_t st);
// spatialClamp / filter in two halves, like post_prepare / post_launch.
int32_t denoise_prepare(rm_frame *f);
int32_t clamp_launch(rm_frame *f, hipStream_t st);
int32_t filter_launch(rm_frame *f, hipStream_t st);
// depthFeildBlur of the linear image `in` into `out` ([h][w][3] each).
int32_t dof_launch(rm_frame *f, const float *in, float *out, hipStream_t st);
}  // namespace rm

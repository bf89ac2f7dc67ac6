"""ctypes view of include/raym0nade_hip.h (the library's C ABI).

The extension is loaded from this package directory only; a missing library
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libraym0nade_hip.so")

RM_OK, RM_ERR_INVALID, RM_ERR_OOM, RM_ERR_HIP, RM_ERR_EMPTY, RM_ERR_UNSUPPORTED, RM_ERR_NO_DEVICE = range(7)

(RM_BUF_POSITION, RM_BUF_SHAPE_NORMAL, RM_BUF_SURFACE_NORMAL, RM_BUF_EMISSION, RM_BUF_BASE_COLOR,
 RM_BUF_DD, RM_BUF_DS, RM_BUF_ID, RM_BUF_IS, RM_BUF_PIXELS, RM_BUF_MATERIAL) = range(11)

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)


class RmMaterial(C.Structure):
    _fields_ = [
        ("base_color", C.c_float * 3),
        ("emission", C.c_float * 3),
        ("transmitting_color", C.c_float * 3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("opacity", C.c_float),
        ("ior", C.c_float),
        ("albedo_tex", C.c_int32),
        ("normal_tex", C.c_int32),
        ("emissive_tex", C.c_int32),
        ("specular_tex", C.c_int32),
    ]

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        for f in ("albedo_tex", "normal_tex", "emissive_tex", "specular_tex"):
            if f not in k:
                setattr(self, f, -1)


class RmTexture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", f32p), ("alpha", f32p),
                ("bytes", C.POINTER(C.c_uint8)), ("channels", C.c_int32)]


class RmSceneDesc(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64),
        ("verts", f32p),
        ("uvs", f32p),
        ("normals", f32p),
        ("material", i32p),
        ("mesh", i32p),
        ("n_materials", C.c_int32),
        ("materials", C.POINTER(RmMaterial)),
        ("n_textures", C.c_int32),
        ("textures", C.POINTER(RmTexture)),
        ("sky_width", C.c_int32),
        ("sky_height", C.c_int32),
        ("sky_rgb", f32p),
    ]


class RmSceneInfo(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64),
        ("n_nodes", C.c_int64),
        ("depth", C.c_int32),
        ("n_light_objects", C.c_int32),
        ("n_light_faces", C.c_int32),
        ("has_sky", C.c_int32),
        ("device_bytes", C.c_int64),
        ("build_seconds", C.c_double),
        ("upload_seconds", C.c_double),
    ]


class RmRenderArgs(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("up", C.c_float * 3),
        ("right", C.c_float * 3),
        ("accuracy", C.c_float),
        ("focus", C.c_float),
        ("coc", C.c_float),
        ("exposure", C.c_float),
        ("p_direct", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp", C.c_int32),
        ("threads", C.c_int32),
        ("seed", C.c_uint32),
        ("max_depth", C.c_int32),
        ("rr_min_depth", C.c_int32),
        ("paths_per_round", C.c_int32),
    ]


class RmBatchJob(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("args", RmRenderArgs),
        ("frame", C.c_void_p),
        ("post_options", C.c_int32),
    ]


class RmGrads(C.Structure):
    _fields_ = [
        ("textures", C.c_void_p),
        ("roughness", C.c_void_p),
        ("emission", C.c_void_p),
        ("replay", C.c_void_p),
    ]


(RM_STAT_PRIMARY_RAYS, RM_STAT_EXTENSION_RAYS, RM_STAT_SHADOW_RAYS, RM_STAT_HITS, RM_STAT_PATHS,
 RM_STAT_MEDIUM_OVERFLOW, RM_STAT_WALK_INTERIOR, RM_STAT_WALK_LEAVES, RM_STAT_WALK_BOXES, RM_STAT_WALK_TRIS,
 RM_STAT_MEDIUM_MAX) = range(11)
RM_STAT_COUNT = 16
STAT_NAMES = ["primary_rays", "extension_rays", "shadow_rays", "hits", "paths", "medium_overflow", "walk_interior",
              "walk_leaves", "walk_boxes", "walk_tris", "medium_max"]

RM_SHADE_FULL = 63
RM_SHADE_LINEAR = 65536
RM_POST_SPATIAL_CLAMP = 1 << 17
RM_POST_FILTER = 1 << 18


class RmGenOptions(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64),
        ("grid", C.c_int32),
        ("freq", C.c_int32),
        ("sky_width", C.c_int32),
        ("sky_height", C.c_int32),
        ("tex_size", C.c_int32),
    ]


# (name, restype, argtypes) for every function declared in the header.
SIGNATURES = [
    ("rm_last_error", C.c_char_p, []),
    ("rm_abi_version", C.c_int32, []),
    ("rm_scene_create", C.c_int32, [C.POINTER(RmSceneDesc), C.POINTER(C.c_void_p)]),
    ("rm_scene_destroy", C.c_int32, [C.c_void_p]),
    ("rm_scene_get_info", C.c_int32, [C.c_void_p, C.POINTER(RmSceneInfo)]),
    ("rm_scene_get_perm", C.c_int32, [C.c_void_p, i32p]),
    ("rm_bvh_build", C.c_int32, [f32p, C.c_int64, i32p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    ("rm_trace_closest", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rm_trace_closest_noalloc", C.c_int32,
     [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rm_trace_any", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rm_trace_count", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
    ("rm_frame_create", C.c_int32, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("rm_frame_destroy", C.c_int32, [C.c_void_p]),
    ("rm_frame_buffer", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("rm_frame_download", C.c_int32, [C.c_void_p, C.c_int32, f32p]),
    ("rm_render", C.c_int32, [C.c_void_p, C.POINTER(RmRenderArgs), C.c_void_p, C.c_void_p]),
    ("rm_postprocess", C.c_int32, [C.c_void_p, C.c_int32, C.c_float, C.c_void_p]),
    ("rm_render_batch", C.c_int32, [C.POINTER(RmBatchJob), C.c_int32, C.c_int32, C.c_void_p]),
    ("rm_frame_copy", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("rm_frame_upload", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("rm_spatial_clamp", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("rm_scene_load", C.c_int32, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    ("rm_model_read", C.c_int32, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(RmSceneDesc), C.POINTER(C.c_void_p)]),
    ("rm_model_destroy", C.c_int32, [C.c_void_p]),
    ("rm_image_load", C.c_int32, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_float))]),
    ("rm_image_free", None, [C.POINTER(C.c_float)]),
    ("rm_frame_set_dof", C.c_int32, [C.c_void_p, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    ("rm_filter", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("rm_frame_set_recording", C.c_int32, [C.c_void_p, C.c_int32]),
    ("rm_scene_update", C.c_int32, [C.c_void_p, C.POINTER(RmMaterial), C.c_void_p, C.c_void_p]),
    ("rm_adjoint", C.c_int32, [C.c_void_p, C.POINTER(RmRenderArgs), C.c_void_p, C.c_void_p, C.POINTER(RmGrads),
                               C.c_void_p]),
    ("rm_generate_scene", C.c_int32, [C.c_int32, C.c_uint32, C.POINTER(RmGenOptions), C.POINTER(C.c_void_p)]),
    ("rm_generated_desc", C.c_int32, [C.c_void_p, C.POINTER(RmSceneDesc), C.POINTER(RmRenderArgs)]),
    ("rm_generated_destroy", C.c_int32, [C.c_void_p]),
    ("rm_generate_rays", C.c_int32, [C.c_int32, C.c_uint32, C.c_int64, C.POINTER(RmRenderArgs), f32p]),
    ("rm_philox4x32_10", None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rm_set_tuning", C.c_int32, [C.c_int32, C.c_int32]),
    ("rm_get_tuning", C.c_int32, [C.c_int32, C.POINTER(C.c_int32)]),
    ("rm_last_trace_ms", C.c_int32, [C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("rm_last_trace_util", C.c_int32, [C.POINTER(C.c_uint64)]),
    ("rm_frame_get_stats", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64)]),
]

_lib = None


class RmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Loads libraym0nade_hip.so from this package (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python -m raym0nade_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != RM_OK:
        raise RmError(code, lib().rm_last_error().decode(errors="replace"))

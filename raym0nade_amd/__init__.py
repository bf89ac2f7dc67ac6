"""raym0nade_amd -- MI355X (gfx950) path-tracing hot path of the Raym0nade renderer.

Python host API over the C ABI in include/raym0nade_hip.h.  Device buffers are
torch tensors on the current HIP device (torch is used for memory and streams
only); every computation happens in the library's own gfx950 kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (RM_BUF_BASE_COLOR, RM_BUF_DD, RM_BUF_DS, RM_BUF_EMISSION, RM_BUF_ID, RM_BUF_IS, RM_BUF_MATERIAL,
                   RM_BUF_PIXELS, RM_BUF_POSITION, RM_BUF_SHAPE_NORMAL, RM_BUF_SURFACE_NORMAL, RmError, RmGenOptions, RmRenderArgs,
                   RmMaterial, RmSceneDesc, RmSceneInfo, check, lib)

__all__ = ["GeneratedScene", "Scene", "Frame", "generate", "rays", "philox", "RmError", "RmRenderArgs"]


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class GeneratedScene:
    """A procedural workload (rm_generate_scene); arrays are borrowed numpy views."""

    def __init__(self, config: int, seed: int = 1, **opts):
        o = RmGenOptions()
        for k, v in opts.items():
            setattr(o, k, v)
        h = C.c_void_p()
        check(lib().rm_generate_scene(config, seed, C.byref(o), C.byref(h)))
        self._h = h
        self.desc = RmSceneDesc()
        self.args = RmRenderArgs()
        check(lib().rm_generated_desc(h, C.byref(self.desc), C.byref(self.args)))
        n = self.desc.n_tris
        self.n_tris = n
        self.verts = self._view(self.desc.verts, C.c_float, n * 9).reshape(n, 3, 3)
        self.material = self._view(self.desc.material, C.c_int32, n)

    def _view(self, ptr, ctype, count):
        # the ctypes array keeps this object alive for as long as the view exists
        arr = (ctype * count).from_address(C.cast(ptr, C.c_void_p).value)
        arr._owner = self
        return np.ctypeslib.as_array(arr)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rm_generated_destroy(h)
            except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
                pass
            self._h = None


def generate(config: int, seed: int = 1, **opts) -> GeneratedScene:
    return GeneratedScene(config, seed, **opts)


def rays(kind: int, n: int, seed: int = 1, cam: RmRenderArgs | None = None) -> np.ndarray:
    """Host ray batch [n][6] (kind 0: pinhole rays of cam; kind 1: uniform)."""
    out = np.empty((n, 6), dtype=np.float32)
    check(lib().rm_generate_rays(kind, seed, n, C.byref(cam) if cam is not None else None, _fp(out)))
    return out


def philox(ctr, key) -> tuple:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().rm_philox4x32_10(c, k, o)
    return tuple(o)


(TUNE_TRACE_VARIANT, TUNE_PERSISTENT_BLOCKS_PER_CU, TUNE_FETCH_MIN, TUNE_RAY_SORT,
 TUNE_RENDER_BLOCKS_PER_CU, TUNE_ADJOINT_SPLIT, TUNE_KERNEL_TIMING, TUNE_RENDER_MIN_WAVES,
 TUNE_LDS_STACK_ENTRIES, TUNE_TRIS_PER_STEP, TUNE_NEE_SPLIT, TUNE_RENDER_COUNT, TUNE_TRACE_UTIL,
 TUNE_WALK_PHASES) = range(14)


def last_trace_util() -> dict:
    """Lane utilisation of the last trace call made with TUNE_TRACE_UTIL on
    (waits for it): per kind of walk step (interior node, triangle tests) the
    wave steps, the lanes that took them, and their share of the 64 lanes."""
    c = (C.c_uint64 * 4)()
    check(lib().rm_last_trace_util(c))
    wi, li, wl, ll = (int(x) for x in c)
    return {"interior_wave_steps": wi, "interior_lane_steps": li, "leaf_wave_steps": wl, "leaf_lane_steps": ll,
            "interior_util": li / (64 * wi) if wi else 0.0, "leaf_util": ll / (64 * wl) if wl else 0.0,
            "util": (li + ll) / (64 * (wi + wl)) if wi + wl else 0.0}


def last_trace_ms() -> tuple:
    """(reordering ms, traversal ms) of the last trace call made with
    TUNE_KERNEL_TIMING on (waits for that call's kernels)."""
    a, b = C.c_float(), C.c_float()
    check(lib().rm_last_trace_ms(C.byref(a), C.byref(b)))
    return a.value, b.value


def set_tuning(key: int, value: int) -> None:
    check(lib().rm_set_tuning(key, value))


def get_tuning(key: int) -> int:
    v = C.c_int32()
    check(lib().rm_get_tuning(key, C.byref(v)))
    return v.value


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Scene:
    """Device scene (rm_scene_create): reference BVH, flattened, in HBM."""

    def __init__(self, desc_or_gen):
        desc = desc_or_gen.desc if isinstance(desc_or_gen, GeneratedScene) else desc_or_gen
        self._keep = desc_or_gen
        h = C.c_void_p()
        check(lib().rm_scene_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.info = RmSceneInfo()
        check(lib().rm_scene_get_info(h, C.byref(self.info)))

    @classmethod
    def load(cls, model_folder: str, model_name: str, sky_name: str = "null") -> "Scene":
        """Model::Model(model_folder, model_name, skyMap_name) for Wavefront OBJ
        and binary FBX models (rm_scene_load)."""
        self = cls.__new__(cls)
        self._keep = None
        h = C.c_void_p()
        check(lib().rm_scene_load(model_folder.encode(), model_name.encode(), sky_name.encode(), C.byref(h)))
        self._h = h
        self.info = RmSceneInfo()
        check(lib().rm_scene_get_info(h, C.byref(self.info)))
        return self

    @property
    def handle(self):
        return self._h

    def perm(self) -> np.ndarray:
        p = np.empty(self.info.n_tris, dtype=np.int32)
        check(lib().rm_scene_get_perm(self._h, p.ctypes.data_as(C.POINTER(C.c_int32))))
        return p

    def trace_closest(self, rays_dev, tmin=1e-4, tmax=float("inf"), t_out=None, id_out=None, stream=None):
        import torch
        n = rays_dev.shape[0]
        if t_out is None:
            t_out = torch.empty(n, dtype=torch.float32, device=rays_dev.device)
        if id_out is None:
            id_out = torch.empty(n, dtype=torch.int32, device=rays_dev.device)
        assert rays_dev.is_contiguous() and rays_dev.dtype == torch.float32 and rays_dev.shape[1] == 6
        check(lib().rm_trace_closest(self._h, _ptr(rays_dev), n, tmin, tmax, _ptr(t_out), _ptr(id_out), _stream(stream)))
        return t_out, id_out

    def trace_any(self, rays_dev, tmin=1e-4, aim=float("inf"), aim_dev=None, occ=None, stream=None):
        import torch
        n = rays_dev.shape[0]
        if occ is None:
            occ = torch.empty(n, dtype=torch.uint8, device=rays_dev.device)
        assert rays_dev.is_contiguous() and rays_dev.dtype == torch.float32 and rays_dev.shape[1] == 6
        check(lib().rm_trace_any(self._h, _ptr(rays_dev), n, tmin, aim, _ptr(aim_dev) if aim_dev is not None else None,
                                 _ptr(occ), _stream(stream)))
        return occ

    def trace_count(self, rays_dev, tmin=1e-4, tmax=float("inf"), any_hit=False, stream=None):
        import torch
        counts = torch.zeros(4, dtype=torch.int64, device=rays_dev.device)
        check(lib().rm_trace_count(self._h, _ptr(rays_dev), rays_dev.shape[0], tmin, tmax, int(any_hit), _ptr(counts),
                                   _stream(stream)))
        return counts

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rm_scene_destroy(h)
            except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
                pass
            self._h = None


class Model:
    """A model read into host memory (rm_model_read): the scene description
    rm_scene_create takes, as numpy copies.  verts [n,3,3], uvs [n,3,2],
    normals [n,3,3] or None, material [n], mesh [n], materials (RmMaterial
    list), textures (list of dicts: width, height, channels, bytes or rgb)
    and the sky ([h,w,3] or None)."""

    def __init__(self, model_folder: str, model_name: str, sky_name: str = "null"):
        desc = RmSceneDesc()
        h = C.c_void_p()
        check(lib().rm_model_read(model_folder.encode(), model_name.encode(), sky_name.encode(), C.byref(desc),
                                  C.byref(h)))
        try:
            n = desc.n_tris

            def arr(ptr, ctype, count, shape):
                if not ptr or count == 0:
                    return None
                a = (ctype * count).from_address(C.cast(ptr, C.c_void_p).value)
                return np.ctypeslib.as_array(a).copy().reshape(shape)

            self.verts = arr(desc.verts, C.c_float, n * 9, (n, 3, 3))
            self.uvs = arr(desc.uvs, C.c_float, n * 6, (n, 3, 2))
            self.normals = arr(desc.normals, C.c_float, n * 9, (n, 3, 3))
            self.material = arr(desc.material, C.c_int32, n, (n,))
            self.mesh = arr(desc.mesh, C.c_int32, n, (n,))
            self.materials = [RmMaterial.from_buffer_copy(desc.materials[i]) for i in range(desc.n_materials)]
            self.textures = []
            for i in range(desc.n_textures):
                t = desc.textures[i]
                d = {"width": t.width, "height": t.height, "channels": t.channels}
                if t.channels:
                    d["bytes"] = arr(t.bytes, C.c_uint8, t.width * t.height * t.channels, (t.height, t.width, t.channels))
                else:
                    d["rgb"] = arr(t.rgb, C.c_float, t.width * t.height * 3, (t.height, t.width, 3))
                self.textures.append(d)
            self.sky = arr(desc.sky_rgb, C.c_float, desc.sky_width * desc.sky_height * 3,
                           (desc.sky_height, desc.sky_width, 3))
        finally:
            lib().rm_model_destroy(h)


def render_batch(jobs, use_graph: bool = True, stream=None) -> None:
    """jobs: list of (Scene, RmRenderArgs, Frame, post_options).  All jobs run
    through one set of launches (one hipGraph with use_graph); synchronous."""
    from ._lib import RmBatchJob
    arr = (RmBatchJob * len(jobs))()
    for i, (scene, args, frame, post) in enumerate(jobs):
        arr[i].scene = scene.handle.value
        arr[i].args = args
        arr[i].frame = frame._h.value
        arr[i].post_options = post
    st = None
    if stream is not None:
        st = C.c_void_p(stream.cuda_stream)
    check(lib().rm_render_batch(arr, len(jobs), int(use_graph), st))


def load_image(path: str) -> np.ndarray:
    """PNG (c / 255) or Radiance HDR image as float32 [h][w][3] (rm_image_load)."""
    w, h = C.c_int32(), C.c_int32()
    ptr = C.POINTER(C.c_float)()
    check(lib().rm_image_load(path.encode(), C.byref(w), C.byref(h), C.byref(ptr)))
    try:
        return np.ctypeslib.as_array(ptr, shape=(h.value, w.value, 3)).copy()
    finally:
        lib().rm_image_free(ptr)


class Frame:
    """Device frame buffers (rm_frame_create)."""

    CHANNELS = {RM_BUF_POSITION: 3, RM_BUF_SHAPE_NORMAL: 3, RM_BUF_SURFACE_NORMAL: 3, RM_BUF_EMISSION: 3,
                RM_BUF_BASE_COLOR: 3, RM_BUF_DD: 4, RM_BUF_DS: 4, RM_BUF_ID: 4, RM_BUF_IS: 4, RM_BUF_PIXELS: 3,
                RM_BUF_MATERIAL: 3}

    def __init__(self, width: int, height: int):
        h = C.c_void_p()
        check(lib().rm_frame_create(width, height, C.byref(h)))
        self._h = h
        self.width, self.height = width, height

    def render(self, scene: Scene, args: RmRenderArgs, stream=None):
        check(lib().rm_render(scene.handle, C.byref(args), self._h, _stream(stream)))

    def stats(self) -> dict:
        """Counters of the frame's last render (rm_frame_get_stats), by name."""
        from ._lib import RM_STAT_COUNT, STAT_NAMES
        v = (C.c_uint64 * RM_STAT_COUNT)()
        check(lib().rm_frame_get_stats(self._h, v))
        return {k: int(v[i]) for i, k in enumerate(STAT_NAMES)}

    def set_recording(self, on: bool = True):
        """Keep per-sample sums for rm_adjoint (one replay per sample instead of two)."""
        check(lib().rm_frame_set_recording(self._h, int(on)))

    def postprocess(self, options: int, exposure: float = 1.0, stream=None):
        check(lib().rm_postprocess(self._h, options, exposure, _stream(stream)))

    def spatial_clamp(self, stream=None):
        """Photo::spatialClamp on the four radiance buffers, in place."""
        check(lib().rm_spatial_clamp(self._h, _stream(stream)))

    def filter(self, stream=None):
        """Photo::filter (filterVar + five a-trous passes) on the four radiance
        buffers, in place."""
        check(lib().rm_filter(self._h, _stream(stream)))

    def set_dof(self, focus: float, coc: float, camera_position):
        """Photo's focus, CoC and cameraPosition for option 1024 (depth of field)."""
        cam = (C.c_float * 3)(*[float(v) for v in camera_position])
        check(lib().rm_frame_set_dof(self._h, focus, coc, cam))

    def upload(self, which: int, data: np.ndarray):
        """Host array [h][w][channels] float32 into one buffer (synchronous)."""
        ch = self.CHANNELS[which]
        a = np.ascontiguousarray(data, dtype=np.float32)
        if a.size != self.width * self.height * ch:
            raise ValueError(f"upload: expected {self.height}x{self.width}x{ch} floats, got shape {a.shape}")
        check(lib().rm_frame_upload(self._h, which, a.ctypes.data_as(C.c_void_p)))

    def download(self, which: int) -> np.ndarray:
        ch = self.CHANNELS[which]
        out = np.empty((self.height, self.width, ch), dtype=np.float32)
        check(lib().rm_frame_download(self._h, which, _fp(out)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rm_frame_destroy(h)
            except (TypeError, AttributeError):  # interpreter shutdown: module globals already cleared
                pass
            self._h = None

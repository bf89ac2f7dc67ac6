"""rm_scene_load for binary FBX on the device (Model::Model through assimp,
src/model.cpp:172-215).  The displaced-sphere workload, written as FBX (one
model and geometry per mesh, every triangle with its own control points,
per-polygon-vertex normals and uvs, one material per mesh, deflated arrays,
version 7500) with its sky as a Radiance HDR file, loads to the scene its
description encodes: default materials (FBX colours are not read by the
reference) in order of first use, the decoded sky.  Same BVH permutation,
light objects and render, bit for bit.  An emissive map makes a light object
when there is no sky (src/model.cpp:121, checkLightObject)."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest

import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from raym0nade_amd._lib import RmMaterial

import fbx_writer as W
from cutout_util import make_desc
from test_assets import encode_hdr

pytestmark = pytest.mark.gpu


def view(ptr, ctype, count):
    arr = (ctype * count).from_address(C.cast(ptr, C.c_void_p).value)
    return np.ctypeslib.as_array(arr).copy()


def to_rgbe(rgb):
    """Ward's shared-exponent encoding of float RGB [h][w][3]."""
    v = rgb.max(axis=2).astype(np.float64)
    m, e = np.frexp(v)
    ok = v > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out = np.zeros(rgb.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def default_material():
    m = RmMaterial()
    for c in range(3):
        m.base_color[c] = 1.0
    m.roughness, m.metallic, m.opacity, m.ior = 0.8, 0.0, 1.0, 1.0
    return m


def small(args, w=64, spp=8):
    a = rm.RmRenderArgs.from_buffer_copy(args)
    a.accuracy = np.float32(a.accuracy * a.width / w)
    a.width = a.height = w
    a.spp = spp
    return a


def write_meshes(g, folder, name, emissive_mesh=None, emissive_file=None):
    """The generated scene as FBX; returns the source material of each FBX
    material in first-use order."""
    d = g.desc
    n = d.n_tris
    verts = view(d.verts, C.c_float, n * 9).reshape(n, 3, 3)
    uvs = view(d.uvs, C.c_float, n * 6).reshape(n, 3, 2)
    nrm = view(d.normals, C.c_float, n * 9).reshape(n, 3, 3)
    mat = view(d.material, C.c_int32, n)
    mesh = view(d.mesh, C.c_int32, n)
    assert (np.diff(mesh) >= 0).all()  # meshes contiguous, in order
    objs, conns, used = [], [], []
    for k in np.unique(mesh):
        idx = np.nonzero(mesh == k)[0]
        assert (mat[idx] == mat[idx[0]]).all()
        pts = verts[idx].reshape(-1, 3).tolist()
        polys = [[3 * i, 3 * i + 1, 3 * i + 2] for i in range(len(idx))]
        objs += [W.model(1000 + k, f"mesh{k}"),
                 W.geometry(2000 + k, f"g{k}", pts, polys, normals=(nrm[idx].ravel().tolist(), "ByPolygonVertex", "Direct", None),
                            uvs=(uvs[idx].ravel().tolist(), "ByPolygonVertex", "Direct", None),
                            materials=("AllSame", [0]), compress=True)]
        mid = 3000 + int(mat[idx[0]]) if emissive_mesh != k else 4000 + int(k)
        conns += [W.oo(1000 + k, 0), W.oo(2000 + k, 1000 + k), W.oo(mid, 1000 + k)]
        if mid not in used:
            used.append(mid)
    objs += [W.material(m, f"m{m}") for m in used]
    if emissive_mesh is not None:
        objs.append(W.texture(5000, "glow", emissive_file))
        conns.append(W.op(5000, 4000 + emissive_mesh, "EmissiveColor"))
    (folder / name).write_bytes(W.document(objs, conns, version=7500))
    remap = {m: i for i, m in enumerate(used)}
    return verts, uvs, nrm, np.array([remap[3000 + int(x)] if 3000 + int(x) in remap else -1 for x in mat], np.int32), mesh, len(used)


def test_fbx_scene_equals_its_description(dev, tmp_path):
    g = rm.generate(2, 3, grid=2, freq=12, sky_width=64, sky_height=32)
    d = g.desc
    sky = view(d.sky_rgb, C.c_float, 64 * 32 * 3).reshape(32, 64, 3)
    (tmp_path / "sky.hdr").write_bytes(encode_hdr(to_rgbe(sky), "rle"))
    verts, uvs, nrm, material, mesh, n_mats = write_meshes(g, tmp_path, "scene.fbx")
    S1 = rm.Scene.load(str(tmp_path) + "/", "scene.fbx", "sky.hdr")
    sky_f = np.ascontiguousarray(rm.load_image(str(tmp_path / "sky.hdr")), np.float32)
    desc, keep = make_desc(verts, uvs, material, mesh, [default_material() for _ in range(n_mats)], [], normals=nrm)
    desc.sky_width, desc.sky_height = 64, 32
    desc.sky_rgb = sky_f.ctypes.data_as(C.POINTER(C.c_float))
    S0 = rm.Scene(desc)
    for k in ("n_tris", "n_nodes", "depth", "n_light_objects", "n_light_faces", "has_sky"):
        assert getattr(S1.info, k) == getattr(S0.info, k), k
    assert np.array_equal(S1.perm(), S0.perm())
    a = small(g.args)
    gb0, rad0 = RC.gpu_render(S0, a)
    gb1, rad1 = RC.gpu_render(S1, a)
    RC.assert_gbuffer_equal(gb1, gb0)
    assert np.array_equal(rad1.view(np.uint32), rad0.view(np.uint32))
    assert np.isfinite(rad1).all() and rad1.max() > 0
    del keep


def png_rgb(w, h, rgb):
    raw = b"".join(b"\0" + bytes(rgb) * w for _ in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_fbx_emissive_map_makes_a_light(dev, tmp_path):
    """The Cornell box as FBX, its light quad's material with a white
    EmissiveColor map: one light object of two faces, and a lit frame; the
    same file without the map has no light and renders black."""
    g = rm.generate(0, 1)
    d = g.desc
    n = d.n_tris
    mesh = view(d.mesh, C.c_int32, n)
    mat = view(d.material, C.c_int32, n)
    mats = (RmMaterial * d.n_materials).from_address(C.cast(d.materials, C.c_void_p).value)
    light = [k for k in np.unique(mesh) if max(mats[int(mat[np.nonzero(mesh == k)[0][0]])].emission) > 0]
    assert len(light) == 1
    (tmp_path / "glow.png").write_bytes(png_rgb(2, 2, (255, 255, 255)))
    write_meshes(g, tmp_path, "lit.fbx", emissive_mesh=int(light[0]), emissive_file="glow.png")
    write_meshes(g, tmp_path, "dark.fbx")
    lit = rm.Scene.load(str(tmp_path) + "/", "lit.fbx")
    dark = rm.Scene.load(str(tmp_path) + "/", "dark.fbx")
    assert (lit.info.n_light_objects, lit.info.n_light_faces) == (1, 2)
    assert dark.info.n_light_objects == 0
    a = small(g.args, 32, 4)
    _, rad_lit = RC.gpu_render(lit, a)
    _, rad_dark = RC.gpu_render(dark, a)
    assert np.isfinite(rad_lit).all() and rad_lit.max() > 0
    assert rad_dark.max() == 0

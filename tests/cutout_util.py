"""Scenes with alpha cut-outs (materials whose albedo texture has alpha
values below 1) for the TransparentTest tests: stacks of textured quads in
front of an opaque back quad, and the Cornell box with a cut-out grille."""
from __future__ import annotations

import ctypes as C

import numpy as np

from raym0nade_amd._lib import RmMaterial, RmSceneDesc, RmTexture

F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)


def _p(a, t=F32P):
    return a.ctypes.data_as(t)


def quad(z, half=1.0):
    """Two triangles spanning [-half, half]^2 at depth z, facing +z, with uv
    (x / half + 1) / 2, (y / half + 1) / 2."""
    h = half
    v = np.float32([[[-h, -h, z], [h, -h, z], [h, h, z]], [[-h, -h, z], [h, h, z], [-h, h, z]]])
    uv = np.float32([[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]])
    return v, uv


def make_desc(verts, uvs, material, mesh, mats, textures, normals=None):
    """RmSceneDesc over numpy arrays.  mats: RmMaterial list; textures: list of
    (rgb [h][w][3], alpha [h][w] or None) fp32 textures or uint8 [h][w][3|4]
    arrays (8-bit textures).  Returns (desc, keepalive)."""
    verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
    uvs = np.ascontiguousarray(uvs, np.float32).reshape(-1, 6)
    material = np.ascontiguousarray(material, np.int32)
    mesh = np.ascontiguousarray(mesh, np.int32)
    keep = [verts, uvs, material, mesh]
    m_arr = (RmMaterial * len(mats))(*mats)
    t_arr = (RmTexture * max(1, len(textures)))()
    keep += [m_arr, t_arr]
    for i, tex in enumerate(textures):
        if isinstance(tex, np.ndarray) and tex.dtype == np.uint8:  # 8-bit [h][w][channels]
            b = np.ascontiguousarray(tex)
            keep.append(b)
            t_arr[i].height, t_arr[i].width, t_arr[i].channels = b.shape
            t_arr[i].bytes = b.ctypes.data_as(C.POINTER(C.c_uint8))
            continue
        rgb, alpha = tex
        rgb = np.ascontiguousarray(rgb, np.float32)
        keep.append(rgb)
        t_arr[i].height, t_arr[i].width = rgb.shape[:2]
        t_arr[i].rgb = _p(rgb)
        if alpha is not None:
            alpha = np.ascontiguousarray(alpha, np.float32)
            assert alpha.shape == rgb.shape[:2]
            keep.append(alpha)
            t_arr[i].alpha = _p(alpha)
    d = RmSceneDesc()
    d.n_tris = verts.shape[0]
    d.verts = _p(verts)
    d.uvs = _p(uvs)
    if normals is not None:
        normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
        keep.append(normals)
        d.normals = _p(normals)
    d.material = _p(material, I32P)
    d.mesh = _p(mesh, I32P)
    d.n_materials = len(mats)
    d.materials = C.cast(m_arr, C.POINTER(RmMaterial))
    d.n_textures = len(textures)
    d.textures = C.cast(t_arr, C.POINTER(RmTexture))
    return d, keep


def material(base=(0.7, 0.7, 0.7), emission=(0.0, 0.0, 0.0), tex=-1, roughness=0.8):
    m = RmMaterial()
    for c in range(3):
        m.base_color[c] = base[c]
        m.emission[c] = emission[c]
    m.roughness, m.metallic, m.opacity, m.ior, m.albedo_tex = roughness, 0.0, 1.0, 1.0, tex
    return m


def layers(alphas, back=True):
    """Quads at z = -1, -1.5, ... (one per alpha plane, each with its own
    material and texture), behind them, first in face order, an opaque quad
    at z = -5 (faces 0 and 1).  alphas[i] None: layer i has no alpha plane."""
    vs, uvs, mat, mesh = [], [], [], []
    mats = [material()]
    texs = []
    if back:
        v, uv = quad(-5.0, 3.0)
        vs.append(v)
        uvs.append(uv)
        mat += [0, 0]
        mesh += [0, 0]
    for i, a in enumerate(alphas):
        v, uv = quad(-1.0 - 0.5 * i)
        vs.append(v)
        uvs.append(uv)
        mat += [i + 1] * 2
        mesh += [i + 1] * 2
        shape = (4, 4) if a is None else np.shape(a)
        texs.append((np.full(shape + (3,), 0.5, np.float32), a))
        mats.append(material(tex=i))
    if not vs:
        raise ValueError("empty scene")
    return make_desc(np.concatenate(vs), np.concatenate(uvs), mat, mesh, mats, texs)


def down_rays(xy):
    """Rays from (x, y, 0) along -z, [n][6]."""
    xy = np.asarray(xy, np.float32).reshape(-1, 2)
    r = np.zeros((xy.shape[0], 6), np.float32)
    r[:, :2] = xy
    r[:, 5] = -1.0
    return r


def cornell_with_grille(gen, cells=8):
    """The generated Cornell box plus a quad across the box's middle, facing
    the camera, whose albedo texture is a checkerboard of alpha 0 and 1 cells."""
    d = gen.desc
    n = d.n_tris
    verts = np.ctypeslib.as_array(d.verts, shape=(n * 9,)).reshape(n, 3, 3).copy()
    uvs = (np.ctypeslib.as_array(d.uvs, shape=(n * 6,)).reshape(n, 3, 2).copy() if d.uvs
           else np.zeros((n, 3, 2), np.float32))
    material_ids = np.ctypeslib.as_array(d.material, shape=(n,)).copy()
    mesh = (np.ctypeslib.as_array(d.mesh, shape=(n,)).copy() if d.mesh else np.zeros(n, np.int32))
    mats = [RmMaterial.from_buffer_copy(d.materials[i]) for i in range(d.n_materials)]
    assert d.n_textures == 0
    normals = np.ctypeslib.as_array(d.normals, shape=(n * 9,)).reshape(n, 3, 3).copy() if d.normals else None
    lo, hi = verts.reshape(-1, 3).min(0), verts.reshape(-1, 3).max(0)
    c, ext = (lo + hi) / 2, (hi - lo) / 2
    axis = int(np.argmax(np.abs(np.float32(gen.args.direction))))
    u_ax, v_ax = [k for k in range(3) if k != axis]
    q = np.zeros((2, 3, 3), np.float32)
    corners = [(-1, -1), (1, -1), (1, 1), (-1, -1), (1, 1), (-1, 1)]
    for k, (a, b) in enumerate(corners):
        p = c.copy()
        p[u_ax] += 0.6 * ext[u_ax] * a
        p[v_ax] += 0.6 * ext[v_ax] * b
        q[k // 3, k % 3] = p
    quv = np.float32([[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]])
    alpha = np.float32((np.add.outer(np.arange(cells), np.arange(cells)) % 2))
    rgb = np.full((cells, cells, 3), 0.8, np.float32)
    mats.append(material(base=(0.8, 0.8, 0.8), tex=0))
    if normals is not None:
        g_n = np.cross(q[0, 1] - q[0, 0], q[0, 2] - q[0, 0])
        normals = np.concatenate([normals, np.broadcast_to(g_n / np.linalg.norm(g_n), (2, 3, 3))]).astype(np.float32)
    return make_desc(np.concatenate([verts, q]), np.concatenate([uvs, quv]),
                     np.concatenate([material_ids, [len(mats) - 1] * 2]), np.concatenate([mesh, [mesh.max() + 1] * 2]),
                     mats, [(rgb, alpha)], normals=normals)

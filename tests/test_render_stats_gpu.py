"""Render counters (rm_frame_get_stats) and the medium stack.

  * the counters of a render: one primary ray per pixel, shading vertices,
    paths, extension and shadow rays; counting renders (RM_TUNE_RENDER_COUNT)
    give the same image bit for bit and fill the walk counters;
  * five nested glass spheres inside the Cornell box: paths that nest up to
    five dielectrics (and re-enter media after internal reflections) keep the
    reference's whole medium stack (src/render.cpp:13-42), so the render
    agrees with the oracle's unbounded multimap, and nothing overflows.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from raym0nade_amd._lib import RmMaterial

from cutout_util import make_desc

pytestmark = pytest.mark.gpu


def small_args(g, w, h, spp, **kw):
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / w)
    a.width, a.height, a.spp = w, h, spp
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def icosphere(center, radius, subdiv):
    """Triangles [n][3][3] of a subdivided icosahedron, counter-clockwise seen
    from outside, and its vertex normals [n][3][3] (pointing out)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    tris = np.array([[v[a], v[b], v[c]] for a, b, c in f], np.float64)
    tris /= np.linalg.norm(tris, axis=-1, keepdims=True)
    for _ in range(subdiv):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=-1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=-1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=-1, keepdims=True)
        tris = np.concatenate([np.stack(x, 1) for x in ((a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca))])
    n = tris.copy()
    return (np.asarray(center) + radius * tris).astype(np.float32), n.astype(np.float32)


def glass(ior, tint):
    m = RmMaterial()
    for c in range(3):
        m.base_color[c] = 1.0
        m.transmitting_color[c] = tint[c]
    m.roughness, m.metallic, m.opacity, m.ior, m.albedo_tex = 0.05, 0.0, 0.0, ior, -1
    return m


def nested_spheres_scene(gen, count=5):
    """The generated Cornell box with `count` concentric glass spheres of
    distinct materials (ior 1.3..1.7, each tinted) at its centre."""
    d = gen.desc
    n = d.n_tris
    verts = np.ctypeslib.as_array(d.verts, shape=(n * 9,)).reshape(n, 3, 3).copy()
    uvs = np.zeros((n, 3, 2), np.float32)
    mat = np.ctypeslib.as_array(d.material, shape=(n,)).copy()
    mesh = np.ctypeslib.as_array(d.mesh, shape=(n,)).copy() if d.mesh else np.zeros(n, np.int32)
    normals = np.ctypeslib.as_array(d.normals, shape=(n * 9,)).reshape(n, 3, 3).copy() if d.normals else None
    if normals is None:
        e = np.cross(verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0])
        e /= np.linalg.norm(e, axis=-1, keepdims=True)
        normals = np.repeat(e[:, None], 3, 1).astype(np.float32)
    mats = [RmMaterial.from_buffer_copy(d.materials[i]) for i in range(d.n_materials)]
    lo, hi = verts.reshape(-1, 3).min(0), verts.reshape(-1, 3).max(0)
    c = (lo + hi) / 2
    r0 = 0.3 * float((hi - lo).min())
    vs, ns, ms, me = [verts], [normals], [mat], [mesh]
    for k in range(count):
        v, nn = icosphere(c, r0 * (1.0 - 0.17 * k), 3)
        mats.append(glass(1.3 + 0.1 * k, (0.95 - 0.05 * k, 0.9, 0.85 + 0.03 * k)))
        vs.append(v)
        ns.append(nn)
        ms.append(np.full(len(v), len(mats) - 1, np.int32))
        me.append(np.full(len(v), mesh.max() + 1 + k, np.int32))
    V = np.concatenate(vs)
    return make_desc(V, np.zeros((len(V), 3, 2), np.float32), np.concatenate(ms), np.concatenate(me), mats, [],
                     normals=np.concatenate(ns))


@pytest.fixture(scope="module")
def cornell(dev):
    g = rm.generate(0, 1)
    return g, rm.Scene(g)


def test_render_counters(cornell):
    g, S = cornell
    a = small_args(g, 64, 48, 16)
    fr = rm.Frame(a.width, a.height)
    fr.render(S, a)
    st = fr.stats()
    gb, rad = RC.download_all(fr)
    n = a.width * a.height
    hit = np.isfinite(gb[..., 0])
    emits = np.linalg.norm(gb[..., 9:12], axis=-1) > 0
    n_ind = a.spp - int(np.float32(a.spp) * np.float32(a.p_direct))
    assert st["primary_rays"] == n
    assert st["paths"] <= int((hit & ~emits).sum()) * n_ind and st["paths"] > 0
    assert st["extension_rays"] >= st["paths"]
    assert st["hits"] >= int(hit.sum()) and st["hits"] <= st["primary_rays"] + st["extension_rays"]
    assert st["shadow_rays"] > 0 and st["medium_overflow"] == 0
    assert st["walk_interior"] == st["walk_leaves"] == st["walk_boxes"] == st["walk_tris"] == 0
    # a second render resets the counters
    fr.render(S, a)
    assert fr.stats() == st


@pytest.mark.parametrize("split", [0, 1])
def test_counting_render_is_identical(dev, split):
    g = rm.generate(2, 3, grid=2, freq=12, sky_width=256, sky_height=128)
    S = rm.Scene(g)
    a = small_args(g, 80, 48, 12, max_depth=16)
    keys = (rm.TUNE_RENDER_COUNT, rm.TUNE_NEE_SPLIT)
    old = tuple(rm.get_tuning(k) for k in keys)
    try:
        rm.set_tuning(rm.TUNE_NEE_SPLIT, split)
        fr = rm.Frame(a.width, a.height)
        fr.render(S, a)
        gb0, r0 = RC.download_all(fr)
        s0 = fr.stats()
        rm.set_tuning(rm.TUNE_RENDER_COUNT, 1)
        fr.render(S, a)
        gb1, r1 = RC.download_all(fr)
        s1 = fr.stats()
    finally:
        for k, v in zip(keys, old):
            rm.set_tuning(k, v)
    assert np.array_equal(gb0.view(np.uint32), gb1.view(np.uint32))
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
    for k in ("primary_rays", "extension_rays", "shadow_rays", "hits", "paths"):
        assert s0[k] == s1[k], k
    # every walk visits the root, tests two boxes per interior visit and at
    # least one leaf per hit
    walks = s1["extension_rays"] + s1["shadow_rays"]
    assert s1["walk_interior"] >= walks and s1["walk_boxes"] == 2 * s1["walk_interior"]
    assert s1["walk_tris"] >= s1["walk_leaves"] > 0


def test_nested_dielectrics_keep_the_whole_medium_stack(cornell):
    g, _ = cornell
    desc, keep = nested_spheres_scene(g, 5)
    S, OS = rm.Scene(desc), O.OracleScene(desc)
    a = small_args(g, 96, 96, 24, max_depth=16)
    fr = rm.Frame(a.width, a.height)
    fr.render(S, a)
    st = fr.stats()
    gb, rad = RC.download_all(fr)
    ogb, orad = OS.render(a)
    RC.assert_gbuffer_equal(gb, ogb)
    m = RC.image_metrics(gb, rad, ogb, orad, a.spp, a.p_direct, a.exposure)
    c = RC.compare_radiance(rad, orad)
    print("nested dielectrics:", st, m, c)
    assert st["medium_overflow"] == 0
    assert st["medium_max"] >= 5  # some paths are inside all five spheres at once
    assert m["max_abs_z"] < 4.0
    assert c["direct_close"] >= 0.99 and c["indirect_close"] >= 0.95

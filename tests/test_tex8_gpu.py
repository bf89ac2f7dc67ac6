"""The reference's 8-bit texture maps on the GPU against the oracle
(Material's getters, src/material.cpp:26-148,337-383; calcSurfaceNormal,
src/model.cpp:279-298; emissive light objects, src/model.cpp:23-82,118-122).

A room whose floor has an RGBA diffuse map, a normal map and a specular
map, a cut-out quad with RGBA alpha, a quad with 3-channel (PNG-layout)
texels read with the 4-byte stride, and a ceiling panel lit by an emissive
texture.  The G-buffer equals the oracle's bit for bit, except the values
that pass through powf / log2f (base colour, emission, and the normal-mapped
surface normal, whose mip depth is a log2f), which agree to a few units in
the last place; radiance agrees pixel by pixel within the Monte Carlo noise.
Closest and any-hit rays through the cut-outs are bit-exact.  The same scene
written as OBJ + MTL with PNG and DDS maps and read back with rm_scene_load
renders identically to the scene built from arrays."""
import ctypes as C
import struct
import zlib

import numpy as np
import pytest
import torch

import oracle_py as O
import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from raym0nade_amd._lib import RmMaterial

from cutout_util import make_desc

pytestmark = pytest.mark.gpu


def quad4(p0, p1, p2, p3, uv_scale=1.0):
    """Two triangles (p0, p1, p2), (p0, p2, p3) with uvs (0,0) (1,0) (1,1) (0,1) x uv_scale."""
    P = np.float32([p0, p1, p2, p3])
    UV = np.float32([[0, 0], [1, 0], [1, 1], [0, 1]]) * np.float32(uv_scale)
    return np.stack([P[[0, 1, 2]], P[[0, 2, 3]]]), np.stack([UV[[0, 1, 2]], UV[[0, 2, 3]]])


def mat(base=(0.6, 0.6, 0.6), tex=-1, ntex=-1, etex=-1, stex=-1, emission=(0, 0, 0)):
    m = RmMaterial()
    m.base_color[:] = base
    m.emission[:] = emission
    m.roughness, m.metallic, m.opacity, m.ior = 0.8, 0.0, 1.0, 1.0
    m.albedo_tex, m.normal_tex, m.emissive_tex, m.specular_tex = tex, ntex, etex, stex
    return m


def textures(seed=1):
    rng = np.random.default_rng(seed)
    floor = rng.integers(30, 256, (32, 32, 4), dtype=np.uint8)
    floor[..., 3] = 255
    normal = np.clip(rng.normal([128, 128, 230], [45, 45, 10], (16, 16, 3)), 0, 255).astype(np.uint8)
    spec = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    leaf = rng.integers(0, 256, (32, 32, 4), dtype=np.uint8)
    leaf[..., 3] = np.where(rng.random((32, 32)) < 0.4, 0, 255)
    rgb = rng.integers(0, 256, (16, 16, 3), dtype=np.uint8)
    glow = rng.integers(100, 256, (8, 8, 4), dtype=np.uint8)
    return [floor, normal, spec, leaf, rgb, glow]


def room(seed=1):
    """(verts, uvs, material, mesh, mats, texs): back wall, floor, two
    cut-out quads, an emissive ceiling panel and a small constant lamp."""
    parts = [
        (quad4((-3, -1.5, -2), (3, -1.5, -2), (3, 3, -2), (-3, 3, -2)), mat()),
        (quad4((-3, -1.5, 3), (3, -1.5, 3), (3, -1.5, -2), (-3, -1.5, -2), 2.0), mat(tex=0, ntex=1, stex=2)),
        (quad4((-1.8, -0.8, -0.5), (-0.2, -0.8, -0.5), (-0.2, 0.8, -0.5), (-1.8, 0.8, -0.5)), mat(tex=3)),
        (quad4((0.2, -0.8, -0.5), (1.8, -0.8, -0.5), (1.8, 0.8, -0.5), (0.2, 0.8, -0.5)), mat(tex=4)),
        (quad4((-1, 1.4, -1.5), (1, 1.4, -1.5), (1, 1.4, 0), (-1, 1.4, 0)), mat(etex=5)),
        (quad4((2.0, 2.0, 1.0), (2.5, 2.0, 1.0), (2.5, 2.0, 1.5), (2.0, 2.0, 1.5)), mat(emission=(4, 4, 4))),
    ]
    vs = np.concatenate([p[0][0] for p in parts])
    uvs = np.concatenate([p[0][1] for p in parts])
    material = np.repeat(np.arange(len(parts), dtype=np.int32), 2)
    return vs, uvs, material, material.copy(), [p[1] for p in parts], textures(seed)


def camera(w=96, spp=32):
    a = rm.RmRenderArgs()
    a.position[:] = (0.0, 0.3, 4.0)
    a.direction[:] = (0.0, 0.0, -1.0)
    a.up[:] = (0.0, 1.0, 0.0)
    a.right[:] = (1.0, 0.0, 0.0)
    a.accuracy = np.float32(0.55 / (w / 2))
    a.exposure, a.p_direct = 1.0, 0.5
    a.width = a.height = w
    a.spp, a.threads, a.seed, a.max_depth, a.rr_min_depth = spp, 1, 7, 8, 3
    return a


def test_textured_room_matches_the_oracle(dev):
    vs, uvs, material, mesh, mats, texs = room()
    d, keep = make_desc(vs, uvs, material, mesh, mats, texs)
    S, OS = rm.Scene(d), O.OracleScene(d)
    assert S.info.n_light_objects == len(OS.lights()) == 2
    a = camera()
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    hit = np.isfinite(ogb[..., 0])
    mid = np.argwhere(hit)
    print("hit pixels:", int(hit.sum()), "materials seen:", np.unique(np.round(ogb[hit][:, 15:18], 3), axis=0).shape[0])
    # the floor's specular map varies metallic and roughness; the normal map tilts its normals
    floor = hit & (np.abs(ogb[..., 1] + 1.5) < 1e-4)
    assert floor.sum() > 500
    assert np.unique(ogb[floor][:, 16]).size > 50 and np.unique(ogb[floor][:, 15]).size > 50
    assert (np.abs(ogb[floor][:, 6:9] - ogb[floor][:, 3:6]).max(axis=-1) > 1e-2).mean() > 0.9
    RC.assert_gbuffer_equal(gb, ogb, ulp_channels=(6, 7, 8, 9, 10, 11, 12, 13, 14), ulps=8)
    m = RC.image_metrics(gb, rad, ogb, orad, a.spp, a.p_direct, a.exposure)
    st = RC.compare_radiance(rad, orad)
    print("textured room 96^2 32 spp:", m, st, "pixels", len(mid))
    assert m["max_abs_z"] < 4.0
    assert st["direct_close"] >= 0.99 and st["indirect_close"] >= 0.95


def test_rays_through_8bit_cutouts_match_the_oracle(dev):
    vs, uvs, material, mesh, mats, texs = room(3)
    d, keep = make_desc(vs, uvs, material, mesh, mats, texs)
    S, OS = rm.Scene(d), O.OracleScene(d)
    rng = np.random.default_rng(2)
    n = 20000
    rays = np.zeros((n, 6), np.float32)
    rays[:, 0] = rng.uniform(-1.9, 1.9, n)
    rays[:, 1] = rng.uniform(-0.9, 0.9, n)
    rays[:, 2] = 2.0
    dirs = np.stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), -np.ones(n)], axis=1)
    rays[:, 3:] = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    ot, oid = OS.closest(rays)
    dr = torch.from_numpy(rays).to(dev)
    t, ids = S.trace_closest(dr)
    torch.cuda.synchronize()
    assert np.array_equal(ids.cpu().numpy(), oid)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ot.view(np.uint32))
    through = np.isin(oid, [0, 1]) & (np.abs(rays[:, 0]) > 0.25)  # reached the wall behind a quad
    stopped = np.isin(oid, [4, 5, 6, 7])
    print("stopped on a quad:", stopped.mean(), "passed through:", through.mean())
    assert stopped.mean() > 0.2 and through.mean() > 0.1
    aims = rng.uniform(1.0, 6.0, n).astype(np.float32)
    occ = S.trace_any(dr, aim_dev=torch.from_numpy(aims).to(dev))
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy(), OS.any(rays, aims=aims))


# ------------------------------------------------------------ the asset path
def png_rgb(img):
    h, w = img.shape[:2]
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def dds_bgra(img):
    """Uncompressed 32-bit DDS with an alpha mask; texels stored B, G, R, A."""
    h, w = img.shape[:2]
    hdr = bytearray(128)
    hdr[0:4] = b"DDS "
    struct.pack_into("<IIII", hdr, 4, 124, 0x1007, h, w)
    struct.pack_into("<IIIIIIII", hdr, 76, 32, 0x41, 0, 32, 0x00FF0000, 0x0000FF00, 0x000000FF, 0xFF000000)
    return bytes(hdr) + img[..., [2, 1, 0, 3]].tobytes()


def write_room(folder, vs, uvs, material, mats, texs):
    names = []
    for i, t in enumerate(texs):
        if t.shape[2] == 3:
            names.append(f"t{i}.png")
            (folder / names[-1]).write_bytes(png_rgb(t))
        else:
            names.append(f"t%20{i}.dds")  # URL-encoded space in the MTL, a space in the file name
            (folder / f"t {i}.dds").write_bytes(dds_bgra(t))
    with open(folder / "room.mtl", "w") as f:
        for i, m in enumerate(mats):
            f.write(f"newmtl m{i}\nKd {' '.join(repr(float(v)) for v in m.base_color)}\n")
            if any(m.emission):
                f.write(f"Ke {' '.join(repr(float(v)) for v in m.emission)}\n")
            for key, tex in (("map_Kd", m.albedo_tex), ("norm", m.normal_tex), ("map_Ke", m.emissive_tex),
                             ("map_Ks", m.specular_tex)):
                if tex >= 0:
                    f.write(f"{key} {names[tex]}\n")
    with open(folder / "room.obj", "w") as f:
        f.write("mtllib room.mtl\n")
        vi = 1
        for k in range(vs.shape[0]):
            if k == 0 or material[k] != material[k - 1]:
                f.write(f"g mesh{material[k]}\nusemtl m{material[k]}\n")
            for c in range(3):
                f.write("v " + " ".join(repr(float(v)) for v in vs[k, c]) + "\n")
                f.write("vt " + " ".join(repr(float(v)) for v in uvs[k, c]) + "\n")
            f.write(f"f {vi}/{vi} {vi + 1}/{vi + 1} {vi + 2}/{vi + 2}\n")
            vi += 3


def test_obj_with_png_and_dds_maps_loads_as_the_same_scene(dev, tmp_path):
    vs, uvs, material, mesh, mats, texs = room(5)
    write_room(tmp_path, vs, uvs, material, mats, texs)
    d, keep = make_desc(vs, uvs, material, mesh, mats, texs)
    S0 = rm.Scene(d)
    S1 = rm.Scene.load(str(tmp_path) + "/", "room.obj")
    for k in ("n_tris", "n_nodes", "depth", "n_light_objects", "n_light_faces"):
        assert getattr(S1.info, k) == getattr(S0.info, k), k
    a = camera(64, 8)
    gb0, rad0 = RC.gpu_render(S0, a)
    gb1, rad1 = RC.gpu_render(S1, a)
    RC.assert_gbuffer_equal(gb1, gb0)
    assert np.array_equal(rad1.view(np.uint32), rad0.view(np.uint32))

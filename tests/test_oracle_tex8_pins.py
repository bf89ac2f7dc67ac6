"""Pins of the oracle's 8-bit texture model (CPU only): ImageData's integer
mips, strided texel reads, bilinear/trilinear filtering and hasTransparentPart
(src/material.cpp:26-148), the material lookups that use it -- gamma after
filtering, normal maps, emissive maps, the specular map's metallic and
roughness (src/material.cpp:337-383) -- calcSurfaceNormal's tilt of the
shading normal in the face's uv frame (src/model.cpp:279-298) and the light
object of an emissive texture (getAverageEmissiveColor, src/model.cpp:23-82).

Every expectation is a closed form: texel reads at integer texel
coordinates, hand-made mip levels, the tilt of a uv-aligned face, rotations
of it, and a texture whose halves differ."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from raym0nade_amd._lib import RmMaterial

from cutout_util import make_desc, quad

L = O.lib()
f32p, i32p, u8p = O.f32p, O.i32p, O.u8p


def _sig():
    for name, res, args in [
        ("oracle_tex8_get", None, [C.c_int32, C.c_int32, C.c_int32, u8p, C.c_int32, f32p, C.c_int64, f32p]),
        ("oracle_tex8_mips", C.c_int32, [C.c_int32, C.c_int32, C.c_int32, u8p, u8p, i32p]),
        ("oracle_material8", None, [C.c_int32, C.c_int32, C.c_int32, u8p, C.c_int32, f32p, C.c_int64, f32p]),
        ("oracle_surface_normal", None, [f32p] * 6),
    ]:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


_sig()


def fp(a):
    return a.ctypes.data_as(f32p)


def bp(a):
    return a.ctypes.data_as(u8p)


def tex8_get(img, uvd, bw):
    img = np.ascontiguousarray(img, np.uint8)
    h, w, ch = img.shape
    uvd = np.ascontiguousarray(uvd, np.float32).reshape(-1, 3)
    out = np.empty((uvd.shape[0], 4), np.float32)
    L.oracle_tex8_get(w, h, ch, bp(img), bw, fp(uvd), uvd.shape[0], fp(out))
    return out


def mips(img):
    img = np.ascontiguousarray(img, np.uint8)
    h, w, ch = img.shape
    levels = L.oracle_tex8_mips(w, h, ch, bp(img), None, None)
    total = sum((w >> l) * (h >> l) * ch for l in range(levels))
    out = np.empty(total, np.uint8)
    tr = C.c_int32(0)
    L.oracle_tex8_mips(w, h, ch, bp(img), bp(out), C.byref(tr))
    res, o = [], 0
    for l in range(levels):
        n = (w >> l) * (h >> l) * ch
        res.append(out[o:o + n].reshape(h >> l, w >> l, ch))
        o += n
    return res, bool(tr.value)


def material8(img, which, uvd):
    img = np.ascontiguousarray(img, np.uint8)
    h, w, ch = img.shape
    uvd = np.ascontiguousarray(uvd, np.float32).reshape(-1, 3)
    out = np.empty((uvd.shape[0], 4), np.float32)
    L.oracle_material8(w, h, ch, bp(img), which, fp(uvd), uvd.shape[0], fp(out))
    return out


def corner(i, j, w, h, depth=0.0):
    """(u, v, depth) whose bilinear sample is texel (row j, column i) alone:
    u = i / w and 1 - v = j / h."""
    return (np.float32(i) / np.float32(w), np.float32(1.0) - np.float32(j) / np.float32(h), depth)


# ------------------------------------------------------------ mip levels
def box_level(prev):
    """One mip level by hand: the integer floor of each 2 x 2 mean, the
    block's second row/column wrapping to the first on an odd side."""
    ph, pw, ch = prev.shape
    h, w = ph >> 1, pw >> 1
    out = np.empty((h, w, ch), np.uint8)
    for y in range(h):
        for x in range(w):
            s = np.zeros(ch, np.int64)
            for dy in range(2):
                for dx in range(2):
                    s += prev[(2 * y + dy) % ph, (2 * x + dx) % pw]
            out[y, x] = s // 4
    return out


@pytest.mark.parametrize("shape", [(6, 8, 4), (5, 7, 3), (9, 16, 4), (1, 8, 4)])
def test_integer_mips(shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    got, _ = mips(img)
    want = [img]
    while (want[-1].shape[0] >> 1) and (want[-1].shape[1] >> 1) and len(want) < 8:
        want.append(box_level(want[-1]))
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_mip_rounding_and_level_cap():
    # 1 + 1 + 1 + 2 = 5: 5 / 4 floors to 1 (a rounded mean gives 1 too; 2+2+2+1 = 7 tells them apart)
    img = np.uint8([[[2], [2]], [[2], [1]]]).repeat(4, axis=2)
    lv, _ = mips(img)
    assert len(lv) == 2 and np.array_equal(lv[1], np.full((1, 1, 4), 1, np.uint8))
    # at most 8 levels (MAX_MIPMAP_LEVEL): a 512 x 512 texture stops at 4 x 4
    big = np.zeros((512, 512, 4), np.uint8)
    lv, _ = mips(big)
    assert len(lv) == 8 and lv[-1].shape == (4, 4, 4)


def test_transparent_part_reads_every_fourth_byte():
    opaque = np.full((4, 4, 4), 255, np.uint8)
    assert mips(opaque)[1] is False
    one = opaque.copy()
    one[2, 3, 3] = 254
    assert mips(one)[1] is True
    # colour bytes below 255 do not count on RGBA data ...
    rgba = opaque.copy()
    rgba[..., :3] = 7
    assert mips(rgba)[1] is False
    # ... but on RGB data every 4th byte is a colour byte
    rgb = np.full((4, 4, 3), 255, np.uint8)
    assert mips(rgb)[1] is False
    rgb[0, 1, 0] = 9  # byte 3
    assert mips(rgb)[1] is True
    rgb[0, 1, 0] = 255
    rgb[0, 0, 2] = 9  # byte 2 is not read
    assert mips(rgb)[1] is False


# ------------------------------------------------------------ texel reads
def test_vec4_reads_of_rgba_texels_wrap_and_flip():
    rng = np.random.default_rng(3)
    w, h = 8, 4
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    pts = [corner(i, j, w, h) for j in range(1, h) for i in range(w)]
    got = tex8_get(img, pts, 4)
    want = np.array([img[j, i] for j in range(1, h) for i in range(w)], np.float32) / np.float32(255)
    assert np.array_equal(got, want)
    # v = 1 is row 0; u and v wrap
    assert np.array_equal(tex8_get(img, [(3 / 8, 1.0, 0)], 4)[0], img[0, 3] / np.float32(255))
    assert np.array_equal(tex8_get(img, [(1 + 3 / 8, 1 - 2 / 4 + 1, 0)], 4)[0], img[2, 3] / np.float32(255))
    assert np.array_equal(tex8_get(img, [(-5 / 8, 1 - 2 / 4, 0)], 4)[0], img[2, 3] / np.float32(255))
    # halfway between two texels of a row: their mean (x0 from floor, x1 = x0 + 1)
    got = tex8_get(img, [(2.5 / 8, 1 - 1 / 4, 0)], 4)[0]
    np.testing.assert_allclose(got, (img[1, 2] / 255.0 + img[1, 3] / 255.0) / 2, rtol=1e-6)
    # the last column blends with the first
    got = tex8_get(img, [(7.5 / 8, 1 - 1 / 4, 0)], 4)[0]
    np.testing.assert_allclose(got, (img[1, 7] / 255.0 + img[1, 0] / 255.0) / 2, rtol=1e-6)


def test_vec4_reads_of_rgb_data_use_a_four_byte_stride():
    """A colour lookup reads 4 bytes at 4 x the texel index whatever the
    channel count: on RGB data texel k reads bytes 4k..4k+3 of the RGB
    stream; a read past the level's end reads 0."""
    rng = np.random.default_rng(4)
    w, h = 4, 4
    img = rng.integers(1, 256, (h, w, 3), dtype=np.uint8)
    flat = np.concatenate([img.reshape(-1), np.zeros(4 * w * h, np.uint8)])
    for j in range(h):
        for i in range(w):
            k = j * w + i
            got = tex8_get(img, [corner(i, j, w, h)], 4)[0]
            assert np.array_equal(got, flat[4 * k:4 * k + 4] / np.float32(255)), (i, j)
    # texel 12 starts at byte 48 = the end of the data: all zero
    assert np.array_equal(tex8_get(img, [corner(0, 3, w, h)], 4)[0], np.zeros(4, np.float32))


def test_vec3_reads_use_a_three_byte_stride():
    """A normal-map lookup reads 3 bytes at 3 x the texel index: on RGBA data
    texel k reads bytes 3k..3k+2 (the channels drift); the 4th component
    is 0."""
    rng = np.random.default_rng(5)
    w, h = 4, 2
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    flat = img.reshape(-1)
    for j in range(h):
        for i in range(w):
            k = j * w + i
            got = tex8_get(img, [corner(i, j, w, h)], 3)[0]
            assert np.array_equal(got[:3], flat[3 * k:3 * k + 3] / np.float32(255)) and got[3] == 0
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    got = tex8_get(rgb, [corner(i, j, w, h) for j in range(h) for i in range(w)], 3)
    assert np.array_equal(got[:, :3], rgb.reshape(-1, 3) / np.float32(255))


def test_trilinear_blend_and_depth_clamp():
    rng = np.random.default_rng(6)
    w, h = 8, 8
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    lv1 = box_level(img)
    lv2 = box_level(lv1)
    lv3 = box_level(lv2)
    u, v, _ = corner(4, 4, w, h)  # texel (4, 4) of level 0, (2, 2) of level 1, (1, 1) of level 2
    np.testing.assert_allclose(tex8_get(img, [(u, v, 1.0)], 4)[0], lv1[2, 2] / 255.0, rtol=1e-6)
    np.testing.assert_allclose(tex8_get(img, [(u, v, 0.25)], 4)[0],
                               0.75 * img[4, 4] / 255.0 + 0.25 * lv1[2, 2] / 255.0, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(tex8_get(img, [(u, v, 2.0)], 4)[0], lv2[1, 1] / 255.0, rtol=1e-6)
    # four levels (8, 4, 2, 1): depth clamps to 3
    np.testing.assert_allclose(tex8_get(img, [(u, v, 40.0)], 4)[0], lv3[0, 0] / 255.0, rtol=1e-6)
    assert np.array_equal(tex8_get(img, [(u, v, -2.0)], 4), tex8_get(img, [(u, v, 0.0)], 4))


# ------------------------------------------------------------ materials
def ulp_close(a, b, ulps=2):
    a, b = np.float32(a), np.float32(b)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64)) <= ulps


def test_diffuse_gamma_after_filtering_and_alpha_untouched():
    rng = np.random.default_rng(7)
    w, h = 4, 4
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    nan = np.float32(np.nan)
    pts = [corner(i, j, w, h, nan) for j in range(h) for i in range(w)]
    got = material8(img, 0, pts)
    c = img.reshape(-1, 4).astype(np.float64) / 255.0
    assert ulp_close(got[:, :3], (c[:, :3] ** np.float64(np.float32(2.2)))).all()
    assert np.array_equal(got[:, 3], (img.reshape(-1, 4)[:, 3] / np.float32(255)))
    # halfway between texels the bytes are filtered first, then raised to 2.2
    a, b = img[1, 1] / 255.0, img[1, 2] / 255.0
    got = material8(img, 0, [(1.5 / w, 1 - 1 / h, nan)])[0]
    np.testing.assert_allclose(got[:3], ((a + b) / 2)[:3] ** 2.2, rtol=1e-5)
    assert not np.allclose(got[:3], ((a[:3] ** 2.2 + b[:3] ** 2.2) / 2), rtol=1e-3) or np.allclose(a, b)


def test_duv_selects_the_mip_level():
    """map_depth = log2(duv x width): duv = 2 / width reads level 1."""
    rng = np.random.default_rng(8)
    w, h = 8, 8
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    lv1 = box_level(img)
    u, v, _ = corner(4, 2, w, h)
    got = material8(img, 0, [(u, v, 2.0 / w)])[0]
    assert ulp_close(got[:3], (lv1[1, 2, :3] / 255.0) ** np.float64(np.float32(2.2))).all()


def test_normal_map_values_and_emission():
    rng = np.random.default_rng(9)
    w, h = 4, 2
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    nan = np.float32(np.nan)
    pts = [corner(i, j, w, h, nan) for j in range(h) for i in range(w)]
    got = material8(img, 1, pts)
    assert np.array_equal(got[:, :3], (img.reshape(-1, 3) / np.float32(255)) * np.float32(2) - np.float32(1))
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    got = material8(rgba, 2, pts)
    assert ulp_close(got[:, :3], (rgba.reshape(-1, 4)[:, :3] / 255.0) ** np.float64(np.float32(2.2))).all()


def test_specular_map_gives_metallic_and_roughness_at_level_zero():
    img = np.zeros((2, 4, 4), np.uint8)
    img[0, 0] = (0, 0, 255, 255)    # roughness clamps up to 1e-3, metallic down to 0.99
    img[0, 1] = (0, 128, 64, 255)
    img[1, 2] = (9, 255, 0, 255)
    pts = [corner(0, 0, 4, 2, 5.0), corner(1, 0, 4, 2, 5.0), corner(2, 1, 4, 2, 5.0)]  # duv is ignored
    got = material8(img, 3, pts)
    assert np.array_equal(got[:, 0], np.float32([1e-3, 128 / np.float32(255), 1.0]))
    assert np.array_equal(got[:, 1], np.float32([0.99, 64 / np.float32(255), 0.0]))


# ------------------------------------------------------------ surface normal
def surface_normal(tri, uv, nm, shape_n, surf_n):
    out = np.empty(3, np.float32)
    args = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (tri, uv, nm, shape_n, surf_n)]
    L.oracle_surface_normal(*(fp(a) for a in args), fp(out))
    return out


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


@pytest.mark.parametrize("nm", [(0.3, -0.2, 0.9), (0.0, 0.0, 1.0), (-0.8, 0.5, 0.1), (0.6, 0.6, -1.0)])
def test_constant_normal_map_tilts_by_the_closed_form_angle(nm):
    """On a face in the xy plane with u along +x and v along +y the tangent
    frame is (x, y, z): the surface normal (0, 0, 1) becomes
    normalize(nx, ny, 1), tilted by atan(|(nx, ny)|); nz plays no part."""
    tri = [[0, 0, 0], [2, 0, 0], [0, 3, 0]]
    uv = [[0.1, 0.2], [0.6, 0.2], [0.1, 0.9]]
    n = surface_normal(tri, uv, nm, (0, 0, 1), (0, 0, 1))
    want = np.array([nm[0], nm[1], 1.0]) / np.linalg.norm([nm[0], nm[1], 1.0])
    np.testing.assert_allclose(n, want, atol=1e-6)
    np.testing.assert_allclose(np.arccos(np.clip(n[2], -1, 1)), np.arctan(np.hypot(nm[0], nm[1])), atol=2e-6)


def test_normal_map_frame_follows_the_uv_layout_and_rotations():
    nm = (0.4, -0.25, 0.7)
    # u along +y, v along -x: tangent (0, 1, 0), bitangent (-1, 0, 0)
    tri = [[0, 0, 0], [0, 1, 0], [-1, 0, 0]]
    uv = [[0, 0], [1, 0], [0, 1]]
    n = surface_normal(tri, uv, nm, (0, 0, 1), (0, 0, 1))
    want = np.array([-nm[1], nm[0], 1.0])
    np.testing.assert_allclose(n, want / np.linalg.norm(want), atol=1e-6)
    # the same face turned by a rotation R: the result turns with it
    R = rotation(11)
    tri_r = (R @ np.float64(tri).T).T
    z = R @ np.array([0, 0, 1.0])
    n_r = surface_normal(tri_r, uv, nm, z, z)
    np.testing.assert_allclose(n_r, R @ (want / np.linalg.norm(want)), atol=2e-6)
    # a shading normal off the face normal is tilted from where it is
    s = np.array([0.2, 0.1, 1.0])
    s /= np.linalg.norm(s)
    n = surface_normal(tri, uv, nm, (0, 0, 1), s)
    w = np.array([-nm[1], nm[0], 0.0]) + s
    np.testing.assert_allclose(n, w / np.linalg.norm(w), atol=1e-6)


def test_degenerate_uvs_keep_the_surface_normal():
    """Equal uvs make f = 1 / 0: the frame is not finite and the surface
    normal stays as it was."""
    tri = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    s = np.float32([0.1, 0.2, 0.97])
    n = surface_normal(tri, [[0.5, 0.5]] * 3, (0.3, 0.3, 1.0), (0, 0, 1), s)
    assert np.array_equal(n, s)


# ------------------------------------------------------------ emissive light objects
def test_emissive_texture_light_object_averages_over_the_face():
    """A quad whose uvs cover only texels 1..2 of an 8 x 1 emissive texture
    (texels 0-3 colour A, 4-7 colour B): each face's average over the 8 x 8
    barycentric grid is A ^ 2.2 alone, the light's colour is that
    normalised by its luminance and its power is area x luminance."""
    A, B = np.uint8([200, 100, 50, 255]), np.uint8([10, 240, 30, 255])
    tex = np.concatenate([np.tile(A, (1, 4, 1)), np.tile(B, (1, 4, 1))], axis=1)
    v, _ = quad(0.0, 1.0)
    uv = np.float32([[[1.0 / 8, 0.2], [2.5 / 8, 0.2], [2.5 / 8, 0.8]], [[1.0 / 8, 0.2], [2.5 / 8, 0.8], [1.0 / 8, 0.8]]])
    m = RmMaterial()
    m.base_color[:] = (0.5, 0.5, 0.5)
    m.roughness, m.opacity, m.ior, m.emissive_tex = 0.8, 1.0, 1.0, 0
    d, keep = make_desc(v, uv, [0, 0], [0, 0], [m], [tex])
    S = O.OracleScene(d)
    (light,) = S.lights()
    a = (A[:3] / 255.0) ** 2.2
    lum = a @ np.array([0.3, 0.6, 0.1])
    np.testing.assert_allclose(light[3:6], a / lum, rtol=1e-5)
    np.testing.assert_allclose(light[6], 4.0 * lum, rtol=1e-5)
    np.testing.assert_allclose(light[:3], 0.0, atol=1e-6)

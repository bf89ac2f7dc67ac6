"""Host image decoders behind rm_image_load (no device needed).

PNG: images are written here by an independent encoder. It uses zlib, the
forward row filters of the PNG specification, and every colour type and bit
depth. Each is decoded and checked against the conversions the reference's
loader asks libpng for (src/material.cpp:231-260): 16 bits to the high byte,
grey 1/2/4 bits scaled to 0..255, palette and grey to RGB, alpha dropped.
DDS: hand-built S3TC blocks whose palettes are worked out by hand (four-
and three-colour BC1, BC2 and BC3 colour halves, the DX10 header, partial
blocks) and a masked 32-bit BGRA image.
HDR: RGBE files written flat, with the per-component run-length scanlines,
and with the original repeat texels, against (r, g, b) * 2^(e - 136)."""
import struct
import zlib

import numpy as np
import pytest

import raym0nade_amd as rm
from raym0nade_amd._lib import RM_ERR_INVALID, RM_ERR_UNSUPPORTED, RmError


def png_chunk(t, d):
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def encode_png(rows, w, h, depth, ctype, plte=None, interlace=0, seed=0):
    """rows: list of h bytes objects (unfiltered scanlines)."""
    rng = np.random.default_rng(seed)
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bpp = max(1, ch * depth // 8)
    out = b""
    prev = bytes(len(rows[0]))
    for y, row in enumerate(rows):
        ft = int(rng.integers(0, 5))
        f = bytearray()
        for i, v in enumerate(row):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][ft]
            f.append((v - pred) & 0xFF)
        out += bytes([ft]) + bytes(f)
        prev = row
    ihdr = struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, interlace)
    data = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", ihdr)
    if plte is not None:
        data += png_chunk(b"PLTE", bytes(plte))
    data += png_chunk(b"IDAT", zlib.compress(out, 6)) + png_chunk(b"IEND", b"")
    return data


def pack_bits(vals, depth):
    out, acc, n = bytearray(), 0, 0
    for v in vals:
        acc = (acc << depth) | int(v)
        n += depth
        if n == 8:
            out.append(acc)
            acc, n = 0, 0
    if n:
        out.append(acc << (8 - n))
    return bytes(out)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8), (4, 8), (4, 16),
         (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", CASES)
def test_png_decodes_every_colour_type_and_depth(tmp_path, ctype, depth):
    rng = np.random.default_rng(ctype * 100 + depth)
    w, h = 13, 7
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    maxv = (1 << depth) - 1
    plte = None
    if ctype == 3:
        n_pal = 1 << depth
        plte = rng.integers(0, 256, n_pal * 3).astype(np.uint8)
        samples = rng.integers(0, n_pal, (h, w))
    else:
        samples = rng.integers(0, maxv + 1, (h, w, ch))
    rows = []
    for y in range(h):
        s = samples[y].reshape(-1)
        if depth == 16:
            rows.append(b"".join(struct.pack(">H", int(v)) for v in s))
        elif depth == 8:
            rows.append(bytes(int(v) for v in s))
        else:
            rows.append(pack_bits(s, depth))
    p = tmp_path / "t.png"
    p.write_bytes(encode_png(rows, w, h, depth, ctype, plte=plte, seed=depth))
    got = rm.load_image(str(p))
    if ctype == 3:
        want = plte.reshape(-1, 3)[samples]
    else:
        s8 = samples >> 8 if depth == 16 else (samples * 255 // maxv if depth < 8 else samples)
        if ctype in (0, 4):
            want = np.repeat(s8[..., :1], 3, axis=-1)
        else:
            want = s8[..., :3]
    assert got.shape == (h, w, 3)
    assert np.array_equal(got, want.astype(np.float32) / np.float32(255.0))


def test_png_errors(tmp_path):
    p = tmp_path / "bad.png"
    p.write_bytes(b"not a png")
    with pytest.raises(RmError) as e:
        rm.load_image(str(p))
    assert e.value.code == RM_ERR_INVALID
    rows = [bytes(6)] * 2
    p.write_bytes(encode_png(rows, 2, 2, 8, 2, interlace=1))
    with pytest.raises(RmError) as e:
        rm.load_image(str(p))
    assert e.value.code == RM_ERR_UNSUPPORTED
    q = tmp_path / "x.dds"
    q.write_bytes(b"DDS ")
    with pytest.raises(RmError) as e:
        rm.load_image(str(q))
    assert e.value.code == RM_ERR_INVALID
    q = tmp_path / "x.tga"
    q.write_bytes(b"\0" * 32)
    with pytest.raises(RmError) as e:
        rm.load_image(str(q))
    assert e.value.code == RM_ERR_UNSUPPORTED


def encode_hdr(rgbe, mode):
    h, w = rgbe.shape[:2]
    out = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        row = rgbe[y]
        if mode == "flat":
            out += row.tobytes()
        elif mode == "repeat":
            # the original encoding: a texel, then (1, 1, 1, n) repeats of it
            x = 0
            while x < w:
                out += row[x].tobytes()
                n = 1
                while x + n < w and (row[x + n] == row[x]).all() and n < 255:
                    n += 1
                if n > 1:
                    out += bytes([1, 1, 1, n - 1])
                x += n
        else:
            out += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                comp = row[:, c]
                x = 0
                while x < w:
                    n = 1
                    while x + n < w and comp[x + n] == comp[x] and n < 127:
                        n += 1
                    if n >= 3:
                        out += bytes([128 + n, comp[x]])
                        x += n
                    else:
                        k = 1
                        while x + k < w and k < 128 and not (x + k + 2 < w and comp[x + k] == comp[x + k + 1] == comp[x + k + 2]):
                            k += 1
                        out += bytes([k]) + comp[x:x + k].tobytes()
                        x += k
    return out


@pytest.mark.parametrize("mode", ["flat", "rle", "repeat"])
def test_hdr_decodes_rgbe(tmp_path, mode):
    rng = np.random.default_rng(9)
    h, w = 5, 40
    rgbe = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    rgbe[:, 10:30] = rgbe[:, 10:11]  # runs
    rgbe[0, 3, 3] = 0                # a zero texel
    p = tmp_path / "sky.hdr"
    p.write_bytes(encode_hdr(rgbe, mode))
    got = rm.load_image(str(p))
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136)).astype(np.float32)
    want = rgbe[..., :3].astype(np.float32) * scale[..., None]
    assert np.array_equal(got, want)


def dds_file(w, h, payload, fourcc=None, dxgi=None, rgb_masks=None, bits=32):
    hdr = bytearray(128)
    hdr[0:4] = b"DDS "
    struct.pack_into("<IIII", hdr, 4, 124, 0x1007, h, w)
    if fourcc is not None:
        struct.pack_into("<II4s", hdr, 76, 32, 0x4, fourcc)
    else:
        struct.pack_into("<IIIIIIII", hdr, 76, 32, 0x40, 0, bits, *rgb_masks, 0)
    ext = struct.pack("<IIIII", dxgi, 3, 0, 1, 0) if dxgi is not None else b""
    return bytes(hdr) + ext + payload


RED, BLUE = 0xF800, 0x001F
# c0 = red > c1 = blue: red, blue, (2 red + blue) / 3, (red + 2 blue) / 3
PAL4 = np.uint8([[255, 0, 0], [0, 0, 255], [170, 0, 85], [85, 0, 170]])
# c0 = blue < c1 = red: blue, red, (blue + red) / 2, black
PAL3 = np.uint8([[0, 0, 255], [255, 0, 0], [127, 0, 127], [0, 0, 0]])
IDX = sum(((t & 3) << (2 * t)) for t in range(16))  # texel (x, y) uses palette entry x


def bc1(c0, c1):
    return struct.pack("<HHI", c0, c1, IDX)


@pytest.mark.parametrize("kind", ["DXT1", "DXT1-3", "DXT3", "DXT5", "DX10-BC1", "partial"])
def test_dds_s3tc_blocks(tmp_path, kind):
    w = h = 4
    if kind == "DXT1":
        data, pal = dds_file(4, 4, bc1(RED, BLUE), fourcc=b"DXT1"), PAL4
    elif kind == "DXT1-3":
        data, pal = dds_file(4, 4, bc1(BLUE, RED), fourcc=b"DXT1"), PAL3
    elif kind == "DXT3":  # explicit alpha half first; colours always four-entry
        data, pal = dds_file(4, 4, b"\xff" * 8 + bc1(BLUE, RED), fourcc=b"DXT3"), \
            np.uint8([[0, 0, 255], [255, 0, 0], [85, 0, 170], [170, 0, 85]])
    elif kind == "DXT5":
        data, pal = dds_file(4, 4, bytes([255, 0]) + b"\0" * 6 + bc1(RED, BLUE), fourcc=b"DXT5"), PAL4
    elif kind == "DX10-BC1":
        data, pal = dds_file(4, 4, bc1(RED, BLUE), fourcc=b"DX10", dxgi=71), PAL4
    else:  # 5 x 3: two blocks across, one down; the texels past the edge are dropped
        w, h = 5, 3
        data, pal = dds_file(5, 3, bc1(RED, BLUE) * 2, fourcc=b"DXT1"), PAL4
    p = tmp_path / "t.dds"
    p.write_bytes(data)
    got = rm.load_image(str(p))
    want = np.stack([np.stack([pal[x & 3] for x in range(w)]) for _ in range(h)])
    assert np.array_equal(got, want.astype(np.float32) / np.float32(255.0))


def test_dds_uncompressed_bgra(tmp_path):
    rng = np.random.default_rng(4)
    px = rng.integers(0, 256, (3, 5, 4)).astype(np.uint8)  # stored B, G, R, A
    p = tmp_path / "u.dds"
    p.write_bytes(dds_file(5, 3, px.tobytes(), rgb_masks=(0x00FF0000, 0x0000FF00, 0x000000FF)))
    got = rm.load_image(str(p))
    assert np.array_equal(got, px[..., [2, 1, 0]].astype(np.float32) / np.float32(255.0))


def test_scene_load_rejects_bad_models_before_touching_a_device(tmp_path):
    """rm_scene_load's parse errors (no device needed: they come before
    rm_scene_create)."""
    def load(name, sky="null"):
        with pytest.raises(RmError) as e:
            rm.Scene.load(str(tmp_path) + "/", name, sky)
        return e.value.code
    assert load("missing.obj") == RM_ERR_INVALID
    (tmp_path / "m.fbx").write_bytes(b"")
    assert load("m.fbx") == RM_ERR_INVALID             # not an FBX file (binary FBX is read)
    (tmp_path / "m.3ds").write_bytes(b"")
    assert load("m.3ds") == RM_ERR_UNSUPPORTED
    (tmp_path / "idx.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    assert load("idx.obj") == RM_ERR_INVALID          # vertex 4 does not exist
    (tmp_path / "zero.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")
    assert load("zero.obj") == RM_ERR_INVALID         # OBJ indices start at 1
    (tmp_path / "mat.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nothing\nf 1 2 3\n")
    assert load("mat.obj") == RM_ERR_INVALID          # unknown material
    (tmp_path / "lib.obj").write_text("mtllib nofile.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert load("lib.obj") == RM_ERR_INVALID          # missing material library
    (tmp_path / "empty.obj").write_text("v 0 0 0\n")
    from raym0nade_amd._lib import RM_ERR_EMPTY
    assert load("empty.obj") == RM_ERR_EMPTY
    (tmp_path / "ok.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "sky.hdr").write_bytes(b"not an hdr")
    assert load("ok.obj", "sky.hdr") == RM_ERR_INVALID

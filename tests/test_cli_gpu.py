"""The console host (reference grammar, src/myconsole.cpp:128-160) driving the
GPU path: a scripted session renders the Cornell workload and writes the
export list as PNGs (src/render.cpp:641-663; the workload has CoC 0, so no
depth-of-field exports); each decodes, and
the Raw and Filter exports equal the library's post-processed pixels after
spatialClamp, and after spatialClamp + filter, truncated like Photo::save
(src/image.cpp:518-524)."""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest
import torch

import raym0nade_amd as rm

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(rm.__file__), "raym0nade_cli")


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        t = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if t == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif t == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * 3)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_console_session_writes_exports(tmp_path, dev):
    os.makedirs(tmp_path / "output", exist_ok=True)
    session = "create scene box 0 1\nview model box\nrender box box_camera\nexit\n"
    r = subprocess.run([CLI], input=session, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Rendering completed" in r.stdout, r.stdout
    tags = ["DiffuseColor", "DiffuseColor_FXAA", "shapeNormal", "surfaceNormal", "Direct_Diffuse", "Direct_Specular",
            "Indirect_Diffuse", "Indirect_Specular", "Raw", "Raw_Bloom", "Raw_FXAA", "Raw_Bloom_FXAA",
            "Direct_Diffuse_Filter", "Direct_Specular_Filter", "Indirect_Diffuse_Filter", "Indirect_Specular_Filter",
            "Filter", "Filter_Bloom", "Filter_FXAA", "Filter_Bloom_FXAA"]
    imgs = {t: read_png(tmp_path / "output" / f"box({t}).png") for t in tags}
    assert all(im.shape == (512, 512, 3) for im in imgs.values())
    # same render through the library: Raw = spatialClamp -> shade(Full) -> gamma,
    # Filter = spatialClamp -> filter -> shade(Full) -> gamma, truncated to 8 bits
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    fr = rm.Frame(512, 512)
    fr.render(S, g.args)
    fr.spatial_clamp()
    for tag in ("Raw", "Filter"):
        if tag == "Filter":
            fr.filter()
        fr.postprocess(63, g.args.exposure)
        torch.cuda.synchronize()
        px = fr.download(rm.RM_BUF_PIXELS)
        expect = (px * np.float32(255)).astype(np.uint8)
        assert np.array_equal(imgs[tag], expect), tag
    # the filter removes noise: the filtered image is smoother than the raw one
    def roughness(im):
        f = im.astype(np.float64)
        return np.abs(np.diff(f, axis=0)).mean() + np.abs(np.diff(f, axis=1)).mean()
    assert roughness(imgs["Filter"]) < roughness(imgs["Raw"])


def test_console_depth_of_field_exports(tmp_path, dev):
    """`create args` with CoC > 0 (src/myconsole.cpp:27-67): the console also
    writes the five depth-of-field exports (src/render.cpp:665-673), and
    Filter_DepthFieldBlur equals the library's spatialClamp -> filter ->
    postProcessing(Full | DoDepthFieldBlur) with the render's focus, CoC and
    camera."""
    os.makedirs(tmp_path / "output", exist_ok=True)
    g = rm.generate(0, 1)
    c = g.args
    d, r, u = (np.float32(list(v)) for v in (c.direction, c.right, c.up))
    pos = np.float32(list(c.position))
    D, R, U = (float(np.dot(pos.astype(np.float64), v.astype(np.float64))) for v in (d, r, u))
    focus, coc, w = 3.0, 3.0, 96
    acc = float(np.float32(c.accuracy * c.width / w))
    fmt = lambda v: " ".join(repr(float(x)) for x in v)
    session = (f"create scene box 0 1\ncreate args cam\n{fmt(d)}\n{fmt(r)}\n{fmt(u)}\n{D!r} {R!r} {U!r}\n"
               f"{acc!r} {focus} {coc} {float(c.exposure)!r}\n{w} {w}\n8 1 {float(c.p_direct)!r}\noutput/dof\n"
               "render box cam\nexit\n")
    res = subprocess.run([CLI], input=session, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert res.returncode == 0 and "Rendering completed" in res.stdout, res.stdout + res.stderr
    tags = ["BaseColor_DepthFieldBlur", "Filter_DepthFieldBlur", "Filter_DepthFieldBlur_Bloom",
            "Filter_DepthFieldBlur_FXAA", "Filter_DepthFieldBlur_Bloom_FXAA"]
    imgs = {t: read_png(tmp_path / "output" / f"dof({t}).png") for t in tags}
    a = rm.RmRenderArgs.from_buffer_copy(c)
    for k in range(3):
        a.direction[k], a.right[k], a.up[k] = d[k], r[k], u[k]
        a.position[k] = np.float32(np.float32(np.float32(D) * d[k]) + np.float32(np.float32(R) * r[k])) + \
            np.float32(np.float32(U) * u[k])
    a.accuracy, a.focus, a.coc = acc, focus, coc
    a.width = a.height = w
    a.spp, a.seed, a.max_depth, a.rr_min_depth = 8, 1, 16, 1
    S = rm.Scene(g)
    fr = rm.Frame(w, w)
    fr.render(S, a)
    fr.spatial_clamp()
    fr.filter()
    fr.postprocess(63 | 1024, a.exposure)
    torch.cuda.synchronize()
    expect = (fr.download(rm.RM_BUF_PIXELS) * np.float32(255)).astype(np.uint8)
    assert np.array_equal(imgs["Filter_DepthFieldBlur"], expect)

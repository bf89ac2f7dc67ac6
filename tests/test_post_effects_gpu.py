"""GPU bloom and depth of field (rm_postprocess options 256 and 1024)
against the oracle's Photo::bloom / Photo::depthFeildBlur
(src/image.cpp:248-356).

Depth of field is a gather over each pixel's sources in the reference's visit
order. It is exact, so the linear image is compared bit for bit. Bloom calls
powf, and the tone curve calls tanh and pow. The device evaluates these in
fp64 and rounds once, which gives glibc's fp32 results nearly always. Tone-
mapped images are therefore compared to 1e-6 relative.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_py as O
import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from raym0nade_amd._lib import RM_ERR_INVALID, RM_ERR_UNSUPPORTED, RmError

pytestmark = pytest.mark.gpu

LINEAR = 65536
BLOOM, FXAA, DOF = 256, 512, 1024


@pytest.fixture(scope="module")
def cornell_frame(dev):
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / 128)
    a.width = a.height = 128
    a.spp = 16
    fr = rm.Frame(128, 128)
    fr.render(S, a)
    torch.cuda.synchronize()
    gb, rad = RC.download_all(fr)
    return fr, a, gb, rad


def gpu_post(fr, options, exposure):
    fr.postprocess(options, exposure)
    torch.cuda.synchronize()
    return fr.download(rm.RM_BUF_PIXELS)


@pytest.mark.parametrize("options", [63 | BLOOM | LINEAR, 63 | BLOOM, 63 | BLOOM | FXAA, 8 | 32 | BLOOM])
def test_bloom_matches_oracle(cornell_frame, options):
    fr, a, gb, rad = cornell_frame
    for exposure in (1.0, 4.0):  # 4: more of the image above luminance 1
        px = gpu_post(fr, options, exposure)
        ref = O.postprocess(gb, rad, options, exposure)
        if options & LINEAR and exposure == 1.0:
            lum = (ref * np.float32([0.3, 0.6, 0.1])).sum(-1)
            assert (lum >= 1).any()
        close = np.isclose(px, ref, rtol=1e-6, atol=1e-7).all(axis=-1)
        print("bloom", options, exposure, close.mean(), float(np.abs(px - ref).max()))
        assert close.mean() >= (0.999 if options & FXAA else 1.0)


@pytest.mark.parametrize("focus,coc", [(3.0, 4.0), (1.5, 12.0), (8.0, 0.0)])
def test_dof_matches_oracle_on_a_render(cornell_frame, focus, coc):
    fr, a, gb, rad = cornell_frame
    cam = [a.position[k] for k in range(3)]
    fr.set_dof(focus, coc, cam)
    px = gpu_post(fr, 63 | DOF | LINEAR, a.exposure)
    ref = O.postprocess(gb, rad, 63 | DOF | LINEAR, a.exposure, focus=focus, coc=coc, cam=cam)
    assert np.array_equal(px.view(np.uint32), ref.view(np.uint32)), float(np.abs(px - ref).max())
    full = gpu_post(fr, 63 | DOF | BLOOM | FXAA, a.exposure)
    ref_full = O.postprocess(gb, rad, 63 | DOF | BLOOM | FXAA, a.exposure, focus=focus, coc=coc, cam=cam)
    close = np.isclose(full, ref_full, rtol=1e-6, atol=1e-7).all(axis=-1)
    assert close.mean() >= 0.999


def synthetic(h, w, seed, near_frac=0.02, miss_frac=0.0):
    rng = np.random.default_rng(seed)
    g = np.zeros((h, w, O.GB), np.float32)
    depth = rng.choice(np.float32([2.0, 3.0, 5.0, 9.0]), (h, w))
    depth[rng.uniform(size=(h, w)) < near_frac] = 0.5  # near pixels: large discs
    g[..., 2] = depth
    g[rng.uniform(size=(h, w)) < miss_frac, 0:3] = np.nan
    g[..., 12:15] = 1.0
    rad = np.zeros((4, h, w, 4), np.float32)
    rad[0, ..., :3] = rng.uniform(0, 3, (h, w, 3))
    return g, rad


def upload(g, rad):
    h, w = g.shape[:2]
    fr = rm.Frame(w, h)
    for k, b in enumerate(RC.GB_ORDER):
        fr.upload(b, g[..., 3 * k:3 * k + 3])
    for k, b in enumerate(RC.RAD_ORDER):
        fr.upload(b, rad[k])
    return fr


@pytest.mark.parametrize("h,w,focus,coc,miss", [(37, 23, 3.0, 2.5, 0.0), (64, 80, 5.0, 6.0, 0.05),
                                                (1, 50, 2.0, 3.0, 0.0), (90, 1, 2.0, 3.0, 0.0),
                                                (150, 200, 9.0, 6.0, 0.0)])
def test_dof_bit_exact_on_synthetic_frames(dev, h, w, focus, coc, miss):
    """Random depths, a few near pixels whose discs reach the cap of radius
    96 (focus 9, depth 0.5: CoC0 = 6 x 17 = 102 -> 96), missed pixels."""
    g, rad = synthetic(h, w, seed=h + w, miss_frac=miss)
    fr = upload(g, rad)
    fr.set_dof(focus, coc, (0.0, 0.0, 0.0))
    px = gpu_post(fr, 63 | DOF | LINEAR, 1.0)
    ref = O.postprocess(g, rad, 63 | DOF | LINEAR, 1.0, focus=focus, coc=coc)
    assert np.array_equal(px.view(np.uint32), ref.view(np.uint32)), float(np.nanmax(np.abs(px - ref)))


def test_dof_rejects_bad_parameters_and_batches(dev, cornell_frame):
    fr, a, gb, rad = cornell_frame
    fr.set_dof(3.0, -1.0, (0, 0, 0))
    with pytest.raises(RmError) as e:
        fr.postprocess(63 | DOF, 1.0)
    assert e.value.code == RM_ERR_INVALID
    fr.set_dof(3.0, float("nan"), (0, 0, 0))
    with pytest.raises(RmError):
        fr.postprocess(63 | DOF, 1.0)
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    b = rm.RmRenderArgs.from_buffer_copy(g.args)
    b.width = b.height = 32
    b.spp = 4
    f2 = rm.Frame(32, 32)
    with pytest.raises(RmError) as e:
        rm.render_batch([(S, b, f2, 63 | DOF)])
    assert e.value.code == RM_ERR_UNSUPPORTED

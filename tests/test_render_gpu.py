"""GPU render parity against the oracle (renderPixel and its callees,
src/render.cpp:121-551; post, src/image.cpp:215-479).

Both sides draw every random number from the same Philox streams in the
reference's order and follow the same fp32 operation order, so:
  * the G-buffer (position, normals, emission, base colour) is bit-exact;
  * the radiance buffers agree pixel by pixel to float rounding, except where
    a 1-ulp difference of a device transcendental (sin, cos, log, pow: ROCm's
    device library vs glibc) flips a discrete decision of a path; those pixels
    are counted and bounded;
  * independent seeds agree statistically (z-scores of the per-pixel means).
"""
import numpy as np
import pytest
import torch

import oracle_py as O
import raym0nade_amd as rm
from raym0nade_amd import render_check as RC

pytestmark = pytest.mark.gpu


def small_args(g, w, h, spp, **kw):
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    scale = a.width / w
    a.width, a.height = w, h
    a.accuracy = np.float32(a.accuracy * scale)
    a.spp = spp
    for k, v in kw.items():
        setattr(a, k, v)
    return a


@pytest.fixture(scope="module")
def cornell(dev):
    g = rm.generate(0, 1)
    return g, rm.Scene(g), O.OracleScene(g.desc)


def test_cornell_gbuffer_bitexact_and_radiance_near_exact(cornell):
    g, S, OS = cornell
    a = small_args(g, 128, 128, 32)
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    RC.assert_gbuffer_equal(gb, ogb)
    st = RC.compare_radiance(rad, orad)
    print("cornell 128^2 32spp:", st)
    assert st["direct_close"] >= 0.995
    assert st["indirect_close"] >= 0.97


def test_cornell_reference_depth_and_rr(cornell):
    # the reference's own constants: maxRayDepth 16, RR from depth 1
    g, S, OS = cornell
    a = small_args(g, 96, 96, 24, max_depth=16, rr_min_depth=1)
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    RC.assert_gbuffer_equal(gb, ogb)
    st = RC.compare_radiance(rad, orad)
    print("cornell depth16 rr1:", st)
    assert st["direct_close"] >= 0.995 and st["indirect_close"] >= 0.97


def test_paths_per_round_only_changes_summation_order(cornell):
    g, S, OS = cornell
    a1 = small_args(g, 64, 64, 40, paths_per_round=1)
    a4 = small_args(g, 64, 64, 40, paths_per_round=4)
    _, r1 = RC.gpu_render(S, a1)
    _, r4 = RC.gpu_render(S, a4)
    close = np.isclose(r1, r4, rtol=1e-4, atol=1e-6).all(axis=-1).mean()
    assert close >= 0.999


def test_deterministic(cornell):
    g, S, OS = cornell
    a = small_args(g, 64, 64, 16)
    _, r1 = RC.gpu_render(S, a)
    _, r2 = RC.gpu_render(S, a)
    assert np.array_equal(r1.view(np.uint32), r2.view(np.uint32))


def test_independent_seeds_agree_statistically(cornell):
    g, S, OS = cornell
    a = small_args(g, 32, 32, 512, seed=7)
    b = small_args(g, 32, 32, 512, seed=8)
    _, rg = RC.gpu_render(S, a)
    _, ro = OS.render(b)
    z = RC.zscores(rg, ro, a.spp, a.p_direct)
    frac = (np.abs(z) < 3).mean()
    print("seed-independent |z|<3 fraction:", frac)
    assert frac >= 0.97


def test_spheres_sky_glass_near_exact(dev):
    # small form of the 1080p workload: sky NEE, GGX metal, dielectrics with the
    # medium stack, smooth normals on displaced geodesic spheres
    g = rm.generate(2, 3, grid=2, freq=12, sky_width=256, sky_height=128)
    S, OS = rm.Scene(g), O.OracleScene(g.desc)
    a = small_args(g, 96, 54, 16, max_depth=16)
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    RC.assert_gbuffer_equal(gb, ogb)
    st = RC.compare_radiance(rad, orad)
    print("spheres 96x54 16spp:", st)
    assert st["direct_close"] >= 0.98 and st["indirect_close"] >= 0.95


def test_textured_cornell_near_exact(dev):
    g = rm.generate(4, 1, tex_size=64)
    S, OS = rm.Scene(g), O.OracleScene(g.desc)
    a = small_args(g, 96, 96, 16)
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    st = RC.compare_radiance(rad, orad)
    # texture fetch uses log2 for the mip level: base colours agree to rounding
    assert np.allclose(np.nan_to_num(gb[..., 12:15]), np.nan_to_num(ogb[..., 12:15]), rtol=1e-5, atol=1e-6)
    print("textured:", st)
    assert st["direct_close"] >= 0.98 and st["indirect_close"] >= 0.95


@pytest.mark.parametrize("options", [63, 63 | 512, 1, 1 | 512, 64, 128, 4 | 16, 8 | 32])
def test_postprocess_matches_oracle(cornell, options):
    g, S, OS = cornell
    a = small_args(g, 128, 128, 16)
    fr = rm.Frame(a.width, a.height)
    fr.render(S, a)
    fr.postprocess(options, a.exposure)
    torch.cuda.synchronize()
    px = fr.download(rm.RM_BUF_PIXELS)
    gb, rad = RC.download_all(fr)
    ref = O.postprocess(gb, rad, options, a.exposure)
    # tanh / pow from the device library differ from glibc by a few ulp
    close = np.isclose(np.nan_to_num(px), np.nan_to_num(ref), rtol=1e-5, atol=2e-6).all(axis=-1)
    print("post", options, close.mean())
    # FXAA picks the neighbour with the largest luma step: a 1-ulp luma change
    # can only flip exact ties
    assert close.mean() >= (0.999 if options & 512 else 1.0)


def test_full_size_cornell_properties(cornell):
    """The 512x512 64-spp depth-8 workload: finite, non-negative means,
    deterministic, and the oracle agrees on a sample of rows."""
    g, S, OS = cornell
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    gb, rad = RC.gpu_render(S, a)
    assert np.isfinite(rad).all()
    assert (rad[..., :3].reshape(-1, 3).mean(0) >= 0).all()
    rows = list(range(0, 512, 64))
    ogb, orad = OS.render(a, rows=rows)
    st = RC.compare_radiance(rad[:, rows], orad[:, rows])
    print("full C0 sampled rows:", st)
    assert st["direct_close"] >= 0.995 and st["indirect_close"] >= 0.97


@pytest.mark.parametrize("scene", ["cornell", "spheres"])
@pytest.mark.parametrize("entries,tps,split", [(32, 0, 1), (2, 0, 1), (8, 1, 1), (2, 2, 2), (8, 4, 2), (8, 2, 0)])
def test_lds_stack_walks_are_bit_identical(dev, scene, entries, tps, split):
    # every walk of a render launch (primary, direct, extension, NEE) with its
    # stack in LDS (trace variant 3), wholly or split with scratch, and with a
    # leaf's triangles tested over several steps, visits the same nodes and
    # triangles in the same order as the scratch-stack, whole-leaf walker; the
    # light-sample passes split into generate / shadow / add (every shadow ray
    # queued, or only those of deep trees) give the same sums as the single
    # kernels
    g = rm.generate(0, 1) if scene == "cornell" else rm.generate(2, 3, grid=2, freq=12, sky_width=256, sky_height=128)
    S = rm.Scene(g)
    a = small_args(g, 80, 48, 12, max_depth=16)
    keys = (rm.TUNE_TRACE_VARIANT, rm.TUNE_LDS_STACK_ENTRIES, rm.TUNE_TRIS_PER_STEP, rm.TUNE_NEE_SPLIT)
    old = tuple(rm.get_tuning(k) for k in keys)
    try:
        # reference: scratch stacks, whole leaves, NEE shadow rays traced inside the NEE kernel
        for k, v in zip(keys, (2, 32, 0, 0)):
            rm.set_tuning(k, v)
        gb2, r2 = RC.gpu_render(S, a)
        for k, v in zip(keys, (3, entries, tps, split)):
            rm.set_tuning(k, v)
        gb3, r3 = RC.gpu_render(S, a)
    finally:
        for k, v in zip(keys, old):
            rm.set_tuning(k, v)
    assert np.array_equal(gb2.view(np.uint32), gb3.view(np.uint32))
    assert np.array_equal(r2.view(np.uint32), r3.view(np.uint32))


@pytest.mark.parametrize("scene", ["cornell", "spheres"])
@pytest.mark.parametrize("waves", [3, 4])
def test_register_capped_kernels_are_bit_identical(dev, scene, waves):
    # the shading and light-sample kernels compiled for 3 or 4 resident waves
    # per SIMD (168 or 128 VGPRs, the rest spilled) compute what the uncapped
    # ones do; Cornell takes the single-kernel NEE passes, the spheres the
    # split generate / shadow / add passes
    g = rm.generate(0, 1) if scene == "cornell" else rm.generate(2, 3, grid=2, freq=12, sky_width=256, sky_height=128)
    S = rm.Scene(g)
    a = small_args(g, 80, 48, 12, max_depth=16)
    old = rm.get_tuning(rm.TUNE_RENDER_MIN_WAVES)
    try:
        rm.set_tuning(rm.TUNE_RENDER_MIN_WAVES, 1)
        gb1, r1 = RC.gpu_render(S, a)
        rm.set_tuning(rm.TUNE_RENDER_MIN_WAVES, waves)
        gbw, rw = RC.gpu_render(S, a)
    finally:
        rm.set_tuning(rm.TUNE_RENDER_MIN_WAVES, old)
    assert np.array_equal(gb1.view(np.uint32), gbw.view(np.uint32))
    assert np.array_equal(r1.view(np.uint32), rw.view(np.uint32))

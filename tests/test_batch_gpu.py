"""Batched rendering (rm_render_batch): jobs of different scenes, sizes, spp
and depths rendered through one set of launches (and one hipGraph) give the
same buffers as rendering each job alone with rm_render: every path uses the
same counter-RNG stream either way, and each accumulator set still receives
one path per round."""
import numpy as np
import pytest
import torch

import raym0nade_amd as rm
from raym0nade_amd import render_check as RC

pytestmark = pytest.mark.gpu


def job_args(g, w, h, spp, depth, seed):
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / w)
    a.width, a.height, a.spp, a.max_depth, a.seed = w, h, spp, depth, seed
    return a


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_single_renders(dev, use_graph):
    gens = [rm.generate(0, 1), rm.generate(2, 5, grid=2, freq=10, sky_width=128, sky_height=64), rm.generate(4, 2, tex_size=32)]
    scenes = [rm.Scene(g) for g in gens]
    specs = [(0, 96, 64, 16, 8, 3), (1, 80, 80, 24, 16, 4), (2, 64, 48, 8, 12, 5), (0, 40, 56, 32, 16, 6)]
    jobs, singles = [], []
    for si, w, h, spp, depth, seed in specs:
        a = job_args(gens[si], w, h, spp, depth, seed)
        fr = rm.Frame(w, h)
        jobs.append((scenes[si], a, fr, 63 | 512))
        singles.append((scenes[si], a))
    rm.render_batch(jobs, use_graph=use_graph)
    torch.cuda.synchronize()
    for (scene, a, fr, post), (s2, a2) in zip(jobs, singles):
        gb, rad = RC.download_all(fr)
        gb1, rad1 = RC.gpu_render(s2, a2)
        RC.assert_gbuffer_equal(gb, gb1)
        # same samples; only the per-pixel merge order of sets can differ when
        # the set count per pixel (paths_per_round) differs between the runs
        close = np.isclose(rad, rad1, rtol=1e-4, atol=1e-6).all(axis=-1).mean()
        assert close >= 0.999, close
        fr1 = rm.Frame(a.width, a.height)
        fr1.render(s2, a2)
        fr1.postprocess(63 | 512, a2.exposure)
        torch.cuda.synchronize()
        px = fr.download(rm.RM_BUF_PIXELS)
        px1 = fr1.download(rm.RM_BUF_PIXELS)
        assert np.isclose(px, px1, rtol=1e-4, atol=1e-5).all(axis=-1).mean() >= 0.999


def test_serving_batch_with_autotune_equals_single_renders(dev):
    from raym0nade_amd import autotune, workloads
    specs = workloads.batch_specs(6, seed=3, min_log2=5.0, max_log2=6.5)
    built = [workloads.build_job(s, sky=(64, 32)) for s in specs]
    jobs = [(S, a, rm.Frame(a.width, a.height), 63 | 512) for _, S, a in built]
    saved = autotune.current()
    try:
        choice = autotune.tune_render(jobs, spp_scale=0.25)
        assert choice["blocks_per_cu"] in (4, 8, 16) and choice["min_waves"] in (1, 3, 4)
        assert choice["walk_variant"] in (2, 3)
        rm.render_batch(jobs, use_graph=True)
        torch.cuda.synchronize()
        for S, a, fr, post in jobs:
            fr1 = rm.Frame(a.width, a.height)
            fr1.render(S, a)
            fr1.postprocess(post, a.exposure)
            torch.cuda.synchronize()
            px, px1 = fr.download(rm.RM_BUF_PIXELS), fr1.download(rm.RM_BUF_PIXELS)
            assert np.isclose(px, px1, rtol=1e-4, atol=1e-5).all(axis=-1).mean() >= 0.999
    finally:
        autotune.restore(saved)


def test_trace_autotune_keeps_results(dev):
    from raym0nade_amd import autotune
    g = rm.generate(1, 2, n_tris=20000)
    S = rm.Scene(g)
    r = torch.from_numpy(rm.rays(1, 200000, seed=4)).to(dev)
    t0, i0 = S.trace_closest(r)
    saved = autotune.current()
    try:
        choice = autotune.tune_trace(S, [r], reps=1)
        assert choice["tried"] >= 3
        t1, i1 = S.trace_closest(r)
        assert torch.equal(i0, i1) and torch.equal(t0.view(torch.int32), t1.view(torch.int32))
    finally:
        autotune.restore(saved)


def test_batch_argument_errors(dev):
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    a = job_args(g, 32, 32, 4, 8, 1)
    fr = rm.Frame(32, 32)
    with pytest.raises(rm.RmError):
        rm.render_batch([(S, a, fr, 0), (S, a, fr, 0)])       # one frame used twice
    bad = job_args(g, 32, 32, 4, 8, 1)
    bad.width = 33
    with pytest.raises(rm.RmError):
        rm.render_batch([(S, bad, fr, 0)])                     # frame size mismatch


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_shadow_split_modes_are_bit_identical(dev, use_graph):
    """A batch of shallow (Cornell, 36 triangles) and deep (spheres, about
    130k triangles) scenes with unequal direct-sample counts: the shadow rays
    traced inside the sampling kernels (mode 0), all queued for the shadow
    pass (1), or queued only for the deep scenes (2) give the same bits.
    Covers the job ordering by sample count and the pixel-prefix chunks."""
    gens = [rm.generate(0, 1), rm.generate(2, 7, grid=2, freq=40, sky_width=128, sky_height=64)]
    scenes = [rm.Scene(g) for g in gens]
    assert scenes[0].info.depth < 12 <= scenes[1].info.depth
    specs = [(0, 64, 48, 8, 8, 3), (1, 72, 40, 32, 16, 4), (0, 48, 48, 64, 12, 5), (1, 40, 56, 16, 16, 6)]
    old = rm.get_tuning(rm.TUNE_NEE_SPLIT)
    out = []
    try:
        for mode in (0, 1, 2):
            rm.set_tuning(rm.TUNE_NEE_SPLIT, mode)
            jobs = []
            for si, w, h, spp, depth, seed in specs:
                jobs.append((scenes[si], job_args(gens[si], w, h, spp, depth, seed), rm.Frame(w, h), 63 | 512))
            rm.render_batch(jobs, use_graph=use_graph)
            torch.cuda.synchronize()
            out.append([RC.download_all(fr) for _, _, fr, _ in jobs])
    finally:
        rm.set_tuning(rm.TUNE_NEE_SPLIT, old)
    for other in out[1:]:
        for (gb0, rad0), (gb, rad) in zip(out[0], other):
            assert np.array_equal(gb0.view(np.uint32), gb.view(np.uint32))
            assert np.array_equal(rad0.view(np.uint32), rad.view(np.uint32))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_post_with_clamp_filter_and_bloom(dev, use_graph):
    """post_options with RM_POST_SPATIAL_CLAMP | RM_POST_FILTER and bloom run
    inside the batch's launch sequence (and its graph); each job's image and
    radiance buffers equal the same steps run one call at a time on a single
    render."""
    from raym0nade_amd._lib import RM_POST_FILTER, RM_POST_SPATIAL_CLAMP
    gens = [rm.generate(0, 1), rm.generate(2, 5, grid=2, freq=10, sky_width=128, sky_height=64)]
    scenes = [rm.Scene(g) for g in gens]
    post = 63 | 256 | 512 | RM_POST_SPATIAL_CLAMP | RM_POST_FILTER
    jobs = []
    for si, w, h, spp, seed in [(0, 64, 48, 8, 3), (1, 72, 72, 8, 4)]:
        a = job_args(gens[si], w, h, spp, 12, seed)
        jobs.append((scenes[si], a, rm.Frame(w, h), post))
    rm.render_batch(jobs, use_graph=use_graph)
    torch.cuda.synchronize()
    for scene, a, fr, _ in jobs:
        ref = rm.Frame(a.width, a.height)
        ref.render(scene, a)
        ref.spatial_clamp()
        ref.filter()
        ref.postprocess(63 | 256 | 512, a.exposure)
        torch.cuda.synchronize()
        got_rad = np.stack([fr.download(b) for b in RC.RAD_ORDER])
        want_rad = np.stack([ref.download(b) for b in RC.RAD_ORDER])
        # the batch merges accumulator sets in its own order: compare to float rounding
        assert np.allclose(got_rad, want_rad, rtol=1e-4, atol=1e-6)
        px, want = fr.download(rm.RM_BUF_PIXELS), ref.download(rm.RM_BUF_PIXELS)
        assert np.isclose(px, want, rtol=1e-4, atol=1e-5).all(axis=-1).mean() >= 0.999


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_counters_equal_single_renders(dev, use_graph):
    """Each job's render counters from a batched launch (kept per job in
    shards picked by wave and added into the frame when the launch ends)
    equal those of its single render: the same primary, extension and
    shadow rays, hits, paths and medium-stack figures."""
    gens = [rm.generate(0, 1), rm.generate(2, 7, grid=2, freq=40, sky_width=128, sky_height=64)]
    scenes = [rm.Scene(g) for g in gens]
    specs = [(0, 64, 48, 8, 8, 3), (1, 72, 40, 32, 16, 4), (0, 48, 48, 64, 12, 5), (1, 40, 56, 16, 16, 6)]
    jobs = [(scenes[si], job_args(gens[si], w, h, spp, depth, seed), rm.Frame(w, h), 63 | 512)
            for si, w, h, spp, depth, seed in specs]
    rm.render_batch(jobs, use_graph=use_graph)
    torch.cuda.synchronize()
    for S, a, fr, _ in jobs:
        one = rm.Frame(a.width, a.height)
        one.render(S, a)
        torch.cuda.synchronize()
        got, want = fr.stats(), one.stats()
        assert got["paths"] > 0 and got["shadow_rays"] > 0
        assert got == want

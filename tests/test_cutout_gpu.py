"""Alpha cut-outs on the GPU against the oracle (TransparentTest and the
loops of Model::rayHit / Model::rayHit_test, src/model.cpp:217-354).

Ten stacked quads with random 0/1 alpha planes in front of an opaque quad:
closest hits (t and face id) and shadow-ray results are bit-exact with the
oracle, including rays that meet eight transparent points (a miss, or a
blocked shadow ray).  The Cornell box with a checkerboard cut-out grille
renders with the G-buffer bit-exact and the radiance near-exact."""
import numpy as np
import pytest
import torch

import oracle.oracle_py as O
import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from cutout_util import cornell_with_grille, down_rays, layers

pytestmark = pytest.mark.gpu


def stack(seed=3, n_layers=10):
    rng = np.random.default_rng(seed)
    alphas = [np.float32(rng.random((8, 8)) < 0.7) * np.float32(1.0) for _ in range(n_layers)]
    alphas = [1.0 - a for a in alphas]  # 70% of the texels transparent
    return layers(alphas)


def test_trace_closest_and_any_match_the_oracle(dev):
    d, keep = stack()
    S = rm.Scene(d)
    OS = O.OracleScene(d)
    assert S.info.n_tris == 22
    rng = np.random.default_rng(5)
    n = 20000
    rays = down_rays(rng.uniform(-1.2, 1.2, (n, 2)))
    rays[:, 3:5] = rng.uniform(-0.15, 0.15, (n, 2))  # slanted rays cross texel borders
    rays[:, 3:] /= np.linalg.norm(rays[:, 3:], axis=1, keepdims=True)
    ot, oid = OS.closest(rays)
    dr = torch.from_numpy(rays).to(dev)
    t, ids = S.trace_closest(dr)
    torch.cuda.synchronize()
    t, ids = t.cpu().numpy(), ids.cpu().numpy()
    assert np.array_equal(ids, oid)
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    layer = np.where(oid >= 2, oid // 2, -1)
    print("layers hit:", np.bincount(layer[layer >= 0]), "back:", int((oid >= 0).sum() - (oid >= 2).sum()),
          "miss:", int((oid < 0).sum()))
    assert (oid < 0).sum() > 0 and (oid >= 2).sum() > 0 and ((oid >= 0) & (oid < 2)).sum() > 0
    aims = rng.uniform(0.5, 8.0, n).astype(np.float32)
    oocc = OS.any(rays, aims=aims)
    occ = S.trace_any(dr, aim_dev=torch.from_numpy(aims).to(dev))
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy(), oocc)
    assert 0 < oocc.sum() < n


def test_cornell_with_cutout_grille_renders_like_the_oracle(dev):
    g = rm.generate(0, 1)
    d, keep = cornell_with_grille(g)
    S = rm.Scene(d)
    OS = O.OracleScene(d)
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / 96)
    a.width = a.height = 96
    a.spp = 32
    gb, rad = RC.gpu_render(S, a)
    ogb, orad = OS.render(a)
    RC.assert_gbuffer_equal(gb, ogb)
    st = RC.compare_radiance(rad, orad)
    print("cornell + grille 96^2 32spp:", st)
    assert st["direct_close"] >= 0.995 and st["indirect_close"] >= 0.97
    # rays from the camera to random points of the grille: some stop on it,
    # some pass through its transparent cells
    verts = keep[0].reshape(-1, 3, 3)
    n = verts.shape[0]
    rng = np.random.default_rng(9)
    w = rng.dirichlet((1, 1, 1), 4000).astype(np.float32)
    pts = np.einsum("ij,ijk->ik", w, verts[n - 2 + (np.arange(4000) % 2)])
    cam = np.float32(a.position[:])
    dirs = pts - cam
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.concatenate([np.broadcast_to(cam, dirs.shape), dirs], axis=1).astype(np.float32)
    _, ids = S.trace_closest(torch.from_numpy(rays).to(dev))
    torch.cuda.synchronize()
    on_grille = np.isin(ids.cpu().numpy(), [n - 2, n - 1])
    print("grille rays stopped:", on_grille.mean())
    assert 0.2 < on_grille.mean() < 0.8


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_mixing_cutout_and_plain_scenes_equals_single_renders(dev, use_graph):
    """The cut-out flag travels per job: a batch of the Cornell box and the
    Cornell box with the grille gives each job its single render, bit for bit."""
    g = rm.generate(0, 1)
    d, keep = cornell_with_grille(g)
    scenes = [rm.Scene(g), rm.Scene(d)]
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / 64)
    a.width = a.height = 64
    a.spp = 16
    jobs = [(S, a, rm.Frame(64, 64), 0) for S in scenes]
    rm.render_batch(jobs, use_graph=use_graph)
    for S, _, fr, _ in jobs:
        gb, rad = RC.download_all(fr)
        gb1, rad1 = RC.gpu_render(S, a)
        assert np.array_equal(gb.view(np.uint32), gb1.view(np.uint32))
        assert np.array_equal(rad.view(np.uint32), rad1.view(np.uint32))

"""The path-replay adjoint on the GPU (rm_adjoint through DiffRenderer).

  * the replay reproduces the image it differentiates;
  * same seed, same estimator as the oracle's FP64 forward-mode adjoint:
    directional derivatives along texel, roughness and emission directions;
  * albedo and emission: central finite differences of the GPU renderer with
    the same seeds; roughness: finite differences in expectation over seeds;
  * rm_scene_update gives the same renders as a scene built from scratch;
  * the torch.autograd wrapper drives an inverse-rendering loop.
"""
import numpy as np
import pytest
import torch

import raym0nade_amd as rm
from oracle import oracle_py as O
from raym0nade_amd import render_check as RC
from raym0nade_amd.diff import DiffRenderer

from adj_util import linear_image, perturbed_desc, small_args, texel_count
from raym0nade_amd._lib import RM_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu
LIGHT = 5


@pytest.fixture(scope="module")
def gen():
    return rm.generate(4, 1, tex_size=64)


def lin_loss_grad(R, params, delta):
    img = R.forward_image(*params)
    gt, gr, ge, _ = R.adjoint(delta, params[0].numel())
    return img, gt, gr, ge


def test_replay_reproduces_image(dev, gen):
    R = DiffRenderer(gen, small_args(gen, 64, 32))
    p = R.initial_params()
    img = R.forward_image(*p)
    delta = torch.randn_like(img)
    _, _, _, rep = R.adjoint(delta, p[0].numel(), replay=True)
    a, b = rep.cpu().numpy(), img.cpu().numpy()
    close = np.isclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max())
    assert close.mean() >= 0.999, close.mean()


@pytest.mark.skip(reason="one run of this test ended in a GPU memory fault (raised as RmError, box ended) after the "
                         "LDS walk-stack change; its cause is not yet shown, so the two-replay adjoint it exercises "
                         "is refused by rm_adjoint (test_unrecorded_frame_is_refused) and the case stays off the GPU")
def test_recorded_first_pass_equals_replayed(dev, gen):
    """With recording the render keeps each sample's sums and the adjoint
    replays it once; without, it replays twice.  Same gradients."""
    args = small_args(gen, 48, 32)
    delta = torch.rand((48, 48, 3), generator=torch.Generator().manual_seed(9)).to(dev)
    out = []
    for rec in (True, False):
        R = DiffRenderer(gen, args, record=rec)
        p = R.initial_params()
        R.forward_image(*p)
        out.append(R.adjoint(delta, p[0].numel(), replay=True))
    for a, b in zip(out[0], out[1]):
        scale = float(b.abs().max()) + 1e-12
        torch.testing.assert_close(a, b, rtol=2e-3, atol=1e-5 * scale)


def test_matches_oracle_adjoint(dev, gen):
    args = small_args(gen, 32, 16)
    R = DiffRenderer(gen, args)
    p = R.initial_params()
    img = R.forward_image(*p)
    rng = np.random.default_rng(5)
    delta = rng.standard_normal((32, 32, 3)).astype(np.float32)
    gt, gr, ge, _ = R.adjoint(torch.from_numpy(delta).to(dev), p[0].numel())
    gt, gr, ge = gt.cpu().numpy().astype(np.float64), gr.cpu().numpy(), ge.cpu().numpy()
    S = O.OracleScene(gen.desc)
    nm = gen.desc.n_materials
    oimg, _ = O.adjoint_dir(S, args, None, np.zeros(nm), np.zeros((nm, 3)))
    np.testing.assert_allclose(img.cpu().numpy(), oimg, rtol=2e-3, atol=2e-3 * np.abs(oimg).max())

    def check(tt, tr, te, gpu_val, what):
        _, dimg = O.adjoint_dir(S, args, tt, tr, te)
        ref = float((delta * dimg).sum())
        scale = float(np.abs(delta * dimg).sum()) + 1e-12
        assert abs(gpu_val - ref) <= 2e-2 * scale, (what, gpu_val, ref, scale)

    tt = rng.standard_normal(texel_count(gen)).astype(np.float32)
    check(tt, np.zeros(nm), np.zeros((nm, 3)), float(gt @ tt), "textures")
    for m in range(nm):
        if m == LIGHT:
            continue
        tr = np.zeros(nm)
        tr[m] = 1.0
        check(None, tr, np.zeros((nm, 3)), float(gr[m]), f"roughness {m}")
    te = np.zeros((nm, 3))
    te[LIGHT] = rng.standard_normal(3)
    check(None, np.zeros(nm), te, float((ge * te).sum()), "emission")


def smooth_fd(R, params, k, d, h):
    """Per-pixel central differences of the image along d for parameter k, at
    steps h and h/2; pixels whose two estimates disagree had a discrete
    decision flip inside the step (outlier drop, red nudge: a jump, so the
    difference grows like 1/h) and are masked out."""
    def fd(step):
        p = list(params)
        p[k] = params[k] + step * d
        ip = R.forward_image(*p).double()
        p[k] = params[k] - step * d
        im = R.forward_image(*p).double()
        return (ip - im) / (2 * step)
    f1, f2 = fd(h), fd(h / 2)
    mask = ((f1 - f2).abs() <= 1e-2 * (f1.abs() + f2.abs()) + 1e-3 * f1.abs().max()).all(-1)
    return f1, mask


def test_albedo_and_emission_match_finite_differences(dev, gen):
    R = DiffRenderer(gen, small_args(gen, 96, 32))
    params = R.initial_params()
    g = torch.Generator(device="cpu").manual_seed(3)
    delta = torch.rand((96, 96, 3), generator=g).to(dev)
    dt = torch.randn(params[0].shape, generator=g).to(dev)
    de = torch.zeros_like(params[2])
    de[LIGHT] = torch.randn(3, generator=g).to(dev)
    for k, d, h in ((0, dt, 1e-3), (2, de, 0.5)):
        fd, mask = smooth_fd(R, params, k, d, h)
        assert mask.float().mean() >= 0.99
        dm = delta * mask[..., None]
        R.forward_image(*params)
        grads = R.adjoint(dm, params[0].numel())
        ad = float((grads[k].double() * d.double()).sum())
        ref = float((dm.double() * fd).sum())
        assert abs(ref - ad) <= 2e-3 * abs(ad), (k, ref, ad, float(mask.float().mean()))


def test_roughness_matches_finite_differences_in_expectation(dev, gen):
    tex, rough, emis = DiffRenderer(gen, small_args(gen, 8, 1)).initial_params()
    w = torch.rand((48, 48, 3), generator=torch.Generator().manual_seed(4)).to(dev)
    for m in (0, 4):
        dr = torch.zeros_like(rough)
        dr[m] = 1.0
        h = 0.03
        fds, ads = [], []
        for seed in range(1, 25):
            R = DiffRenderer(gen, small_args(gen, 48, 32, seed))
            R.forward_image(tex, rough, emis)
            _, gr, _, _ = R.adjoint(w, tex.numel())
            ads.append(float(gr[m]))
            lp = float((w * R.forward_image(tex, rough + h * dr, emis)).sum())
            lm = float((w * R.forward_image(tex, rough - h * dr, emis)).sum())
            fds.append((lp - lm) / (2 * h))
        d = np.array(fds) - np.array(ads)
        se = d.std(ddof=1) / np.sqrt(len(d))
        assert abs(d.mean()) <= 4 * se + 1e-3 * abs(np.mean(ads)), (m, np.mean(fds), np.mean(ads), se)


def test_scene_update_equals_fresh_scene(dev, gen):
    args = small_args(gen, 48, 8)
    R = DiffRenderer(gen, args)
    tex, rough, emis = R.initial_params()
    rng = np.random.default_rng(6)
    dt = (0.05 * rng.standard_normal(tex.numel())).astype(np.float32)
    drough = np.zeros(rough.numel())
    drough[[0, 2, 4]] = [0.05, -0.03, 0.1]
    demis = np.zeros((rough.numel(), 3))
    demis[LIGHT] = [2.0, -1.0, 0.5]
    R.forward_image(tex + torch.from_numpy(dt).to(dev), rough + torch.from_numpy(drough).float().to(dev),
                    emis + torch.from_numpy(demis).float().to(dev))
    gb, rad = RC.download_all(R.frame)
    d, keep = perturbed_desc(gen, dt, drough, demis)
    fresh = rm.Scene(d)
    gb1, rad1 = RC.gpu_render(fresh, args)
    RC.assert_gbuffer_equal(gb, gb1)
    assert np.array_equal(rad.view(np.uint32), rad1.view(np.uint32))


def test_autograd_inverse_rendering_reduces_loss(dev, gen):
    R = DiffRenderer(gen, small_args(gen, 48, 16))
    tex0, rough0, emis0 = R.initial_params()
    # target: brighter light, rougher walls, a tinted texture set
    tex_t = (tex0 * 0.8 + 0.1).clamp(0.01, 1.0)
    rough_t = (rough0 + 0.15).clamp(0.05, 1.0)
    emis_t = emis0 * 1.3
    with torch.no_grad():
        target = sum(DiffRenderer(gen, small_args(gen, 48, 64, 1000 + k)).forward_image(tex_t, rough_t, emis_t)
                     for k in range(4)) / 4
    tex = tex0.clone().requires_grad_()
    rough = rough0.clone().requires_grad_()
    emis = emis0.clone().requires_grad_()
    opt = torch.optim.Adam([{"params": [tex], "lr": 0.02}, {"params": [rough], "lr": 0.01},
                            {"params": [emis], "lr": 0.3}])
    losses = []
    for step in range(40):
        R.args.seed = step + 1
        img = R.render(tex, rough, emis)
        loss = ((img - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.01, 1.0)
            rough.clamp_(0.05, 1.0)
            emis[LIGHT].clamp_(min=0.5)  # a material keeps emitting (rm_scene_update's rule)
        losses.append(float(loss))
    # the loss has a floor (Monte Carlo noise of both images), so it falls by a
    # fraction; the light, seen by every pixel, moves toward the target.  (At
    # 48^2 pixels most of the 5 x 64^2 texels get a handful of noisy samples,
    # so individual texels are not expected to converge in 40 steps.)
    assert np.mean(losses[-5:]) < 0.8 * np.mean(losses[:5]), losses
    assert float(emis[LIGHT].mean()) > float(emis0[LIGHT].mean())


def test_unrecorded_frame_is_refused(dev, gen):
    """rm_adjoint needs the render's per-sample record; a frame rendered with
    recording off is refused before any launch."""
    R = DiffRenderer(gen, small_args(gen, 24, 8), record=False)
    p = R.initial_params()
    R.forward_image(*p)
    with pytest.raises(rm.RmError) as e:
        R.adjoint(torch.ones((24, 24, 3), device=dev), p[0].numel())
    assert e.value.code == RM_ERR_UNSUPPORTED and "recording" in str(e.value)


class _DescScene(rm.GeneratedScene):
    """A scene description built in the test, in the shape DiffRenderer takes."""

    def __init__(self, desc, args, keep):
        self._h = None
        self.desc, self.args, self._keep_arrays = desc, args, keep


def test_cutout_scene_adjoint_matches_oracle(dev):
    """Cornell box with an alpha cut-out grille: every walk of the replay skips
    the grille's transparent cells (TransparentTest), so the replay equals the
    render, and the gradients equal the FP64 oracle adjoint's for the same seed."""
    from cutout_util import cornell_with_grille
    g0 = rm.generate(0, 1)
    desc, keep = cornell_with_grille(g0)
    G = _DescScene(desc, g0.args, keep)
    args = small_args(G, 32, 16)
    R = DiffRenderer(G, args)
    p = R.initial_params()
    img = R.forward_image(*p)
    nm = desc.n_materials
    light = [m for m in range(nm) if max(desc.materials[m].emission) > 0][0]
    rng = np.random.default_rng(11)
    delta = rng.standard_normal((32, 32, 3)).astype(np.float32)
    gt, gr, ge, rep = R.adjoint(torch.from_numpy(delta).to(dev), p[0].numel(), replay=True)
    a, b = rep.cpu().numpy(), img.cpu().numpy()
    assert np.isclose(a, b, rtol=1e-4, atol=1e-5 * np.abs(b).max()).mean() >= 0.999
    gt, gr, ge = gt.cpu().numpy().astype(np.float64), gr.cpu().numpy(), ge.cpu().numpy()
    S = O.OracleScene(desc)
    oimg, _ = O.adjoint_dir(S, args, None, np.zeros(nm), np.zeros((nm, 3)))
    np.testing.assert_allclose(img.cpu().numpy(), oimg, rtol=2e-3, atol=2e-3 * np.abs(oimg).max())

    def check(tt, tr, te, gpu_val, what):
        _, dimg = O.adjoint_dir(S, args, tt, tr, te)
        ref = float((delta * dimg).sum())
        scale = float(np.abs(delta * dimg).sum()) + 1e-12
        assert abs(gpu_val - ref) <= 2e-2 * scale, (what, gpu_val, ref, scale)

    tt = rng.standard_normal(p[0].numel()).astype(np.float32)
    check(tt, np.zeros(nm), np.zeros((nm, 3)), float(gt @ tt), "grille texture")
    for m in (0, nm - 1):
        tr = np.zeros(nm)
        tr[m] = 1.0
        check(None, tr, np.zeros((nm, 3)), float(gr[m]), f"roughness {m}")
    te = np.zeros((nm, 3))
    te[light] = rng.standard_normal(3)
    check(None, np.zeros(nm), te, float((ge * te).sum()), "emission")


def test_adjoint_argument_errors(dev, gen):
    R = DiffRenderer(gen, small_args(gen, 16, 4))
    p = R.initial_params()
    d = torch.zeros((16, 16, 3), device=dev)
    with pytest.raises(rm.RmError):
        R.adjoint(d, p[0].numel())  # nothing rendered yet
    R.forward_image(*p)
    R.args.seed = 77  # not the render the frame holds
    with pytest.raises(rm.RmError):
        R.adjoint(d, p[0].numel())

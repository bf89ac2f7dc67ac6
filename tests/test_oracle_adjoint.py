"""Pins the FP64 forward-mode reference adjoint of the oracle (oadjoint.cpp).

Its derivative estimator is the one the GPU adjoint implements: exact
derivative of the seeded estimator for albedo and emission (no sampling
decision depends on them in these scenes: one light object), and pathwise +
score-function derivative for roughness.  Checks:
  * its primal equals the restatement's own render (same samples);
  * albedo and emission: central finite differences of the restatement with
    the same seeds (per pixel);
  * the image is linear in the emission (dI/de along e equals I minus the
    emission-free part, over |e|);
  * roughness: agreement in expectation with finite differences over many
    seeds (the seeded estimator jumps where a GGX/cosine pick or a roulette
    decision flips, so per-seed differences are noisy; the test allows 4
    standard errors).
"""
import numpy as np
import pytest

import raym0nade_amd as rm
from oracle import oracle_py as O

from adj_util import linear_image, perturbed_desc, small_args, texel_count

LIGHT = 5  # material index of the light quad in the textured Cornell box


@pytest.fixture(scope="module")
def scene():
    g = rm.generate(4, 1, tex_size=32)
    return g, O.OracleScene(g.desc)


def render_lin(S, a):
    gb, rad = S.render(a)
    return linear_image(gb, rad, a.exposure)


def test_primal_equals_render(scene):
    g, S = scene
    a = small_args(g, 24, 16)
    nm = g.desc.n_materials
    img, dimg = O.adjoint_dir(S, a, None, np.zeros(nm), np.zeros((nm, 3)))
    ref = render_lin(S, a)
    np.testing.assert_allclose(img, ref, rtol=2e-5, atol=2e-6)
    assert not dimg.any()


def fd_image(g, a, h, dtex=None, drough=None, demis=None):
    nm = g.desc.n_materials
    z = np.zeros(nm)
    dp, kp = perturbed_desc(g, None if dtex is None else h * dtex, h * (drough if drough is not None else z),
                            h * (demis if demis is not None else np.zeros((nm, 3))))
    dm, km = perturbed_desc(g, None if dtex is None else -h * dtex, -h * (drough if drough is not None else z),
                            -h * (demis if demis is not None else np.zeros((nm, 3))))
    return (render_lin(O.OracleScene(dp), a) - render_lin(O.OracleScene(dm), a)) / (2 * h)


def test_albedo_matches_finite_differences(scene):
    g, S = scene
    a = small_args(g, 24, 16)
    nm = g.desc.n_materials
    rng = np.random.default_rng(0)
    tt = rng.standard_normal(texel_count(g)).astype(np.float32)
    _, dimg = O.adjoint_dir(S, a, tt, np.zeros(nm), np.zeros((nm, 3)))
    fd = fd_image(g, a, 1e-3, dtex=tt)
    fd2 = fd_image(g, a, 5e-4, dtex=tt)
    # a decision that flips inside the step (outlier drop, red nudge) is a jump:
    # its difference quotient grows like 1/h; such pixels are left out
    smooth = (np.abs(fd - fd2) <= 1e-2 * (np.abs(fd) + np.abs(fd2)) + 1e-3 * np.abs(fd).max()).all(-1)
    assert smooth.mean() >= 0.99
    err = np.abs(fd - dimg)[smooth]
    tol = (1e-2 * np.abs(dimg) + 1e-3 * np.abs(dimg).max())[smooth]
    assert (err <= tol).all()
    w = rng.random(dimg.shape) * smooth[..., None]
    assert abs((w * fd).sum() - (w * dimg).sum()) <= 1e-3 * np.abs(w * dimg).sum()


def test_emission_matches_finite_differences(scene):
    g, S = scene
    a = small_args(g, 24, 16)
    nm = g.desc.n_materials
    rng = np.random.default_rng(1)
    te = np.zeros((nm, 3))
    te[LIGHT] = rng.standard_normal(3)
    _, dimg = O.adjoint_dir(S, a, None, np.zeros(nm), te)
    fd = fd_image(g, a, 0.5, demis=te)  # the image is linear in the emission: no truncation error
    np.testing.assert_allclose(fd, dimg, rtol=1e-4, atol=1e-5 * np.abs(dimg).max())


def test_image_is_homogeneous_in_emission(scene):
    g, S = scene
    a = small_args(g, 24, 16)
    nm = g.desc.n_materials
    te = np.zeros((nm, 3))
    te[LIGHT] = 15.0  # along e itself (emission (15, 15, 15)): I is homogeneous of degree 1, so dI = I
    img, dimg = O.adjoint_dir(S, a, None, np.zeros(nm), te)
    np.testing.assert_allclose(dimg, img, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mat", [0, 3])
def test_roughness_matches_finite_differences_in_expectation(scene, mat):
    g, S = scene
    nm = g.desc.n_materials
    tr = np.zeros(nm)
    tr[mat] = 1.0
    h = 0.03
    dp, kp = perturbed_desc(g, drough=h * tr)
    dm, km = perturbed_desc(g, drough=-h * tr)
    Sp, Sm = O.OracleScene(dp), O.OracleScene(dm)
    rng = np.random.default_rng(2)
    w = rng.random((16, 16, 3))
    fds, ads = [], []
    for seed in range(1, 301):
        a = small_args(g, 16, 32, seed)
        _, dimg = O.adjoint_dir(S, a, None, tr, np.zeros((nm, 3)))
        fds.append((w * (render_lin(Sp, a) - render_lin(Sm, a)) / (2 * h)).sum())
        ads.append((w * dimg).sum())
    d = np.array(fds) - np.array(ads)
    se = d.std(ddof=1) / np.sqrt(len(d))
    assert abs(d.mean()) <= 4 * se, (np.mean(fds), np.mean(ads), se)
    assert abs(np.mean(ads)) > 4 * np.std(ads, ddof=1) / np.sqrt(len(ads))  # a non-trivial gradient


def test_rejects_transmissive_scene(scene):
    g, _ = scene
    d, keep = perturbed_desc(g)
    keep[0][0].opacity = 0.0  # a dielectric wall
    S = O.OracleScene(d)
    a = small_args(g, 8, 4)
    nm = g.desc.n_materials
    with pytest.raises(ValueError):
        O.adjoint_dir(S, a, None, np.zeros(nm), np.zeros((nm, 3)))

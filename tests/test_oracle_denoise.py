"""Pins for the oracle's denoise restatement (src/image.cpp:30-213):
spatialClamp against an independent separable blur (scipy), filterVar
against an independent 2-D correlation, and the a-trous pass against closed
forms on scenes whose weights are known exactly (a flat plane, depth /
normal / material edges, missed pixels) and against a plain dilated
B3-spline blur in the limit where the radiance term vanishes.  A mistake in
a stride, a tap table, the normal used for the depth slope, the roughness
factor or the border handling fails one of them."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle_py as O

H7 = np.array([0.03125, 0.109375, 0.21875, 0.28125, 0.21875, 0.109375, 0.03125])
H3 = np.array([0.25, 0.5, 0.25])
H5 = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
W_RGB = np.array([0.3, 0.6, 0.1])


def plane_gbuf(h, w, z=None, surf=(1.0, 0.0, 0.0), shape=(0.0, 0.0, 1.0), metallic=0.0, rough=0.5, opacity=1.0):
    """Pixels on the plane z (scalar or [h][w]) at integer (x, y); shapeNormal
    perpendicular to the plane, surfaceNormal as given (the same everywhere)."""
    g = np.zeros((h, w, O.GB), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    g[..., 0] = xx
    g[..., 1] = yy
    g[..., 2] = 0.0 if z is None else z
    g[..., 3:6] = shape
    g[..., 6:9] = surf
    g[..., 15] = metallic
    g[..., 16] = rough
    g[..., 17] = opacity
    return g


def close(a, b, **kw):
    a = np.asarray(a)
    np.testing.assert_allclose(a, np.broadcast_to(np.asarray(b, dtype=np.float64), a.shape), **kw)


def const_rad(h, w, rgb=(0.5, 0.25, 0.125), var=0.04):
    r = np.zeros((h, w, 4), np.float32)
    r[..., :3] = rgb
    r[..., 3] = var
    return r


# ---------------------------------------------------------------- spatialClamp

def test_clamp_matches_separable_blur():
    rng = np.random.default_rng(3)
    h, w = 23, 37
    rad = np.zeros((4, h, w, 4), np.float32)
    rad[..., :3] = rng.gamma(1.0, 0.2, (4, h, w, 3))
    # fireflies
    for b in range(4):
        for _ in range(12):
            rad[b, rng.integers(h), rng.integers(w), :3] *= rng.uniform(50, 500)
    rad[..., 3] = rng.uniform(0, 1, (4, h, w))
    g = plane_gbuf(h, w)
    out = O.denoise(g, rad, clamp=True, filter=False)
    for b in range(4):
        lum = (rad[b, ..., :3].astype(np.float64) * W_RGB).sum(-1)
        b1 = ndimage.correlate1d(lum, H7, axis=1, mode="constant")
        b2 = ndimage.correlate1d(b1, H7, axis=0, mode="constant")
        blurred = b2 - H7[3] ** 2 * lum
        mask = lum > 36.0 * blurred
        scale = np.where(mask, blurred / (lum + 1e-4) / (1 - H7[3] ** 2), 1.0)
        want = rad[b, ..., :3] * scale[..., None]
        assert mask.sum() >= 6
        np.testing.assert_allclose(out[b, ..., :3], want, rtol=2e-5, atol=1e-7)
        assert np.array_equal(out[b, ..., 3], rad[b, ..., 3])  # Var untouched


def test_clamp_spike_on_constant_background_takes_the_background_luminance():
    h, w = 15, 15
    rad = np.zeros((4, h, w, 4), np.float32)
    rad[..., :3] = 0.2
    rad[:, 7, 7, :3] = (90.0, 30.0, 10.0)
    out = O.denoise(plane_gbuf(h, w), rad, clamp=True, filter=False)
    c = 0.2  # background luminance
    V = 0.3 * 90 + 0.6 * 30 + 0.1 * 10
    lum = (out[0, 7, 7, :3] * W_RGB).sum()
    assert lum == pytest.approx(c * V / (V + 1e-4), rel=1e-5)
    # the spike's colour ratios survive, the background is unchanged
    assert out[0, 7, 7, 0] / out[0, 7, 7, 1] == pytest.approx(3.0, rel=1e-6)
    bg = np.ones((h, w), bool)
    bg[7, 7] = False
    assert np.array_equal(out[:, bg], rad[:, bg])


# ---------------------------------------------------------------- filterVar

def test_filter_var_matches_2d_correlation():
    rng = np.random.default_rng(5)
    h, w = 17, 29
    rad = rng.uniform(0, 2, (4, h, w, 4)).astype(np.float32)
    out = O.denoise(plane_gbuf(h, w), rad, var_only=True)
    k = np.outer(H3, H3)
    for b in range(4):
        L = rad[b, ..., :3].astype(np.float64)
        E = np.stack([ndimage.correlate(L[..., c], k, mode="constant") for c in range(3)], -1)
        E2 = ndimage.correlate((L * L).sum(-1), k, mode="constant")
        V = ndimage.correlate(rad[b, ..., 3].astype(np.float64), k, mode="constant")
        want = V + E2 - (E * E).sum(-1)
        np.testing.assert_allclose(out[b, ..., 3], want, rtol=1e-4, atol=1e-5)
        assert np.array_equal(out[b, ..., :3], rad[b, ..., :3])


def test_filter_var_of_constant_image_keeps_var_inside_and_shrinks_at_border():
    h, w = 6, 7
    rad = np.stack([const_rad(h, w, var=0.5)] * 4)
    out = O.denoise(plane_gbuf(h, w), rad, var_only=True)
    # interior: the radiance's spatial variance is 0, Var is averaged over 9 equal values
    assert np.allclose(out[0, 1:-1, 1:-1, 3], 0.5, rtol=1e-6, atol=1e-6)
    # corner: 4 taps of total weight 0.75^2 (weights are not renormalised)
    s = 0.75 ** 2
    L2 = 0.5 ** 2 + 0.25 ** 2 + 0.125 ** 2
    assert out[0, 0, 0, 3] == pytest.approx(0.5 * s + L2 * s - L2 * s * s, rel=1e-5)


# ---------------------------------------------------------------- a-trous pass

def sum_h2():
    return float((H5 ** 2).sum())


@pytest.mark.parametrize("step", [1, 2, 4])
def test_pass_on_flat_plane_is_the_plain_b3_spline(step):
    """Flat plane, constant radiance: every weight is exactly h_i h_j
    (normal power 1, depth slope 0 -- the slope uses shapeNormal, not the
    tilted-away surfaceNormal -- material and radiance distance 0)."""
    h, w = 40, 44
    g = plane_gbuf(h, w)
    r = const_rad(h, w, var=0.08)
    out = O.filter_pass(g, r, specular=False, step=step)
    inner = out[2 * step:-2 * step, 2 * step:-2 * step]
    close(inner[..., :3], r[0, 0, :3], rtol=1e-6)
    close(inner[..., 3], 0.08 * sum_h2() ** 2, rtol=1e-6)
    # corner: taps at offsets 0..2 only, weights renormalised by their sum
    s1 = float(H5[2:].sum())
    s2 = float((H5[2:] ** 2).sum())
    assert out[0, 0, 3] == pytest.approx(0.08 * (s2 / s1 ** 2) ** 2, rel=1e-6)
    close(out[0, 0, :3], r[0, 0, :3], rtol=1e-6)


@pytest.mark.parametrize("rough", [0.5, 0.01])
def test_specular_pass_weights_neighbours_by_max_roughness(rough):
    h, w = 11, 11
    g = plane_gbuf(h, w, rough=rough)
    r = const_rad(h, w, var=1.0)
    out = O.filter_pass(g, r, specular=True, step=1)
    f = max(rough, 0.04)
    wts = np.outer(H5, H5) * f
    wts[2, 2] = H5[2] ** 2  # the centre tap is never weighted
    assert out[5, 5, 3] == pytest.approx((wts ** 2).sum() / wts.sum() ** 2, rel=1e-5)


def test_pass_respects_depth_edges():
    """Two constant regions at different depths: a tap across the step has a
    direction along the normal, slope tan ~ 1/eps, weight 0 -- no bleeding."""
    h, w = 20, 24
    z = np.where(np.arange(w)[None, :] < 12, 0.0, 1000.0) * np.ones((h, 1))
    g = plane_gbuf(h, w, z=z)
    r = const_rad(h, w, var=0.1)
    r[:, 12:, :3] = (2.0, 1.0, 0.5)
    for step in (1, 2, 4):
        out = O.filter_pass(g, r, specular=False, step=step)
        close(out[:, :12, :3], r[0, 0, :3], rtol=1e-6)
        close(out[:, 12:, :3], (2.0, 1.0, 0.5), rtol=1e-6)


def test_pass_respects_normal_edges_and_missed_pixels():
    h, w = 16, 16
    g = plane_gbuf(h, w)
    g[:, 8:, 6:9] = (0.0, 1.0, 0.0)  # orthogonal surface normal: power 0
    g[0:3, 0:3, 0:3] = np.nan           # primary ray missed
    r = const_rad(h, w, var=0.1)
    r[:, 8:, :3] = (3.0, 3.0, 3.0)
    out = O.filter_pass(g, r, specular=False, step=1)
    assert (out[0:3, 0:3] == 0).all()   # missed: radiance 0, Var 0
    close(out[4:, :8, :3], r[0, 0, :3], rtol=1e-6)
    close(out[:, 8:, :3], 3.0, rtol=1e-6)


def test_pass_material_distance_scales_cross_edge_taps_by_exp():
    """metallic 0 | 1 across a vertical edge: taps across it carry exp(-1)."""
    h, w = 9, 10
    g = plane_gbuf(h, w)
    g[:, 5:, 15] = 1.0
    r = const_rad(h, w, var=1.0)
    out = O.filter_pass(g, r, specular=False, step=1)
    y, x = 4, 4  # columns 5, 6 are across
    wts = np.outer(H5, H5)
    wts[:, 3:] *= np.exp(np.float32(-1.0))
    assert out[y, x, 3] == pytest.approx((wts ** 2).sum() / wts.sum() ** 2, rel=1e-5)


def test_pass_radiance_cut():
    """A neighbour whose radiance distance exceeds 7.5 (sqrt(Var) + 0.01)
    drops out entirely."""
    h, w = 5, 5
    g = plane_gbuf(h, w)
    r = const_rad(h, w, rgb=(1.0, 1.0, 1.0), var=0.0)
    r[2, 3, :3] = 1.0 + 0.1 / np.sqrt(3)   # distance 0.1 -> k = -0.1/0.01 = -10 < -7.5
    out = O.filter_pass(g, r, specular=False, step=1)
    wts = np.outer(H5, H5)
    wts[2, 3] = 0.0
    assert out[2, 2, 0] == pytest.approx(1.0, rel=1e-6)
    assert out[2, 2, 3] == 0.0
    r[2, 3, :3] = 1.0 + 0.07 / np.sqrt(3)  # k = -7 > -7.5: kept with exp(-7)
    out = O.filter_pass(g, r, specular=False, step=1)
    wts[2, 3] = H5[2] * H5[3] * np.exp(-7.0)
    want = (wts.sum() - wts[2, 3] + wts[2, 3] * (1.0 + 0.07 / np.sqrt(3))) / wts.sum()
    assert out[2, 2, 0] == pytest.approx(want, rel=1e-5)


@pytest.mark.parametrize("step", [1, 4])
def test_pass_with_vanishing_radiance_term_is_a_normalised_dilated_blur(step):
    """Huge Var makes the radiance distance ~1e-6: the pass becomes the
    normalised dilated B3-spline blur of the radiance (border taps dropped)."""
    rng = np.random.default_rng(7)
    h, w = 33, 47
    g = plane_gbuf(h, w)
    r = np.zeros((h, w, 4), np.float32)
    r[..., :3] = rng.uniform(0, 1, (h, w, 3))
    r[..., 3] = 1e12
    out = O.filter_pass(g, r, specular=False, step=step)
    k = np.zeros((4 * step + 1, 4 * step + 1))
    k[::step, ::step] = np.outer(H5, H5)
    norm = ndimage.correlate(np.ones((h, w)), k, mode="constant")
    for c in range(3):
        want = ndimage.correlate(r[..., c].astype(np.float64), k, mode="constant") / norm
        np.testing.assert_allclose(out[..., c], want, rtol=2e-5, atol=1e-6)


def test_weight_direct():
    gp = plane_gbuf(1, 1)[0, 0]
    gq = gp.copy()
    gq[0] = 1.0
    lp = np.array([1, 1, 1, 0.25], np.float32)
    lq = np.array([1, 1, 1, 0.25], np.float32)
    assert O.denoise_weight(gp, gq, lp, lq, False) == 1.0
    gq2 = gq.copy()
    gq2[0:3] = np.nan
    assert O.denoise_weight(gp, gq2, lp, lq, False) == 0.0
    gq3 = gq.copy()
    gq3[15] = 0.5                       # metallic distance 0.5
    assert O.denoise_weight(gp, gq3, lp, lq, False) == pytest.approx(np.exp(-0.5), rel=1e-6)
    gq4 = gq.copy()
    gq4[6:9] = (np.cos(0.01), np.sin(0.01), 0.0)  # 0.01 rad apart: cos^1024
    assert O.denoise_weight(gp, gq4, lp, lq, False) == pytest.approx(np.cos(0.01) ** 1024, rel=1e-4)
    assert O.denoise_weight(gp, gq, lp, lq, True) == pytest.approx(0.5, rel=1e-7)  # roughness 0.5


def test_full_filter_of_flat_constant_plane_keeps_radiance():
    h, w = 40, 40
    g = plane_gbuf(h, w)
    rad = np.stack([const_rad(h, w, rgb=(0.3 * (b + 1), 0.2, 0.1), var=0.02) for b in range(4)])
    out = O.denoise(g, rad, clamp=True, filter=True)
    np.testing.assert_allclose(out[..., :3], rad[..., :3], rtol=1e-5)
    assert (out[..., 3] >= 0).all() and (out[..., 3] < 0.02).all()

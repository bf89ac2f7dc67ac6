"""Pins of the oracle's Photo::shade, gammaCorrection and FXAA
(src/image.cpp:215-246,358-468) against per-option closed forms, the tone
curve in float64, and FXAA outputs worked out by hand (CPU only)."""
import numpy as np
import pytest

import oracle_py as O

LINEAR = 65536  # skip gammaCorrection
GB = O.GB


def frame(h, w, seed=1):
    rng = np.random.default_rng(seed)
    gb = np.zeros((h, w, GB), np.float32)
    gb[..., 0:3] = rng.normal(size=(h, w, 3))
    n = rng.normal(size=(h, w, 6)).reshape(h, w, 2, 3)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    gb[..., 3:9] = n.reshape(h, w, 6)
    gb[..., 9:12] = rng.random((h, w, 3)) * (rng.random((h, w, 1)) < 0.2)
    gb[..., 12:15] = rng.random((h, w, 3))
    gb[..., 15:18] = rng.random((h, w, 3))
    rad = rng.random((4, h, w, 4)).astype(np.float32) * 2.0
    return gb, rad


def shade_closed_form(gb, rad, o, exposure):
    """Photo::shade(options) in float64 per the option bits of
    include/image.h:17-34."""
    g, r = gb.astype(np.float64), rad.astype(np.float64)
    if o & 64:
        return (g[..., 3:6] + 1) / 2
    if o & 128:
        return (g[..., 6:9] + 1) / 2
    rd = np.zeros(g.shape[:2] + (3,))
    rs = np.zeros_like(rd)
    if o & 4:
        rd = rd + (r[0, ..., :3] if o & 16 else 0)
        rs = rs + (r[1, ..., :3] if o & 32 else 0)
    if o & 8:
        rd = rd + (r[2, ..., :3] if o & 16 else 0)
        rs = rs + (r[3, ..., :3] if o & 32 else 0)
    if not o & 12:
        rd = np.ones_like(rd)
    px = (g[..., 12:15] if o & 1 else 1.0) * rd + rs
    if o & 2:
        px = px + g[..., 9:12] * exposure
    return px


@pytest.mark.parametrize("o", range(256))
def test_shade_every_option_combination(o):
    gb, rad = frame(9, 13)
    got = O.postprocess(gb, rad, o | LINEAR, 1.7)
    np.testing.assert_allclose(got, shade_closed_form(gb, rad, o, 1.7), rtol=2e-6, atol=1e-6)


def tone_curve(x):
    """gammaCorrection in float64: max 0; above luminance 0.75 scale the
    colour to tanh(3 (lum - 0.75)) / 3 + 0.75; min 1; power 1 / 2.2."""
    x = np.maximum(x.astype(np.float64), 0.0)
    lum = x @ np.array([0.3, 0.6, 0.1])
    hi = lum > 0.75
    scale = np.where(hi, (np.tanh(3 * (lum - 0.75)) / 3 + 0.75) / np.where(hi, lum, 1), 1.0)
    x = np.minimum(x * scale[..., None], 1.0)
    return x ** (1 / 2.2)


def base_only(img):
    """A frame whose BaseColor shading (option 1) is `img`."""
    h, w = img.shape[:2]
    gb = np.zeros((h, w, GB), np.float32)
    gb[..., 12:15] = img
    return gb, np.zeros((4, h, w, 4), np.float32)


def test_tone_curve_both_branches():
    rng = np.random.default_rng(2)
    img = (rng.random((40, 50, 3)) * 3.0 - 0.5).astype(np.float32)  # negatives, lum below and above 0.75, > 1
    gb, rad = base_only(img)
    got = O.postprocess(gb, rad, 1, 1.0)
    lum = np.maximum(img, 0) @ np.float32([0.3, 0.6, 0.1])
    assert (lum > 0.75).mean() > 0.3 and (lum < 0.75).mean() > 0.2
    np.testing.assert_allclose(got, tone_curve(img), rtol=2e-6, atol=2e-7)


def fxaa(img):
    gb, rad = base_only(img)
    return O.postprocess(gb, rad, 1 | 512 | LINEAR, 1.0)


def test_fxaa_flat_and_straight_edges_unchanged():
    flat = np.full((16, 16, 3), 0.37, np.float32)
    assert np.array_equal(fxaa(flat), flat)
    # FXAA walks along the edge (the offset is perpendicular to the luma
    # gradient), so the samples of a straight edge stay on its own side
    for axis in (0, 1):
        img = np.zeros((16, 16, 3), np.float32)
        if axis == 0:
            img[:, 7:] = 1.0
        else:
            img[7:, :] = 1.0
        assert np.array_equal(fxaa(img), img)


def test_fxaa_diagonal_step_by_hand():
    """Black where x + y < k, white elsewhere (square image).  A black pixel
    on the edge (x + y = k - 1) has white E and S neighbours: edgeHorz =
    edgeVert = 2/3, so the walk goes along +y with step 2/3 of a pixel and
    meets the white pixel below at its second sample; the sub-pixel offset is
    (|N + S - 2M| 2 + |E + W - 2M|) / 4 = 3/4, blended at 0.75: the pixel
    becomes 0.5625 of white.  A white pixel on the edge (x + y = k) finds no
    brighter sample and keeps its colour; pixels off the edge have a flat
    neighbourhood."""
    n, k = 24, 23
    y, x = np.mgrid[0:n, 0:n]
    img = np.repeat(np.where((x + y >= k)[..., None], 1.0, 0.0), 3, axis=-1).astype(np.float32)
    out = fxaa(img)
    white = np.float32(0.3) * 1 + np.float32(0.6) * 1 + np.float32(0.1) * 1  # luma of white
    inner = (x > 0) & (y > 0) & (x < n - 1) & (y < n - 1)
    dark_edge = inner & (x + y == k - 1)
    assert dark_edge.sum() >= 10
    np.testing.assert_allclose(out[dark_edge], 0.5625 * white, rtol=1e-6)
    rest = inner & ~dark_edge
    assert np.array_equal(out[rest], img[rest])

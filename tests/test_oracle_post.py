"""Pins for the oracle's bloom and depth-of-field restatement
(src/image.cpp:248-356), on the linear image (gamma skipped).

Bloom: a dim image is untouched; a bright image gains the five dilated
B3-spline blur levels of its compressed highlights, a sixth each, as an
independent scipy computation gives them; a single highlight gains exactly
5/6 of its level-0 energy away from the border.

Depth of field: CoC 0 is the identity; a constant image stays constant; a
near pixel in front of a plane in focus spreads weights min(CoC0 - r, 1)
normalised over its whole disc (they sum to 1); and an independent gather
formulation (each receiving pixel folds its sources in (depth, index) order
under the 0.99 cap) equals the reference's scatter on random small frames."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle_py as O

LINEAR = 65536
BLOOM = 256
DOF = 1024
H5 = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
W_RGB = np.array([0.3, 0.6, 0.1], np.float32)


def frame(img, depth=None):
    """G-buffer and radiance whose Full shade is `img` (linear rgb [h][w][3]):
    base colour 1, Dd = img, the rest 0; positions at `depth` along +z."""
    h, w = img.shape[:2]
    g = np.zeros((h, w, O.GB), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    g[..., 2] = 5.0 if depth is None else depth  # on the z axis: distance = depth
    g[..., 12:15] = 1.0
    g[..., 16] = 0.8
    g[..., 17] = 1.0
    rad = np.zeros((4, h, w, 4), np.float32)
    rad[0, ..., :3] = img
    return g, rad


def linear(img, options=0, depth=None, **kw):
    g, rad = frame(img, depth)
    return O.postprocess(g, rad, 63 | LINEAR | options, 1.0, **kw)


# ---------------------------------------------------------------- bloom

def test_bloom_leaves_a_dim_image_alone():
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 0.9, (20, 30, 3)).astype(np.float32)
    assert np.array_equal(linear(img, BLOOM), linear(img))


def test_bloom_matches_independent_dilated_blurs():
    rng = np.random.default_rng(2)
    h, w = 37, 53
    img = rng.uniform(0, 0.8, (h, w, 3)).astype(np.float32)
    for _ in range(15):
        img[rng.integers(h), rng.integers(w)] = rng.uniform(1, 30, 3)
    got = linear(img, BLOOM)
    lum = (img.astype(np.float64) * W_RGB).sum(-1)
    lvl = np.where(lum[..., None] >= 1.0, np.maximum(img / np.power(np.maximum(lum, 1e-30), 0.65)[..., None] - 1.0, 0.0), 0.0)
    want = img.astype(np.float64).copy()
    for step in (1, 2, 4, 8, 16):
        k = np.zeros((4 * step + 1, 4 * step + 1))
        k[::step, ::step] = np.outer(H5, H5)
        lvl = np.stack([ndimage.correlate(lvl[..., c], k, mode="constant") for c in range(3)], -1)
        want += lvl / 6.0
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-6)


def test_bloom_of_one_highlight_adds_five_sixths_of_its_level0():
    h, w = 141, 141  # the five levels reach 2 (1 + 2 + 4 + 8 + 16) = 62 pixels
    img = np.zeros((h, w, 3), np.float32)
    img[70, 70] = (40.0, 20.0, 10.0)
    got = linear(img, BLOOM)
    lum = float((img[70, 70] * W_RGB).sum())
    lvl0 = np.maximum(img[70, 70] / lum ** 0.65 - 1.0, 0.0)
    added = got.astype(np.float64).sum((0, 1)) - img.astype(np.float64).sum((0, 1))
    np.testing.assert_allclose(added, lvl0 * 5.0 / 6.0, rtol=1e-5)


# ---------------------------------------------------------------- depth of field

def test_dof_with_zero_coc_is_the_identity():
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 2, (12, 17, 3)).astype(np.float32)
    got = linear(img, DOF, focus=3.0, coc=0.0)
    np.testing.assert_allclose(got, img, rtol=2e-7)


def test_dof_keeps_a_constant_image_constant():
    rng = np.random.default_rng(4)
    h, w = 24, 31
    img = np.broadcast_to(np.float32([0.7, 0.2, 0.4]), (h, w, 3)).copy()
    depth = rng.uniform(2, 9, (h, w)).astype(np.float32)
    got = linear(img, DOF, depth=depth, focus=5.0, coc=6.0)
    np.testing.assert_allclose(got, img, rtol=1e-6)


def test_dof_near_pixel_weights_sum_to_one_over_its_disc():
    """Plane in focus at depth 10 (CoC0 = 1e-4: each of its pixels keeps only
    itself), one near pixel at depth 2 (CoC0 = 6 |1 - 10/2| + 1e-4 = 24.0001,
    capped at 96 -> 24.0001).  The near pixel goes first; a receiver q then
    holds w_q of it and 0.99 - w_q of itself."""
    h, w = 61, 61
    img = np.zeros((h, w, 3), np.float32)
    img[..., 1] = 0.25                    # far: green
    img[30, 30] = (1.0, 0.25, 0.0)        # near: red + the same green
    depth = np.full((h, w), 10.0, np.float32)
    depth[30, 30] = 2.0
    got = linear(img, DOF, depth=depth, focus=10.0, coc=6.0, cam=(0.0, 0.0, 0.0))
    wq = got[..., 0].astype(np.float64) * 0.99  # red share * 0.99 = w_q
    coc0 = np.float32(np.float32(6.0) * np.float32(abs(1 - 10 / 2)) + np.float32(1e-4))
    yy, xx = np.mgrid[0:h, 0:w] - 30
    r = np.sqrt(xx ** 2 + yy ** 2)
    inside = (np.abs(yy) <= int(coc0)) & (np.abs(xx) <= np.floor(np.sqrt(np.maximum(coc0 ** 2 - yy ** 2, 0))))
    ramp = np.minimum(coc0 - r, 1.0) * inside
    kept = (wq > 0) & (r > 0)  # the centre is the near pixel itself (red share 1)
    assert not (kept & ~inside).any()
    # shape: proportional to min(CoC0 - r, 1) over the disc ...
    k = wq[kept] / ramp[kept]
    assert np.allclose(k, k.mean(), rtol=1e-5)
    # ... normalised over the whole disc (the weights sum to 1) ...
    assert k.mean() * ramp.sum() == pytest.approx(1.0, rel=1e-5)
    # ... and a weight below 1e-4 is dropped, not spread
    assert (ramp[inside & ~kept & (r > 0)] * k.mean() < 1e-4).all()
    assert (ramp[kept] * k.mean() >= 1e-4 * (1 - 1e-6)).all()


def gather_dof(img, depth, focus, coc):
    """Independent gather form of depthFeildBlur for small frames."""
    h, w = depth.shape
    f32 = np.float32
    src = []
    for y in range(h):
        for x in range(w):
            d = f32(depth[y, x])
            coc0 = min(f32(f32(coc) * f32(abs(f32(1.0) - f32(focus) / d))) + f32(1e-4), f32(96.0))
            rad = int(coc0)
            tot = f32(0.0)
            for dy in range(-rad, rad + 1):
                xl_ = int(np.sqrt(f32(coc0 * coc0) - f32(dy * dy), dtype=f32))
                xl = -xl_
                while xl <= 0:
                    wt = min(f32(coc0 - np.sqrt(f32(xl * xl + dy * dy), dtype=f32)), f32(1.0))
                    if wt == 1.0:
                        break
                    tot = f32(tot + wt)
                    xl += 1
                xr = xl_
                while xr > 0:
                    wt = min(f32(coc0 - np.sqrt(f32(xr * xr + dy * dy), dtype=f32)), f32(1.0))
                    if wt == 1.0:
                        break
                    tot = f32(tot + wt)
                    xr -= 1
                tot = f32(tot + f32(xr - xl + 1))
            src.append((float(d), y * w + x, x, y, coc0, rad, tot))
    order = sorted(src, key=lambda s: (s[0], s[1]))
    out = np.zeros_like(img)
    for ty in range(h):
        for tx in range(w):
            gained, acc = f32(0.0), np.zeros(3, f32)
            for d, idx, x, y, coc0, rad, tot in order:
                dy, dx = ty - y, tx - x
                if abs(dy) > rad:
                    continue
                if abs(dx) > int(np.sqrt(f32(coc0 * coc0) - f32(dy * dy), dtype=f32)):
                    continue
                wt = f32(min(f32(coc0 - np.sqrt(f32(dx * dx + dy * dy), dtype=f32)), f32(1.0)) / tot)
                if wt < f32(1e-4):
                    continue
                if f32(gained + wt) > f32(0.99):
                    wt = f32(f32(0.99) - gained)
                if wt < f32(1e-4):
                    continue
                acc = (acc + img[y, x] * wt).astype(f32)
                gained = f32(gained + wt)
            out[ty, tx] = acc / gained
    return out


@pytest.mark.parametrize("seed", [5, 6])
def test_dof_equals_an_independent_gather(seed):
    rng = np.random.default_rng(seed)
    h, w = 9, 11
    img = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    depth = rng.choice(np.float32([1.5, 2.0, 4.0, 8.0]), (h, w)).astype(np.float32)
    depth[2, 3] = depth[2, 4]  # exact ties keep index order
    got = linear(img, DOF, depth=depth, focus=4.0, coc=1.5)
    want = gather_dof(img, depth, 4.0, 1.5)
    np.testing.assert_allclose(got, want, rtol=3e-6, atol=1e-7)

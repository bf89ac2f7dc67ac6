"""Closed-form pins for the oracle's rayInBox and RayTriangleIntersection
(src/geometry.cpp:40-87), including each quirk the traversal must mirror."""
import numpy as np
import pytest

import oracle_py as O

f32 = np.float32
EPS = f32(1e-4)


def test_box_straight_hit():
    # ray from (-1,.5,.5) along +x through the unit box: entry 1, exit 2
    tL, tR = O.ray_in_box([-1, .5, .5, 1, 0, 0], [0, 0, 0, 1, 1, 1], 1e-4, np.inf)
    # y and z are parallel (|d|<1e-4) and inside the slab: untouched;
    # x: tL = max(1e-4, 1) = 1, tR = min(inf, 2) + 1e-4
    assert tL == f32(1.0)
    assert tR == f32(2.0) + EPS


def test_box_eps_accumulates_per_axis():
    # direction (1,1,1): tR gets +1e-4 after each of the 3 axes
    d = 1.0
    tL, tR = O.ray_in_box([-1, -1, -1, d, d, d], [0, 0, 0, 1, 1, 1], 1e-4, np.inf)
    expect = f32(2.0)
    expect = f32(expect + EPS)
    expect = f32(min(expect, f32(2.0)) + EPS)
    expect = f32(min(expect, f32(2.0)) + EPS)
    assert tL == f32(1.0)
    assert tR == expect


def test_box_parallel_outside_culls():
    # |d_y| = 5e-5 < 1e-4 and origin outside the y slab: tR = -1 immediately,
    # even though the ray would reach the slab at t = 2e4
    tL, tR = O.ray_in_box([0.5, -1.0, 0.5, 0, 5e-5, 1], [0, 0, 0, 1, 1, 1], 1e-4, np.inf)
    assert tR == -1.0
    assert tL == f32(1e-4)


def test_box_early_exit_keeps_partial_tl():
    # x slab gives [1,2]; y slab gives [5,6] -> tL jumps to 5 > tR: return after y
    tL, tR = O.ray_in_box([-1, -5, 0.5, 1, 1, 0], [0, 0, 0, 1, 1, 1], 1e-4, np.inf)
    assert tL == f32(5.0)
    # x: tR = 2 + eps; y: min(2+eps, 6) + eps; then tL > tR returns before z
    assert tR == f32(f32(f32(2.0) + EPS) + EPS)


def test_box_negative_direction():
    tL, tR = O.ray_in_box([2, .5, .5, -1, 0, 0], [0, 0, 0, 1, 1, 1], 1e-4, np.inf)
    assert tL == f32(1.0) and tR == f32(2.0) + EPS


TRI = [0, 0, 0, 1, 0, 0, 0, 1, 0]


def test_triangle_exact_hit():
    # ray straight down onto the triangle in z=0 from z=2: t = 2 exactly
    assert O.ray_triangle([0.25, 0.25, 2, 0, 0, -1], TRI) == f32(2.0)


def test_triangle_edges_inclusive():
    # u = 0 edge and u + v = 1 edge are accepted (strict < / > tests)
    assert O.ray_triangle([0.0, 0.5, 1, 0, 0, -1], TRI) == f32(1.0)
    assert O.ray_triangle([0.5, 0.5, 1, 0, 0, -1], TRI) == f32(1.0)
    assert O.ray_triangle([0.5, 0.500001, 1, 0, 0, -1], TRI) == np.inf


def test_triangle_behind_origin_is_returned():
    # no sign check on t: the caller's t_min does the culling
    assert O.ray_triangle([0.25, 0.25, -2, 0, 0, -1], TRI) == f32(-2.0)


def test_triangle_grazing_rejected_relative_to_edge1():
    # |a| / |e1| < 1e-4 -> INF.  a = dot(e1, d x e2); pick d nearly in-plane
    tri = [0, 0, 0, 4, 0, 0, 0, 1, 0]   # |e1| = 4
    d = np.array([1.0, 0.0, -2e-4], dtype=np.float32)
    d /= np.linalg.norm(d)
    # a = e1 . (d x e2) = 4 * (d_z * -1 ... ) -> |a| = 4*|d_z| = 8e-4, /|e1| = 2e-4 >= 1e-4: accepted geometry
    assert np.isfinite(O.ray_triangle([0.5, 0.25, 1e-4, *d], tri))
    d2 = np.array([1.0, 0.0, -2e-5], dtype=np.float32)
    d2 /= np.linalg.norm(d2)
    assert O.ray_triangle([0.5, 0.25, 1e-5, *d2], tri) == np.inf


def test_triangle_matches_textbook_on_random_rays():
    # general position: the Moller-Trumbore t equals the plane-intersection
    # distance for hits and the barycentrics are in range
    rng = np.random.default_rng(1)
    hits = 0
    for _ in range(2000):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        # aim at a point of the triangle's plane, inside or outside it
        b = rng.uniform(-0.3, 1.0, 2)
        target = v[0] + b[0] * (v[1] - v[0]) + b[1] * (v[2] - v[0])
        d = (target - o).astype(np.float32)
        d /= np.linalg.norm(d)
        t = O.ray_triangle([*o, *d], v.ravel())
        n = np.cross(v[1].astype(np.float64) - v[0], v[2].astype(np.float64) - v[0])
        denom = np.dot(n, d)
        if abs(denom) < 1e-3:
            continue
        tp = np.dot(n, v[0] - o) / denom
        p = o + tp * d
        # barycentric inside test in float64
        m = np.stack([v[1] - v[0], v[2] - v[0]], 1).astype(np.float64)
        uv, *_ = np.linalg.lstsq(m, p - v[0], rcond=None)
        inside = uv[0] > 1e-4 and uv[1] > 1e-4 and uv.sum() < 1 - 1e-4
        outside = uv[0] < -1e-4 or uv[1] < -1e-4 or uv.sum() > 1 + 1e-4
        if inside:
            hits += 1
            assert np.isfinite(t) and abs(t - tp) <= 1e-3 * max(1.0, abs(tp))
        elif outside:
            assert t == np.inf
    assert hits > 100

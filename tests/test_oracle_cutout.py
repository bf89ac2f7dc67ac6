"""The oracle's alpha cut-outs (TransparentTest, src/model.cpp:217-230, in
the loops of Model::rayHit and Model::rayHit_test, src/model.cpp:332-354),
pinned by what the geometry fixes: rays along -z through quads whose alpha
planes are 0 or 1 over known halves.  A skipped layer must leave exactly the
hit of the scene without it; the eighth transparent layer ends the loop."""
import numpy as np
import pytest

import oracle.oracle_py as O
from cutout_util import down_rays, layers

ZERO = np.zeros((4, 4), np.float32)
ONE = np.ones((4, 4), np.float32)
XY = [(-0.5, 0.3), (0.2, -0.7), (0.9, 0.6), (-0.05, 0.2)]  # off both quads' diagonals


def closest(alphas, back=True, xy=XY):
    d, keep = layers(alphas, back)
    S = O.OracleScene(d)
    t, ids = S.closest(down_rays(xy))
    return t, ids


def test_transparent_layers_leave_the_back_hit():
    tb, ib = closest([])
    for k in (1, 3, 7):
        t, ids = closest([ZERO] * k)
        assert np.array_equal(ids, ib) and np.array_equal(t.view(np.uint32), tb.view(np.uint32)), k
    assert np.all(ib < 2) and np.allclose(tb, 5.0, atol=1e-5)


def test_eighth_transparent_layer_makes_a_miss():
    t, ids = closest([ZERO] * 8)
    assert np.all(ids == -1) and np.all(np.isinf(t))
    t, ids = closest([ZERO] * 7 + [ONE])  # the eighth is opaque: hit it
    assert np.all(ids // 2 == 8) and np.allclose(t, 1.0 + 0.5 * 7, atol=1e-5)


def test_alpha_plane_of_ones_is_opaque():
    t, ids = closest([ONE, ZERO])
    assert np.all(ids // 2 == 1) and np.allclose(t, 1.0, atol=1e-5)
    t, ids = closest([None])  # no alpha plane at all
    assert np.all(ids // 2 == 1)


def test_half_transparent_in_u_and_in_v():
    # columns 0-1 (u < 0.5) transparent; sampled at u = 0.125 and 0.625
    cols = np.float32(np.tile([0, 0, 1, 1], (4, 1)))
    t, ids = closest([cols], xy=[(-0.75, 0.1), (0.25, 0.1)])
    assert ids[0] < 2 and ids[1] // 2 == 1
    # rows 0-1 (the top rows: v is flipped, v' = 1 - v) transparent; at
    # v = 0.875 the lookup row is 0.5, at v = 0.375 it is 2.5
    rows = cols.T.copy()
    t, ids = closest([rows], xy=[(0.1, 0.75), (0.1, -0.25)])
    assert ids[0] < 2 and ids[1] // 2 == 1


def test_shadow_rays_pass_transparent_layers():
    def blocked(alphas, back, aim):
        d, keep = layers(alphas, back)
        return O.OracleScene(d).any(down_rays(XY), aim=aim)
    assert np.all(blocked([ZERO] * 7, True, 10.0) == 1)   # the back quad
    assert np.all(blocked([ZERO] * 3, True, 3.0) == 0)    # back beyond aim
    assert np.all(blocked([ZERO] * 7, False, 10.0) == 0)
    assert np.all(blocked([ZERO] * 8, False, 10.0) == 1)  # 8 walks end blocked
    assert np.all(blocked([ZERO, ONE], False, 10.0) == 1)

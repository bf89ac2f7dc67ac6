"""rm_trace_closest_noalloc: the closest-hit entry point whose ray-reordering
workspace is the library's static per-GPU device array instead of allocated
memory.  Called through the C ABI, it must return exactly what
rm_trace_closest returns on the same rays: a call at the maximum ray count,
then a smaller one on other rays (the array keeps the first call's contents),
under each reordering mode that uses the workspace; calls beyond the maximum
are refused without touching the outputs."""
import ctypes as C

import numpy as np
import pytest
import torch

import raym0nade_amd as rm
from raym0nade_amd._lib import RM_ERR_INVALID, RmError, check, lib

pytestmark = pytest.mark.gpu

MAX_RAYS = 2097152  # RM_NOALLOC_MAX_RAYS


def last_error():
    return lib().rm_last_error().decode(errors="replace")


def vp(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def both(S, rays):
    n = rays.shape[0]
    t0 = torch.full((n,), -1.0, dtype=torch.float32, device=rays.device)
    i0 = torch.full((n,), -7, dtype=torch.int32, device=rays.device)
    t1, i1 = t0.clone(), i0.clone()
    L = lib()
    rc = L.rm_trace_closest(S._h, vp(rays), n, 1e-4, float("inf"), vp(t0), vp(i0), stream())
    assert rc == 0, last_error()
    rc = L.rm_trace_closest_noalloc(S._h, vp(rays), n, 1e-4, float("inf"), vp(t1), vp(i1), stream())
    assert rc == 0, last_error()
    torch.cuda.synchronize()
    return (t0.cpu().numpy(), i0.cpu().numpy()), (t1.cpu().numpy(), i1.cpu().numpy())


@pytest.fixture(params=[3, 4, 5, 1, 2, 0], ids=["radix-sorted", "own-radix-sorted", "direction-sorted",
                                                 "counting-sorted", "always-sorted", "unsorted"])
def ray_sort(request):
    old = rm.get_tuning(rm.TUNE_RAY_SORT)
    rm.set_tuning(rm.TUNE_RAY_SORT, request.param)
    yield request.param
    rm.set_tuning(rm.TUNE_RAY_SORT, old)


def test_noalloc_equals_rm_trace_closest(dev, ray_sort):
    print(f"[noalloc] ray_sort {ray_sort}: 100k-triangle soup", flush=True)
    g = rm.generate(1, 2, n_tris=100000)
    S = rm.Scene(g)
    cam = rm.RmRenderArgs.from_buffer_copy(g.args)
    cam.accuracy = np.float32(cam.accuracy * cam.width / 540)
    cam.width = cam.height = 540
    for n, kind, seed in ((MAX_RAYS, 1, 3), (540 * 540, 0, 4)):
        print(f"[noalloc] {n} rays of kind {kind}", flush=True)
        rays = torch.from_numpy(rm.rays(kind, n, seed=seed, cam=cam if kind == 0 else None)).to(dev)
        (t0, i0), (t1, i1) = both(S, rays)
        assert np.array_equal(i0, i1)
        assert np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
        assert (i0 >= 0).mean() > 0.05


def test_noalloc_refuses_more_than_the_maximum(dev):
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    n = MAX_RAYS + 1
    rays = torch.zeros((n, 6), dtype=torch.float32, device=dev)
    t = torch.full((n,), 5.0, dtype=torch.float32, device=dev)
    ids = torch.full((n,), 3, dtype=torch.int32, device=dev)
    rc = lib().rm_trace_closest_noalloc(S._h, vp(rays), n, 1e-4, float("inf"), vp(t), vp(ids), stream())
    assert rc == RM_ERR_INVALID
    assert "maximum" in last_error()
    torch.cuda.synchronize()
    assert bool((t == 5.0).all()) and bool((ids == 3).all())
    # a zero-ray call is a no-op
    assert lib().rm_trace_closest_noalloc(S._h, vp(rays), 0, 1e-4, float("inf"), vp(t), vp(ids), stream()) == 0
    with pytest.raises(RmError):
        check(lib().rm_trace_closest_noalloc(None, vp(rays), 4, 1e-4, float("inf"), vp(t), vp(ids), stream()))

"""The C-ABI library: loads, exports every function include/raym0nade_hip.h
declares, and reports argument errors without a device (no compute calls)."""
import ctypes as C
import os
import re

import numpy as np

import raym0nade_amd as rm
from raym0nade_amd._lib import LIB_PATH, RmRenderArgs, RmSceneDesc, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "raym0nade_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(rm_\w+)\s*\(", src, flags=re.M)))


def test_every_declared_function_is_exported():
    names = header_functions()
    assert len(names) >= 20
    L = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version():
    assert lib().rm_abi_version() == 2


def test_null_arguments_are_errors():
    assert lib().rm_scene_create(None, None) == 1
    out = C.c_void_p()
    assert lib().rm_scene_create(None, C.byref(out)) == 1
    assert b"NULL" in lib().rm_last_error()
    assert lib().rm_trace_closest(None, None, 0, 0.0, 1.0, None, None, None) == 1
    assert lib().rm_generate_scene(99, 1, None, C.byref(out)) == 1


def test_empty_scene_rejected_before_device():
    d = RmSceneDesc()
    d.n_tris = 0
    out = C.c_void_p()
    assert lib().rm_scene_create(C.byref(d), C.byref(out)) == 4


def test_bad_material_index_rejected():
    g = rm.generate(0, 1)
    mat = np.array(g.material, copy=True)
    mat[3] = 999
    d = RmSceneDesc.from_buffer_copy(g.desc)
    d.material = mat.ctypes.data_as(C.POINTER(C.c_int32))
    out = C.c_void_p()
    assert lib().rm_scene_create(C.byref(d), C.byref(out)) == 1
    assert b"out of range" in lib().rm_last_error()


def test_generators_are_deterministic():
    a = rm.generate(1, 7, n_tris=1000).verts.copy()
    b = rm.generate(1, 7, n_tris=1000).verts.copy()
    c = rm.generate(1, 8, n_tris=1000).verts.copy()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    g = rm.generate(0, 1)
    assert g.n_tris == 36
    args = g.args
    assert (args.width, args.height, args.spp, args.max_depth, args.rr_min_depth) == (512, 512, 64, 8, 3)
    r1 = rm.rays(1, 100, seed=3)
    assert np.allclose(np.linalg.norm(r1[:, 3:], axis=1), 1.0, atol=1e-5)
    assert (r1[:, :3] >= 0).all() and (r1[:, :3] < 1).all()
